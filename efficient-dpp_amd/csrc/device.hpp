// device.hpp — gfx950 building blocks shared by every kernel of the library.
//
// Wave64 scans use DPP row shifts plus the GFX9 row broadcasts; ranks inside a
// wave use ballot + mbcnt; cross-workgroup state uses agent-scope atomics so
// that all 8 XCDs (whose L2s are not coherent with each other) observe it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace edpp {

constexpr int kWave = 64;

typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));
typedef uint64_t v2u64 __attribute__((ext_vector_type(2)));
typedef int32_t  v4i32 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------- randomness
// SplitMix64 (Steele, Lea, Flood, OOPSLA 2014).  x_i is the (i+1)-th output of
// the generator seeded with `seed`, so any thread can produce element i.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ uint64_t splitmix_at(uint64_t seed, uint64_t i) {
    return mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull);
}

// Keyed 4-round balanced Feistel network on `2*half` bits: a bijection of
// [0, 2^(2*half)).  Round function: low `half` bits of mix64(R ^ round key).
__host__ __device__ __forceinline__ uint64_t feistel(uint64_t x, int half, uint64_t key) {
    const uint64_t m = (1ull << half) - 1;
    uint64_t l = x >> half, r = x & m;
    #pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint64_t f = mix64(r ^ (key + 0xD1B54A32D192ED03ull * (uint64_t)(k + 1))) & m;
        uint64_t nl = r;
        r = l ^ f;
        l = nl;
    }
    return (l << half) | r;
}
// Half width for a domain of size n >= 1: smallest h >= 1 with 4^h >= n.
__host__ __device__ __forceinline__ int feistel_half(uint64_t n) {
    int h = 1;
    while ((1ull << (2 * h)) < n) ++h;
    return h;
}
// Permutation of [0,n) by cycle walking the Feistel bijection of [0,4^h).
__host__ __device__ __forceinline__ uint64_t perm_at(uint64_t i, uint64_t n, int half, uint64_t key) {
    uint64_t x = feistel(i, half, key);
    while (x >= n) x = feistel(x, half, key);
    return x;
}

// ---------------------------------------------------------------- lanes
__device__ __forceinline__ int lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
__device__ __forceinline__ uint64_t ballot(bool p) {
    return __builtin_amdgcn_ballot_w64(p);
}
// number of set bits of `mask` in lanes strictly below the calling lane
__device__ __forceinline__ uint32_t count_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_or_zero(uint32_t x) {
    // bound_ctrl = true: lanes whose source is outside the row read 0;
    // rows disabled by ROW_MASK keep `old`, which is 0.
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROW_MASK, 0xf, true);
}

// Inclusive wave64 scan: row_shr 1,2,4,8 scans each 16-lane row, row_bcast:15
// carries row totals into rows 1 and 3, row_bcast:31 carries lanes 0-31 into
// rows 2 and 3.
__device__ __forceinline__ uint32_t wave_inclusive_add(uint32_t x) {
    x += dpp_or_zero<0x111, 0xf>(x);
    x += dpp_or_zero<0x112, 0xf>(x);
    x += dpp_or_zero<0x114, 0xf>(x);
    x += dpp_or_zero<0x118, 0xf>(x);
    x += dpp_or_zero<0x142, 0xa>(x);
    x += dpp_or_zero<0x143, 0xc>(x);
    return x;
}
__device__ __forceinline__ uint64_t wave_inclusive_add(uint64_t x) {
    // same network on both halves; each step adds the shifted 64-bit value
#define EDPP_STEP64(C, M)                                                       \
    {                                                                           \
        uint32_t lo = dpp_or_zero<C, M>((uint32_t)x);                           \
        uint32_t hi = dpp_or_zero<C, M>((uint32_t)(x >> 32));                   \
        x += ((uint64_t)hi << 32) | lo;                                         \
    }
    EDPP_STEP64(0x111, 0xf)
    EDPP_STEP64(0x112, 0xf)
    EDPP_STEP64(0x114, 0xf)
    EDPP_STEP64(0x118, 0xf)
    EDPP_STEP64(0x142, 0xa)
    EDPP_STEP64(0x143, 0xc)
#undef EDPP_STEP64
    return x;
}

__device__ __forceinline__ uint32_t read_lane(uint32_t x, int l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, l);
}
__device__ __forceinline__ uint64_t read_lane(uint64_t x, int l) {
    uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, l);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}
template <class T>
__device__ __forceinline__ T wave_sum(T x) {
    return read_lane(wave_inclusive_add(x), kWave - 1);
}

// ---------------------------------------------------------------- atomics
template <class T>
__device__ __forceinline__ T ld_relaxed(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T ld_acquire(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_relaxed(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_release(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T fetch_add(T* p, T v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------- streaming memory
template <class V>
__device__ __forceinline__ V ld_stream(const V* p) { return __builtin_nontemporal_load(p); }
template <class V>
__device__ __forceinline__ void st_stream(V v, V* p) { __builtin_nontemporal_store(v, p); }

// 16-byte vector of T
template <class T> struct vec16;
template <> struct vec16<uint32_t> { typedef v4u32 type; static constexpr int n = 4; };
template <> struct vec16<uint64_t> { typedef v2u64 type; static constexpr int n = 2; };

}  // namespace edpp
