// join.hip — partitioned hash join (count of equal-key pairs) on the stable
// split.
//
// Reference: obsolete/OpenCL/phjImpl.cpp:14-66 (partition R and S with
// WG_split into 2^13 buckets, then one workgroup per bucket pair builds an LDS
// hash table of one side and probes it with the other,
// obsolete/OpenCL/hj_partitioned_kernel.cl:7-115) and its count-only result
// (phjImpl.cpp:13 "only count the number of outputs").  The count is
//   sum over keys k of |{i: R[i] == k}| * |{j: S[j] == k}|.
// Here the keys are replaced by a bijective mix of their bits (mix32), so that
// the partitions stay balanced for structured keys; equal keys keep equal
// mixes.  The split's first pass mixes each key as it loads it (MixOnLoad),
// so the mixed keys are never written and read back by a pass of their own.
// The partition width targets 4096 build keys (up to 16 partition bits); a
// workgroup builds its table from the smaller side of its pair in chunks of at
// most kJoinChunk keys (load factor <= 1/2), so any skew ends, at the cost of
// re-probing.  Every
// key is valid (no sentinel value, unlike the reference's -1).
#include <algorithm>
#include "ctx.hpp"
#include "device.hpp"
#include "split_tile.hpp"

namespace edpp {

namespace {

constexpr int kJoinBlock = 1024;
constexpr int kJoinSlots = 16384;  // 64-bit LDS slots (128 KiB): (count << 32) | key; count 0 = empty
constexpr int kJoinChunk = kJoinSlots / 2;  // build keys per table fill
constexpr int kJoinTargetBuild = 4096;
constexpr int kJoinPre = kJoinChunk / kJoinBlock;  // keys per lane per side held in registers (8)
// Partition bits at most: the split runs up to 16 bucket bits as two stable
// passes, so relations up to 2^28 keys keep about kJoinTargetBuild build keys
// per partition and one table fill each.
constexpr int kJoinMaxBits = 16;

__global__ void k_join_zero(unsigned long long* p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0ull;
}

// Slots are probed from the mixed key's bits above the partition bits.
__device__ __forceinline__ uint32_t join_slot(uint32_t hk, int pbits) {
    return (pbits >= 32 ? 0u : (hk >> pbits)) & (uint32_t)(kJoinSlots - 1);
}

// A claim is tried first: on an empty slot (most of them, at a load factor
// of at most 1/2) it is the only LDS operation; a failed claim returns the
// slot's word, whose key never changes once set.  Returns the slot this
// call claimed, or kNoClaim when the key was counted into a slot that
// another call had claimed.
constexpr uint32_t kNoClaim = 0xFFFFFFFFu;
__device__ __forceinline__ uint32_t join_insert(uint64_t* table, uint32_t hk, int pbits) {
    uint32_t h = join_slot(hk, pbits);
    while (true) {
        uint64_t w = 0;  // empty: claim it for this key with count 1
        if (__hip_atomic_compare_exchange_strong(table + h, &w, (1ull << 32) | hk, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP))
            return h;
        if ((uint32_t)w == hk) {
            __hip_atomic_fetch_add(table + h, 1ull << 32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            return kNoClaim;
        }
        h = (h + 1) & (uint32_t)(kJoinSlots - 1);
    }
}

__device__ __forceinline__ uint32_t join_lookup(const uint64_t* table, uint32_t hk, int pbits) {
    uint32_t h = join_slot(hk, pbits);
    while (true) {
        const uint64_t w = table[h];
        if ((w >> 32) == 0) return 0;
        if ((uint32_t)w == hk) return (uint32_t)(w >> 32);
        h = (h + 1) & (uint32_t)(kJoinSlots - 1);
    }
}

// Rows of partition pair b: R rows [rs[b], rs[b+1]) and S rows [ss[b],
// ss[b+1]) of the partitioned mixed keys, the smaller side built into the
// table.  The offsets are clamped to the relations: a split whose device
// wait timed out (reported as EDPP_ETIMEOUT) leaves offsets that must not
// steer reads.
struct JoinPair {
    const uint32_t* bkeys;
    const uint32_t* pkeys;
    int64_t b0, b1, p0, p1;
    // both sides fit the registers of one table fill (the balanced case)
    __device__ __forceinline__ bool held() const {
        return b1 > b0 && p1 > p0 && b1 - b0 <= kJoinPre * kJoinBlock && p1 - p0 <= kJoinPre * kJoinBlock;
    }
};
__device__ __forceinline__ JoinPair join_pair(int b, const uint32_t* rk, int64_t rn, const int32_t* rs,
                                              const uint32_t* sk, int64_t sn, const int32_t* ss, int B) {
    auto clamp = [](int64_t x, int64_t lo, int64_t hi) { return x < lo ? lo : (x > hi ? hi : x); };
    const int64_t r0 = clamp(rs[b], 0, rn), r1 = clamp(b + 1 < B ? rs[b + 1] : rn, r0, rn);
    const int64_t s0 = clamp(ss[b], 0, sn), s1 = clamp(b + 1 < B ? ss[b + 1] : sn, s0, sn);
    const bool build_r = (r1 - r0) <= (s1 - s0);
    JoinPair jp;
    jp.bkeys = build_r ? rk : sk;
    jp.pkeys = build_r ? sk : rk;
    jp.b0 = build_r ? r0 : s0;
    jp.b1 = build_r ? r1 : s1;
    jp.p0 = build_r ? s0 : r0;
    jp.p1 = build_r ? s1 : r1;
    return jp;
}
// lane's keys of a held pair's two sides
__device__ __forceinline__ void join_load(const JoinPair& jp, uint32_t (&bk)[kJoinPre], uint32_t (&pk)[kJoinPre]) {
    #pragma unroll
    for (int q = 0; q < kJoinPre; ++q) {
        const int64_t i = jp.b0 + threadIdx.x + (int64_t)q * kJoinBlock;
        bk[q] = i < jp.b1 ? jp.bkeys[i] : 0u;
    }
    #pragma unroll
    for (int q = 0; q < kJoinPre; ++q) {
        const int64_t i = jp.p0 + threadIdx.x + (int64_t)q * kJoinBlock;
        pk[q] = i < jp.p1 ? jp.pkeys[i] : 0u;
    }
}

// Persistent: workgroup g (one per CU, as its table fills the CU's LDS)
// takes partition pairs g, g + gridDim.x, ...  While a pair's table is
// cleared, filled and probed, the next pair's offsets and (when both of its
// sides fit the registers) keys are loaded, so their global latency is hidden
// behind the LDS work.  A table fill holds at most kJoinChunk distinct keys in
// kJoinSlots slots, so every probe sequence meets an empty slot and ends.
__global__ __launch_bounds__(kJoinBlock) void k_join_probe(const uint32_t* __restrict__ rk, int64_t rn,
                                                           const int32_t* __restrict__ rs,
                                                           const uint32_t* __restrict__ sk, int64_t sn,
                                                           const int32_t* __restrict__ ss, int B, int pbits,
                                                           unsigned long long* d_count) {
    extern __shared__ uint64_t table[];  // kJoinSlots
    unsigned long long local = 0;
    int b = blockIdx.x;
    JoinPair cur = join_pair(b, rk, rn, rs, sk, sn, ss, B);
    uint32_t bk[kJoinPre], pk[kJoinPre];
    if (cur.held()) join_load(cur, bk, pk);
    // after a held pair the only nonzero slots are the ones its lanes
    // claimed (recorded in claimed[]): the next held pair clears just those
    bool sparse = false;
    uint32_t claimed[kJoinPre];
    for (; b < B; b += gridDim.x) {
        const int nb = b + (int)gridDim.x;
        JoinPair nxt = cur;
        if (nb < B) nxt = join_pair(nb, rk, rn, rs, sk, sn, ss, B);
        uint32_t nbk[kJoinPre], npk[kJoinPre];
        __syncthreads();  // the previous pair's lookups are done before the table is cleared
        if (cur.held()) {
            if (sparse) {
                #pragma unroll
                for (int q = 0; q < kJoinPre; ++q)
                    if (claimed[q] != kNoClaim) table[claimed[q]] = 0ull;
            } else {
                for (int i = threadIdx.x; i < kJoinSlots; i += kJoinBlock) table[i] = 0ull;
            }
            if (nb < B && nxt.held()) join_load(nxt, nbk, npk);
            __syncthreads();
            #pragma unroll
            for (int q = 0; q < kJoinPre; ++q)
                claimed[q] = cur.b0 + threadIdx.x + (int64_t)q * kJoinBlock < cur.b1 ? join_insert(table, bk[q], pbits)
                                                                                     : kNoClaim;
            sparse = true;
            __syncthreads();
            #pragma unroll
            for (int q = 0; q < kJoinPre; ++q)
                if (cur.p0 + threadIdx.x + (int64_t)q * kJoinBlock < cur.p1)
                    local += join_lookup(table, pk[q], pbits);
        } else {
            if (nb < B && nxt.held()) join_load(nxt, nbk, npk);
            if (cur.b1 > cur.b0 && cur.p1 > cur.p0) {
                sparse = false;  // the chunk fills leave slots no lane recorded
                for (int64_t c0 = cur.b0; c0 < cur.b1; c0 += kJoinChunk) {
                    const int64_t c1 = std::min<int64_t>(c0 + kJoinChunk, cur.b1);
                    for (int i = threadIdx.x; i < kJoinSlots; i += kJoinBlock) table[i] = 0ull;
                    __syncthreads();
                    for (int64_t i = c0 + threadIdx.x; i < c1; i += kJoinBlock)
                        join_insert(table, cur.bkeys[i], pbits);
                    __syncthreads();
                    for (int64_t i = cur.p0 + threadIdx.x; i < cur.p1; i += kJoinBlock)
                        local += join_lookup(table, cur.pkeys[i], pbits);
                    __syncthreads();
                }
            }
        }
        cur = nxt;
        #pragma unroll
        for (int q = 0; q < kJoinPre; ++q) {
            bk[q] = nbk[q];
            pk[q] = npk[q];
        }
    }
    // workgroup sum: wave sums, then one lane per wave into the table's first slots
    const uint64_t ws = wave_sum((uint64_t)local);
    __syncthreads();
    if (lane_id() == 0) table[threadIdx.x / kWave] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long total = 0;
        for (int q = 0; q < kJoinBlock / kWave; ++q) total += table[q];
        if (total) atomicAdd(d_count, total);
    }
}

}  // namespace

int launch_join_count(edpp_ctx* c, const uint32_t* rk, int64_t rn, const uint32_t* sk, int64_t sn,
                      unsigned long long* d_count) {
    hipLaunchKernelGGL(k_join_zero, dim3(1), dim3(64), 0, c->stream, d_count);
    EDPP_HIP(c, hipGetLastError());
    if (rn == 0 || sn == 0) return EDPP_OK;
    // partition bits: about kJoinTargetBuild keys of the smaller side per partition
    const int64_t small = std::min(rn, sn);
    int pbits = 0;
    while (pbits < kJoinMaxBits && (small >> pbits) > kJoinTargetBuild) ++pbits;
    const int B = 1 << pbits;
    const size_t rb = ((size_t)rn * 4 + 255) & ~(size_t)255, sb = ((size_t)sn * 4 + 255) & ~(size_t)255;
    const size_t stb = ((size_t)B * 4 + 255) & ~(size_t)255;
    void* w;
    int rc = ws_outer_reserve(c, rb + sb + 2 * stb, &w);
    if (rc) return rc;
    char* base = static_cast<char*>(w);
    uint32_t* rpart = reinterpret_cast<uint32_t*>(base);
    uint32_t* spart = reinterpret_cast<uint32_t*>(base + rb);
    int32_t* rst = reinterpret_cast<int32_t*>(base + rb + sb);
    int32_t* sst = reinterpret_cast<int32_t*>(base + rb + sb + stb);
    // the partitioned relations hold the mixed keys
    rc = launch_split<MixOnLoad>(c, reinterpret_cast<const MixOnLoad*>(rk), reinterpret_cast<MixOnLoad*>(rpart),
                                 nullptr, nullptr, rst, rn, B, EDPP_SPLIT_AUTO);
    if (!rc)
        rc = launch_split<MixOnLoad>(c, reinterpret_cast<const MixOnLoad*>(sk), reinterpret_cast<MixOnLoad*>(spart),
                                     nullptr, nullptr, sst, sn, B, EDPP_SPLIT_AUTO);
    if (rc) return rc;
    const size_t lds = (size_t)kJoinSlots * 8;
    EDPP_HIP(c, hipFuncSetAttribute((const void*)k_join_probe, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    // one persistent workgroup per CU (its table holds the CU's LDS)
    const unsigned grid = (unsigned)std::max(1, std::min(B, std::max(c->cus, 1)));
    hipLaunchKernelGGL(k_join_probe, dim3(grid), dim3(kJoinBlock), lds, c->stream, rpart, rn, rst, spart, sn, sst, B,
                       pbits, d_count);
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

}  // namespace edpp
