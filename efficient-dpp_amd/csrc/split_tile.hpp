// split_tile.hpp — stable ranking of bucketed keys inside a workgroup tile,
// shared by the split kernels and the batched executor.
#pragma once
#include <algorithm>
#include "device.hpp"

namespace edpp {

constexpr int kSplitBlock = 256;
constexpr int kSplitWaves = kSplitBlock / kWave;

// Mask of the lanes whose bucket equals the caller's, among valid lanes: one
// ballot per bucket bit (cheap while B is small).
__device__ __forceinline__ uint64_t match_bucket(uint32_t b, int lg, uint64_t valid) {
    uint64_t m = valid;
    for (int bit = 0; bit < lg; ++bit) {
        bool on = (b >> bit) & 1u;
        uint64_t bal = ballot(on);
        m &= on ? bal : ~bal;
    }
    return m;
}

// Up to this many bucket bits the ballot match is cheaper than LDS matching,
// whose same-address atomics serialize when many lanes share few buckets.
constexpr int kBallotMatchBits = 4;

// The lanes whose bucket equals that of the lowest valid lane, from one
// ballot, and that bucket in *bf (0xFFFFFFFF when no lane is valid).  Under
// skewed keys most lanes of a wave share one bucket; taking that group out of
// the per-lane LDS atomics keeps them from serializing on one address.
__device__ __forceinline__ uint64_t first_bucket_group(uint32_t b, bool ok, uint32_t* bf) {
    const uint64_t okm = ballot(ok);
    if (okm == 0) {
        *bf = 0xFFFFFFFFu;
        return 0;
    }
    const uint32_t f = read_lane(b, (int)__builtin_ctzll(okm));
    *bf = f;
    return ballot(ok && b == f);
}

// XCD-aware tile order: workgroups are dispatched round-robin over the XCDs,
// so workgroup g runs on XCD g % xcds.  Give every XCD a contiguous range of
// tiles, so that the write frontier of adjacent tiles (adjacent runs of the
// same bucket) meets in one L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t g, int64_t tiles, int xcds) {
    if (xcds <= 1 || tiles < xcds) return g;
    int64_t per = tiles / xcds, rem = tiles % xcds;  // first `rem` XCDs get per+1 tiles
    int64_t x = g % xcds, k = g / xcds;
    // rounds where every XCD still has a tile map by the split; the ragged last
    // round (k == per) only exists for XCDs x < rem
    int64_t start = x * per + (x < rem ? x : rem);
    return start + k;
}
// LDS layout of the WG scatter (32-bit words): wmask[waves][B] (64-bit),
// whist[waves][B], run[B], sub[B], off[B], wtmp[16], then the key and value
// reorder buffers from a 16-byte boundary (the value buffer aliases the key
// buffer when values are staged after keys)
__host__ __device__ constexpr int whist_word_offset(int B, int waves = kSplitWaves) { return 2 * waves * B; }
__host__ __device__ constexpr int kbuf_word_offset(int B, int waves = kSplitWaves) {
    return ((whist_word_offset(B, waves) + (waves + 3) * B + 16) + 3) & ~3;
}

// ------------------------------------------------------------- WG strategy
// Tile = SUBS sub-tiles of S = 256*IPT elements.  In a sub-tile, wave w owns
// [w*64*IPT, (w+1)*64*IPT) and item j of lane l is element w*64*IPT + j*64 + l,
// so ranks follow input order: sub-tile, wave, item, lane.

struct SplitGeom {
    int64_t n;
    int lg;          // log2 B
    int shift;       // bucket = ((key >> shift) & (B-1)) ^ flip
    int64_t tile;    // elements per tile (a multiple of the sub-tile)
    int64_t tiles;
    int xcds;
    uint32_t flip;   // 0, or the top bucket bit: buckets of a signed top digit in order
};

// Bucket of a key: bits [shift, shift + log2 B) of its low 32 bits, with the
// flip bits inverted (a radix sort's signed top digit orders negative keys
// first when its top bit is flipped).
template <class K>
__device__ __forceinline__ uint32_t bucket_of(K key, int shift, uint32_t mask, uint32_t flip) {
    return ((((uint32_t)key) >> shift) & mask) ^ flip;
}

// murmur3's 32-bit finalizer: a bijection of 32-bit words
__host__ __device__ inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x85EBCA6Bu;
    x ^= x >> 13;
    x *= 0xC2B2AE35u;
    x ^= x >> 16;
    return x;
}

// A 32-bit key replaced by mix32(key) as it is loaded: a split of MixOnLoad
// keys is the split of the mixed keys, read straight from the unmixed ones.
// The hash join partitions its relations so, which keeps structured keys
// spread evenly over the partitions without a separate pass that mixes them.
struct MixOnLoad {
    uint32_t v;
    MixOnLoad() = default;
    __host__ __device__ constexpr explicit MixOnLoad(uint32_t x) : v(x) {}
};
__device__ __forceinline__ uint32_t bucket_of(MixOnLoad key, int shift, uint32_t mask, uint32_t flip) {
    return ((key.v >> shift) & mask) ^ flip;
}
// A key as the split sees it once loaded: itself, or mixed for MixOnLoad.
template <class K>
__device__ __forceinline__ K on_load(K k) {
    return k;
}
__device__ __forceinline__ MixOnLoad on_load(MixOnLoad k) {
    return MixOnLoad(mix32(k.v));
}

// Raw buffer descriptor of n elements at p (gfx9 word 3: 32-bit data
// format), and a load through one of the element at byte offset off, which
// reads 0 past the descriptor's end (the memory unit checks the bound).
template <class T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const T* p, int64_t n) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, 0, (int)(n * (int64_t)sizeof(T)), 0x00020000);
}
template <class T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff = 0) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, (int)soff, 0));
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, (int)soff, 0));
    }
}

// Buffer loads of the calling wave's chunk of a sub-tile (item j of lane l is
// element sbase + w*64*IPT + j*64 + l) through one descriptor from the chunk's
// start to n: past n they read 0, with no branch per load, and the lane's
// loads share one address register with scalar offsets.  The wave index is
// read as wave-uniform, so the descriptor stays in scalar registers.
template <class T, int IPT>
__device__ __forceinline__ void load_chunk_buf(const T* __restrict__ p, int64_t sbase, int64_t n, T (&x)[IPT]) {
    const int64_t wb = sbase + (int64_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave)) * kWave * IPT;
    const __amdgpu_buffer_rsrc_t r =
        buf_rsrc(p + wb, std::min<int64_t>(std::max<int64_t>(n - wb, 0), (int64_t)kWave * IPT));
    #pragma unroll
    for (int j = 0; j < IPT; ++j)
        x[j] = buf_load<T>(r, (uint32_t)lane_id() * (uint32_t)sizeof(T), (uint32_t)(j * kWave * (int)sizeof(T)));
}

// Loads the calling wave's chunk of a sub-tile (item j of lane l is element
// sbase + w*64*IPT + j*64 + l), and its values when vin is given, then
// extracts buckets (0xFFFFFFFF past n).  Every load is issued before any is
// used (the scheduling barrier keeps the compiler from consuming, and so
// waiting for, each load before issuing the next).  BUF: each wave loads
// through raw buffer descriptors that start at its chunk and end at n, so
// that a load past n reads 0 without a branch per load.
template <class K, int IPT, bool BUF = false>
__device__ __forceinline__ void load_subtile(const K* __restrict__ kin, const int32_t* __restrict__ vin, int64_t sbase,
                                             int64_t n, int lg, int shift, uint32_t (&bk)[IPT], K (&key)[IPT],
                                             int32_t (&val)[IPT], uint32_t flip = 0) {
    const int lane = lane_id();
    const int w = threadIdx.x / kWave;
    const uint32_t mask = (1u << lg) - 1u;
    const int64_t wbase = sbase + (int64_t)w * kWave * IPT;
    if constexpr (BUF) {
        load_chunk_buf<K, IPT>(kin, sbase, n, key);
        #pragma unroll
        for (int j = 0; j < IPT; ++j) key[j] = on_load(key[j]);
        if (vin) load_chunk_buf<int32_t, IPT>(vin, sbase, n, val);
    } else {
        #pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const int64_t e = wbase + (int64_t)j * kWave + lane;
            key[j] = e < n ? on_load(kin[e]) : K(0);
        }
        if (vin) {
            #pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int64_t e = wbase + (int64_t)j * kWave + lane;
                val[j] = e < n ? vin[e] : 0;
            }
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    #pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const int64_t e = wbase + (int64_t)j * kWave + lane;
        bk[j] = e < n ? bucket_of(key[j], shift, mask, flip) : 0xFFFFFFFFu;
    }
}

// Bucket counts of the wave's chunk added to whist_w[b] (order-free).
// BALLOT: one lane per group adds the group's size (few buckets, where LDS
// atomics would collide); otherwise per-lane LDS atomics, except for the
// lowest valid lane's bucket group, which adds its size at once.
template <class K, int IPT, bool BALLOT>
__device__ __forceinline__ void count_subtile(const K* __restrict__ kin, int64_t sbase, int64_t n, int lg, int shift,
                                              uint32_t* whist_w) {
    uint32_t bk[IPT];
    K key[IPT];
    int32_t val[IPT];
    load_subtile<K, IPT>(kin, nullptr, sbase, n, lg, shift, bk, key, val);
    #pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const bool ok = bk[j] != 0xFFFFFFFFu;
        if (BALLOT) {
            const uint64_t m = match_bucket(bk[j], lg, ballot(ok));
            if (ok && count_below(m) == 0)
                __hip_atomic_fetch_add(whist_w + bk[j], (uint32_t)__popcll(m), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WAVEFRONT);
        } else {
            uint32_t bf;
            const uint64_t mf = first_bucket_group(bk[j], ok, &bf);
            if (ok && bk[j] == bf) {
                if (count_below(mf) == 0)
                    __hip_atomic_fetch_add(whist_w + bf, (uint32_t)__popcll(mf), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WAVEFRONT);
            } else if (ok) {
                __hip_atomic_fetch_add(whist_w + bk[j], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            }
        }
    }
}

// Histogram counting of the WG strategy's histogram kernel.  Up to this many
// bucket bits every lane keeps its counts as 8-bit fields packed four to a
// register (REGW = B/4 registers, at least 1), and the wave adds the fields up
// once per sub-tile: no LDS traffic per key and no dependence on key skew
// (LDS atomics serialize when many lanes hit few addresses).  Above it,
// per-lane LDS atomics.
constexpr int kRegCountBits = 4;
__host__ __device__ constexpr int reg_count_words(int lg) { return lg <= 2 ? 1 : (1 << (lg - 2)); }

// Counts one lane's IPT keys of a sub-tile, item j being element
// e0 + j*64 (elements at or past n ignored), into the wave's row: row[b] +=
// the wave's count of bucket b.  REGW > 0 selects the register counters
// (B <= 4*REGW), REGW == 0 LDS atomics.  Wave-uniform control flow.
template <class K, int IPT, int REGW>
__device__ __forceinline__ void count_keys(const K (&key)[IPT], int64_t e0, int64_t n, int lg, int shift,
                                           uint32_t flip, uint32_t* row) {
    static_assert(IPT <= 255, "8-bit per-lane counters");
    const int lane = lane_id();
    const int B = 1 << lg;
    const uint32_t mask = (uint32_t)B - 1u;
    if constexpr (REGW >= 2) {
        // 8 or 16 buckets: 4-bit counters, bucket b at bits [4b, 4b+4) of one
        // 64-bit word per item parity (at most IPT/2 <= 15 items each), so a
        // key costs a shift and an add instead of a compare-and-select per
        // packed word (a wave64 instruction holds a SIMD for 4 cycles)
        static_assert(IPT / 2 <= 15, "4-bit counters");
        uint64_t acc0 = 0, acc1 = 0;
        #pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const int64_t e = e0 + (int64_t)j * kWave;
            const uint32_t b = bucket_of(key[j], shift, mask, flip);
            const uint64_t inc = e < n ? 1ull << (4 * b) : 0ull;
            if (j & 1) acc1 += inc;
            else acc0 += inc;
        }
        constexpr uint64_t kNib = 0x0F0F0F0F0F0F0F0Full;
        const uint64_t ev = (acc0 & kNib) + (acc1 & kNib);                // byte k: bucket 2k
        const uint64_t od = ((acc0 >> 4) & kNib) + ((acc1 >> 4) & kNib);  // byte k: bucket 2k+1
        #pragma unroll
        for (int h = 0; h < REGW / 2; ++h) {  // buckets 8h .. 8h+7
            const uint32_t we = (uint32_t)(ev >> (32 * h)), wo = (uint32_t)(od >> (32 * h));
            // two 16-bit fields per sum: at most 64*IPT per bucket
            const uint32_t e04 = wave_sum(we & 0x00FF00FFu);         // buckets 8h, 8h+4
            const uint32_t e26 = wave_sum((we >> 8) & 0x00FF00FFu);  // buckets 8h+2, 8h+6
            const uint32_t o15 = wave_sum(wo & 0x00FF00FFu);         // buckets 8h+1, 8h+5
            const uint32_t o37 = wave_sum((wo >> 8) & 0x00FF00FFu);  // buckets 8h+3, 8h+7
            if (lane == 0) {  // the row belongs to this wave alone
                const int b0 = 8 * h;
                row[b0 + 0] += e04 & 0xFFFFu;
                row[b0 + 4] += e04 >> 16;
                row[b0 + 2] += e26 & 0xFFFFu;
                row[b0 + 6] += e26 >> 16;
                row[b0 + 1] += o15 & 0xFFFFu;
                row[b0 + 5] += o15 >> 16;
                row[b0 + 3] += o37 & 0xFFFFu;
                row[b0 + 7] += o37 >> 16;
            }
        }
    } else if constexpr (REGW > 0) {
        uint32_t packed[REGW];  // bucket b: word b/4, bits [8(b%4), 8(b%4)+8)
        #pragma unroll
        for (int q = 0; q < REGW; ++q) packed[q] = 0;
        #pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const int64_t e = e0 + (int64_t)j * kWave;
            const uint32_t b = bucket_of(key[j], shift, mask, flip);
            const uint32_t inc = e < n ? 1u << (8 * (b & 3u)) : 0u;
            #pragma unroll
            for (int q = 0; q < REGW; ++q) packed[q] += (b >> 2) == (uint32_t)q ? inc : 0u;
        }
        #pragma unroll
        for (int q = 0; q < REGW; ++q) {
            // two 16-bit fields per word: at most 64*IPT per bucket
            const uint32_t even = wave_sum(packed[q] & 0x00FF00FFu);        // buckets 4q, 4q+2
            const uint32_t odd = wave_sum((packed[q] >> 8) & 0x00FF00FFu);  // buckets 4q+1, 4q+3
            if (lane == 0) {  // the row belongs to this wave alone
                const int b0 = 4 * q;
                if (b0 < B) row[b0] += even & 0xFFFFu;
                if (b0 + 1 < B) row[b0 + 1] += odd & 0xFFFFu;
                if (b0 + 2 < B) row[b0 + 2] += even >> 16;
                if (b0 + 3 < B) row[b0 + 3] += odd >> 16;
            }
        }
    } else {
        // skewed keys (an eighth of the wave or more in the first lane's
        // bucket on the first or the middle item) take the grouped path for
        // the whole sub-tile; the test is wave-uniform
        uint32_t bf0, bf1;
        const uint64_t mf0 = first_bucket_group(bucket_of(key[0], shift, mask, flip), e0 < n, &bf0);
        const uint64_t mf1 =
            first_bucket_group(bucket_of(key[IPT / 2], shift, mask, flip), e0 + (int64_t)(IPT / 2) * kWave < n, &bf1);
        if (__popcll(mf0) >= kWave / 8 || __popcll(mf1) >= kWave / 8) {
            #pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int64_t e = e0 + (int64_t)j * kWave;
                const bool ok = e < n;
                const uint32_t b = bucket_of(key[j], shift, mask, flip);
                uint32_t bf;
                const uint64_t mf = first_bucket_group(b, ok, &bf);
                if (ok && b == bf) {
                    if (count_below(mf) == 0)
                        __hip_atomic_fetch_add(row + bf, (uint32_t)__popcll(mf), __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WAVEFRONT);
                } else if (ok) {
                    __hip_atomic_fetch_add(row + b, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                }
            }
        } else {
            #pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int64_t e = e0 + (int64_t)j * kWave;
                if (e < n)
                    __hip_atomic_fetch_add(row + bucket_of(key[j], shift, mask, flip), 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WAVEFRONT);
            }
        }
    }
}

// Counts TPW consecutive sub-tiles of S = 256*IPT elements, sub-tile t
// starting at sbase + t*S (elements at or past n ignored), into the rows of
// tile t/spt: whist + ((t/spt)*waves + w)*B.  Element layout as in
// load_subtile.  Every key load of all TPW sub-tiles is issued before any is
// used.  REGW > 0 selects the register counters (B <= 4*REGW), REGW == 0 LDS
// atomics.
template <class K, int IPT, int TPW, int REGW>
__device__ __forceinline__ void count_tiles(const K* __restrict__ kin, int64_t sbase, int64_t n, int lg, int shift,
                                            int spt, uint32_t* whist, uint32_t flip = 0) {
    constexpr int S = kSplitBlock * IPT;
    const int lane = lane_id();
    const int w = threadIdx.x / kWave;
    const int B = 1 << lg;
    const int64_t wbase = sbase + (int64_t)w * kWave * IPT + lane;
    K key[TPW][IPT];
    // one raw buffer descriptor from the wave's first element to n (loads
    // past n read 0, with no branch per load; the wave-uniform start keeps it
    // in scalar registers)
    const int64_t wb = sbase + (int64_t)__builtin_amdgcn_readfirstlane(w) * kWave * IPT;
    const __amdgpu_buffer_rsrc_t rk =
        buf_rsrc(kin + wb, std::min<int64_t>(std::max<int64_t>(n - wb, 0), (int64_t)TPW * S));
    #pragma unroll
    for (int t = 0; t < TPW; ++t)
        #pragma unroll
        for (int j = 0; j < IPT; ++j)
            key[t][j] = on_load(buf_load<K>(rk, (uint32_t)lane * (uint32_t)sizeof(K),
                                            (uint32_t)((t * S + j * kWave) * (int)sizeof(K))));
    __builtin_amdgcn_sched_barrier(0);
    #pragma unroll
    for (int t = 0; t < TPW; ++t)
        count_keys<K, IPT, REGW>(key[t], wbase + (int64_t)t * S, n, lg, shift, flip,
                                 whist + ((t / spt) * kSplitWaves + w) * B);
}

// Stable ranks of the wave's chunk: rk[j] = whist_w[b] (count of earlier keys
// of bucket b) + number of lanes below with the same bucket in item j.
//
// Phase 1 finds, for every item, the mask of lanes sharing the caller's
// bucket: from bucket-bit ballots (BALLOT), or from one ballot for the lowest
// valid lane's bucket plus wmask[b] (per wave, zero on entry and on exit) for
// the other lanes: each of them ORs in its own bit, reads the mask back and
// clears it.  Phase 2 reads whist_w[b] and lets the lowest lane of each group
// add the group size.  A wave's LDS instructions execute in program order, so
// neither phase waits for an earlier item's result before issuing the next
// item: item j's mask read sees exactly item j's ORs (item j-1 cleared its
// slots before them), and item j's whist_w read sees every add of items < j.
template <class K, int IPT, bool BALLOT, bool BUF = false>
__device__ __forceinline__ void rank_subtile(const K* __restrict__ kin, const int32_t* __restrict__ vin,
                                             int64_t sbase, int64_t n, int lg, int shift, uint32_t* whist_w,
                                             uint64_t* wmask, uint32_t (&bk)[IPT], uint32_t (&rk)[IPT],
                                             K (&key)[IPT], int32_t (&val)[IPT], uint32_t flip = 0) {
    const int lane = lane_id();
    load_subtile<K, IPT, BUF>(kin, vin, sbase, n, lg, shift, bk, key, val, flip);
    // Each item's match mask is reduced at once to its rank among the lanes
    // and its group size, packed as below | size << 16, so that 16 32-bit
    // words stay live instead of 16 64-bit masks; an LDS-matched item is
    // reduced kPipe items later, when its mask read has long arrived, so the
    // wave keeps issuing the next items' LDS operations meanwhile.
    constexpr int kPipe = 3;
    uint32_t info[IPT];
    uint64_t m[IPT];
    #pragma unroll
    for (int j = 0; j < IPT + kPipe; ++j) {
        if (j < IPT) {
            const bool ok = bk[j] != 0xFFFFFFFFu;
            if (BALLOT) {
                m[j] = match_bucket(bk[j], lg, ballot(ok));
            } else {
                // the lowest valid lane's bucket group comes from a ballot; the
                // other lanes (none of them in that bucket) match through wmask
                const uint32_t b = ok ? bk[j] : 0u;
                uint32_t bf;
                const uint64_t mf = first_bucket_group(b, ok, &bf);
                const bool lds = ok && b != bf;
                if (lds)
                    __hip_atomic_fetch_or(wmask + b, 1ull << lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                const uint64_t ml =
                    lds ? __hip_atomic_load(wmask + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) : 0ull;
                if (lds) __hip_atomic_store(wmask + b, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                m[j] = lds ? ml : (ok ? mf : 0ull);
            }
        }
        const int r = BALLOT ? j : j - kPipe;  // the item reduced now
        if (r >= 0 && r < IPT) info[r] = count_below(m[r]) | ((uint32_t)__popcll(m[r]) << 16);
    }
    #pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const bool ok = bk[j] != 0xFFFFFFFFu;
        const uint32_t b = ok ? bk[j] : 0u;
        const uint32_t prior =
            ok ? __hip_atomic_load(whist_w + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) : 0u;
        const uint32_t below = info[j] & 0xFFFFu;
        if (ok && below == 0)
            __hip_atomic_fetch_add(whist_w + b, info[j] >> 16, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        rk[j] = prior + below;
    }
}

// exclusive scan over the B bucket totals held in cnt[] into off[] (block-wide,
// BLOCK threads)
template <int BLOCK = kSplitBlock>
__device__ __forceinline__ void scan_buckets(const uint32_t* cnt, uint32_t* off, int B, uint32_t* wtmp) {
    constexpr int W = BLOCK / kWave;
    // each lane takes a run of per = ceil(B/BLOCK) consecutive buckets
    const int per = (B + BLOCK - 1) / BLOCK;
    const int b0 = threadIdx.x * per;
    uint32_t s = 0;
    for (int q = 0; q < per; ++q)
        if (b0 + q < B) s += cnt[b0 + q];
    uint32_t inc = wave_inclusive_add(s);
    const int w = threadIdx.x / kWave;
    if (lane_id() == kWave - 1) wtmp[w] = inc;
    __syncthreads();
    uint32_t before = 0;
    for (int q = 0; q < W; ++q) before += q < w ? wtmp[q] : 0u;
    uint32_t x = before + inc - s;
    for (int q = 0; q < per; ++q)
        if (b0 + q < B) {
            uint32_t c = cnt[b0 + q];
            off[b0 + q] = x;
            x += c;
        }
}

// Second half of a WG-strategy sub-tile (S = BLOCK*IPT elements), after
// rank_subtile: per-wave counts become sub-tile offsets, the sub-tile is
// reordered by bucket in LDS, and slot i of bucket b is written to
// run[b] + (i - off[b]), so the consecutive slots of a bucket leave as
// consecutive stores.  run[b] (each bucket's next output position) advances by
// the sub-tile's count of b; sub[] is reused for run[b] - off[b].  cnt is the
// number of valid elements (S for every sub-tile but the last).
//
// hook.publish() runs right after the sub-tile's bucket totals are in sub[],
// and hook.resolve() right before run[] is read (followed by a barrier when
// Hook::kSync), on every thread: the chained scatter publishes its counts in
// the first and sets run[] from its look-back in the second.
//
// SEP: the values follow the keys through the same buffer (vbuf == kbuf), so
// a sub-tile twice as long fits the LDS of one staged with its keys; longer
// bucket runs leave as fewer partial-line stores.  Otherwise keys and values
// are staged together and streamed out by a plain loop (each store leaves as
// soon as its slot is read, which measured faster than one burst).  Ends
// without a barrier.
struct NoPlaceHook {
    static constexpr bool kSync = false;  // resolve() writes nothing another thread reads
    __device__ __forceinline__ void publish() const {}
    __device__ __forceinline__ void resolve() const {}
    __device__ __forceinline__ void keys_placed() const {}
};

// Where place_and_store writes element i of its output (i < n; the scatter
// never forms a larger i from a consistent histogram, and an empty slot's
// 0xFFFFFFFF is dropped).  PtrSink tests the bound and stores through a
// global pointer.  BufSink stores through a raw buffer descriptor of n
// elements, whose bound the memory unit checks (an out-of-range store is
// dropped), so no per-store branch is needed; it needs n * sizeof(T) < 2^31,
// since byte offsets travel as signed 32-bit operands.
template <class T>
struct PtrSink {
    T* __restrict__ p;
    int64_t n;
    __device__ __forceinline__ void put(uint32_t i, T v) const {
        if ((int64_t)i < n) p[i] = v;
    }
};
template <class T>
struct BufSink {
    __amdgpu_buffer_rsrc_t r;
    uint32_t n;
    __device__ __forceinline__ void put(uint32_t i, T v) const {
        const int off = (int)((i < n ? i : n) * (uint32_t)sizeof(T));  // n: past the end, dropped
        if constexpr (sizeof(T) == 4) {
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, off, 0, 0);
        } else {
            typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, off, 0, 0);
        }
    }
};
template <class T>
__device__ __forceinline__ BufSink<T> buf_sink(T* p, int64_t n) {
    return {buf_rsrc(p, n), (uint32_t)n};
}

template <class K, int IPT, bool KV, int BLOCK, bool SEP, class Hook, class KS, class VS>
__device__ __forceinline__ void place_and_store_to(uint32_t* whist, uint32_t* run, uint32_t* sub, uint32_t* off,
                                                   uint32_t* wtmp, K* kbuf, int32_t* vbuf,
                                                   const uint32_t (&bk)[IPT], const uint32_t (&rk)[IPT],
                                                   const K (&key)[IPT], const int32_t (&val)[IPT], int B,
                                                   int shift, int cnt, const KS& kout, const VS& vout,
                                                   uint32_t flip, const Hook& hook) {
    constexpr int W = BLOCK / kWave;
    constexpr int S = BLOCK * IPT;
    const int w = threadIdx.x / kWave;
    const uint32_t mask = (uint32_t)B - 1u;
    // wave prefixes per bucket (in place) and sub-tile totals
    for (int b = threadIdx.x; b < B; b += BLOCK) {
        uint32_t x[W];
        #pragma unroll
        for (int q = 0; q < W; ++q) x[q] = whist[q * B + b];
        uint32_t a = 0;
        #pragma unroll
        for (int q = 0; q < W; ++q) {
            whist[q * B + b] = a;
            a += x[q];
        }
        sub[b] = a;
    }
    __syncthreads();
    hook.publish();
    scan_buckets<BLOCK>(sub, off, B, wtmp);
    __syncthreads();
    uint32_t pos[IPT];
    #pragma unroll
    for (int j = 0; j < IPT; ++j)
        pos[j] = bk[j] != 0xFFFFFFFFu ? off[bk[j]] + whist[w * B + bk[j]] + rk[j] : (uint32_t)S;
    #pragma unroll
    for (int j = 0; j < IPT; ++j) {
        if (pos[j] < (uint32_t)S) {  // false only past n, for consistent ranks
            kbuf[pos[j]] = key[j];
            if (KV && !SEP) vbuf[pos[j]] = val[j];
        }
    }
    hook.keys_placed();
    // each thread owns its buckets here: nobody reads sub[] or run[] again
    // before the barrier below
    hook.resolve();
    if constexpr (Hook::kSync) __syncthreads();
    for (int b = threadIdx.x; b < B; b += BLOCK) {
        const uint32_t c = sub[b];
        sub[b] = run[b] - off[b];
        run[b] += c;
    }
    __syncthreads();
    const uint32_t* dlt = sub;
    if constexpr (SEP && KV) {
        // slot tid + j*BLOCK; its output position is kept for the value round
        uint32_t d[IPT];
        #pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const int i = threadIdx.x + j * BLOCK;
            const K k = kbuf[i < cnt ? i : 0];
            d[j] = i < cnt ? dlt[bucket_of(k, shift, mask, flip)] + (uint32_t)i : 0xFFFFFFFFu;
            kout.put(d[j], k);  // dropped only for empty slots
        }
        __syncthreads();  // every key read before the values overwrite the buffer
        #pragma unroll
        for (int j = 0; j < IPT; ++j)
            if (pos[j] < (uint32_t)S) vbuf[pos[j]] = val[j];
        __syncthreads();
        #pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const int i = threadIdx.x + j * BLOCK;
            vout.put(d[j], vbuf[i]);
        }
    } else {
        for (int i = threadIdx.x; i < cnt; i += BLOCK) {
            const K k = kbuf[i];
            const uint32_t p = dlt[bucket_of(k, shift, mask, flip)] + (uint32_t)i;
            kout.put(p, k);
            if constexpr (KV) vout.put(p, vbuf[i]);
        }
    }
}

template <class K, int IPT, bool KV, int BLOCK = kSplitBlock, bool SEP = false, class Hook = NoPlaceHook>
__device__ __forceinline__ void place_and_store(uint32_t* whist, uint32_t* run, uint32_t* sub, uint32_t* off,
                                                uint32_t* wtmp, K* kbuf, int32_t* vbuf, const uint32_t (&bk)[IPT],
                                                const uint32_t (&rk)[IPT], const K (&key)[IPT],
                                                const int32_t (&val)[IPT], int B, int shift, int cnt,
                                                K* __restrict__ kout, int32_t* __restrict__ vout, int64_t n,
                                                uint32_t flip = 0, const Hook& hook = Hook()) {
    place_and_store_to<K, IPT, KV, BLOCK, SEP, Hook>(whist, run, sub, off, wtmp, kbuf, vbuf, bk, rk, key, val, B,
                                                     shift, cnt, PtrSink<K>{kout, n}, PtrSink<int32_t>{vout, n},
                                                     flip, hook);
}

// ------------------------------------------------------------- reorder variants
// The reference's WG split offers three ways to write a tile's elements
// (opencl/types.h:43-45, opencl/kernels/split_kernel.cl:455-665): straight
// to their output positions (NO_REORDER), through per-bucket staging buffers
// of a fixed length flushed whenever they fill (FIXED_REORDER), or through a
// buffer holding each bucket's whole run of the tile (VARIED_REORDER, which
// place_and_store above implements).  Both functions below give the same
// stable output.

// NO_REORDER: each element goes from registers to run[b] + (elements of b in
// earlier waves of the sub-tile) + rk.  whist holds the per-wave counts on
// entry; run[b] advances by the sub-tile's count of b.  Ends without a
// barrier after the run update (the caller's next barrier orders it).
template <class K, int IPT, bool KV, int BLOCK = kSplitBlock>
__device__ __forceinline__ void store_direct(uint32_t* whist, uint32_t* run, uint32_t* sub,
                                             const uint32_t (&bk)[IPT], const uint32_t (&rk)[IPT],
                                             const K (&key)[IPT], const int32_t (&val)[IPT], int B,
                                             K* __restrict__ kout, int32_t* __restrict__ vout, int64_t n) {
    constexpr int W = BLOCK / kWave;
    const int w = threadIdx.x / kWave;
    for (int b = threadIdx.x; b < B; b += BLOCK) {
        uint32_t x[W];
        #pragma unroll
        for (int q = 0; q < W; ++q) x[q] = whist[q * B + b];
        uint32_t a = 0;
        #pragma unroll
        for (int q = 0; q < W; ++q) {
            whist[q * B + b] = a;
            a += x[q];
        }
        sub[b] = a;
    }
    __syncthreads();
    #pragma unroll
    for (int j = 0; j < IPT; ++j) {
        if (bk[j] != 0xFFFFFFFFu) {
            const uint32_t p = run[bk[j]] + whist[w * B + bk[j]] + rk[j];
            if ((int64_t)p < n) {  // always true for a consistent histogram scan
                kout[p] = key[j];
                if (KV) vout[p] = val[j];
            }
        }
    }
    __syncthreads();  // every read of run[] before it advances
    for (int b = threadIdx.x; b < B; b += BLOCK) run[b] += sub[b];
}

// FIXED_REORDER: per-bucket staging buffers of kFixedCap elements in LDS
// (ck/cv, [B][kFixedCap], ccnt[b] elements held).  The sub-tile is first
// reordered by bucket in kbuf/vbuf as in place_and_store; then, for every
// bucket, the held elements followed by the sub-tile's run form one run of
// which the largest multiple of kFixedCap is written to run[b] onwards and
// the rest is held for the next sub-tile.  flush_fixed writes what is held
// at the end of the tile.  run[b] is the output position of the first held
// element.  Ends with the per-bucket updates after a barrier (the caller's
// next barrier orders them).
constexpr int kFixedCap = 16;

template <class K, int IPT, bool KV, int BLOCK = kSplitBlock>
__device__ __forceinline__ void place_fixed(uint32_t* whist, uint32_t* run, uint32_t* sub, uint32_t* off,
                                            uint32_t* wtmp, K* kbuf, int32_t* vbuf, K* ck, int32_t* cv,
                                            uint32_t* ccnt, const uint32_t (&bk)[IPT], const uint32_t (&rk)[IPT],
                                            const K (&key)[IPT], const int32_t (&val)[IPT], int B, int shift,
                                            int cnt, K* __restrict__ kout, int32_t* __restrict__ vout, int64_t n,
                                            uint32_t flip = 0) {
    constexpr int W = BLOCK / kWave;
    constexpr int S = BLOCK * IPT;
    const int w = threadIdx.x / kWave;
    const uint32_t mask = (uint32_t)B - 1u;
    for (int b = threadIdx.x; b < B; b += BLOCK) {
        uint32_t x[W];
        #pragma unroll
        for (int q = 0; q < W; ++q) x[q] = whist[q * B + b];
        uint32_t a = 0;
        #pragma unroll
        for (int q = 0; q < W; ++q) {
            whist[q * B + b] = a;
            a += x[q];
        }
        sub[b] = a;
    }
    __syncthreads();
    scan_buckets<BLOCK>(sub, off, B, wtmp);
    __syncthreads();
    #pragma unroll
    for (int j = 0; j < IPT; ++j) {
        if (bk[j] != 0xFFFFFFFFu) {
            const uint32_t pos = off[bk[j]] + whist[w * B + bk[j]] + rk[j];
            if (pos < (uint32_t)S) {
                kbuf[pos] = key[j];
                if (KV) vbuf[pos] = val[j];
            }
        }
    }
    __syncthreads();
    // flush: combined index ci of a bucket's element (held ones first);
    // fl = whole staging lengths of the combined run, written now
    for (int i = threadIdx.x; i < cnt; i += BLOCK) {
        const K k = kbuf[i];
        const uint32_t b = bucket_of(k, shift, mask, flip);
        const uint32_t ci = ccnt[b] + ((uint32_t)i - off[b]);
        const uint32_t fl = (ccnt[b] + sub[b]) / kFixedCap * kFixedCap;
        if (ci < fl) {
            const uint32_t p = run[b] + ci;
            if ((int64_t)p < n) {
                kout[p] = k;
                if (KV) vout[p] = vbuf[i];
            }
        }
    }
    for (int e = threadIdx.x; e < B * kFixedCap; e += BLOCK) {
        const int b = e / kFixedCap, k = e % kFixedCap;
        const uint32_t fl = (ccnt[b] + sub[b]) / kFixedCap * kFixedCap;
        if ((uint32_t)k < ccnt[b] && (uint32_t)k < fl) {  // fl > 0 flushes every held element
            const uint32_t p = run[b] + (uint32_t)k;
            if ((int64_t)p < n) {
                kout[p] = ck[e];
                if (KV) vout[p] = cv[e];
            }
        }
    }
    __syncthreads();  // held elements read before the new ones take their slots
    for (int i = threadIdx.x; i < cnt; i += BLOCK) {
        const K k = kbuf[i];
        const uint32_t b = bucket_of(k, shift, mask, flip);
        const uint32_t ci = ccnt[b] + ((uint32_t)i - off[b]);
        const uint32_t fl = (ccnt[b] + sub[b]) / kFixedCap * kFixedCap;
        if (ci >= fl) {  // held: ci - fl < kFixedCap
            ck[b * kFixedCap + (ci - fl)] = k;
            if (KV) cv[b * kFixedCap + (ci - fl)] = vbuf[i];
        }
    }
    __syncthreads();  // every read of ccnt, sub and run before they change
    for (int b = threadIdx.x; b < B; b += BLOCK) {
        const uint32_t fl = (ccnt[b] + sub[b]) / kFixedCap * kFixedCap;
        ccnt[b] = ccnt[b] + sub[b] - fl;
        run[b] += fl;
    }
}

// FIXED_REORDER, end of a tile: every held element to run[b] onwards
template <class K, bool KV, int BLOCK = kSplitBlock>
__device__ __forceinline__ void flush_fixed(const uint32_t* run, const K* ck, const int32_t* cv,
                                            const uint32_t* ccnt, int B, K* __restrict__ kout,
                                            int32_t* __restrict__ vout, int64_t n) {
    for (int e = threadIdx.x; e < B * kFixedCap; e += BLOCK) {
        const int b = e / kFixedCap, k = e % kFixedCap;
        if ((uint32_t)k < ccnt[b]) {
            const uint32_t p = run[b] + (uint32_t)k;
            if ((int64_t)p < n) {
                kout[p] = ck[e];
                if (KV) vout[p] = cv[e];
            }
        }
    }
}

}  // namespace edpp
