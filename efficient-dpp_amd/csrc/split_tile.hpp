// split_tile.hpp — stable ranking of bucketed keys inside a workgroup tile,
// shared by the split kernels and the batched executor.
#pragma once
#include <algorithm>
#include "device.hpp"

namespace edpp {

constexpr int kSplitBlock = 256;
constexpr int kSplitWaves = kSplitBlock / kWave;

// Mask of the lanes whose bucket equals the caller's, among valid lanes.
__device__ __forceinline__ uint64_t match_bucket(uint32_t b, int lg, uint64_t valid) {
    uint64_t m = valid;
    for (int bit = 0; bit < lg; ++bit) {
        bool on = (b >> bit) & 1u;
        uint64_t bal = ballot(on);
        m &= on ? bal : ~bal;
    }
    return m;
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
// LDS layout of the WG scatter: whist[waves][B], run[B], sub[B], off[B],
// wtmp[16], then the key and value reorder buffers from a 16-byte boundary
__host__ __device__ __forceinline__ int kbuf_word_offset(int B) {
    return (((kSplitWaves + 3) * B + 16) + 3) & ~3;
}

// ------------------------------------------------------------- WG strategy
// Tile = SUBS sub-tiles of S = 256*IPT elements.  In a sub-tile, wave w owns
// [w*64*IPT, (w+1)*64*IPT) and item j of lane l is element w*64*IPT + j*64 + l,
// so ranks follow input order: sub-tile, wave, item, lane.

struct SplitGeom {
    int64_t n;
    int lg;          // log2 B
    int shift;       // bucket = (key >> shift) & (B-1)
    int64_t tile;    // elements per tile
    int64_t tiles;
    int xcds;
};

// per-wave counts of one sub-tile into whist[w][b] (zeroed by the caller)
template <class K, int IPT>
__device__ __forceinline__ void rank_subtile(const K* __restrict__ kin, int64_t sbase, int64_t n, int lg, int shift,
                                             uint32_t* whist_w, uint32_t (&bk)[IPT], uint32_t (&rk)[IPT],
                                             K (&key)[IPT]) {
    const int lane = lane_id();
    const int w = threadIdx.x / kWave;
    const uint32_t mask = (1u << lg) - 1u;
    const int64_t wbase = sbase + (int64_t)w * kWave * IPT;
    #pragma unroll
    for (int j = 0; j < IPT; ++j) {
        int64_t e = wbase + (int64_t)j * kWave + lane;
        bool ok = e < n;
        key[j] = ok ? kin[e] : K(0);
        bk[j] = ok ? (((uint32_t)key[j] >> shift) & mask) : 0xFFFFFFFFu;
    }
    #pragma unroll
    for (int j = 0; j < IPT; ++j) {
        bool ok = bk[j] != 0xFFFFFFFFu;
        uint64_t valid = ballot(ok);
        uint64_t m = match_bucket(bk[j], lg, valid);
        uint32_t below = count_below(m);
        uint32_t cnt = (uint32_t)__popcll(m);
        uint32_t prior = ok ? whist_w[bk[j]] : 0u;
        rk[j] = prior + below;
        // the highest lane of each group updates the count; the read above
        // precedes this write in the wave's LDS instruction order
        if (ok && below == cnt - 1) whist_w[bk[j]] = prior + cnt;
    }
}

// exclusive scan over the B bucket totals held in cnt[] into off[] (block-wide)
__device__ __forceinline__ void scan_buckets(const uint32_t* cnt, uint32_t* off, int B, uint32_t* wtmp) {
    // each lane takes a run of per = ceil(B/256) consecutive buckets
    const int per = (B + kSplitBlock - 1) / kSplitBlock;
    const int b0 = threadIdx.x * per;
    uint32_t s = 0;
    for (int q = 0; q < per; ++q)
        if (b0 + q < B) s += cnt[b0 + q];
    uint32_t inc = wave_inclusive_add(s);
    const int w = threadIdx.x / kWave;
    if (lane_id() == kWave - 1) wtmp[w] = inc;
    __syncthreads();
    uint32_t before = 0;
    for (int q = 0; q < kSplitWaves; ++q) before += q < w ? wtmp[q] : 0u;
    uint32_t x = before + inc - s;
    for (int q = 0; q < per; ++q)
        if (b0 + q < B) {
            uint32_t c = cnt[b0 + q];
            off[b0 + q] = x;
            x += c;
        }
}

}  // namespace edpp
