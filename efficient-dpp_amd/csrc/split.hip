// split.hip — stable multi-split: bucket = key & (B-1), input order kept
// inside every bucket, optional bucket start offsets.
//
// Reference: opencl/primitives/splitImpl.cpp:21-498 (WI / WG / single host
// code), opencl/kernels/split_kernel.cl:320-801 (kernels),
// cuda/test_split.cu:18-88 (warp-ballot multisplit).  The reference's fast
// GPU splits rank with LDS atomic_inc and are therefore not stable
// (split_kernel.cl:445,488,543); here every rank comes from ballot match
// masks, so the output is the sequential stable partition.
//
// Strategies (all three produce identical output):
//   WG      per-tile LDS histograms -> exclusive scan of the bucket-major
//           [B][tiles] matrix -> per-tile scatter through an LDS reorder
//           buffer so bucket runs leave the workgroup as contiguous stores
//           (the reference's WG_split with VARIED_REORDER).
//   WI      the same with one wave per histogram column and stores straight
//           from registers (WI_split / the CUDA warp multisplit).
//   SINGLE  one lane walks a contiguous chunk (single_split).
#include <algorithm>
#include <type_traits>
#include "ctx.hpp"
#include "split_tile.hpp"
#include "tile_scan.hpp"

namespace edpp {

// Per-tile bucket counts his[b*tiles + t], counted in sub-tiles of
// S = 256*IPT elements (spt of them per tile).  With TPW > 1 a workgroup
// counts TPW consecutive sub-tiles, that is TPW/spt whole tiles (spt divides
// TPW), so that every lane has TPW*IPT loads in flight and the workgroup's
// fixed costs (LDS clearing, barriers, the column writes) are shared; that
// pays where the per-key LDS atomics dominate (many buckets).
template <class K, int IPT, int TPW, int REGW>
__global__ __launch_bounds__(kSplitBlock) void k_split_hist(const K* __restrict__ kin, SplitGeom g, int spt,
                                                            int32_t* __restrict__ his) {
    extern __shared__ uint32_t lds[];
    const int B = 1 << g.lg;
    constexpr int S = kSplitBlock * IPT;
    const int tpw = TPW > 1 ? TPW / spt : 1;  // whole tiles per workgroup
    uint32_t* whist = lds;                    // [tpw][waves][B]
    const int64_t groups = (g.tiles + tpw - 1) / tpw;
    const int64_t t0 = xcd_tile(blockIdx.x, groups, g.xcds) * tpw;
    for (int b = threadIdx.x; b < tpw * kSplitWaves * B; b += kSplitBlock) whist[b] = 0;
    __syncthreads();
    if (TPW > 1) {
        count_tiles<K, IPT, TPW, REGW>(kin, t0 * g.tile, g.n, g.lg, g.shift, spt, whist, g.flip);
    } else {
        const int64_t tbase = t0 * g.tile;
        const int64_t tend = std::min<int64_t>(tbase + g.tile, g.n);
        for (int64_t s = tbase; s < tend; s += S)
            count_tiles<K, IPT, 1, REGW>(kin, s, tend, g.lg, g.shift, 1, whist, g.flip);
    }
    __syncthreads();
    // bucket-major: consecutive threads write consecutive tiles of one bucket
    for (int i = threadIdx.x; i < tpw * B; i += kSplitBlock) {
        const int b = i / tpw, t = i % tpw;
        if (t0 + t < g.tiles) {
            uint32_t a = 0;
            #pragma unroll
            for (int q = 0; q < kSplitWaves; ++q) a += whist[(t * kSplitWaves + q) * B + b];
            his[(int64_t)b * g.tiles + t0 + t] = (int32_t)a;
        }
    }
}

// Reorder variants of the scatter (the reference's ReorderType,
// opencl/types.h:43-45): each bucket's run of a sub-tile staged whole in LDS
// (varied, the default), elements stored straight from registers (none), or
// fixed-length per-bucket staging buffers flushed when they fill (fixed).
enum : int { kReorderVaried = 0, kReorderNone = 1, kReorderFixed = 2 };

// LDS words in front of the fixed variant's staging buffers: the varied
// layout of kbuf_word_offset plus a sub-tile of keys and values
template <class K>
__host__ __device__ constexpr int fixed_ccnt_word_offset(int B, int S) {
    return kbuf_word_offset(B) + (int)((S * (sizeof(K) + 4) + 3) / 4);
}
template <class K>
__host__ __device__ constexpr int fixed_ck_word_offset(int B, int S) {
    return (fixed_ccnt_word_offset<K>(B, S) + B + 3) & ~3;  // 16-byte aligned
}
template <class K>
__host__ __device__ constexpr size_t fixed_lds_bytes(int B, int S) {
    return (size_t)fixed_ck_word_offset<K>(B, S) * 4 + (size_t)B * kFixedCap * (sizeof(K) + 4);
}

// place_and_store hook of the scatter whose values follow its keys through
// LDS: the sub-tile's values are loaded once its keys are placed.
template <int IPT>
struct LateValues {
    static constexpr bool kSync = false;
    int32_t (&val)[IPT];
    const int32_t* __restrict__ vin;
    int64_t s, n;
    __device__ __forceinline__ void publish() const {}
    __device__ __forceinline__ void resolve() const {}
    __device__ __forceinline__ void keys_placed() const { load_chunk_buf<int32_t, IPT>(vin, s, n, val); }
};

// BLOCK threads; sub-tile S = BLOCK*IPT.  Above 256 threads the values are
// staged after the keys through one buffer (place_and_store's SEP), and the
// register budget is held to two workgroups per CU (4 waves per SIMD), which
// their LDS allows.  RO: kReorder*; the none and fixed variants run with
// 256 lanes.  Keys and values are loaded through raw buffer descriptors
// (no branch per load); BST: the varied variant also stores through them,
// for outputs whose byte offsets fit 31 bits.
template <class K, int IPT, bool KV, bool BALLOT, int BLOCK, int RO = kReorderVaried, bool BST = false>
__global__ __launch_bounds__(BLOCK, (BLOCK > kSplitBlock ? 2 * BLOCK / (4 * kWave) : 1)) void k_split_scatter(const K* __restrict__ kin, K* __restrict__ kout,
                                                         const int32_t* __restrict__ vin, int32_t* __restrict__ vout,
                                                         SplitGeom g, const int32_t* __restrict__ scanned) {
    extern __shared__ uint32_t lds[];
    constexpr int W = BLOCK / kWave;
    constexpr bool SEP = KV && BLOCK > kSplitBlock;
    static_assert(RO == kReorderVaried || BLOCK == kSplitBlock, "the none and fixed variants use 256 lanes");
    const int B = 1 << g.lg;
    constexpr int S = BLOCK * IPT;
    uint64_t* wmask = reinterpret_cast<uint64_t*>(lds);  // [waves][B], zero between items
    uint32_t* whist = lds + whist_word_offset(B, W);     // [waves][B]
    uint32_t* run = whist + W * B;               // [B] next output position of each bucket
    uint32_t* sub = run + B;                     // [B] sub-tile bucket counts
    uint32_t* off = sub + B;                     // [B] sub-tile bucket offsets
    uint32_t* wtmp = off + B;                    // [waves]
    K* kbuf = reinterpret_cast<K*>(lds + kbuf_word_offset(B, W));  // [S], 16-byte aligned
    int32_t* vbuf = SEP ? reinterpret_cast<int32_t*>(kbuf) : reinterpret_cast<int32_t*>(kbuf + S);
    const int64_t tile = xcd_tile(blockIdx.x, g.tiles, g.xcds);
    const int64_t tbase = tile * g.tile;
    const int64_t tend = std::min<int64_t>(tbase + g.tile, g.n);
    const int w = threadIdx.x / kWave;
    uint32_t* ccnt = lds + fixed_ccnt_word_offset<K>(B, S);             // fixed: [B] held elements
    K* ck = reinterpret_cast<K*>(lds + fixed_ck_word_offset<K>(B, S));  // fixed: [B][kFixedCap] held keys
    int32_t* cv = reinterpret_cast<int32_t*>(ck + B * kFixedCap);       // fixed: held values
    for (int b = threadIdx.x; b < B; b += BLOCK) {
        run[b] = (uint32_t)scanned[(int64_t)b * g.tiles + tile];
        if (RO == kReorderFixed) ccnt[b] = 0;
    }
    for (int b = threadIdx.x; b < W * B; b += BLOCK) wmask[b] = 0ull;
    for (int64_t s = tbase; s < tend; s += S) {
        for (int b = threadIdx.x; b < W * B; b += BLOCK) whist[b] = 0;
        __syncthreads();
        uint32_t bk[IPT], rk[IPT];
        K key[IPT];
        int32_t val[IPT];
        // values staged after the keys are loaded once the keys are in LDS
        // (LateValues), so they are not held through the ranking
        rank_subtile<K, IPT, BALLOT, true>(kin, KV && !(SEP && BST) ? vin : nullptr, s, tend, g.lg, g.shift,
                                           whist + w * B, wmask + w * B, bk, rk, key, val, g.flip);
        __syncthreads();
        const int cnt = (int)std::min<int64_t>((int64_t)S, tend - s);
        if (RO == kReorderNone)
            store_direct<K, IPT, KV, BLOCK>(whist, run, sub, bk, rk, key, val, B, kout, vout, g.n);
        else if (RO == kReorderFixed)
            place_fixed<K, IPT, KV, BLOCK>(whist, run, sub, off, wtmp, kbuf, vbuf, ck, cv, ccnt, bk, rk, key, val, B,
                                           g.shift, cnt, kout, vout, g.n, g.flip);
        else if constexpr (BST && SEP)
            place_and_store_to<K, IPT, KV, BLOCK, SEP>(whist, run, sub, off, wtmp, kbuf, vbuf, bk, rk, key, val, B,
                                                       g.shift, cnt, buf_sink(kout, g.n), buf_sink(vout, g.n),
                                                       g.flip, LateValues<IPT>{val, vin, s, tend});
        else if constexpr (BST)
            place_and_store_to<K, IPT, KV, BLOCK, SEP>(whist, run, sub, off, wtmp, kbuf, vbuf, bk, rk, key, val, B,
                                                       g.shift, cnt, buf_sink(kout, g.n), buf_sink(vout, g.n),
                                                       g.flip, NoPlaceHook());
        else
            place_and_store<K, IPT, KV, BLOCK, SEP>(whist, run, sub, off, wtmp, kbuf, vbuf, bk, rk, key, val, B,
                                                    g.shift, cnt, kout, vout, g.n, g.flip);
        // the next iteration's first barrier ends every thread's stores
        // before kbuf, sub and run are used again
    }
    if (RO == kReorderFixed) {
        __syncthreads();
        flush_fixed<K, KV, BLOCK>(run, ck, cv, ccnt, B, kout, vout, g.n);
    }
}

// d_start[b] = scanned[b*tiles] (opencl/kernels/split_kernel.cl:320-332)
__global__ void k_split_start(const int32_t* scanned, int64_t tiles, int B, int32_t* d_start) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x)
        d_start[b] = scanned[(int64_t)b * tiles];
}

// ------------------------------------------------------------- chained strategy
// A pass counts its digit per range of tiles (bucket totals of each range,
// not of each tile); then its scatter finds each tile's bucket offsets by a
// look-back over predecessor tiles' per-bucket counts instead of a
// [B][tiles] matrix written by a histogram kernel and scanned.  The tiles
// are cut into one contiguous range per XCD: workgroup g takes tiles in
// order from range g % xcds (the XCD it runs on), so neighbouring tiles,
// whose runs of a bucket share cache lines, are written through one L2, and
// a look-back never leaves its range, whose base offset per bucket the
// counts give.  A workgroup whose range is used up takes a tile from another
// range; since every range hands out its tiles in order and a tile waits
// only for earlier tiles of its own range, each of which belongs to a
// running workgroup, every wait ends.

constexpr int kChainMaxDigits = 4;  // 32 key bits in digits of at most 8
constexpr int kChainMaxRanges = 8;
constexpr int kChainStride = 256;   // bucket stride of the count and base tables
constexpr int kChainWindow = 8;     // predecessors one lane reads per round trip

struct ChainRanges {
    int64_t tiles;
    int ranges;  // 1 .. kChainMaxRanges, at most tiles
    // first tile of range r (r == ranges: tiles); ranges differ by at most one tile
    __host__ __device__ int64_t first(int r) const {
        const int64_t per = tiles / ranges, rem = tiles % ranges;
        return (int64_t)r * per + std::min<int64_t>(r, rem);
    }
};

struct ChainDigits {
    int count;
    int lg[kChainMaxDigits];
    int shift[kChainMaxDigits];
    uint32_t flip[kChainMaxDigits];
};

// counts[(d*ranges + r)*kChainStride + b] += number of keys of range r whose
// digit d is b, for up to kChainMaxDigits digits per read of the keys.
// Workgroup j of range r counts the j-th of wpr contiguous parts of the
// range into per-wave LDS rows [waves][digits][kChainStride]; one atomic per
// nonzero count at the end.
template <class K, int IPT>
__global__ __launch_bounds__(kSplitBlock) void k_chain_hist(const K* __restrict__ kin, int64_t n, int64_t tile,
                                                            ChainRanges cr, int wpr, ChainDigits dg,
                                                            uint32_t* __restrict__ counts) {
    extern __shared__ uint32_t lds[];
    constexpr int S = kSplitBlock * IPT;
    const int r = blockIdx.x / wpr, j = blockIdx.x % wpr;
    const int D = dg.count;
    const int64_t r_lo = cr.first(r) * tile;
    const int64_t r_hi = std::min<int64_t>(cr.first(r + 1) * tile, n);
    const int64_t len = std::max<int64_t>(r_hi - r_lo, 0);
    const int64_t part = ((len + wpr - 1) / wpr + S - 1) / S * S;
    const int64_t lo = r_lo + (int64_t)j * part;
    const int64_t hi = std::min<int64_t>(lo + part, r_hi);
    for (int i = threadIdx.x; i < kSplitWaves * D * kChainStride; i += kSplitBlock) lds[i] = 0;
    __syncthreads();
    const int w = threadIdx.x / kWave;
    for (int64_t s = lo; s < hi; s += S) {
        const int64_t e0 = s + (int64_t)w * kWave * IPT + lane_id();
        K key[IPT];
        #pragma unroll
        for (int q = 0; q < IPT; ++q) {
            const int64_t e = e0 + (int64_t)q * kWave;
            key[q] = e < hi ? kin[e] : K(0);
        }
        __builtin_amdgcn_sched_barrier(0);
        for (int d = 0; d < D; ++d) {
            uint32_t* row = lds + (w * D + d) * kChainStride;
            const int lg = dg.lg[d];
            switch (lg <= kRegCountBits ? reg_count_words(lg) : 0) {
                case 1: count_keys<K, IPT, 1>(key, e0, hi, lg, dg.shift[d], dg.flip[d], row); break;
                case 2: count_keys<K, IPT, 2>(key, e0, hi, lg, dg.shift[d], dg.flip[d], row); break;
                case 4: count_keys<K, IPT, 4>(key, e0, hi, lg, dg.shift[d], dg.flip[d], row); break;
                default: count_keys<K, IPT, 0>(key, e0, hi, lg, dg.shift[d], dg.flip[d], row);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < D * kChainStride; i += kSplitBlock) {
        const int d = i / kChainStride, b = i % kChainStride;
        if (b < (1 << dg.lg[d])) {
            uint32_t a = 0;
            for (int q = 0; q < kSplitWaves; ++q) a += lds[(q * D + d) * kChainStride + b];
            if (a) fetch_add(counts + (d * cr.ranges + r) * kChainStride + b, a);
        }
    }
}

// base[(d*ranges + r)*kChainStride + b] = output position of the first key of
// range r in bucket b of digit d: the keys of lower buckets, plus those of
// bucket b in earlier ranges.  One workgroup per digit; d_start (digit 0)
// receives the bucket starts.
__global__ __launch_bounds__(kSplitBlock) void k_chain_bases(const uint32_t* __restrict__ counts, int ranges,
                                                             ChainDigits dg, uint32_t* __restrict__ base,
                                                             int32_t* __restrict__ d_start) {
    __shared__ uint32_t tot[kChainStride], off[kChainStride], wtmp[kSplitWaves];
    const int d = blockIdx.x, b = threadIdx.x;
    const int B = 1 << dg.lg[d];
    uint32_t c[kChainMaxRanges];
    uint32_t t = 0;
    #pragma unroll
    for (int r = 0; r < kChainMaxRanges; ++r) {
        c[r] = (r < ranges && b < B) ? counts[(d * ranges + r) * kChainStride + b] : 0u;
        t += c[r];
    }
    tot[b] = t;
    __syncthreads();
    scan_buckets<kSplitBlock>(tot, off, B, wtmp);
    __syncthreads();
    if (b < B) {
        uint32_t x = off[b];
        #pragma unroll
        for (int r = 0; r < kChainMaxRanges; ++r) {
            if (r < ranges) base[(d * ranges + r) * kChainStride + b] = x;
            x += c[r];
        }
        if (d_start && d == 0) d_start[b] = (int32_t)off[b];
    }
}

struct ChainPass {
    LookbackState32 st;      // status word of tile t, bucket b at t*B + b
    const uint32_t* base;    // [ranges][kChainStride] of this digit
    uint32_t* tickets;       // [ranges], zero before the pass
    ChainRanges cr;
    int xcds;
    uint32_t spin_limit;
    int32_t* err;
};

// The chained scatter's part of place_and_store.  publish(): thread b < B
// publishes its bucket's count of the tile (as the inclusive prefix for a
// range's first tile, else as an aggregate).  resolve(): L consecutive lanes
// per bucket (L = BLOCK/B, at most a wave) walk back over the range's
// earlier tiles, L*kChainWindow of them per round trip, nearest first,
// adding aggregates until they meet an inclusive prefix (a round whose
// nearest words up to that prefix include an unpublished one is read again);
// the segment's first lane publishes the tile's inclusive prefix and sets
// run[b].  A barrier follows resolve(), since run[b] is read by thread b.
struct ChainHook {
    static constexpr bool kSync = true;
    LookbackState32 st;
    const uint32_t* base;
    uint32_t spin_limit;
    int32_t* err;
    int64_t tile, first;
    int range, B, L;
    const uint32_t* sub;
    uint32_t* run;
    __device__ __forceinline__ void keys_placed() const {}
    __device__ __forceinline__ void publish() const {
        const int b = threadIdx.x;
        if (b < B) st.publish(tile * B + b, tile == first ? kFlagPrefix : kFlagAggregate, sub[b]);
    }
    __device__ __forceinline__ void resolve() const {
        const int x = threadIdx.x;
        if (x >= B * L) return;  // whole waves
        const int b = x / L, j = x % L;
        const int sbase = lane_id() & ~(L - 1);
        const uint64_t smask = L == kWave ? ~0ull : (1ull << L) - 1ull;
        uint32_t excl = 0;
        if (tile > first) {
            int64_t cur = tile - 1;  // nearest predecessor not yet added
            uint32_t spins = 0;
            while (true) {
                uint32_t f[kChainWindow], v[kChainWindow];
                #pragma unroll
                for (int q = 0; q < kChainWindow; ++q) {
                    const int64_t p = cur - j - (int64_t)L * q;
                    if (p >= first) {
                        st.read(p * B + b, f[q], v[q]);
                    } else {
                        f[q] = kFlagPrefix;  // before the range: its base, added below
                        v[q] = 0;
                    }
                }
                bool done = false, stalled = false;
                #pragma unroll
                for (int q = 0; q < kChainWindow; ++q) {
                    if (done || stalled) continue;
                    const uint64_t pm = (ballot(f[q] == kFlagPrefix) >> sbase) & smask;
                    const uint64_t zm = (ballot(f[q] == kFlagNone) >> sbase) & smask;
                    const uint64_t upto = pm ? (pm ^ (pm - 1)) : smask;  // lanes up to the nearest prefix
                    if (zm & upto) {
                        stalled = true;
                        continue;
                    }
                    uint32_t a = ((upto >> j) & 1ull) ? v[q] : 0u;
                    for (int o = 1; o < L; o <<= 1) a += (uint32_t)__shfl_xor((int)a, o);
                    excl += a;
                    if (pm) done = true;
                    else cur -= L;
                }
                if (done) break;
                if (stalled) {
                    if (++spins > spin_limit) {
                        if (j == 0) st_relaxed(err, (int32_t)EDPP_ETIMEOUT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            if (j == 0) st.publish(tile * B + b, kFlagPrefix, excl + sub[b]);
        }
        if (j == 0) run[b] = base[range * kChainStride + b] + excl;
    }
};

// One tile (= one sub-tile of S = BLOCK*IPT elements) per workgroup; the
// grid has one workgroup per tile.  LDS layout and the ranking are those of
// k_split_scatter (varied reorder).
template <class K, int IPT, bool KV, bool BALLOT, int BLOCK>
__global__ __launch_bounds__(BLOCK, (BLOCK > kSplitBlock ? 2 * BLOCK / (4 * kWave) : 1)) void k_chain_scatter(const K* __restrict__ kin, K* __restrict__ kout,
                                                         const int32_t* __restrict__ vin, int32_t* __restrict__ vout,
                                                         SplitGeom g, ChainPass cp) {
    extern __shared__ uint32_t lds[];
    __shared__ int64_t s_tile;
    __shared__ int s_range;
    constexpr int W = BLOCK / kWave;
    constexpr bool SEP = KV && BLOCK > kSplitBlock;
    const int B = 1 << g.lg;
    constexpr int S = BLOCK * IPT;
    uint64_t* wmask = reinterpret_cast<uint64_t*>(lds);
    uint32_t* whist = lds + whist_word_offset(B, W);
    uint32_t* run = whist + W * B;
    uint32_t* sub = run + B;
    uint32_t* off = sub + B;
    uint32_t* wtmp = off + B;
    K* kbuf = reinterpret_cast<K*>(lds + kbuf_word_offset(B, W));
    int32_t* vbuf = SEP ? reinterpret_cast<int32_t*>(kbuf) : reinterpret_cast<int32_t*>(kbuf + S);
    if (threadIdx.x == 0) {
        const int R = cp.cr.ranges;
        const int r0 = (int)(blockIdx.x % (unsigned)cp.xcds) % R;
        int64_t t = -1;
        int rr = 0;
        for (int k = 0; k < R && t < 0; ++k) {
            const int r = r0 + k < R ? r0 + k : r0 + k - R;
            const int64_t f = cp.cr.first(r), len = cp.cr.first(r + 1) - f;
            if ((int64_t)ld_relaxed(cp.tickets + r) >= len) continue;
            const int64_t q = (int64_t)fetch_add(cp.tickets + r, 1u);
            if (q < len) {
                t = f + q;
                rr = r;
            }
        }
        s_tile = t;
        s_range = rr;
    }
    for (int b = threadIdx.x; b < W * B; b += BLOCK) {
        wmask[b] = 0ull;
        whist[b] = 0;
    }
    __syncthreads();
    const int64_t tile = s_tile;
    if (tile < 0) return;  // more workgroups than tiles (never launched so)
    const int range = s_range;
    const int64_t tbase = tile * S;
    const int64_t tend = std::min<int64_t>(tbase + S, g.n);
    const int w = threadIdx.x / kWave;
    uint32_t bk[IPT], rk[IPT];
    K key[IPT];
    int32_t val[IPT];
    rank_subtile<K, IPT, BALLOT>(kin, KV ? vin : nullptr, tbase, tend, g.lg, g.shift, whist + w * B, wmask + w * B, bk,
                                 rk, key, val, g.flip);
    __syncthreads();
    const int L = std::min(kWave, std::max(1, BLOCK >> g.lg));
    const ChainHook hook = {cp.st, cp.base, cp.spin_limit, cp.err, tile, cp.cr.first(range), range, B, L, sub, run};
    place_and_store<K, IPT, KV, BLOCK, SEP, ChainHook>(whist, run, sub, off, wtmp, kbuf, vbuf, bk, rk, key, val, B,
                                                       g.shift, (int)(tend - tbase), kout, vout, g.n, g.flip, hook);
}

// ------------------------------------------------------------- WI strategy
// One histogram column per wave: the wave walks `colen` consecutive elements
// (a multiple of 64*IPT) in rounds, keeping per-bucket counts in its own LDS
// row, so no cross-wave work is needed.  The scatter starts each row at the
// scanned column offsets, which turns the running counts into absolute
// output positions.
template <class K, int IPT, bool BALLOT>
__global__ __launch_bounds__(kSplitBlock) void k_split_wave_hist(const K* __restrict__ kin, int64_t n, int lg,
                                                                 int shift, int64_t cols, int64_t colen,
                                                                 int32_t* __restrict__ his) {
    extern __shared__ uint32_t lds[];
    const int B = 1 << lg;
    const int w = threadIdx.x / kWave;
    uint32_t* h = lds + w * B;
    const int64_t col = (int64_t)blockIdx.x * kSplitWaves + w;
    for (int b = lane_id(); b < B; b += kWave) h[b] = 0;
    __builtin_amdgcn_wave_barrier();
    if (col >= cols) return;
    const int64_t c0 = col * colen, c1 = std::min<int64_t>(c0 + colen, n);
    // count_subtile addresses the calling wave's chunk at sbase + w*64*IPT
    for (int64_t s = c0; s < c1; s += (int64_t)kWave * IPT)
        count_subtile<K, IPT, BALLOT>(kin, s - (int64_t)w * kWave * IPT, c1, lg, shift, h);
    __builtin_amdgcn_wave_barrier();
    for (int b = lane_id(); b < B; b += kWave) his[(int64_t)b * cols + col] = (int32_t)h[b];
}

template <class K, int IPT, bool KV, bool BALLOT>
__global__ __launch_bounds__(kSplitBlock) void k_split_wave_scatter(const K* __restrict__ kin, K* __restrict__ kout,
                                                                    const int32_t* __restrict__ vin,
                                                                    int32_t* __restrict__ vout, int64_t n, int lg,
                                                                    int shift, int64_t cols, int64_t colen,
                                                                    const int32_t* __restrict__ scanned) {
    extern __shared__ uint32_t lds[];
    const int B = 1 << lg;
    const int w = threadIdx.x / kWave;
    const int lane = lane_id();
    uint64_t* wmask = reinterpret_cast<uint64_t*>(lds) + w * B;
    uint32_t* h = lds + 2 * kSplitWaves * B + w * B;
    const int64_t col = (int64_t)blockIdx.x * kSplitWaves + w;
    if (col >= cols) return;
    for (int b = lane; b < B; b += kWave) {
        h[b] = (uint32_t)scanned[(int64_t)b * cols + col];
        wmask[b] = 0ull;
    }
    __builtin_amdgcn_wave_barrier();
    const int64_t c0 = col * colen, c1 = std::min<int64_t>(c0 + colen, n);
    for (int64_t s = c0; s < c1; s += (int64_t)kWave * IPT) {
        uint32_t bk[IPT], rk[IPT];
        K key[IPT];
        int32_t val[IPT];
        rank_subtile<K, IPT, BALLOT>(kin, KV ? vin : nullptr, s - (int64_t)w * kWave * IPT, c1, lg, shift, h, wmask,
                                     bk, rk, key, val);
        #pragma unroll
        for (int j = 0; j < IPT; ++j) {
            if (bk[j] != 0xFFFFFFFFu && rk[j] < (uint64_t)n) {  // rk is the output position
                kout[rk[j]] = key[j];
                if (KV) vout[rk[j]] = val[j];
            }
        }
    }
}

// ------------------------------------------------------------- SINGLE strategy
// Lane g owns [g*len, (g+1)*len) and keeps its histogram column his[b][g].
template <class K>
__global__ void k_split_single_hist(const K* kin, int64_t n, int lg, int shift, int64_t len, int64_t lanes,
                                    int32_t* his) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= lanes) return;
    const int B = 1 << lg;
    for (int b = 0; b < B; ++b) his[(int64_t)b * lanes + g] = 0;
    const int64_t e0 = g * len, e1 = std::min<int64_t>(e0 + len, n);
    for (int64_t e = e0; e < e1; ++e) his[(int64_t)(((uint32_t)kin[e] >> shift) & (B - 1)) * lanes + g] += 1;
}
template <class K, bool KV>
__global__ void k_split_single_scatter(const K* kin, K* kout, const int32_t* vin, int32_t* vout, int64_t n, int lg,
                                       int shift, int64_t len, int64_t lanes, int32_t* scanned) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= lanes) return;
    const int B = 1 << lg;
    const int64_t e0 = g * len, e1 = std::min<int64_t>(e0 + len, n);
    for (int64_t e = e0; e < e1; ++e) {
        K k = kin[e];
        int32_t* slot = scanned + (int64_t)(((uint32_t)k >> shift) & (B - 1)) * lanes + g;
        int32_t p = *slot;
        *slot = p + 1;
        if ((uint32_t)p < (uint64_t)n) {
            kout[p] = k;
            if (KV) vout[p] = vin[e];
        }
    }
}

// d_start[b] = first position of bucket b in the split output kout, which is
// ordered by bucket (a bucket absent from the output starts where the next
// present one does, or at n).  A group of G lanes searches for one bucket,
// kWave/G buckets per wave, (G+1)-way: every round the group's lanes test G
// evenly spaced positions of [lo, hi) at once, and one ballot keeps the gap
// between the last position below the bucket and the first one at or above
// it.  Invariant: every key before lo lies in a smaller bucket, every key
// from hi on in bucket b or a larger one.
template <class K, int G>
__global__ void k_split_start_search(const K* __restrict__ kout, int64_t n, int lg, int32_t* __restrict__ d_start) {
    constexpr int GPW = kWave / G;  // buckets per wave
    constexpr uint64_t kGroupMask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    const int B = 1 << lg;
    const uint32_t mask = (uint32_t)B - 1u;
    const int lane = lane_id();
    const int gl = lane % G, grp = lane / G;
    const int64_t waves = (int64_t)gridDim.x * blockDim.x / kWave;
    for (int64_t b0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave * GPW; b0 < B; b0 += waves * GPW) {
        // a group past the last bucket (B < GPW) searches for b >= B: every
        // key lies below it, every position it loads lies in [0, n), and its
        // result is not stored
        const uint32_t b = (uint32_t)b0 + (uint32_t)grp;
        int64_t lo = 0, hi = n;
        while (ballot(hi - lo > G)) {
            const int64_t len = hi - lo;
            const bool act = len > G;
            bool below = false;
            if (act) {
                // positions lo + (l+1)*len/(G+1), l = 0..G-1: strictly
                // increasing (len > G) and all inside [lo, hi)
                const int64_t p = lo + ((int64_t)(gl + 1) * len) / (G + 1);
                below = bucket_of(kout[p], 0, mask, 0u) < b;
            }
            // positions below b: the group's lanes 0..c-1
            const int c = __popcll((ballot(below) >> (grp * G)) & kGroupMask);
            if (act) {
                const int64_t nlo = c > 0 ? lo + ((int64_t)c * len) / (G + 1) + 1 : lo;
                const int64_t nhi = c < G ? lo + ((int64_t)(c + 1) * len) / (G + 1) : hi;
                lo = nlo;
                hi = nhi;
            }
        }
        // at most G positions left: one test each
        const int64_t e = lo + gl;
        const bool below = e < hi && bucket_of(kout[e], 0, mask, 0u) < b;
        const int c = __popcll((ballot(below) >> (grp * G)) & kGroupMask);
        if (gl == 0 && b < (uint32_t)B) d_start[b] = (int32_t)(lo + c);
    }
}

// launches the bucket-start search with 16 lanes per bucket (17-way rounds).
// Lanes per bucket trade rounds of dependent loads against lines fetched:
// over 2^24 keys 64 lanes take 3 rounds and a final test but fetch about 64
// lines per bucket in their last round, 16 lanes take 5 rounds and fetch at
// most 16 lines per round.  A join of 2^24 x 2^24 keys (4096 partitions)
// measured 450 us with 16 or 8 lanes, 454 with 4 and 457 with 64.
template <class K>
void launch_start_search(edpp_ctx* c, const K* kout, int64_t n, int lg, int32_t* d_start) {
    constexpr int G = 16;
    const int64_t B = (int64_t)1 << lg;
    const unsigned grid = (unsigned)ceil_div(B, (int64_t)(256 / kWave) * (kWave / G));
    hipLaunchKernelGGL((k_split_start_search<K, G>), dim3(grid), dim3(256), 0, c->stream, kout, n, lg, d_start);
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ------------------------------------------------------------- host side
namespace {

constexpr int kHistTilesPerWG = 4;

template <int IPT, class K, int BLOCK = kSplitBlock>
size_t wg_lds_bytes(int B) {
    constexpr bool SEP = BLOCK > kSplitBlock;  // the key-value kernel stages values after keys
    const size_t per = SEP ? std::max<size_t>(sizeof(K), 4) : sizeof(K) + 4;
    return (size_t)kbuf_word_offset(B, BLOCK / kWave) * 4 + (size_t)BLOCK * IPT * per;
}

template <class K>
struct SplitCall {
    const K* kin;
    K* kout;
    const int32_t* vin;
    int32_t* vout;
    int32_t* d_start;  // start offsets of the (lg, shift) buckets, may be null
    int64_t n;
    int lg, shift;
    uint32_t flip;     // WG strategy only: bucket bits inverted (signed top digit of a radix sort)
};

int64_t wg_tile(edpp_ctx* c, int lg, int S) {
    int64_t subs = std::max<int64_t>(1, ((int64_t)c->tune.split_tile_mult * (1 << lg) + S - 1) / S);
    return subs * S;
}
int64_t wi_colen(edpp_ctx* c, int lg, int ipt) {
    const int64_t R = (int64_t)kWave * ipt;
    return std::max<int64_t>(R, ceil_div((int64_t)c->tune.split_tile_mult * (1 << lg), R) * R);
}
int64_t single_lanes(edpp_ctx* c, int64_t n, int lg) {
    int64_t lanes = std::min<int64_t>((int64_t)c->cus * 64, std::max<int64_t>(1, n / 64));
    while (lanes > 1 && ((int64_t)1 << lg) * lanes >= (1ll << 28)) lanes /= 2;
    return lanes;
}

bool reorder_256(edpp_ctx* c) {
    return c->tune.split_reorder == EDPP_REORDER_NONE || c->tune.split_reorder == EDPP_REORDER_FIXED;
}

// 512-lane scatter workgroups with 8 keys per lane (the 4096-pair sub-tile of
// 256 lanes at 16, with values staged after keys and loaded late) at the
// large-input shape: 8-byte keys from 128 buckets (B=256, 2^27 pairs: 0.99
// against 1.04 ms), 4-byte keys at 64 buckets (a 4096-bucket split's two
// passes: 2.15 against 2.21 ms; at 32 buckets and below 256 lanes are faster)
template <class K>
bool wide_8(edpp_ctx* c, int lg) {
    if (reorder_256(c) || c->tune.split_ipt < 16 || c->tune.split_block != 0) return false;  // default shape only
    if (wg_lds_bytes<8, K, 512>(1 << lg) > 160 * 1024) return false;
    return sizeof(K) == 8 ? lg >= 7 : lg == 6;
}

template <class K>
int wg_ipt_base(edpp_ctx* c, int lg) {
    if (c->tune.split_ipt < 16 || wg_lds_bytes<16, K>(1 << lg) > 160 * 1024) return 8;
    if (c->tune.split_reorder == EDPP_REORDER_FIXED && fixed_lds_bytes<K>(1 << lg, kSplitBlock * 16) > 160 * 1024)
        return 8;
    return 16;
}

// Lanes per scatter workgroup.  512 (sub-tiles of 8192 keys, values staged
// after keys) from 128 buckets on: a bucket's run per sub-tile doubles, so
// fewer stores are partial cache lines; only for 4-byte keys at 16 per lane.
// B=128, 2^28 pairs: 1.12 against 1.22 ms uniform, 1.13 against 1.16 ms
// Zipf.  Up to 64 buckets 256 lanes at 16 keys are as fast as 512 at 16 or
// faster (at 64, wide_8's 512 lanes at 8 keys are faster still).
template <class K>
int wg_block_base(edpp_ctx* c, int lg) {
    if (reorder_256(c)) return kSplitBlock;
    if (sizeof(K) != 4 || wg_ipt_base<K>(c, lg) != 16 || wg_lds_bytes<16, K, 512>(1 << lg) > 160 * 1024)
        return kSplitBlock;
    if (c->tune.split_block != 0) return c->tune.split_block;
    return lg >= 7 ? 512 : kSplitBlock;
}
// the WG strategy's scatter shape (the chained strategy keeps the base one)
template <class K>
int wg_ipt(edpp_ctx* c, int lg) {
    return wide_8<K>(c, lg) ? 8 : wg_ipt_base<K>(c, lg);
}
template <class K>
int wg_block(edpp_ctx* c, int lg) {
    return wide_8<K>(c, lg) ? 512 : wg_block_base<K>(c, lg);
}

// histogram entries one pass needs
template <class K>
int64_t his_len(edpp_ctx* c, int strategy, int64_t n, int lg) {
    const int64_t B = (int64_t)1 << lg;
    if (strategy == EDPP_SPLIT_WG) return B * ceil_div(n, wg_tile(c, lg, wg_block<K>(c, lg) * wg_ipt<K>(c, lg)));
    if (strategy == EDPP_SPLIT_WI) return B * ceil_div(n, wi_colen(c, lg, 16));
    return B * single_lanes(c, n, lg);
}

template <class K, int IPT, bool BALLOT, int BLOCK, int RO = kReorderVaried>
int run_wg(edpp_ctx* c, const SplitCall<K>& a, int32_t* his) {
    const int B = 1 << a.lg;
    constexpr int S = BLOCK * IPT;         // scatter sub-tile
    constexpr int SH = kSplitBlock * IPT;  // histogram sub-tile
    SplitGeom g;
    g.n = a.n;
    g.lg = a.lg;
    g.shift = a.shift;
    g.tile = wg_tile(c, a.lg, S);
    g.tiles = ceil_div(a.n, g.tile);
    g.xcds = c->xcds;
    g.flip = a.flip;
    const int64_t hlen = (int64_t)B * g.tiles;
    // histogram: register counters up to 16 buckets; above, LDS atomics with
    // four histogram sub-tiles per workgroup where they make whole tiles and
    // still leave 32 workgroups per CU, two below that (more workgroups then
    // share the CUs)
    const int regw = a.lg <= kRegCountBits ? reg_count_words(a.lg) : 0;
    const int64_t spt = g.tile / SH;
    const int hsub =
        ceil_div(g.tiles * spt, (int64_t)kHistTilesPerWG) >= 32 * (int64_t)c->cus ? kHistTilesPerWG : 2;
    const bool multi = regw == 0 && spt <= hsub && (spt & (spt - 1)) == 0;
    const int tpw = multi ? hsub / (int)spt : 1;
    const size_t lh = (size_t)tpw * kSplitWaves * B * 4;
    const size_t ls = RO == kReorderFixed ? fixed_lds_bytes<K>(B, S)
                      : RO == kReorderNone ? (size_t)kbuf_word_offset(B) * 4
                                           : wg_lds_bytes<IPT, K, BLOCK>(B);
    EDPP_HIP(c, hipFuncSetAttribute((const void*)k_split_scatter<K, IPT, true, BALLOT, BLOCK, RO>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)ls));
    EDPP_HIP(c, hipFuncSetAttribute((const void*)k_split_scatter<K, IPT, false, BALLOT, BLOCK, RO>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)ls));
    const unsigned hgrid = (unsigned)ceil_div(g.tiles, (int64_t)tpw);
    const int ispt = (int)spt;
    switch (regw) {
        case 1: hipLaunchKernelGGL((k_split_hist<K, IPT, 1, 1>), dim3(hgrid), dim3(kSplitBlock), lh, c->stream, a.kin, g, 1, his); break;
        case 2: hipLaunchKernelGGL((k_split_hist<K, IPT, 1, 2>), dim3(hgrid), dim3(kSplitBlock), lh, c->stream, a.kin, g, 1, his); break;
        case 4: hipLaunchKernelGGL((k_split_hist<K, IPT, 1, 4>), dim3(hgrid), dim3(kSplitBlock), lh, c->stream, a.kin, g, 1, his); break;
        default:
            if (multi && hsub == 2)
                hipLaunchKernelGGL((k_split_hist<K, IPT, 2, 0>), dim3(hgrid), dim3(kSplitBlock), lh,
                                   c->stream, a.kin, g, ispt, his);
            else if (multi)
                hipLaunchKernelGGL((k_split_hist<K, IPT, kHistTilesPerWG, 0>), dim3(hgrid), dim3(kSplitBlock), lh,
                                   c->stream, a.kin, g, ispt, his);
            else
                hipLaunchKernelGGL((k_split_hist<K, IPT, 1, 0>), dim3(hgrid), dim3(kSplitBlock), lh, c->stream, a.kin,
                                   g, 1, his);
    }
    EDPP_HIP(c, hipGetLastError());
    int rc = launch_scan<uint32_t>(c, (const uint32_t*)his, (uint32_t*)his, hlen, EDPP_SCAN_LOOKBACK, nullptr);
    if (rc) return rc;
    if (a.d_start)
        hipLaunchKernelGGL(k_split_start, dim3((unsigned)ceil_div(B, 256)), dim3(256), 0, c->stream,
                           (const int32_t*)his, g.tiles, B, a.d_start);
    if constexpr (RO == kReorderVaried) {
        // buffer stores while every output offset fits the descriptor's 31 bits
        if ((uint64_t)a.n * sizeof(K) < (1ull << 31)) {
            EDPP_HIP(c, hipFuncSetAttribute((const void*)k_split_scatter<K, IPT, true, BALLOT, BLOCK, RO, true>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)ls));
            EDPP_HIP(c, hipFuncSetAttribute((const void*)k_split_scatter<K, IPT, false, BALLOT, BLOCK, RO, true>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)ls));
            if (a.vin)
                hipLaunchKernelGGL((k_split_scatter<K, IPT, true, BALLOT, BLOCK, RO, true>), dim3((unsigned)g.tiles),
                                   dim3(BLOCK), ls, c->stream, a.kin, a.kout, a.vin, a.vout, g, (const int32_t*)his);
            else
                hipLaunchKernelGGL((k_split_scatter<K, IPT, false, BALLOT, BLOCK, RO, true>), dim3((unsigned)g.tiles),
                                   dim3(BLOCK), ls, c->stream, a.kin, a.kout, a.vin, a.vout, g, (const int32_t*)his);
            EDPP_HIP(c, hipGetLastError());
            return EDPP_OK;
        }
    }
    if (a.vin)
        hipLaunchKernelGGL((k_split_scatter<K, IPT, true, BALLOT, BLOCK, RO>), dim3((unsigned)g.tiles), dim3(BLOCK),
                           ls, c->stream, a.kin, a.kout, a.vin, a.vout, g, (const int32_t*)his);
    else
        hipLaunchKernelGGL((k_split_scatter<K, IPT, false, BALLOT, BLOCK, RO>), dim3((unsigned)g.tiles), dim3(BLOCK),
                           ls, c->stream, a.kin, a.kout, a.vin, a.vout, g, (const int32_t*)his);
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

template <class K, int IPT, bool BALLOT>
int run_wi(edpp_ctx* c, const SplitCall<K>& a, int32_t* his) {
    const int B = 1 << a.lg;
    const int64_t colen = wi_colen(c, a.lg, IPT);
    const int64_t cols = ceil_div(a.n, colen);
    const int64_t hlen = (int64_t)B * cols;
    const unsigned grid = (unsigned)ceil_div(cols, kSplitWaves);
    const size_t lds = (size_t)kSplitWaves * B * 4;        // histogram kernel: counts
    const size_t lds_s = (size_t)kSplitWaves * B * 12;     // scatter kernel: masks and positions
    EDPP_HIP(c, hipFuncSetAttribute((const void*)k_split_wave_hist<K, IPT, BALLOT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    EDPP_HIP(c, hipFuncSetAttribute((const void*)k_split_wave_scatter<K, IPT, true, BALLOT>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_s));
    EDPP_HIP(c, hipFuncSetAttribute((const void*)k_split_wave_scatter<K, IPT, false, BALLOT>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_s));
    hipLaunchKernelGGL((k_split_wave_hist<K, IPT, BALLOT>), dim3(grid), dim3(kSplitBlock), lds, c->stream, a.kin, a.n, a.lg,
                       a.shift, cols, colen, his);
    EDPP_HIP(c, hipGetLastError());
    int rc = launch_scan<uint32_t>(c, (const uint32_t*)his, (uint32_t*)his, hlen, EDPP_SCAN_LOOKBACK, nullptr);
    if (rc) return rc;
    if (a.d_start)
        hipLaunchKernelGGL(k_split_start, dim3((unsigned)ceil_div(B, 256)), dim3(256), 0, c->stream,
                           (const int32_t*)his, cols, B, a.d_start);
    if (a.vin)
        hipLaunchKernelGGL((k_split_wave_scatter<K, IPT, true, BALLOT>), dim3(grid), dim3(kSplitBlock), lds_s, c->stream, a.kin,
                           a.kout, a.vin, a.vout, a.n, a.lg, a.shift, cols, colen, (const int32_t*)his);
    else
        hipLaunchKernelGGL((k_split_wave_scatter<K, IPT, false, BALLOT>), dim3(grid), dim3(kSplitBlock), lds_s, c->stream,
                           a.kin, a.kout, a.vin, a.vout, a.n, a.lg, a.shift, cols, colen, (const int32_t*)his);
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

template <class K>
int run_single(edpp_ctx* c, const SplitCall<K>& a, int32_t* his) {
    const int B = 1 << a.lg;
    const int64_t lanes = single_lanes(c, a.n, a.lg);
    const int64_t len = ceil_div(a.n, lanes);
    const int64_t hlen = (int64_t)B * lanes;
    const unsigned grid = (unsigned)ceil_div(lanes, 64);
    hipLaunchKernelGGL((k_split_single_hist<K>), dim3(grid), dim3(64), 0, c->stream, a.kin, a.n, a.lg, a.shift, len,
                       lanes, his);
    EDPP_HIP(c, hipGetLastError());
    int rc = launch_scan<uint32_t>(c, (const uint32_t*)his, (uint32_t*)his, hlen, EDPP_SCAN_LOOKBACK, nullptr);
    if (rc) return rc;
    if (a.d_start)
        hipLaunchKernelGGL(k_split_start, dim3((unsigned)ceil_div(B, 256)), dim3(256), 0, c->stream,
                           (const int32_t*)his, lanes, B, a.d_start);
    if (a.vin)
        hipLaunchKernelGGL((k_split_single_scatter<K, true>), dim3(grid), dim3(64), 0, c->stream, a.kin, a.kout, a.vin,
                           a.vout, a.n, a.lg, a.shift, len, lanes, his);
    else
        hipLaunchKernelGGL((k_split_single_scatter<K, false>), dim3(grid), dim3(64), 0, c->stream, a.kin, a.kout,
                           a.vin, a.vout, a.n, a.lg, a.shift, len, lanes, his);
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

template <class K>
int run_pass(edpp_ctx* c, int strategy, const SplitCall<K>& a, int32_t* his) {
    if constexpr (std::is_same_v<K, MixOnLoad>) {
        // the hash join's partition: the WG strategy with the varied reorder
        (void)strategy;
        const bool bal = a.lg <= kBallotMatchBits;
        if (wg_ipt<K>(c, a.lg) == 16) {
            if (wg_block<K>(c, a.lg) == 512)
                return bal ? run_wg<K, 16, true, 512>(c, a, his) : run_wg<K, 16, false, 512>(c, a, his);
            return bal ? run_wg<K, 16, true, kSplitBlock>(c, a, his) : run_wg<K, 16, false, kSplitBlock>(c, a, his);
        }
        return bal ? run_wg<K, 8, true, kSplitBlock>(c, a, his) : run_wg<K, 8, false, kSplitBlock>(c, a, his);
    } else {
        switch (strategy) {
            case EDPP_SPLIT_WG: {
                const bool bal = a.lg <= kBallotMatchBits;
                const int ro = c->tune.split_reorder;
                if (ro == EDPP_REORDER_NONE || ro == EDPP_REORDER_FIXED) {
                    // 256 lanes (wg_block); wg_ipt also checks that the fixed
                    // variant's staging fits beside 16 keys per lane
                    const bool fx = ro == EDPP_REORDER_FIXED;
                    const bool i16 = wg_ipt<K>(c, a.lg) == 16;
                    if (fx) {
                        if (i16) return bal ? run_wg<K, 16, true, kSplitBlock, kReorderFixed>(c, a, his)
                                            : run_wg<K, 16, false, kSplitBlock, kReorderFixed>(c, a, his);
                        return bal ? run_wg<K, 8, true, kSplitBlock, kReorderFixed>(c, a, his)
                                   : run_wg<K, 8, false, kSplitBlock, kReorderFixed>(c, a, his);
                    }
                    if (i16) return bal ? run_wg<K, 16, true, kSplitBlock, kReorderNone>(c, a, his)
                                        : run_wg<K, 16, false, kSplitBlock, kReorderNone>(c, a, his);
                    return bal ? run_wg<K, 8, true, kSplitBlock, kReorderNone>(c, a, his)
                               : run_wg<K, 8, false, kSplitBlock, kReorderNone>(c, a, his);
                }
                if (wide_8<K>(c, a.lg)) return run_wg<K, 8, false, 512>(c, a, his);  // lg >= 6: no ballots
                if (wg_ipt<K>(c, a.lg) == 16) {
                    if constexpr (sizeof(K) == 4) {
                        if (wg_block<K>(c, a.lg) == 512)
                            return bal ? run_wg<K, 16, true, 512>(c, a, his) : run_wg<K, 16, false, 512>(c, a, his);
                    }
                    return bal ? run_wg<K, 16, true, kSplitBlock>(c, a, his) : run_wg<K, 16, false, kSplitBlock>(c, a, his);
                }
                return bal ? run_wg<K, 8, true, kSplitBlock>(c, a, his) : run_wg<K, 8, false, kSplitBlock>(c, a, his);
            }
            case EDPP_SPLIT_WI:
                return a.lg <= kBallotMatchBits ? run_wi<K, 16, true>(c, a, his) : run_wi<K, 16, false>(c, a, his);
            default: return run_single<K>(c, a, his);
        }
    }
}

// Workspace of the chained strategy's small tables: counts and tickets (one
// fill clears both), then the bases.
constexpr size_t kChainTableWords = (size_t)kChainMaxDigits * kChainMaxRanges * kChainStride;
constexpr size_t kChainSmallBytes =
    ((2 * kChainTableWords + (size_t)kChainMaxDigits * kChainMaxRanges) * 4 + 255) & ~(size_t)255;

template <class K>
struct ChainStep {
    const K* kin;
    K* kout;
    const int32_t* vin;
    int32_t* vout;
    int lg, shift;
    uint32_t flip;
};

template <class K, int IPT, bool BALLOT, int BLOCK>
int run_chain_pass(edpp_ctx* c, const ChainStep<K>& a, int64_t n, const ChainPass& cp) {
    SplitGeom g;
    g.n = n;
    g.lg = a.lg;
    g.shift = a.shift;
    g.tile = (int64_t)BLOCK * IPT;
    g.tiles = cp.cr.tiles;
    g.xcds = c->xcds;
    g.flip = a.flip;
    const size_t ls = wg_lds_bytes<IPT, K, BLOCK>(1 << a.lg);
    if (a.vin) {
        EDPP_HIP(c, hipFuncSetAttribute((const void*)k_chain_scatter<K, IPT, true, BALLOT, BLOCK>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)ls));
        hipLaunchKernelGGL((k_chain_scatter<K, IPT, true, BALLOT, BLOCK>), dim3((unsigned)g.tiles), dim3(BLOCK), ls,
                           c->stream, a.kin, a.kout, a.vin, a.vout, g, cp);
    } else {
        EDPP_HIP(c, hipFuncSetAttribute((const void*)k_chain_scatter<K, IPT, false, BALLOT, BLOCK>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)ls));
        hipLaunchKernelGGL((k_chain_scatter<K, IPT, false, BALLOT, BLOCK>), dim3((unsigned)g.tiles), dim3(BLOCK), ls,
                           c->stream, a.kin, a.kout, a.vin, a.vout, g, cp);
    }
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

// Chained split of n elements in `digits` stable LSD passes (each of at most
// 8 bits; pass p reads steps[p].kin and writes steps[p].kout).  Every pass
// uses the tile shape the WG scatter picks for the widest digit, so that all
// passes share one cut into ranges.  Each pass counts its digit per range of
// its own input (a later pass's ranges hold what the pass before wrote
// there), turns the counts into range bases and runs the chained scatter.
// d_start, when given, receives the first pass's bucket starts.  `small`
// holds kChainSmallBytes of workspace.
template <class K>
int launch_chain(edpp_ctx* c, int64_t n, int digits, const ChainStep<K>* steps, int32_t* d_start, uint32_t* small) {
    int maxlg = 0;
    for (int p = 0; p < digits; ++p) maxlg = std::max(maxlg, steps[p].lg);
    const int ipt = wg_ipt_base<K>(c, maxlg), block = wg_block_base<K>(c, maxlg);
    const int64_t S = (int64_t)block * ipt;
    ChainRanges cr;
    cr.tiles = ceil_div(n, S);
    cr.ranges = (int)std::min<int64_t>(std::min<int64_t>(std::max(c->xcds, 1), kChainMaxRanges), cr.tiles);
    uint32_t* counts = small;
    uint32_t* tickets = small + kChainTableWords;
    uint32_t* base = tickets + (size_t)kChainMaxDigits * kChainMaxRanges;
    int rc = zero_fill(c, counts, (kChainTableWords + (size_t)kChainMaxDigits * kChainMaxRanges) * 4);
    if (rc) return rc;
    uint32_t epoch0;
    rc = chain_state_acquire(c, cr.tiles << maxlg, digits, &epoch0);
    if (rc) return rc;
    // about four counting workgroups per CU, none of them without keys
    constexpr int64_t SH = (int64_t)kSplitBlock * 16;
    const int64_t per_range = ceil_div(cr.tiles, (int64_t)cr.ranges) * S;
    const int wpr = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(4 * (int64_t)std::max(c->cus, 1), cr.ranges),
                                                                 ceil_div(per_range, SH)));
    const size_t lh = (size_t)kSplitWaves * kChainStride * 4;
    for (int p = 0; p < digits && rc == EDPP_OK; ++p) {
        ChainDigits dg = {};
        dg.count = 1;
        dg.lg[0] = steps[p].lg;
        dg.shift[0] = steps[p].shift;
        dg.flip[0] = steps[p].flip;
        uint32_t* cnt_p = counts + (size_t)p * cr.ranges * kChainStride;
        uint32_t* base_p = base + (size_t)p * cr.ranges * kChainStride;
        hipLaunchKernelGGL((k_chain_hist<K, 16>), dim3((unsigned)(cr.ranges * wpr)), dim3(kSplitBlock), lh, c->stream,
                           steps[p].kin, n, S, cr, wpr, dg, cnt_p);
        hipLaunchKernelGGL(k_chain_bases, dim3(1), dim3(kSplitBlock), 0, c->stream, (const uint32_t*)cnt_p, cr.ranges,
                           dg, base_p, p == 0 ? d_start : nullptr);
        EDPP_HIP(c, hipGetLastError());
        ChainPass cp;
        cp.st.status = c->chain_status;
        cp.st.epoch = epoch0 + (uint32_t)p;
        cp.base = base_p;
        cp.tickets = tickets + (size_t)p * cr.ranges;
        cp.cr = cr;
        cp.xcds = std::max(c->xcds, 1);
        cp.spin_limit = c->spin_limit;
        cp.err = c->d_err;
        const bool bal = steps[p].lg <= kBallotMatchBits;
        if (ipt == 16) {
            if constexpr (sizeof(K) == 4) {
                if (block == 512) {
                    rc = bal ? run_chain_pass<K, 16, true, 512>(c, steps[p], n, cp)
                             : run_chain_pass<K, 16, false, 512>(c, steps[p], n, cp);
                    continue;
                }
            }
            rc = bal ? run_chain_pass<K, 16, true, kSplitBlock>(c, steps[p], n, cp)
                     : run_chain_pass<K, 16, false, kSplitBlock>(c, steps[p], n, cp);
        } else {
            rc = bal ? run_chain_pass<K, 8, true, kSplitBlock>(c, steps[p], n, cp)
                     : run_chain_pass<K, 8, false, kSplitBlock>(c, steps[p], n, cp);
        }
    }
    return rc;
}

}  // namespace

// Bucket counts above kMaxSplitBitsPerPass bits are split as two stable LSD
// passes (low bits first), whose runs per bucket stay long enough for
// coalesced stores; the composition is the stable split by the full bucket.
constexpr int kMaxSplitBitsPerPass = 8;

template <class K>
int launch_split(edpp_ctx* c, const K* kin, K* kout, const int32_t* vin, int32_t* vout, int32_t* d_start, int64_t n,
                 int buckets, int strategy) {
    int lg = 0;
    while ((1 << lg) < buckets) ++lg;
    if (n == 0) {
        // a fill kernel, not a memset node, so the call stays graph-capturable
        if (d_start)
            hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)ceil_div(buckets, 256)), dim3(256), 0, c->stream,
                               (uint32_t*)d_start, (int64_t)buckets, 0u);
        EDPP_HIP(c, hipGetLastError());
        return EDPP_OK;
    }
    if (strategy == EDPP_SPLIT_AUTO) {
        const edpp_tuning saved = c->tune;
        default_split_shape(c, n, buckets, (int)sizeof(K), &c->tune.split_ipt, &c->tune.split_tile_mult);
        int rc = launch_split<K>(c, kin, kout, vin, vout, d_start, n, buckets, default_split_strategy(c, n, buckets));
        c->tune = saved;
        return rc;
    }
    if (strategy < EDPP_SPLIT_WG || strategy > EDPP_SPLIT_CHAINED)
        return fail(c, EDPP_EINVAL, "unknown split strategy %d", strategy);
    const bool two = strategy != EDPP_SPLIT_SINGLE && lg > kMaxSplitBitsPerPass;
    const int lg1 = two ? lg / 2 : lg, lg2 = lg - lg1;
    if constexpr (std::is_same_v<K, MixOnLoad>) {
        if (strategy != EDPP_SPLIT_WG) return fail(c, EDPP_EINVAL, "mix-on-load splits run the WG strategy");
    } else if (strategy == EDPP_SPLIT_CHAINED) {
        // both digits of a two-pass split are counted by the one up-front read
        const size_t kbytes = two ? (((size_t)n * sizeof(K) + 255) & ~(size_t)255) : 0;
        const size_t vbytes = (two && vin) ? (((size_t)n * 4 + 255) & ~(size_t)255) : 0;
        void* w;
        int rc = ws_reserve(c, kChainSmallBytes + kbytes + vbytes, &w);
        if (rc) return rc;
        char* base = static_cast<char*>(w);
        uint32_t* small = reinterpret_cast<uint32_t*>(base);
        if (!two) {
            const ChainStep<K> st = {kin, kout, vin, vout, lg, 0, 0u};
            return launch_chain<K>(c, n, 1, &st, d_start, small);
        }
        K* tk = reinterpret_cast<K*>(base + kChainSmallBytes);
        int32_t* tv = vin ? reinterpret_cast<int32_t*>(base + kChainSmallBytes + kbytes) : nullptr;
        const ChainStep<K> st[2] = {{kin, tk, vin, tv, lg1, 0, 0u}, {tk, kout, tv, vout, lg2, lg1, 0u}};
        rc = launch_chain<K>(c, n, 2, st, nullptr, small);
        if (rc) return rc;
        if (d_start) {
            launch_start_search<K>(c, kout, n, lg, d_start);
            EDPP_HIP(c, hipGetLastError());
        }
        return EDPP_OK;
    }
    int64_t hlen = his_len<K>(c, strategy, n, lg1);
    if (two) hlen = std::max(hlen, his_len<K>(c, strategy, n, lg2));
    if (hlen >= (1ll << 31)) return fail(c, EDPP_EUNSUP, "histogram of %lld entries too large", (long long)hlen);
    const size_t hbytes = ((size_t)hlen * 4 + 255) & ~(size_t)255;
    const size_t kbytes = two ? (((size_t)n * sizeof(K) + 255) & ~(size_t)255) : 0;
    const size_t vbytes = (two && vin) ? (((size_t)n * 4 + 255) & ~(size_t)255) : 0;
    void* w;
    int rc = ws_reserve(c, hbytes + kbytes + vbytes, &w);
    if (rc) return rc;
    char* base = static_cast<char*>(w);
    int32_t* his = reinterpret_cast<int32_t*>(base);
    if (!two) {
        SplitCall<K> a = {kin, kout, vin, vout, d_start, n, lg, 0, 0u};
        return run_pass<K>(c, strategy, a, his);
    }
    K* tk = reinterpret_cast<K*>(base + hbytes);
    int32_t* tv = vin ? reinterpret_cast<int32_t*>(base + hbytes + kbytes) : nullptr;
    SplitCall<K> a1 = {kin, tk, vin, tv, nullptr, n, lg1, 0, 0u};
    rc = run_pass<K>(c, strategy, a1, his);
    if (rc) return rc;
    // the second pass reads what the first wrote: for MixOnLoad keys, the
    // mixed keys themselves, split as plain 32-bit keys
    using K2 = std::conditional_t<std::is_same_v<K, MixOnLoad>, uint32_t, K>;
    K2* kout2 = reinterpret_cast<K2*>(kout);
    SplitCall<K2> a2 = {reinterpret_cast<const K2*>(tk), kout2, tv, vout, nullptr, n, lg2, lg1, 0u};
    rc = run_pass<K2>(c, strategy, a2, his);
    if (rc) return rc;
    if (d_start) {
        launch_start_search<K2>(c, kout2, n, lg, d_start);
        EDPP_HIP(c, hipGetLastError());
    }
    return EDPP_OK;
}

// Stable LSD radix sort of 32-bit keys (values follow) by key bits
// [begin_bit, end_bit): one stable WG split per digit of at most 8 bits,
// lowest digit first, as the reference's radix sort counts, scans and
// scatters one digit per pass (obsolete/CUDA/kernels/radixSort.cu:29-200).
// Signed keys order as int32: the digit holding bit 31 has its top bit
// flipped.  Passes alternate between the output and workspace copies, so the
// last one lands in the output; keys_in may equal keys_out.
int launch_sort_u32(edpp_ctx* c, const uint32_t* kin, uint32_t* kout, const int32_t* vin, int32_t* vout, int64_t n,
                    int begin_bit, int end_bit, bool is_signed) {
    if (n == 0) return EDPP_OK;
    if (begin_bit == end_bit) {  // no key bits: the input order is the sorted order
        if (kout != kin) EDPP_HIP(c, hipMemcpyAsync(kout, kin, (size_t)n * 4, hipMemcpyDeviceToDevice, c->stream));
        if (vin && vout != vin)
            EDPP_HIP(c, hipMemcpyAsync(vout, vin, (size_t)n * 4, hipMemcpyDeviceToDevice, c->stream));
        return EDPP_OK;
    }
    const int bits = end_bit - begin_bit;
    const int passes = (bits + kMaxSplitBitsPerPass - 1) / kMaxSplitBitsPerPass;
    const int per = (bits + passes - 1) / passes;  // balanced digits of at most 8 bits
    const edpp_tuning saved = c->tune;
    default_split_shape(c, n, 1 << per, 4, &c->tune.split_ipt, &c->tune.split_tile_mult);
    int64_t hlen = 0;
    for (int p = 0; p < passes; ++p) {
        const int lo = begin_bit + p * per, hi = std::min(end_bit, lo + per);
        hlen = std::max(hlen, his_len<uint32_t>(c, EDPP_SPLIT_WG, n, hi - lo));
    }
    // destinations from the last pass back: output, copy A, output, copy A, ...
    // An in-place call whose first pass would write the output sends that pass
    // to copy B instead; with a single pass it goes to copy A, which is then
    // copied to the output.
    const uint32_t* dk[8];
    const int32_t* dv[8];
    const bool alias = kin == kout || (vin && vin == vout);
    const bool first_out = (passes % 2) == 1;  // the first pass writes the output
    const bool need_a = passes > 1 || (alias && first_out);
    const bool need_b = alias && first_out && passes > 1;
    const bool chained = c->tune.sort_strategy == EDPP_SPLIT_CHAINED;
    const size_t hbytes = chained ? kChainSmallBytes : ((size_t)hlen * 4 + 255) & ~(size_t)255;
    const size_t kb = ((size_t)n * 4 + 255) & ~(size_t)255;
    const size_t vb = vin ? kb : 0;
    const size_t copies = (need_a ? 1 : 0) + (need_b ? 1 : 0);
    void* w;
    int rc = ws_reserve(c, hbytes + copies * (kb + vb), &w);
    if (rc) {
        c->tune = saved;
        return rc;
    }
    char* base = static_cast<char*>(w);
    int32_t* his = reinterpret_cast<int32_t*>(base);
    uint32_t* ak = need_a ? reinterpret_cast<uint32_t*>(base + hbytes) : nullptr;
    int32_t* av = (need_a && vin) ? reinterpret_cast<int32_t*>(base + hbytes + kb) : nullptr;
    uint32_t* bk = need_b ? reinterpret_cast<uint32_t*>(base + hbytes + kb + vb) : nullptr;
    int32_t* bv = (need_b && vin) ? reinterpret_cast<int32_t*>(base + hbytes + kb + vb + kb) : nullptr;
    {
        bool out = true;
        for (int p = passes - 1; p >= 0; --p, out = !out) {
            dk[p] = out ? kout : ak;
            dv[p] = out ? vout : av;
        }
        if (alias && first_out) {
            dk[0] = passes > 1 ? bk : ak;
            dv[0] = passes > 1 ? bv : av;
        }
    }
    const uint32_t* sk = kin;
    const int32_t* sv = vin;
    if (chained) {
        // every digit counted in one read of the keys, then one chained scatter per digit
        ChainStep<uint32_t> st[kChainMaxDigits];
        for (int p = 0; p < passes; ++p) {
            const int lo = begin_bit + p * per, hi = std::min(end_bit, lo + per);
            st[p] = {sk, const_cast<uint32_t*>(dk[p]), sv, const_cast<int32_t*>(dv[p]), hi - lo, lo,
                     (is_signed && hi == 32) ? (1u << (hi - lo - 1)) : 0u};
            sk = dk[p];
            sv = dv[p];
        }
        rc = launch_chain<uint32_t>(c, n, passes, st, nullptr, reinterpret_cast<uint32_t*>(his));
    }
    for (int p = 0; p < passes && rc == EDPP_OK && !chained; ++p) {
        const int lo = begin_bit + p * per, hi = std::min(end_bit, lo + per), lg = hi - lo;
        const uint32_t flip = (is_signed && hi == 32) ? (1u << (lg - 1)) : 0u;
        SplitCall<uint32_t> a = {sk, const_cast<uint32_t*>(dk[p]), sv, const_cast<int32_t*>(dv[p]), nullptr, n, lg, lo,
                                 flip};
        rc = run_pass<uint32_t>(c, EDPP_SPLIT_WG, a, his);
        sk = dk[p];
        sv = dv[p];
    }
    c->tune = saved;
    if (rc == EDPP_OK && sk != kout) {  // single in-place pass: copy A to the output
        EDPP_HIP(c, hipMemcpyAsync(kout, sk, (size_t)n * 4, hipMemcpyDeviceToDevice, c->stream));
        if (vin) EDPP_HIP(c, hipMemcpyAsync(vout, sv, (size_t)n * 4, hipMemcpyDeviceToDevice, c->stream));
    }
    return rc;
}

template int launch_split<MixOnLoad>(edpp_ctx*, const MixOnLoad*, MixOnLoad*, const int32_t*, int32_t*, int32_t*,
                                     int64_t, int, int);
template int launch_split<uint32_t>(edpp_ctx*, const uint32_t*, uint32_t*, const int32_t*, int32_t*, int32_t*, int64_t,
                                    int, int);
template int launch_split<uint64_t>(edpp_ctx*, const uint64_t*, uint64_t*, const int32_t*, int32_t*, int32_t*,
                                    int64_t, int, int);

}  // namespace edpp
