// batch.hip — persistent executor for many scans, splits and gathers of
// uneven length in one launch, fed by a device ticket queue.
//
// Queue order (tickets):  [split histogram tiles of every split job of
//                          more than one tile]
//                         [tiles of the in-place scans of those histograms]
//                         [scan tiles and gather tiles, in job order]
//                         [split scatter tiles]
// (the PHASES order; the JOBS order puts each split's scatter tiles at its
// job's place among the scan and gather tiles).  A split job of one tile has
// only its scatter tile, which takes its offsets from its own counts.
// Every dependency points to smaller tickets: a histogram scan waits for its
// job's histogram tiles, a scatter tile for its job's histogram scan, a scan
// tile (and a histogram-scan tile) for its predecessors in the look-back.
// Tickets are handed out in order, so whatever a workgroup waits for is held
// by a workgroup that is already running; every wait ends.  In the XCD order the
// jobs are dealt to one queue per XCD (each queue in the JOBS order over its
// own jobs, with its own ticket counter); workgroup g starts on queue
// g % xcds, the queue of the XCD it runs on, so a job's tiles share one L2,
// and moves to the next queue when its own is used up.  Every job lies in one
// queue, so every wait still points to a smaller ticket of a queue that hands
// out its tickets in order.  The reference has no counterpart (its primitives
// are one synchronous call each).
#include <algorithm>
#include <vector>
#include "ctx.hpp"
#include "edpp_batch.h"
#include "split_tile.hpp"
#include "tile_scan.hpp"

namespace edpp {

namespace {

constexpr int kPhaseHist = 0, kPhaseHistScan = 1, kPhaseMain = 2, kPhaseScatter = 3;

struct DJob {
    int32_t kind, wide, lg, pad;
    int64_t n, tiles, hscan_tiles;
    const void* in;
    void* out;
    const int32_t* idx;
    const int32_t* vin;
    int32_t* vout;
    void* aux;
    uint64_t* status;  // u32 scan: 1 word per tile; u64 scan: 2 per tile; split: 1 per histogram-scan tile
    uint64_t* his;     // split: [B][tiles] tagged words {tag:32 | value:32}
};

constexpr int kMaxQueues = 8;
constexpr int kTicketStride = 32;  // 32-bit words between the queues' ticket counters (one 128-byte line each)

// One 64-bit word per ticket names its tile: job (bits 34..63), phase (32..33)
// and the tile's index within the job and phase (0..31), so a workgroup finds
// its work with one load.
__host__ __device__ constexpr uint64_t pack_ticket(uint32_t job, uint32_t phase, uint32_t tile) {
    return ((uint64_t)job << 34) | ((uint64_t)phase << 32) | tile;
}
constexpr int64_t kMaxBatchTiles = 1ll << 27;  // a 1 GiB ticket table

// The queues of a launch: queue q's tickets are entries [tix_first[q],
// tix_first[q] + total[q]) of the ticket table.
struct QTable {
    int64_t total[kMaxQueues];
    int64_t tix_first[kMaxQueues];
    int32_t nq, xcds;
};

// Tile shapes of the persistent launch: lanes per workgroup, 16-byte vectors
// per lane of a 32-bit and of a 64-bit scan tile, keys per lane of a split
// tile, 16-byte index vectors per lane of a gather tile.  The autotuner
// (edpp_batch_autotune) times every shape at several occupancies.
struct BatchShape {
    int block, scan_vpt32, scan_vpt64, split_ipt, gather_u;
};
constexpr BatchShape kShapes[] = {
    {256, 8, 4, 8, 4},   // 0: scan 8192 / 2048, split 2048, gather 4096
    {256, 4, 2, 4, 2},   // 1: half-size tiles: more tiles for small jobs
    {256, 8, 4, 8, 8},   // 2: gather tiles of 8192
    {512, 4, 2, 8, 4},   // 3: 8 waves, scan 8192 / 2048, split 4096, gather 8192
    {512, 8, 4, 4, 2},   // 4: 8 waves, scan 16384 / 4096, split 2048, gather 4096
    {256, 16, 8, 8, 8},  // 5: scan 16384 / 4096, split 2048, gather 8192 (the default)
};
constexpr int kNumShapes = (int)(sizeof(kShapes) / sizeof(kShapes[0]));
constexpr int kDefaultShape = 5;  // fastest on the 10 000-job mix (1.83 against 1.92 ms for shape 0)
constexpr int kOccupancies[] = {2, 4, 8};
constexpr int kNumOcc = 3;
constexpr int kNumOrders = EDPP_BATCH_ORDER_XCD + 1;

__host__ __device__ inline int64_t tile_elems(const BatchShape& s, int kind, int wide) {
    if (kind == EDPP_JOB_SCAN) return (int64_t)s.block * (wide ? 2 * s.scan_vpt64 : 4 * s.scan_vpt32);
    if (kind == EDPP_JOB_SPLIT) return (int64_t)s.block * s.split_ipt;
    return (int64_t)s.block * 4 * s.gather_u;
}
inline int64_t hscan_tile_elems(const BatchShape& s) { return (int64_t)s.block * 16; }  // TileScan<u32,BLOCK,4>

// Cross-workgroup data of one launch moves only in self-describing words:
// look-back status words {flag | value} for the scans, and for the splits
// histogram words {tag | value} whose tag says what the value is (a tile's
// bucket count, then its scanned offset).  A reader polls each word it needs
// with relaxed agent-scope atomic loads until the word shows the tag it
// waits for, so it never infers one location's contents from another's and
// needs no fence; the words are zeroed by the init kernel of every run.
enum : uint32_t { kTagCount = 1, kTagScanned = 2 };

__device__ __forceinline__ uint64_t tagged(uint32_t tag, uint32_t v) { return ((uint64_t)tag << 32) | v; }

// polls *p until its tag is `tag`; on watchdog expiry sets *err and returns
// what it read (the run then reports EDPP_ETIMEOUT; consumers stay in bounds)
__device__ __forceinline__ uint32_t wait_tagged(const uint64_t* p, uint32_t tag, uint32_t spin_limit, int32_t* err) {
    uint64_t w = ld_relaxed(p);
    uint32_t spins = 0;
    while ((uint32_t)(w >> 32) != tag) {
        if (++spins > spin_limit) {
            st_relaxed(err, (int32_t)EDPP_ETIMEOUT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        w = ld_relaxed(p);
    }
    return (uint32_t)w;
}

// ---------------------------------------------------------------- scan tiles
template <class T, int BLOCK, int VPT>
__device__ void scan_tile(const T* in, T* out, int64_t n, int64_t t, typename LookbackFor<T>::type st, T* total,
                          char* smem, uint32_t spin_limit, int32_t* err) {
    typedef TileScan<T, BLOCK, VPT> TS;
    typename TS::Smem& sm = *reinterpret_cast<typename TS::Smem*>(smem);
    const int64_t base = t * TS::TILE;
    TS ts;
    ts.load(in, base, n);
    ts.scan(sm);
    if (threadIdx.x < kWave) {
        T prefix = 0;
        if (t == 0) {
            if (threadIdx.x == 0) st.publish(0, kFlagPrefix, ts.total);
        } else {
            if (threadIdx.x == 0) st.publish(t, kFlagAggregate, ts.total);
            prefix = lookback<T>(st, t, spin_limit, err);
            if (threadIdx.x == 0) st.publish(t, kFlagPrefix, prefix + ts.total);
        }
        if (threadIdx.x == 0) {
            sm.prefix = prefix;
            if (total && base + TS::TILE >= n) *total = prefix + ts.total;
        }
    }
    __syncthreads();
    ts.store(out, base, n, sm.prefix);
}

// In-place scan of a split job's histogram words: tile t of the bucket-major
// matrix waits for the count words of its entries, scans them with the
// look-back over the job's histogram-scan tiles, and rewrites them as
// scanned words.  Only this tile reads or writes its entries' count words.
template <int BLOCK>
__device__ void hist_scan_tile(const DJob& J, int64_t t, char* smem, uint32_t spin_limit, int32_t* err) {
    typedef TileScan<uint32_t, BLOCK, 4> TS;
    typename TS::Smem& sm = *reinterpret_cast<typename TS::Smem*>(smem);
    const int64_t hlen = ((int64_t)1 << J.lg) * J.tiles;
    const int64_t base = t * TS::TILE;
    TS ts;
    // all 16 words are read at once; only those whose count has not arrived
    // yet are polled again
    uint64_t w[4][TS::EPV];
    #pragma unroll
    for (int j = 0; j < 4; ++j)
        #pragma unroll
        for (int k = 0; k < TS::EPV; ++k) {
            const int64_t e = base + (int64_t)j * TS::SUB + (int64_t)threadIdx.x * TS::EPV + k;
            w[j][k] = e < hlen ? ld_relaxed(J.his + e) : tagged(kTagCount, 0u);
        }
    #pragma unroll
    for (int j = 0; j < 4; ++j)
        #pragma unroll
        for (int k = 0; k < TS::EPV; ++k) {
            const int64_t e = base + (int64_t)j * TS::SUB + (int64_t)threadIdx.x * TS::EPV + k;
            ts.v[j][k] = (uint32_t)(w[j][k] >> 32) == kTagCount ? (uint32_t)w[j][k]
                                                                 : wait_tagged(J.his + e, kTagCount, spin_limit, err);
        }
    ts.scan(sm);
    LookbackState32 st;
    st.status = J.status;
    st.epoch = 1;
    if (threadIdx.x < kWave) {
        uint32_t prefix = 0;
        if (t == 0) {
            if (threadIdx.x == 0) st.publish(0, kFlagPrefix, ts.total);
        } else {
            if (threadIdx.x == 0) st.publish(t, kFlagAggregate, ts.total);
            prefix = lookback<uint32_t>(st, t, spin_limit, err);
            if (threadIdx.x == 0) st.publish(t, kFlagPrefix, prefix + ts.total);
        }
        if (threadIdx.x == 0) sm.prefix = prefix;
    }
    __syncthreads();
    #pragma unroll
    for (int j = 0; j < 4; ++j) {
        uint32_t x = sm.prefix + ts.off[j];
        #pragma unroll
        for (int k = 0; k < TS::EPV; ++k) {
            const int64_t e = base + (int64_t)j * TS::SUB + (int64_t)threadIdx.x * TS::EPV + k;
            if (e < hlen) st_relaxed(J.his + e, tagged(kTagScanned, x));
            x += ts.v[j][k];
        }
    }
}

// ---------------------------------------------------------------- split tiles
template <class K, int IPT, int BLOCK>
__device__ void split_hist_tile(const DJob& J, int64_t t, uint32_t* lds) {
    constexpr int W = BLOCK / kWave;
    const int B = 1 << J.lg;
    const int w = threadIdx.x / kWave;
    constexpr int S = BLOCK * IPT;
    uint32_t* whist = lds;
    for (int b = threadIdx.x; b < W * B; b += BLOCK) whist[b] = 0;
    __syncthreads();
    const int64_t base = t * S, end = std::min<int64_t>(base + S, J.n);
    if (J.lg <= kBallotMatchBits) count_subtile<K, IPT, true>((const K*)J.in, base, end, J.lg, 0, whist + w * B);
    else count_subtile<K, IPT, false>((const K*)J.in, base, end, J.lg, 0, whist + w * B);
    __syncthreads();
    for (int b = threadIdx.x; b < B; b += BLOCK) {
        uint32_t a = 0;
        for (int q = 0; q < W; ++q) a += whist[q * B + b];
        st_relaxed(J.his + (int64_t)b * J.tiles + t, tagged(kTagCount, a));
    }
}

// place_and_store hook of a split job of one tile: its output offsets are the
// tile's own exclusive bucket prefix (off[]), so the job needs no histogram
// or histogram-scan tile.  Thread b sets run[b] (and the job's bucket start)
// and then reads it itself.
struct SelfOffsets {
    static constexpr bool kSync = false;
    const uint32_t* off;
    uint32_t* run;
    int32_t* aux;
    int B;
    __device__ __forceinline__ void publish() const {}
    __device__ __forceinline__ void keys_placed() const {}
    __device__ __forceinline__ void resolve() const {
        for (int b = threadIdx.x; b < B; b += blockDim.x) {
            run[b] = off[b];
            if (aux) aux[b] = (int32_t)off[b];
        }
    }
};

template <class K, int IPT, int BLOCK>
__device__ void split_scatter_tile(const DJob& J, int64_t t, uint32_t* lds, uint32_t spin_limit, int32_t* err) {
    constexpr int W = BLOCK / kWave;
    const int B = 1 << J.lg;
    constexpr int S = BLOCK * IPT;
    uint64_t* wmask = reinterpret_cast<uint64_t*>(lds);
    uint32_t* whist = lds + whist_word_offset(B, W);
    uint32_t* run = whist + W * B;
    uint32_t* sub = run + B;
    uint32_t* off = sub + B;
    uint32_t* wtmp = off + B;
    K* kbuf = reinterpret_cast<K*>(lds + kbuf_word_offset(B, W));
    int32_t* vbuf = reinterpret_cast<int32_t*>(kbuf + S);
    const int w = threadIdx.x / kWave;
    const bool kv = J.vin != nullptr;
    const bool single = J.tiles == 1;
    if (!single) {
        for (int b = threadIdx.x; b < B; b += BLOCK) {
            const uint32_t r = wait_tagged(J.his + (int64_t)b * J.tiles + t, kTagScanned, spin_limit, err);
            run[b] = r;
            if (t == 0 && J.aux) static_cast<int32_t*>(J.aux)[b] = (int32_t)r;
        }
    }
    for (int b = threadIdx.x; b < W * B; b += BLOCK) {
        whist[b] = 0;
        wmask[b] = 0ull;
    }
    __syncthreads();
    const int64_t s = t * S, tend = std::min<int64_t>(s + S, J.n);
    uint32_t bk[IPT], rk[IPT];
    K key[IPT];
    int32_t val[IPT];
    if (J.lg <= kBallotMatchBits)
        rank_subtile<K, IPT, true>((const K*)J.in, J.vin, s, tend, J.lg, 0, whist + w * B, wmask + w * B, bk, rk, key,
                                   val);
    else
        rank_subtile<K, IPT, false>((const K*)J.in, J.vin, s, tend, J.lg, 0, whist + w * B, wmask + w * B, bk, rk,
                                    key, val);
    __syncthreads();
    const int cnt = (int)(tend - s);
    if (single) {
        const SelfOffsets hook{off, run, static_cast<int32_t*>(J.aux), B};
        if (kv)
            place_and_store<K, IPT, true, BLOCK, false, SelfOffsets>(whist, run, sub, off, wtmp, kbuf, vbuf, bk, rk,
                                                                     key, val, B, 0, cnt, static_cast<K*>(J.out),
                                                                     J.vout, J.n, 0u, hook);
        else
            place_and_store<K, IPT, false, BLOCK, false, SelfOffsets>(whist, run, sub, off, wtmp, kbuf, vbuf, bk,
                                                                      rk, key, val, B, 0, cnt, static_cast<K*>(J.out),
                                                                      nullptr, J.n, 0u, hook);
    } else if (kv) {
        place_and_store<K, IPT, true, BLOCK>(whist, run, sub, off, wtmp, kbuf, vbuf, bk, rk, key, val, B, 0, cnt,
                                             static_cast<K*>(J.out), J.vout, J.n);
    } else {
        place_and_store<K, IPT, false, BLOCK>(whist, run, sub, off, wtmp, kbuf, vbuf, bk, rk, key, val, B, 0, cnt,
                                              static_cast<K*>(J.out), nullptr, J.n);
    }
}

// ---------------------------------------------------------------- gather tiles
// A tile of BLOCK*4*U elements: U rounds of one 16-byte index vector per lane
template <class T, int BLOCK, int U>
__device__ void gather_tile(const DJob& J, int64_t t) {
    const T* in = static_cast<const T*>(J.in);
    T* out = static_cast<T*>(J.out);
    const uint32_t n = (uint32_t)J.n;
    const int64_t base = t * (int64_t)BLOCK * 4 * U;
    if (base + (int64_t)BLOCK * 4 * U <= J.n) {
        // a whole tile: every index vector, then every random read, then the
        // stores, so that each group of loads is in flight at once; four
        // valid indices leave as one 16-byte store per 16 bytes of output
        // (batch buffers are 16-byte aligned)
        v4u32 x[U];
        #pragma unroll
        for (int u = 0; u < U; ++u)
            x[u] = ld_stream(reinterpret_cast<const v4u32*>(J.idx + base + (int64_t)u * BLOCK * 4 + threadIdx.x * 4));
        __builtin_amdgcn_sched_barrier(0);
        T v[U][4];
        #pragma unroll
        for (int u = 0; u < U; ++u)
            #pragma unroll
            for (int k = 0; k < 4; ++k) v[u][k] = (uint32_t)x[u][k] < n ? in[(uint32_t)x[u][k]] : T(0);
        __builtin_amdgcn_sched_barrier(0);
        #pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t e0 = base + (int64_t)u * BLOCK * 4 + threadIdx.x * 4;
            const bool all = (uint32_t)x[u][0] < n && (uint32_t)x[u][1] < n && (uint32_t)x[u][2] < n &&
                             (uint32_t)x[u][3] < n;
            if (all) {
                typedef typename vec16<T>::type V;
                constexpr int EPV = vec16<T>::n;
                #pragma unroll
                for (int h = 0; h < 4 / EPV; ++h) {
                    V o;
                    #pragma unroll
                    for (int k = 0; k < EPV; ++k) o[k] = v[u][h * EPV + k];
                    st_stream(o, reinterpret_cast<V*>(out + e0 + h * EPV));
                }
            } else {
                #pragma unroll
                for (int k = 0; k < 4; ++k)
                    if ((uint32_t)x[u][k] < n) out[e0 + k] = v[u][k];
            }
        }
        return;
    }
    #pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t e0 = base + (int64_t)u * BLOCK * 4 + threadIdx.x * 4;
        if (e0 + 4 <= J.n) {
            v4u32 x = ld_stream(reinterpret_cast<const v4u32*>(J.idx + e0));
            T v[4];
            #pragma unroll
            for (int k = 0; k < 4; ++k) v[k] = (uint32_t)x[k] < n ? in[(uint32_t)x[k]] : T(0);
            #pragma unroll
            for (int k = 0; k < 4; ++k)
                if ((uint32_t)x[k] < n) out[e0 + k] = v[k];
        } else {
            for (int64_t e = e0; e < J.n && e < e0 + 4; ++e) {
                uint32_t p = (uint32_t)J.idx[e];
                if (p < n) out[e] = in[p];
            }
        }
    }
}

// ---------------------------------------------------------------- kernels
__global__ void k_batch_init(uint32_t* state, int64_t words, const DJob* jobs, int32_t njobs) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = tid; i < words; i += stride) state[i] = 0u;
    // outputs of empty jobs: scan total 0, split starts 0
    for (int64_t j = tid; j < njobs; j += stride) {
        const DJob& J = jobs[j];
        if (J.n != 0 || !J.aux) continue;
        if (J.kind == EDPP_JOB_SCAN) {
            if (J.wide) *static_cast<uint64_t*>(J.aux) = 0;
            else *static_cast<uint32_t*>(J.aux) = 0;
        } else if (J.kind == EDPP_JOB_SPLIT) {
            for (int b = 0; b < (1 << J.lg); ++b) static_cast<int32_t*>(J.aux)[b] = 0;
        }
    }
}

template <int BLOCK, int SV32, int SV64, int IPT, int GU>
__global__ __launch_bounds__(BLOCK) void k_batch(const DJob* __restrict__ jobs, const uint64_t* __restrict__ tix,
                                                 QTable qt, uint32_t* tickets, uint32_t spin_limit, int32_t* err) {
    extern __shared__ __align__(16) char smem[];
    __shared__ uint32_t s_ticket;
    int q = qt.nq == 1 ? 0 : (int)(blockIdx.x % (unsigned)qt.xcds) % qt.nq;  // this XCD's queue
    int spent = 0;  // queues found used up
    while (true) {
        if (threadIdx.x == 0) s_ticket = fetch_add(tickets + q * kTicketStride, 1u);
        __syncthreads();
        const int64_t tk = (uint32_t)__builtin_amdgcn_readfirstlane(s_ticket);
        if (tk >= qt.total[q]) {
            __syncthreads();  // every lane has read s_ticket before lane 0 writes it again
            if (++spent >= qt.nq) break;
            q = q + 1 == qt.nq ? 0 : q + 1;
            continue;
        }
        // the same word in every lane: made uniform so that the job's fields
        // come through scalar loads
        const uint64_t ev = tix[qt.tix_first[q] + tk];
        const uint64_t e = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(ev >> 32)) << 32) |
                           (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)ev);
        const int32_t phase = (int32_t)((e >> 32) & 3u);
        const DJob& J = jobs[e >> 34];
        const int64_t t = (int64_t)(uint32_t)e;
        uint32_t* lds = reinterpret_cast<uint32_t*>(smem);
        if (phase == kPhaseHist) {
            if (J.wide) split_hist_tile<uint64_t, IPT, BLOCK>(J, t, lds);
            else split_hist_tile<uint32_t, IPT, BLOCK>(J, t, lds);
        } else if (phase == kPhaseHistScan) {
            hist_scan_tile<BLOCK>(J, t, smem, spin_limit, err);
        } else if (phase == kPhaseScatter) {
            if (J.wide) split_scatter_tile<uint64_t, IPT, BLOCK>(J, t, lds, spin_limit, err);
            else split_scatter_tile<uint32_t, IPT, BLOCK>(J, t, lds, spin_limit, err);
        } else if (J.kind == EDPP_JOB_SCAN) {
            if (J.wide) {
                LookbackState64 st;
                st.status = J.status;
                st.epoch = 1;
                scan_tile<uint64_t, BLOCK, SV64>((const uint64_t*)J.in, (uint64_t*)J.out, J.n, t, st, (uint64_t*)J.aux,
                                                 smem, spin_limit, err);
            } else {
                LookbackState32 st;
                st.status = J.status;
                st.epoch = 1;
                scan_tile<uint32_t, BLOCK, SV32>((const uint32_t*)J.in, (uint32_t*)J.out, J.n, t, st, (uint32_t*)J.aux,
                                                 smem, spin_limit, err);
            }
        } else {
            if (J.wide) gather_tile<uint64_t, BLOCK, GU>(J, t);
            else gather_tile<uint32_t, BLOCK, GU>(J, t);
        }
        __syncthreads();  // LDS of this tile is dead before the next one starts
    }
}

// LDS of the largest tile of a shape: the split scatter with 64-bit keys at
// 256 buckets (ranking masks, counts and offsets, then the keys and values of
// one tile), or a scan tile's scratch
constexpr size_t shape_lds(const BatchShape& s) {
    return std::max<size_t>((size_t)kbuf_word_offset(256, s.block / kWave) * 4 +
                                (size_t)s.block * s.split_ipt * (sizeof(uint64_t) + sizeof(int32_t)),
                            1024);
}

typedef void (*BatchKernel)(const DJob*, const uint64_t*, QTable, uint32_t*, uint32_t, int32_t*);

template <int I>
constexpr BatchKernel kernel_of() {
    return &k_batch<kShapes[I].block, kShapes[I].scan_vpt32, kShapes[I].scan_vpt64, kShapes[I].split_ipt,
                    kShapes[I].gather_u>;
}
const BatchKernel kKernels[kNumShapes] = {kernel_of<0>(), kernel_of<1>(), kernel_of<2>(),
                                          kernel_of<3>(), kernel_of<4>(), kernel_of<5>()};
static_assert(kNumShapes == 6, "one kernel per shape");

}  // namespace
}  // namespace edpp

using namespace edpp;

struct edpp_batch {
    edpp_ctx* ctx = nullptr;
    std::vector<edpp_job> jobs;
    int shape = kDefaultShape;
    int wgs_per_cu = 4;
    int order = EDPP_BATCH_ORDER_XCD;  // EDPP_BATCH_ORDER_*
    // the plan of the current shape
    DJob* d_jobs = nullptr;
    uint64_t* d_tix = nullptr;  // one word per ticket: job, phase, tile
    QTable qt = {};
    int64_t total = 0, hist_tiles = 0, hscan_tiles = 0;
    uint32_t* state = nullptr;  // zeroed each run: ticket, look-back status, split histogram words
    int64_t state_words = 0;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static void drop_graph(edpp_batch* b) {
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

static void drop_plan(edpp_batch* b) {
    drop_graph(b);
    hipFree(b->state);
    hipFree(b->d_jobs);
    hipFree(b->d_tix);
    b->state = nullptr;
    b->d_jobs = nullptr;
    b->d_tix = nullptr;
}

static int batch_launch(edpp_batch* b, hipStream_t s) {
    edpp_ctx* c = b->ctx;
    const int64_t init_grid = std::min<int64_t>(std::max<int64_t>(1, ceil_div(b->state_words, 256)), 4096);
    hipLaunchKernelGGL(k_batch_init, dim3((unsigned)init_grid), dim3(256), 0, s, b->state, b->state_words,
                       (const DJob*)b->d_jobs, (int32_t)b->jobs.size());
    if (b->total > 0) {
        const BatchShape& sh = kShapes[b->shape];
        const int64_t grid = std::min<int64_t>(b->total, (int64_t)c->cus * b->wgs_per_cu);
        hipLaunchKernelGGL(kKernels[b->shape], dim3((unsigned)grid), dim3(sh.block), shape_lds(sh), s,
                           (const DJob*)b->d_jobs, (const uint64_t*)b->d_tix, b->qt, b->state, c->spin_limit,
                           c->d_err);
    }
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

static bool a16(const void* p) { return p == nullptr || aligned16(p); }

// Validates the jobs (shape-independent); fills the shape-independent fields
static int validate_jobs(edpp_ctx* c, const edpp_job* jobs, int32_t njobs) {
    for (int32_t j = 0; j < njobs; ++j) {
        const edpp_job& q = jobs[j];
        if (q.kind < EDPP_JOB_SCAN || q.kind > EDPP_JOB_GATHER || (q.elem_bytes != 4 && q.elem_bytes != 8) ||
            q.n < 0 || q.n >= (1ll << 31))
            return fail(c, EDPP_EINVAL, "batch job %d: bad kind, width or length", j);
        if (q.n > 0 && (!q.in || !q.out || (q.kind == EDPP_JOB_GATHER && !q.idx)))
            return fail(c, EDPP_EINVAL, "batch job %d: missing buffer", j);
        if (q.kind == EDPP_JOB_SPLIT) {
            if (q.buckets < 1 || q.buckets > 256 || (q.buckets & (q.buckets - 1)))
                return fail(c, EDPP_EINVAL, "batch job %d: buckets must be a power of two <= 256", j);
            if ((q.vals_in == nullptr) != (q.vals_out == nullptr))
                return fail(c, EDPP_EINVAL, "batch job %d: values need both vals_in and vals_out", j);
        }
        if (!a16(q.in) || !a16(q.out) || !a16(q.idx) || !a16(q.vals_in) || !a16(q.vals_out))
            return fail(c, EDPP_EINVAL, "batch job %d: buffers must be 16-byte aligned", j);
    }
    return EDPP_OK;
}

// Builds the work queue and state of b->jobs for b->shape and uploads them.
static int plan(edpp_batch* b) {
    edpp_ctx* c = b->ctx;
    const BatchShape& sh = kShapes[b->shape];
    const int32_t njobs = (int32_t)b->jobs.size();
    drop_plan(b);
    std::vector<DJob> dj(njobs);
    // state layout in 32-bit words: the queues' ticket counters, then per job
    // its look-back status words and, for a split, its tagged histogram words
    int64_t words = (int64_t)kMaxQueues * kTicketStride;
    std::vector<int64_t> soff(njobs, -1);
    for (int32_t j = 0; j < njobs; ++j) {
        const edpp_job& q = b->jobs[j];
        DJob& d = dj[j];
        d = DJob();
        d.kind = q.kind;
        d.wide = q.elem_bytes == 8;
        d.n = q.n;
        d.in = q.in;
        d.out = q.out;
        d.idx = q.idx;
        d.vin = q.vals_in;
        d.vout = q.vals_out;
        d.aux = q.aux;
        if (q.kind == EDPP_JOB_SPLIT)
            while ((1 << d.lg) < q.buckets) ++d.lg;
        d.tiles = ceil_div(q.n, tile_elems(sh, q.kind, d.wide));
        soff[j] = words;
        if (q.kind == EDPP_JOB_SCAN) {
            words += (d.wide ? 4 : 2) * d.tiles;  // one or two u64 status words per tile
        } else if (q.kind == EDPP_JOB_SPLIT && d.tiles > 1) {
            // (a job of one tile takes its offsets from its own counts)
            const int64_t hlen = ((int64_t)1 << d.lg) * d.tiles;
            d.hscan_tiles = ceil_div(hlen, hscan_tile_elems(sh));
            words += 2 * d.hscan_tiles + 2 * hlen;  // u64 status, then u64 histogram words
        }
        words = (words + 3) & ~(int64_t)3;  // 16-byte alignment of every job's state
    }
    // jobs to queues: one queue, or (XCD order) one per XCD, each job to the
    // least loaded queue, largest jobs first, by the bytes it moves (a random
    // gather read counted as 16 bytes)
    const int nq = b->order == EDPP_BATCH_ORDER_XCD ? std::max(1, std::min(c->xcds, kMaxQueues)) : 1;
    std::vector<int32_t> qof(njobs, 0);
    if (nq > 1) {
        std::vector<int32_t> idx(njobs);
        std::vector<double> cost(njobs), load(nq, 0.0);
        for (int32_t j = 0; j < njobs; ++j) {
            idx[j] = j;
            const double w = dj[j].wide ? 8.0 : 4.0, n = (double)dj[j].n;
            cost[j] = dj[j].kind == EDPP_JOB_SCAN ? 2 * w * n : (dj[j].kind == EDPP_JOB_SPLIT ? (3 * w + 8) * n
                                                                                              : (4 + w + 16) * n);
        }
        std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b2) { return cost[a] > cost[b2]; });
        for (int32_t j : idx) {
            const int qq = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            qof[j] = qq;
            load[qq] += cost[j];
        }
    }
    // the ticket table, queue after queue
    int64_t all_tiles = 0;
    for (int32_t j = 0; j < njobs; ++j)  // a split: histogram, histogram-scan and scatter tiles
        all_tiles += dj[j].tiles + (dj[j].kind == EDPP_JOB_SPLIT && dj[j].tiles > 1 ? dj[j].tiles + dj[j].hscan_tiles : 0);
    if (all_tiles > kMaxBatchTiles)
        return fail(c, EDPP_EUNSUP, "batch of %lld tiles is too large (at most %lld)", (long long)all_tiles,
                    (long long)kMaxBatchTiles);
    std::vector<uint64_t> tix;
    tix.reserve((size_t)all_tiles);
    QTable qt = {};
    qt.nq = nq;
    qt.xcds = std::max(1, c->xcds);
    int64_t tk = 0, total = 0, hist = 0, hscan = 0;
    auto add = [&](int32_t j, int32_t phase, int64_t count) {
        for (int64_t i = 0; i < count; ++i) tix.push_back(pack_ticket((uint32_t)j, (uint32_t)phase, (uint32_t)i));
        tk += std::max<int64_t>(count, 0);
    };
    for (int q = 0; q < nq; ++q) {
        tk = 0;
        qt.tix_first[q] = (int64_t)tix.size();
        for (int32_t j = 0; j < njobs; ++j)
            if (qof[j] == q && dj[j].kind == EDPP_JOB_SPLIT && dj[j].tiles > 1) {
                add(j, kPhaseHist, dj[j].tiles);
                hist += dj[j].tiles;
            }
        for (int32_t j = 0; j < njobs; ++j)
            if (qof[j] == q && dj[j].kind == EDPP_JOB_SPLIT && dj[j].tiles > 1) {
                add(j, kPhaseHistScan, dj[j].hscan_tiles);
                hscan += dj[j].hscan_tiles;
            }
        if (b->order != EDPP_BATCH_ORDER_PHASES) {
            // split scatters in job order among the scan and gather tiles, so
            // the streaming tiles overlap the gathers' line-fetch-bound ones
            for (int32_t j = 0; j < njobs; ++j)
                if (qof[j] == q) add(j, dj[j].kind == EDPP_JOB_SPLIT ? kPhaseScatter : kPhaseMain, dj[j].tiles);
        } else {
            for (int32_t j = 0; j < njobs; ++j)
                if (qof[j] == q && dj[j].kind != EDPP_JOB_SPLIT) add(j, kPhaseMain, dj[j].tiles);
            for (int32_t j = 0; j < njobs; ++j)
                if (qof[j] == q && dj[j].kind == EDPP_JOB_SPLIT) add(j, kPhaseScatter, dj[j].tiles);
        }
        qt.total[q] = tk;
        total += tk;
    }
    b->qt = qt;
    b->total = total;
    b->hist_tiles = hist;
    b->hscan_tiles = hscan;
    b->state_words = words;
    if (hipMalloc(&b->state, (size_t)words * 4) != hipSuccess) return fail(c, EDPP_ENOMEM, "batch state");
    for (int32_t j = 0; j < njobs; ++j) {
        DJob& d = dj[j];
        uint32_t* st = b->state + soff[j];
        if (d.kind == EDPP_JOB_SCAN) {
            d.status = reinterpret_cast<uint64_t*>(st);
        } else if (d.kind == EDPP_JOB_SPLIT) {
            d.status = reinterpret_cast<uint64_t*>(st);
            d.his = reinterpret_cast<uint64_t*>(st + 2 * d.hscan_tiles);
        }
    }
    if (hipMalloc(&b->d_jobs, sizeof(DJob) * std::max<size_t>(1, dj.size())) != hipSuccess ||
        hipMalloc(&b->d_tix, sizeof(uint64_t) * std::max<size_t>(1, tix.size())) != hipSuccess)
        return fail(c, EDPP_ENOMEM, "batch job table");
    if ((njobs && hipMemcpy(b->d_jobs, dj.data(), sizeof(DJob) * dj.size(), hipMemcpyHostToDevice) != hipSuccess) ||
        (!tix.empty() &&
         hipMemcpy(b->d_tix, tix.data(), sizeof(uint64_t) * tix.size(), hipMemcpyHostToDevice) != hipSuccess))
        return fail(c, EDPP_EHIP, "batch: upload of the job table failed");
    EDPP_HIP(c, hipFuncSetAttribute((const void*)kKernels[b->shape], hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)shape_lds(sh)));
    return EDPP_OK;
}

extern "C" {

int edpp_batch_create(edpp_ctx* c, const edpp_job* jobs, int32_t njobs, edpp_batch** out) {
    if (!c || !out || njobs < 0 || (njobs > 0 && !jobs)) return c ? fail(c, EDPP_EINVAL, "batch: bad arguments") : EDPP_EINVAL;
    *out = nullptr;
    int rc = validate_jobs(c, jobs, njobs);
    if (rc) return rc;
    edpp_batch* b = new edpp_batch();
    b->ctx = c;
    b->jobs.assign(jobs, jobs + njobs);
    rc = plan(b);
    if (rc) {
        edpp_batch_destroy(b);
        return rc;
    }
    *out = b;
    return EDPP_OK;
}

int edpp_batch_set_occupancy(edpp_batch* b, int32_t wgs_per_cu) {
    if (!b || wgs_per_cu < 1 || wgs_per_cu > 8) return EDPP_EINVAL;
    if (wgs_per_cu != b->wgs_per_cu) drop_graph(b);
    b->wgs_per_cu = wgs_per_cu;
    return EDPP_OK;
}

int edpp_batch_num_shapes(void) { return kNumShapes; }

int edpp_batch_shape_info(int32_t shape, int32_t* block, int32_t* scan_tile32, int32_t* scan_tile64,
                          int32_t* split_tile, int32_t* gather_tile) {
    if (shape < 0 || shape >= kNumShapes) return EDPP_EINVAL;
    const BatchShape& s = kShapes[shape];
    if (block) *block = s.block;
    if (scan_tile32) *scan_tile32 = (int32_t)tile_elems(s, EDPP_JOB_SCAN, 0);
    if (scan_tile64) *scan_tile64 = (int32_t)tile_elems(s, EDPP_JOB_SCAN, 1);
    if (split_tile) *split_tile = (int32_t)tile_elems(s, EDPP_JOB_SPLIT, 0);
    if (gather_tile) *gather_tile = (int32_t)tile_elems(s, EDPP_JOB_GATHER, 0);
    return EDPP_OK;
}

int edpp_batch_set_shape(edpp_batch* b, int32_t shape) {
    if (!b) return EDPP_EINVAL;
    if (shape < 0 || shape >= kNumShapes) return fail(b->ctx, EDPP_EINVAL, "batch shape %d outside [0, %d)", shape,
                                                      kNumShapes);
    if (shape == b->shape && b->state) return EDPP_OK;
    EDPP_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));  // a run may still use the old plan
    b->shape = shape;
    return plan(b);
}

int edpp_batch_set_order(edpp_batch* b, int32_t order) {
    if (!b) return EDPP_EINVAL;
    if (order < EDPP_BATCH_ORDER_PHASES || order > EDPP_BATCH_ORDER_XCD)
        return fail(b->ctx, EDPP_EINVAL, "batch order %d unknown", order);
    if (order == b->order && b->state) return EDPP_OK;
    EDPP_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));  // a run may still use the old plan
    b->order = order;
    return plan(b);
}

int edpp_batch_get_shape(edpp_batch* b, int32_t* shape, int32_t* wgs_per_cu, int32_t* order) {
    if (!b) return EDPP_EINVAL;
    if (shape) *shape = b->shape;
    if (wgs_per_cu) *wgs_per_cu = b->wgs_per_cu;
    if (order) *order = b->order;
    return EDPP_OK;
}

int edpp_batch_info(edpp_batch* b, int64_t* tiles, int64_t* hist_tiles, int64_t* hscan_tiles) {
    if (!b) return EDPP_EINVAL;
    if (tiles) *tiles = b->total;
    if (hist_tiles) *hist_tiles = b->hist_tiles;
    if (hscan_tiles) *hscan_tiles = b->hscan_tiles;
    return EDPP_OK;
}

int edpp_batch_run(edpp_batch* b, double* ms) {
    if (!b) return EDPP_EINVAL;
    edpp_ctx* c = b->ctx;
    int rc = timing_begin(c, ms);
    if (!rc) rc = batch_launch(b, c->stream);
    if (!rc) rc = timing_end(c, ms);
    return rc;
}

int edpp_batch_run_graph(edpp_batch* b, double* ms) {
    if (!b) return EDPP_EINVAL;
    edpp_ctx* c = b->ctx;
    if (!b->exec) {
        // capture on a private stream (the null stream cannot be captured), replay on the context's
        if (!b->cap_stream) EDPP_HIP(c, hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
        EDPP_HIP(c, hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
        int rc = batch_launch(b, b->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
        if (rc) return rc;
        if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
        b->graph = g;
        EDPP_HIP(c, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    int rc = timing_begin(c, ms);
    if (rc) return rc;
    EDPP_HIP(c, hipGraphLaunch(b->exec, c->stream));
    return timing_end(c, ms);
}

// Exhaustive autotune of the batch's launch: every queue order, every tile
// shape, every occupancy of kOccupancies, one warm-up run and `reps` timed
// runs each, the median kept; the fastest is left set.  Each run recomputes
// every job's output from its inputs, so no job may write its own input.
int edpp_batch_autotune(edpp_batch* b, int32_t reps, double* times) {
    if (!b || reps < 1) return b ? fail(b->ctx, EDPP_EINVAL, "batch autotune: reps < 1") : EDPP_EINVAL;
    edpp_ctx* c = b->ctx;
    for (size_t j = 0; j < b->jobs.size(); ++j) {
        const edpp_job& q = b->jobs[j];
        if (q.n > 0 && (q.in == q.out || (q.vals_in && q.vals_in == q.vals_out)))
            return fail(c, EDPP_EINVAL, "batch autotune: job %zu works in place, repeated runs would change it", j);
    }
    double best = -1.0;
    int best_shape = b->shape, best_occ = b->wgs_per_cu, best_order = b->order;
    std::vector<double> ts((size_t)reps);
    for (int ord = 0; ord < kNumOrders; ++ord) {
        int rc = edpp_batch_set_order(b, ord);
        if (rc) return rc;
        for (int s = 0; s < kNumShapes; ++s) {
            rc = edpp_batch_set_shape(b, s);
            if (rc) return rc;
            for (int o = 0; o < kNumOcc; ++o) {
                edpp_batch_set_occupancy(b, kOccupancies[o]);
                double ms = 0.0;
                rc = edpp_batch_run(b, &ms);  // warm-up
                for (int r = 0; r < reps && !rc; ++r) rc = edpp_batch_run(b, &ts[(size_t)r]);
                if (rc) return rc;
                std::sort(ts.begin(), ts.end());
                const double med = ts[(size_t)reps / 2];
                if (times) times[(ord * kNumShapes + s) * kNumOcc + o] = med;
                if (best < 0 || med < best) {
                    best = med;
                    best_order = ord;
                    best_shape = s;
                    best_occ = kOccupancies[o];
                }
            }
        }
    }
    int rc = edpp_batch_set_order(b, best_order);
    if (!rc) rc = edpp_batch_set_shape(b, best_shape);
    if (!rc) rc = edpp_batch_set_occupancy(b, best_occ);
    return rc;
}

int edpp_batch_destroy(edpp_batch* b) {
    if (!b) return EDPP_OK;
    if (b->ctx && b->ctx->stream) hipStreamSynchronize(b->ctx->stream);
    drop_plan(b);
    if (b->cap_stream) hipStreamDestroy(b->cap_stream);
    delete b;
    return EDPP_OK;
}

int edpp_batch_run_plain(edpp_ctx* c, const edpp_job* jobs, int32_t njobs, double* ms) {
    if (!c || njobs < 0 || (njobs > 0 && !jobs)) return EDPP_EINVAL;
    int rc = timing_begin(c, ms);
    for (int32_t j = 0; j < njobs && !rc; ++j) {
        const edpp_job& q = jobs[j];
        const bool w = q.elem_bytes == 8;
        switch (q.kind) {
            case EDPP_JOB_SCAN:
                rc = w ? launch_scan<uint64_t>(c, (const uint64_t*)q.in, (uint64_t*)q.out, q.n, EDPP_SCAN_AUTO,
                                               (uint64_t*)q.aux)
                       : launch_scan<uint32_t>(c, (const uint32_t*)q.in, (uint32_t*)q.out, q.n, EDPP_SCAN_AUTO,
                                               (uint32_t*)q.aux);
                break;
            case EDPP_JOB_SPLIT:
                rc = w ? launch_split<uint64_t>(c, (const uint64_t*)q.in, (uint64_t*)q.out, q.vals_in, q.vals_out,
                                                (int32_t*)q.aux, q.n, q.buckets, EDPP_SPLIT_AUTO)
                       : launch_split<uint32_t>(c, (const uint32_t*)q.in, (uint32_t*)q.out, q.vals_in, q.vals_out,
                                                (int32_t*)q.aux, q.n, q.buckets, EDPP_SPLIT_AUTO);
                break;
            case EDPP_JOB_GATHER:
                rc = w ? launch_gather<uint64_t>(c, (const uint64_t*)q.in, (uint64_t*)q.out, q.idx, q.n, 1, false)
                       : launch_gather<uint32_t>(c, (const uint32_t*)q.in, (uint32_t*)q.out, q.idx, q.n, 1, false);
                break;
            default:
                rc = fail(c, EDPP_EINVAL, "plain batch job %d: unknown kind", j);
        }
    }
    if (!rc) rc = timing_end(c, ms);
    return rc;
}

}  // extern "C"
