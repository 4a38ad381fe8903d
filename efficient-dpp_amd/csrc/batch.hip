// batch.hip — persistent executor for many scans, splits and gathers of
// uneven length in one launch, fed by a device ticket queue.
//
// Queue order (tickets):  [split histogram tiles of every split job]
//                         [tiles of the in-place scans of those histograms]
//                         [scan tiles and gather tiles, in job order]
//                         [split scatter tiles]
// Every dependency points to smaller tickets: a histogram scan waits for its
// job's histogram tiles, a scatter tile for its job's histogram scan, a scan
// tile (and a histogram-scan tile) for its predecessors in the look-back.
// Tickets are handed out in order, so whatever a workgroup waits for is held
// by a workgroup that is already running; every wait ends.  The reference has
// no counterpart (its primitives are one synchronous call each).
#include <algorithm>
#include <vector>
#include "ctx.hpp"
#include "edpp_batch.h"
#include "split_tile.hpp"
#include "tile_scan.hpp"

namespace edpp {

namespace {

constexpr int kBatchBlock = 256;
static_assert(kBatchBlock == kSplitBlock, "split tiles assume the split block size");
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
    int32_t* his;      // split: [B][tiles], scanned in place (accessed with agent-scope atomics only)
    uint32_t* ctr;     // split: [0] histogram tiles done, [1] histogram-scan tiles done
};

struct Seg {
    int64_t first;  // first ticket
    int32_t job, phase;
};

// keys per lane of a split tile: 8 keeps the persistent kernel at about 100
// registers (16 for 32-bit keys needed 200, two workgroups per CU)
constexpr int kBatchSplitIpt = 8;

// tile lengths per kind and width
__host__ __device__ inline int64_t tile_elems(int kind, int wide) {
    if (kind == EDPP_JOB_SCAN) return wide ? 2048 : 8192;  // TileScan<u64,256,4>, TileScan<u32,256,8>
    if (kind == EDPP_JOB_SPLIT) return 256 * kBatchSplitIpt;  // 256 lanes x kBatchSplitIpt keys
    return 4096;                                           // gather: 256 lanes x 16
}
constexpr int64_t kHistScanTile = 4096;  // TileScan<uint32_t,256,4>

// Cross-workgroup data of one launch (histograms, counters) moves through
// agent-scope atomic loads and stores, which stay coherent across the XCDs'
// L2s.  A finished tile is counted with a release increment and a waiter
// that has seen its count takes one acquire fence, so the counts written
// before the increment are ordered before the waiter's reads of them by the
// memory model, not only by the hardware.  Polling stays relaxed: no spin
// iteration invalidates a cache, only the one successful wait does.

// thread 0 waits for *p >= target, then acquires (agent scope) what the
// release increments up to target published; the barrier passes that on to
// the workgroup
__device__ bool wait_count(uint32_t* p, uint32_t target, uint32_t spin_limit, int32_t* err, uint32_t* s_ok) {
    if (threadIdx.x == 0) {
        uint32_t spins = 0, ok = 1;
        while (ld_relaxed(p) < target) {
            if (++spins > spin_limit) {
                st_relaxed(err, (int32_t)EDPP_ETIMEOUT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        if (ok) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok != 0;
}

// every lane's stores of this tile complete and reach thread 0 through the
// barrier; thread 0 then counts the tile with a release increment
__device__ void signal_count(uint32_t* p) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------- scan tiles
template <class T, int VPT, bool COHERENT>
__device__ void scan_tile(const T* in, T* out, int64_t n, int64_t t, typename LookbackFor<T>::type st, T* total,
                          char* smem, uint32_t spin_limit, int32_t* err) {
    typedef TileScan<T, kBatchBlock, VPT> TS;
    typename TS::Smem& sm = *reinterpret_cast<typename TS::Smem*>(smem);
    const int64_t base = t * TS::TILE;
    TS ts;
    if (COHERENT) ts.load_coherent(in, base, n);
    else ts.load(in, base, n);
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
    if (COHERENT) ts.store_coherent(out, base, n, sm.prefix);
    else ts.store(out, base, n, sm.prefix);
}

// ---------------------------------------------------------------- split tiles
template <class K, int IPT>
__device__ void split_hist_tile(const DJob& J, int64_t t, uint32_t* lds) {
    const int B = 1 << J.lg;
    const int w = threadIdx.x / kWave;
    constexpr int S = kBatchBlock * IPT;
    uint32_t* whist = lds;
    for (int b = threadIdx.x; b < kSplitWaves * B; b += kBatchBlock) whist[b] = 0;
    __syncthreads();
    const int64_t base = t * S, end = std::min<int64_t>(base + S, J.n);
    if (J.lg <= kBallotMatchBits) count_subtile<K, IPT, true>((const K*)J.in, base, end, J.lg, 0, whist + w * B);
    else count_subtile<K, IPT, false>((const K*)J.in, base, end, J.lg, 0, whist + w * B);
    __syncthreads();
    for (int b = threadIdx.x; b < B; b += kBatchBlock) {
        uint32_t a = 0;
        for (int q = 0; q < kSplitWaves; ++q) a += whist[q * B + b];
        st_relaxed(J.his + (int64_t)b * J.tiles + t, (int32_t)a);
    }
    signal_count(J.ctr + 0);
}

template <class K, int IPT>
__device__ void split_scatter_tile(const DJob& J, int64_t t, uint32_t* lds) {
    const int B = 1 << J.lg;
    constexpr int S = kBatchBlock * IPT;
    uint64_t* wmask = reinterpret_cast<uint64_t*>(lds);
    uint32_t* whist = lds + whist_word_offset(B);
    uint32_t* run = whist + kSplitWaves * B;
    uint32_t* sub = run + B;
    uint32_t* off = sub + B;
    uint32_t* wtmp = off + B;
    K* kbuf = reinterpret_cast<K*>(lds + kbuf_word_offset(B));
    int32_t* vbuf = reinterpret_cast<int32_t*>(kbuf + S);
    const int w = threadIdx.x / kWave;
    const bool kv = J.vin != nullptr;
    for (int b = threadIdx.x; b < B; b += kBatchBlock) {
        uint32_t r = (uint32_t)ld_relaxed(J.his + (int64_t)b * J.tiles + t);
        run[b] = r;
        if (t == 0 && J.aux) static_cast<int32_t*>(J.aux)[b] = (int32_t)r;
    }
    for (int b = threadIdx.x; b < kSplitWaves * B; b += kBatchBlock) {
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
    if (kv)
        place_and_store<K, IPT, true>(whist, run, sub, off, wtmp, kbuf, vbuf, bk, rk, key, val, B, 0, cnt,
                                      static_cast<K*>(J.out), J.vout, J.n);
    else
        place_and_store<K, IPT, false>(whist, run, sub, off, wtmp, kbuf, vbuf, bk, rk, key, val, B, 0, cnt,
                                       static_cast<K*>(J.out), nullptr, J.n);
}

// ---------------------------------------------------------------- gather tiles
template <class T>
__device__ void gather_tile(const DJob& J, int64_t t) {
    const T* in = static_cast<const T*>(J.in);
    T* out = static_cast<T*>(J.out);
    const uint32_t n = (uint32_t)J.n;
    const int64_t base = t * 4096;
    #pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int64_t e0 = base + (int64_t)u * 1024 + threadIdx.x * 4;
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

__global__ __launch_bounds__(kBatchBlock) void k_batch(const DJob* __restrict__ jobs, const Seg* __restrict__ segs,
                                                       int32_t nseg, int64_t total, uint32_t* ticket,
                                                       uint32_t spin_limit, int32_t* err) {
    extern __shared__ __align__(16) char smem[];
    __shared__ uint32_t s_ticket, s_ok;
    while (true) {
        if (threadIdx.x == 0) s_ticket = fetch_add(ticket, 1u);
        __syncthreads();
        const int64_t tk = s_ticket;
        __syncthreads();  // all lanes hold the ticket before lane 0 may take the next
        if (tk >= total) break;
        int lo = 0, hi = nseg - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (segs[mid].first <= tk) lo = mid;
            else hi = mid - 1;
        }
        const Seg sg = segs[lo];
        const DJob& J = jobs[sg.job];
        const int64_t t = tk - sg.first;
        uint32_t* lds = reinterpret_cast<uint32_t*>(smem);
        if (sg.phase == kPhaseHist) {
            if (J.wide) split_hist_tile<uint64_t, kBatchSplitIpt>(J, t, lds);
            else split_hist_tile<uint32_t, kBatchSplitIpt>(J, t, lds);
        } else if (sg.phase == kPhaseHistScan) {
            wait_count(J.ctr + 0, (uint32_t)J.tiles, spin_limit, err, &s_ok);
            LookbackState32 st;
            st.status = J.status;
            st.epoch = 1;
            const int64_t hlen = ((int64_t)1 << J.lg) * J.tiles;
            scan_tile<uint32_t, 4, true>((const uint32_t*)J.his, (uint32_t*)J.his, hlen, t, st, nullptr, smem,
                                         spin_limit, err);
            signal_count(J.ctr + 1);
        } else if (sg.phase == kPhaseScatter) {
            wait_count(J.ctr + 1, (uint32_t)J.hscan_tiles, spin_limit, err, &s_ok);
            if (J.wide) split_scatter_tile<uint64_t, kBatchSplitIpt>(J, t, lds);
            else split_scatter_tile<uint32_t, kBatchSplitIpt>(J, t, lds);
        } else if (J.kind == EDPP_JOB_SCAN) {
            if (J.wide) {
                LookbackState64 st;
                st.status = J.status;
                st.epoch = 1;
                scan_tile<uint64_t, 4, false>((const uint64_t*)J.in, (uint64_t*)J.out, J.n, t, st, (uint64_t*)J.aux,
                                              smem, spin_limit, err);
            } else {
                LookbackState32 st;
                st.status = J.status;
                st.epoch = 1;
                scan_tile<uint32_t, 8, false>((const uint32_t*)J.in, (uint32_t*)J.out, J.n, t, st, (uint32_t*)J.aux,
                                              smem, spin_limit, err);
            }
        } else {
            if (J.wide) gather_tile<uint64_t>(J, t);
            else gather_tile<uint32_t>(J, t);
        }
        __syncthreads();  // LDS of this tile is dead before the next one starts
    }
}

// LDS of the largest tile: the split scatter with 64-bit keys at 256 buckets
// (ranking masks, counts and offsets, then 2048 keys and 2048 values): 40000
// bytes, so four workgroups share a CU
constexpr size_t kBatchLds =
    (size_t)kbuf_word_offset(256) * 4 + (size_t)kBatchBlock * kBatchSplitIpt * (sizeof(uint64_t) + sizeof(int32_t));

}  // namespace
}  // namespace edpp

using namespace edpp;

struct edpp_batch {
    edpp_ctx* ctx = nullptr;
    std::vector<edpp_job> jobs;
    DJob* d_jobs = nullptr;
    Seg* d_segs = nullptr;
    int32_t nseg = 0;
    int64_t total = 0, hist_tiles = 0, hscan_tiles = 0;
    uint32_t* state = nullptr;  // zeroed each run: ticket, counters, look-back status and flags
    int64_t state_words = 0;
    void* rest = nullptr;       // split histograms (no reset needed)
    int wgs_per_cu = 4;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static int batch_launch(edpp_batch* b, hipStream_t s) {
    edpp_ctx* c = b->ctx;
    const int64_t init_grid = std::min<int64_t>(std::max<int64_t>(1, ceil_div(b->state_words, 256)), 4096);
    hipLaunchKernelGGL(k_batch_init, dim3((unsigned)init_grid), dim3(256), 0, s, b->state, b->state_words,
                       (const DJob*)b->d_jobs, (int32_t)b->jobs.size());
    if (b->total > 0) {
        const int64_t grid = std::min<int64_t>(b->total, (int64_t)c->cus * b->wgs_per_cu);
        hipLaunchKernelGGL(k_batch, dim3((unsigned)grid), dim3(kBatchBlock), kBatchLds, s, (const DJob*)b->d_jobs,
                           (const Seg*)b->d_segs, b->nseg, b->total, b->state, c->spin_limit, c->d_err);
    }
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

static bool a16(const void* p) { return p == nullptr || aligned16(p); }

extern "C" {

int edpp_batch_create(edpp_ctx* c, const edpp_job* jobs, int32_t njobs, edpp_batch** out) {
    if (!c || !out || njobs < 0 || (njobs > 0 && !jobs)) return c ? fail(c, EDPP_EINVAL, "batch: bad arguments") : EDPP_EINVAL;
    *out = nullptr;
    std::vector<DJob> dj(njobs);
    // state layout in 32-bit words: [0] ticket, then per job its status/flags/counters
    int64_t words = 64;  // ticket and padding
    int64_t rest_bytes = 0;
    std::vector<int64_t> soff(njobs, -1), roff(njobs, -1);
    for (int32_t j = 0; j < njobs; ++j) {
        const edpp_job& q = jobs[j];
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
            while ((1 << d.lg) < q.buckets) ++d.lg;
        }
        if (!a16(q.in) || !a16(q.out) || !a16(q.idx) || !a16(q.vals_in) || !a16(q.vals_out))
            return fail(c, EDPP_EINVAL, "batch job %d: buffers must be 16-byte aligned", j);
        d.tiles = ceil_div(q.n, tile_elems(q.kind, d.wide));
        soff[j] = words;
        if (q.kind == EDPP_JOB_SCAN) {
            words += (d.wide ? 4 : 2) * d.tiles;  // one or two u64 status words per tile
        } else if (q.kind == EDPP_JOB_SPLIT) {
            const int64_t hlen = ((int64_t)1 << d.lg) * d.tiles;
            d.hscan_tiles = ceil_div(hlen, kHistScanTile);
            words += 2 + 2 * d.hscan_tiles;  // counters, then u64 status
            roff[j] = rest_bytes;
            rest_bytes += ((hlen * 4 + 255) / 256) * 256;
        }
        words = (words + 3) & ~(int64_t)3;  // 16-byte alignment of every job's state
    }
    // segments in queue order
    std::vector<Seg> segs;
    int64_t tk = 0, hist = 0, hscan = 0;
    auto add = [&](int32_t j, int32_t phase, int64_t count) {
        if (count <= 0) return;
        Seg s;
        s.first = tk;
        s.job = j;
        s.phase = phase;
        segs.push_back(s);
        tk += count;
    };
    for (int32_t j = 0; j < njobs; ++j)
        if (dj[j].kind == EDPP_JOB_SPLIT) { add(j, kPhaseHist, dj[j].tiles); hist += dj[j].tiles; }
    for (int32_t j = 0; j < njobs; ++j)
        if (dj[j].kind == EDPP_JOB_SPLIT) { add(j, kPhaseHistScan, dj[j].hscan_tiles); hscan += dj[j].hscan_tiles; }
    for (int32_t j = 0; j < njobs; ++j)
        if (dj[j].kind != EDPP_JOB_SPLIT) add(j, kPhaseMain, dj[j].tiles);
    for (int32_t j = 0; j < njobs; ++j)
        if (dj[j].kind == EDPP_JOB_SPLIT) add(j, kPhaseScatter, dj[j].tiles);
    if (tk >= (1ll << 32)) return fail(c, EDPP_EUNSUP, "batch of %lld tiles is too large", (long long)tk);

    edpp_batch* b = new edpp_batch();
    b->ctx = c;
    b->jobs.assign(jobs, jobs + njobs);
    b->total = tk;
    b->hist_tiles = hist;
    b->hscan_tiles = hscan;
    b->nseg = (int32_t)segs.size();
    b->state_words = words;
    int rc = EDPP_OK;
    do {
        if (hipMalloc(&b->state, (size_t)words * 4) != hipSuccess) { rc = EDPP_ENOMEM; break; }
        if (rest_bytes && hipMalloc(&b->rest, (size_t)rest_bytes) != hipSuccess) { rc = EDPP_ENOMEM; break; }
        for (int32_t j = 0; j < njobs; ++j) {
            DJob& d = dj[j];
            uint32_t* st = b->state + soff[j];
            char* rs = roff[j] >= 0 ? static_cast<char*>(b->rest) + roff[j] : nullptr;
            if (d.kind == EDPP_JOB_SCAN) {
                d.status = reinterpret_cast<uint64_t*>(st);
            } else if (d.kind == EDPP_JOB_SPLIT) {
                d.ctr = st;
                d.status = reinterpret_cast<uint64_t*>(st + 2);
                d.his = reinterpret_cast<int32_t*>(rs);
            }
        }
        if (hipMalloc(&b->d_jobs, sizeof(DJob) * std::max<size_t>(1, dj.size())) != hipSuccess ||
            hipMalloc(&b->d_segs, sizeof(Seg) * std::max<size_t>(1, segs.size())) != hipSuccess) {
            rc = EDPP_ENOMEM;
            break;
        }
        if ((njobs && hipMemcpy(b->d_jobs, dj.data(), sizeof(DJob) * dj.size(), hipMemcpyHostToDevice) != hipSuccess) ||
            (!segs.empty() &&
             hipMemcpy(b->d_segs, segs.data(), sizeof(Seg) * segs.size(), hipMemcpyHostToDevice) != hipSuccess)) {
            rc = EDPP_EHIP;
            break;
        }
        if (hipFuncSetAttribute((const void*)k_batch, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kBatchLds) !=
            hipSuccess) {
            rc = EDPP_EHIP;
            break;
        }
    } while (0);
    if (rc) {
        edpp_batch_destroy(b);
        return fail(c, rc, "batch: device allocation or upload failed");
    }
    *out = b;
    return EDPP_OK;
}

int edpp_batch_set_occupancy(edpp_batch* b, int32_t wgs_per_cu) {
    if (!b || wgs_per_cu < 1 || wgs_per_cu > 8) return EDPP_EINVAL;
    if (wgs_per_cu != b->wgs_per_cu && b->exec) {
        hipGraphExecDestroy(b->exec);
        hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
    }
    b->wgs_per_cu = wgs_per_cu;
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

int edpp_batch_destroy(edpp_batch* b) {
    if (!b) return EDPP_OK;
    if (b->ctx && b->ctx->stream) hipStreamSynchronize(b->ctx->stream);
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->cap_stream) hipStreamDestroy(b->cap_stream);
    hipFree(b->state);
    hipFree(b->rest);
    hipFree(b->d_jobs);
    hipFree(b->d_segs);
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
