// scan.hip — global exclusive scans: single-pass decoupled look-back,
// reduce-then-scan (RTS) and scan-scan-add (SSA).
//
// Reference: opencl/primitives/scanImpl.cpp:15-236 (host),
// opencl/kernels/scan_global_chain_kernel.cl:12-106 (chained scan),
// openmp/test_scan_CPU.cpp:84-163 (SSA / RTS structure).  The reference's
// chained scan uses a static tile schedule and an in-band -1 sentinel; here
// tiles are claimed from a ticket counter and status words carry a flag and
// a call epoch, so wrapped prefixes equal to -1 are legal values.
#include "ctx.hpp"
#include "tile_scan.hpp"

namespace edpp {

// ---------------------------------------------------------------- look-back scan
// `in` and `out` may be the same array (an in-place scan): neither pointer is
// __restrict__, and every load of a tile precedes a barrier that precedes its
// stores.  The tile is held as folded prefixes through the look-back.
// (Holding 1024x8 tiles to 64 registers for two workgroups per CU measured
// slower from 2^22 to 2^26 elements: 35.6 against 33.4 us at 2^24.)
template <class T, int BLOCK, int VPT>
__global__ __launch_bounds__(BLOCK) void k_scan_lookback(const T* in, T* out,
                                                         int64_t n, typename LookbackFor<T>::type st,
                                                         uint32_t* tickets, T* d_total,
                                                         uint32_t spin_limit, int32_t* err, int64_t tiles) {
    typedef TileScan<T, BLOCK, VPT> TS;
    __shared__ typename TS::Smem sm;
    if (threadIdx.x == 0) {
        uint32_t t = fetch_add(tickets + (st.epoch & 1u), 1u);
        // the next call uses the other counter; it starts after this grid ends
        if (t == 0) st_relaxed(tickets + ((st.epoch + 1u) & 1u), 0u);
        sm.tile = t;
    }
    __syncthreads();
    const int64_t tile = sm.tile;
    if (tile >= tiles) {  // a stale ticket counter: report instead of reading past the state
        if (threadIdx.x == 0) st_relaxed(err, (int32_t)EDPP_EHIP);
        return;
    }
    const int64_t base = tile * TS::TILE;
    TS ts;
    ts.load(in, base, n);
    ts.scan(sm);
    ts.fold();
    if (threadIdx.x < kWave) {
        T prefix = 0;
        if (tile == 0) {
            if (threadIdx.x == 0) st.publish(0, kFlagPrefix, ts.total);
        } else {
            if (threadIdx.x == 0) st.publish(tile, kFlagAggregate, ts.total);
            prefix = lookback<T>(st, tile, spin_limit, err);
            if (threadIdx.x == 0) st.publish(tile, kFlagPrefix, prefix + ts.total);
        }
        if (threadIdx.x == 0) {
            sm.prefix = prefix;
            if (d_total && base + TS::TILE >= n) *d_total = prefix + ts.total;
        }
    }
    __syncthreads();
    ts.store_folded(out, base, n, sm.prefix);
}

// Persistent form of the look-back scan, for calls with several tiles per
// workgroup: each workgroup loops over tickets and loads its next tile while
// the current one looks back and stores, so the look-back's round trips and
// the ticket's atomic overlap memory traffic.  Per step: take the next
// ticket, issue its loads, look back for the current tile and publish its
// prefix, scan the next tile and publish its aggregate, store the current
// tile.  A workgroup holds two tickets at once, but the smaller one has
// published its prefix before the larger one is waited on, and the larger
// one's aggregate needs nothing that is not finished, so every wait ends.
template <class T, int BLOCK, int VPT>
__global__ __launch_bounds__(BLOCK) void k_scan_pipe(const T* in, T* out, int64_t n,
                                                     typename LookbackFor<T>::type st, uint32_t* tickets,
                                                     T* d_total, uint32_t spin_limit, int32_t* err, int64_t tiles) {
    typedef TileScan<T, BLOCK, VPT> TS;
    __shared__ typename TS::Smem sm;
    __shared__ uint32_t s_ticket;
    if (threadIdx.x == 0) {
        const uint32_t t = fetch_add(tickets + (st.epoch & 1u), 1u);
        // the next call uses the other counter; it starts after this grid ends
        if (t == 0) st_relaxed(tickets + ((st.epoch + 1u) & 1u), 0u);
        s_ticket = t;
    }
    __syncthreads();
    int64_t cur = s_ticket;
    if (cur >= tiles) return;  // more workgroups than tiles
    TS a;
    a.load(in, cur * TS::TILE, n);
    a.scan(sm);
    a.fold();
    if (threadIdx.x == 0) st.publish(cur, cur == 0 ? kFlagPrefix : kFlagAggregate, a.total);
    while (true) {
        if (threadIdx.x == 0) {
            const uint32_t t = fetch_add(tickets + (st.epoch & 1u), 1u);
            if (t == 0) st_relaxed(tickets + ((st.epoch + 1u) & 1u), 0u);
            s_ticket = t;
        }
        __syncthreads();
        const int64_t nxt = s_ticket;
        TS b;
        if (nxt < tiles) b.load(in, nxt * TS::TILE, n);  // in flight through the look-back below
        if (threadIdx.x < kWave) {
            T prefix = 0;
            if (cur > 0) {
                prefix = lookback<T>(st, cur, spin_limit, err);
                if (threadIdx.x == 0) st.publish(cur, kFlagPrefix, prefix + a.total);
            }
            if (threadIdx.x == 0) {
                sm.prefix = prefix;
                if (d_total && cur * TS::TILE + TS::TILE >= n) *d_total = prefix + a.total;
            }
        }
        __syncthreads();
        const T prefix = sm.prefix;
        if (nxt < tiles) {
            b.scan(sm);  // leaves sm.prefix alone; its barriers order every read of it before the next write
            b.fold();
            if (threadIdx.x == 0) st.publish(nxt, nxt == 0 ? kFlagPrefix : kFlagAggregate, b.total);
        }
        a.store_folded(out, cur * TS::TILE, n, prefix);
        if (nxt >= tiles) break;
        a = b;
        cur = nxt;
    }
}

// ---------------------------------------------------------------- RTS / SSA pieces
// partials[tile] = sum of the tile
template <class T, int BLOCK, int VPT>
__global__ __launch_bounds__(BLOCK) void k_tile_reduce(const T* __restrict__ in, int64_t n, T* __restrict__ partials) {
    typedef TileScan<T, BLOCK, VPT> TS;
    __shared__ T lds[BLOCK / kWave];
    TS ts;
    const int64_t base = (int64_t)blockIdx.x * TS::TILE;
    ts.load(in, base, n);
    T a = 0;
    #pragma unroll
    for (int j = 0; j < VPT; ++j)
        #pragma unroll
        for (int k = 0; k < TS::EPV; ++k) a += ts.v[j][k];
    T r = block_sum<T, BLOCK>(a, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// exclusive scan of m partials by one workgroup of 1024 lanes; *d_total = sum
template <class T>
__global__ __launch_bounds__(1024) void k_scan_partials(T* p, int64_t m, T* d_total) {
    constexpr int NW = 1024 / kWave;
    __shared__ T wt[NW];
    __shared__ T carry_s;
    const int lane = lane_id(), w = threadIdx.x / kWave;
    T carry = 0;
    for (int64_t b = 0; b < m; b += 1024) {
        int64_t i = b + threadIdx.x;
        T x = i < m ? p[i] : T(0);
        T inc = wave_inclusive_add(x);
        if (lane == kWave - 1) wt[w] = inc;
        __syncthreads();
        T before = 0, all = 0;
        #pragma unroll
        for (int q = 0; q < NW; ++q) {
            T y = wt[q];
            before += q < w ? y : T(0);
            all += y;
        }
        if (i < m) p[i] = carry + before + inc - x;
        carry += all;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        carry_s = carry;
        if (d_total) *d_total = carry;
    }
}

// RTS final pass: scan each tile from its offset
template <class T, int BLOCK, int VPT>
__global__ __launch_bounds__(BLOCK) void k_tile_scan_offset(const T* in, T* out, int64_t n,
                                                            const T* __restrict__ offsets) {
    typedef TileScan<T, BLOCK, VPT> TS;
    __shared__ typename TS::Smem sm;
    TS ts;
    const int64_t base = (int64_t)blockIdx.x * TS::TILE;
    ts.load(in, base, n);
    ts.scan(sm);
    ts.store(out, base, n, offsets[blockIdx.x]);
}

// SSA first pass: local exclusive scan of each tile, tile sums to partials
template <class T, int BLOCK, int VPT>
__global__ __launch_bounds__(BLOCK) void k_tile_scan_local(const T* in, T* out, int64_t n,
                                                           T* __restrict__ partials) {
    typedef TileScan<T, BLOCK, VPT> TS;
    __shared__ typename TS::Smem sm;
    TS ts;
    const int64_t base = (int64_t)blockIdx.x * TS::TILE;
    ts.load(in, base, n);
    ts.scan(sm);
    ts.store(out, base, n, T(0));
    if (threadIdx.x == 0) partials[blockIdx.x] = ts.total;
}

// SSA last pass: out[i] += offsets[tile of i]
template <class T, int BLOCK, int VPT>
__global__ __launch_bounds__(BLOCK) void k_tile_add(T* __restrict__ out, int64_t n, const T* __restrict__ offsets) {
    typedef TileScan<T, BLOCK, VPT> TS;
    typedef typename TS::V V;
    const int64_t base = (int64_t)blockIdx.x * TS::TILE;
    const T o = offsets[blockIdx.x];
    const int t = threadIdx.x;
    if (base + TS::TILE <= n) {
        #pragma unroll
        for (int j = 0; j < VPT; ++j) {
            V* p = reinterpret_cast<V*>(out + base + (int64_t)j * TS::SUB) + t;
            V x = ld_stream(p);
            #pragma unroll
            for (int k = 0; k < TS::EPV; ++k) x[k] += o;
            st_stream(x, p);
        }
    } else {
        for (int j = 0; j < VPT; ++j)
            for (int k = 0; k < TS::EPV; ++k) {
                int64_t e = base + (int64_t)j * TS::SUB + (int64_t)t * TS::EPV + k;
                if (e < n) out[e] += o;
            }
    }
}

// scalar fallback for pointers that are not 16-byte aligned: one lane scans
// a chunk, chunks are chained through the same look-back protocol
template <class T>
__global__ __launch_bounds__(64) void k_scan_unaligned(const T* in, T* out, int64_t n, typename LookbackFor<T>::type st,
                                                       uint32_t* tickets, T* d_total, uint32_t spin_limit,
                                                       int32_t* err, int64_t tiles) {
    constexpr int CH = 64 * 16;
    __shared__ uint32_t tile_s;
    if (threadIdx.x == 0) {
        uint32_t t = fetch_add(tickets + (st.epoch & 1u), 1u);
        if (t == 0) st_relaxed(tickets + ((st.epoch + 1u) & 1u), 0u);
        tile_s = t;
    }
    __syncthreads();
    const int64_t tile = tile_s;
    if (tile >= tiles) {
        if (threadIdx.x == 0) st_relaxed(err, (int32_t)EDPP_EHIP);
        return;
    }
    const int64_t base = tile * CH;
    T v[16];
    T s = 0;
    #pragma unroll
    for (int k = 0; k < 16; ++k) {
        int64_t e = base + threadIdx.x * 16 + k;
        v[k] = e < n ? in[e] : T(0);
        s += v[k];
    }
    T inc = wave_inclusive_add(s);
    T total = read_lane(inc, 63);
    T prefix = 0;
    if (tile == 0) {
        if (threadIdx.x == 0) st.publish(0, kFlagPrefix, total);
    } else {
        if (threadIdx.x == 0) st.publish(tile, kFlagAggregate, total);
        prefix = lookback<T>(st, tile, spin_limit, err);
        if (threadIdx.x == 0) st.publish(tile, kFlagPrefix, prefix + total);
    }
    if (threadIdx.x == 0 && d_total && base + CH >= n) *d_total = prefix + total;
    T x = prefix + inc - s;
    #pragma unroll
    for (int k = 0; k < 16; ++k) {
        int64_t e = base + threadIdx.x * 16 + k;
        if (e < n) out[e] = x;
        x += v[k];
    }
}

// ---------------------------------------------------------------- host side
namespace {

template <class T> typename LookbackFor<T>::type make_state(edpp_ctx* c, uint32_t epoch);
template <> LookbackState32 make_state<uint32_t>(edpp_ctx* c, uint32_t epoch) {
    LookbackState32 s;
    s.status = c->scan_status;
    s.epoch = epoch;
    return s;
}
template <> LookbackState64 make_state<uint64_t>(edpp_ctx* c, uint32_t epoch) {
    LookbackState64 s;
    s.status = c->scan_status64;
    s.epoch = epoch;
    return s;
}

constexpr int kBlock = 256;

template <class T, int BLOCK, int VPT>
int run_lookback(edpp_ctx* c, const T* in, T* out, int64_t n, T* d_total) {
    typedef TileScan<T, BLOCK, VPT> TS;
    int64_t tiles = ceil_div(n, TS::TILE);
    uint32_t epoch;
    int rc = scan_state_acquire(c, tiles, &epoch);
    if (rc) return rc;
    hipLaunchKernelGGL((k_scan_lookback<T, BLOCK, VPT>), dim3((unsigned)tiles), dim3(BLOCK), 0, c->stream, in, out,
                       n, make_state<T>(c, epoch), c->tickets, d_total, c->spin_limit, c->d_err, tiles);
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

// Persistent look-back scan: 1024-lane workgroups holding two tiles of
// registers each, as many as fit on the chip at once (two per CU for 4
// vectors of 32-bit values per lane, else one).
template <class T, int VPT>
int run_pipe(edpp_ctx* c, const T* in, T* out, int64_t n, T* d_total) {
    typedef TileScan<T, 1024, VPT> TS;
    const int64_t tiles = ceil_div(n, TS::TILE);
    uint32_t epoch;
    int rc = scan_state_acquire(c, tiles, &epoch);
    if (rc) return rc;
    const int per_cu = (sizeof(T) == 4 && VPT <= 4) ? 2 : 1;
    const int64_t grid = std::min<int64_t>(tiles, (int64_t)std::max(c->cus, 1) * per_cu);
    hipLaunchKernelGGL((k_scan_pipe<T, 1024, VPT>), dim3((unsigned)grid), dim3(1024), 0, c->stream, in, out, n,
                       make_state<T>(c, epoch), c->tickets, d_total, c->spin_limit, c->d_err, tiles);
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

template <class T>
int run_unaligned(edpp_ctx* c, const T* in, T* out, int64_t n, T* d_total) {
    int64_t tiles = ceil_div(n, 64 * 16);
    uint32_t epoch;
    int rc = scan_state_acquire(c, tiles, &epoch);
    if (rc) return rc;
    hipLaunchKernelGGL((k_scan_unaligned<T>), dim3((unsigned)tiles), dim3(64), 0, c->stream, in, out, n,
                       make_state<T>(c, epoch), c->tickets, d_total, c->spin_limit, c->d_err, tiles);
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

template <class T, int VPT>
int run_rts_ssa(edpp_ctx* c, const T* in, T* out, int64_t n, T* d_total, bool ssa) {
    typedef TileScan<T, kBlock, VPT> TS;
    int64_t tiles = ceil_div(n, TS::TILE);
    void* w;
    int rc = ws_reserve(c, (size_t)tiles * sizeof(T) + 256, &w);
    if (rc) return rc;
    T* partials = static_cast<T*>(w);
    if (ssa) {
        hipLaunchKernelGGL((k_tile_scan_local<T, kBlock, VPT>), dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, in,
                           out, n, partials);
    } else {
        hipLaunchKernelGGL((k_tile_reduce<T, kBlock, VPT>), dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, in, n,
                           partials);
    }
    hipLaunchKernelGGL((k_scan_partials<T>), dim3(1), dim3(1024), 0, c->stream, partials, tiles, d_total);
    if (ssa) {
        hipLaunchKernelGGL((k_tile_add<T, kBlock, VPT>), dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, out, n,
                           (const T*)partials);
    } else {
        hipLaunchKernelGGL((k_tile_scan_offset<T, kBlock, VPT>), dim3((unsigned)tiles), dim3(kBlock), 0, c->stream,
                           in, out, n, (const T*)partials);
    }
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

}  // namespace

template <class T>
int launch_scan(edpp_ctx* c, const T* in, T* out, int64_t n, int scheme, T* d_total) {
    if (n == 0) {
        if (d_total) return zero_fill(c, d_total, sizeof(T));
        return EDPP_OK;
    }
    if (!aligned16(in) || !aligned16(out)) return run_unaligned<T>(c, in, out, n, d_total);
    if (scheme == EDPP_SCAN_AUTO) {
        scheme = default_scan_scheme(c, n, (int)sizeof(T));
        if (scheme == EDPP_SCAN_LOOKBACK) {
            const edpp_tuning saved = c->tune;
            default_scan_shape(c, n, &c->tune.scan_block, &c->tune.scan_vpt);
            int rc = launch_scan<T>(c, in, out, n, EDPP_SCAN_LOOKBACK, d_total);
            c->tune = saved;
            return rc;
        }
    }
    // look-back tiles of lanes x 4..16 vectors per lane (c->tune); 2048-lane
    // tiles of 8 vectors for the three-kernel schemes
    switch (scheme) {
        case EDPP_SCAN_LOOKBACK: {
            int v = c->tune.scan_vpt;
            const int b = c->tune.scan_block;
            if (v == 0) {  // vectors per lane by size
                int unused_block;
                default_scan_shape(c, n, &unused_block, &v);
            }

            if (b == 512) {
                if (v <= 4) return run_lookback<T, 512, 4>(c, in, out, n, d_total);
                if (v == 8) return run_lookback<T, 512, 8>(c, in, out, n, d_total);
                return run_lookback<T, 512, 16>(c, in, out, n, d_total);
            }
            if (b == 1024) {
                if (v <= 4) return run_lookback<T, 1024, 4>(c, in, out, n, d_total);
                if constexpr (sizeof(T) == 4) {
                    if (v == 16) return run_lookback<T, 1024, 16>(c, in, out, n, d_total);
                }
                return run_lookback<T, 1024, 8>(c, in, out, n, d_total);
            }
            if (v == 2) return run_lookback<T, 256, 2>(c, in, out, n, d_total);
            if (v == 4) return run_lookback<T, 256, 4>(c, in, out, n, d_total);
            if (v == 8) return run_lookback<T, 256, 8>(c, in, out, n, d_total);
            return run_lookback<T, 256, 16>(c, in, out, n, d_total);
        }
        case EDPP_SCAN_PERSIST:  // 1024 lanes, 4 or 8 vectors per lane (8 unless 4 is asked)
            if (c->tune.scan_vpt == 4) return run_pipe<T, 4>(c, in, out, n, d_total);
            return run_pipe<T, 8>(c, in, out, n, d_total);
        case EDPP_SCAN_RTS: return run_rts_ssa<T, 8>(c, in, out, n, d_total, false);
        case EDPP_SCAN_SSA: return run_rts_ssa<T, 8>(c, in, out, n, d_total, true);
        default: return fail(c, EDPP_EINVAL, "unknown scan scheme %d", scheme);
    }
}

// Look-back scan of 32-bit values (the default tile shape, or the scalar
// chunks for unaligned buffers) on state the caller provides: `status` holds
// status_cap words and `tickets` two counters.  Both are cleared here by a
// fill kernel and the epoch is fixed, so nothing is allocated and nothing is
// kept from one call to the next.
int launch_scan_lookback_u32_on(edpp_ctx* c, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* d_total,
                                uint64_t* status, int64_t status_cap, uint32_t* tickets) {
    if (n == 0) return d_total ? zero_fill(c, d_total, sizeof(uint32_t)) : EDPP_OK;
    const bool vec = aligned16(in) && aligned16(out);
    int block = 1024, vpt = 8;
    default_scan_shape(c, n, &block, &vpt);  // 1024 lanes, 4, 8 or 16 vectors per lane
    const int64_t tile = vec ? (int64_t)block * 4 * vpt : 64 * 16;
    const int64_t tiles = ceil_div(n, tile);
    if (tiles > status_cap)
        return fail(c, EDPP_EINVAL, "n=%lld needs %lld look-back words, more than the %lld available", (long long)n,
                    (long long)tiles, (long long)status_cap);
    int rc = zero_fill(c, status, (size_t)tiles * sizeof(uint64_t));
    if (!rc) rc = zero_fill(c, tickets, 2 * sizeof(uint32_t));
    if (rc) return rc;
    LookbackState32 st;
    st.status = status;
    st.epoch = 1;
    if (!vec)
        hipLaunchKernelGGL((k_scan_unaligned<uint32_t>), dim3((unsigned)tiles), dim3(64), 0, c->stream, in, out, n, st,
                           tickets, d_total, c->spin_limit, c->d_err, tiles);
    else if (vpt <= 4)
        hipLaunchKernelGGL((k_scan_lookback<uint32_t, 1024, 4>), dim3((unsigned)tiles), dim3(1024), 0, c->stream, in,
                           out, n, st, tickets, d_total, c->spin_limit, c->d_err, tiles);
    else if (vpt == 8)
        hipLaunchKernelGGL((k_scan_lookback<uint32_t, 1024, 8>), dim3((unsigned)tiles), dim3(1024), 0, c->stream, in,
                           out, n, st, tickets, d_total, c->spin_limit, c->d_err, tiles);
    else
        hipLaunchKernelGGL((k_scan_lookback<uint32_t, 1024, 16>), dim3((unsigned)tiles), dim3(1024), 0, c->stream, in,
                           out, n, st, tickets, d_total, c->spin_limit, c->d_err, tiles);
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

template int launch_scan<uint32_t>(edpp_ctx*, const uint32_t*, uint32_t*, int64_t, int, uint32_t*);
template int launch_scan<uint64_t>(edpp_ctx*, const uint64_t*, uint64_t*, int64_t, int, uint64_t*);

}  // namespace edpp
