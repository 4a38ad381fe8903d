// gather.hip — multi-pass gather and scatter.
//
// Reference: opencl/primitives/gatherImpl.cpp:7-55, scatterImpl.cpp:9-54
// (host pass loop), opencl/kernels/gather_kernel.cl:6-19,
// scatter_kernel.cl:6-21 (kernels).  Semantics kept: pass k of P handles the
// indices inside [k*ceil(n/P), (k+1)*ceil(n/P)), so the random-side working
// set of one pass is n/P elements and can stay in L2 / Infinity Cache while
// the sequential index stream is re-read every pass.
//
// MI355X layout: each lane owns U groups of 4 consecutive positions; the
// index (and, for scatter, value) groups are read as 16-byte vectors with
// non-temporal hints so the streaming side does not evict the random side.
#include <algorithm>
#include "ctx.hpp"
#include "device.hpp"

namespace edpp {

constexpr int kGsBlock = 256;

template <class T>
__device__ __forceinline__ void load4(const T* p, T (&v)[4]) {
    if (sizeof(T) == 4) {
        v4u32 x = ld_stream(reinterpret_cast<const v4u32*>(p));
        #pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = (T)x[k];
    } else {
        v2u64 a = ld_stream(reinterpret_cast<const v2u64*>(p));
        v2u64 b = ld_stream(reinterpret_cast<const v2u64*>(p) + 1);
        v[0] = (T)a[0]; v[1] = (T)a[1]; v[2] = (T)b[0]; v[3] = (T)b[1];
    }
}
template <class T>
__device__ __forceinline__ void store4(T* p, const T (&v)[4]) {
    if (sizeof(T) == 4) {
        v4u32 x;
        #pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = (uint32_t)v[k];
        st_stream(x, reinterpret_cast<v4u32*>(p));
    } else {
        v2u64 a, b;
        a[0] = v[0]; a[1] = v[1]; b[0] = v[2]; b[1] = v[3];
        st_stream(a, reinterpret_cast<v2u64*>(p));
        st_stream(b, reinterpret_cast<v2u64*>(p) + 1);
    }
}

// Window staging.  A workgroup's chunk is 1024*U consecutive positions.  When
// the in-slice indices of the chunk fall inside a window of kGsWin elements
// (a block-local permutation: shuffled 4096-element blocks, shuffled inside),
// the random side goes through LDS: a gather loads the window with
// consecutive loads and reads its values from LDS; a scatter writes its values
// into the window and stores the covered positions as consecutive runs, so
// the lines of the block leave whole instead of as 4-byte pieces.  Otherwise
// the random side is addressed directly.
//
// With several passes the window test runs in the first pass over all of
// the chunk's indices in [0, n); a chunk that fits is done there at once
// (each position is still written exactly once) and marked in done[], and the
// later passes skip it without reading its indices.  A block-local
// permutation so costs one pass whatever the pass count.
constexpr int kGsWin = 4096;
enum GsPhase { kGsOnly = 0, kGsFirst = 1, kGsLater = 2 };

// Smallest and largest index in [0, nall) of the workgroup's chunk (lo > hi
// when there is none), the same in every lane.  Ends with a barrier.
template <int U>
__device__ __forceinline__ void chunk_window(const uint32_t (&pos)[U][4], uint32_t nall, uint32_t* wlo, uint32_t* whi,
                                             uint32_t& lo, uint32_t& hi) {
    lo = 0xFFFFFFFFu;
    hi = 0;
    #pragma unroll
    for (int u = 0; u < U; ++u)
        #pragma unroll
        for (int k = 0; k < 4; ++k)
            if (pos[u][k] < nall) {
                lo = std::min(lo, pos[u][k]);
                hi = std::max(hi, pos[u][k]);
            }
    #pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        lo = std::min(lo, (uint32_t)__shfl_xor((int)lo, o));
        hi = std::max(hi, (uint32_t)__shfl_xor((int)hi, o));
    }
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) {
        wlo[w] = lo;
        whi[w] = hi;
    }
    __syncthreads();
    #pragma unroll
    for (int q = 0; q < kGsBlock / kWave; ++q) {
        lo = std::min(lo, wlo[q]);
        hi = std::max(hi, whi[q]);
    }
}

// out[i] = in[idx[i]] for idx[i] in [from, from+span).  The host clamps
// from+span to n, so an index outside [0,n) is skipped instead of read.
// Chunks of the first pass (kGsBlock*UF index vectors each) that its window
// path finished, from done[]: skip[u] for the workgroup's u-th slice of
// kGsBlock index vectors, the same in the whole workgroup.  Returns whether
// every slice is skipped.  A later pass may run with more vectors per lane
// (U > UF) than the first.
// ufs = log2 of the first pass's vectors per lane, which divides U; nchunk =
// the first pass's chunks (flags).  Shifts only: the test runs in every
// workgroup of every later pass.
template <int U>
__device__ __forceinline__ bool later_skips(const uint8_t* __restrict__ done, int64_t ngrp, int ufs, uint32_t nchunk,
                                            bool (&skip)[U]) {
    // the workgroup's slices belong to the U >> ufs consecutive first-pass
    // chunks from blockIdx.x * (U >> ufs), one flag each
    const int per = U >> ufs;
    const uint32_t c0 = blockIdx.x * (uint32_t)per;
    uint32_t mask = 0;
    for (int f = 0; f < per; ++f)
        if (c0 + f < nchunk && done[c0 + f]) mask |= 1u << f;
    bool all = true;
    #pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t slice = (int64_t)blockIdx.x * U + u;  // its first vector: slice * kGsBlock
        skip[u] = slice * kGsBlock >= ngrp || ((mask >> (u >> ufs)) & 1u) != 0;
        all = all && skip[u];
    }
    return all;
}

template <class T, int U, bool RNT>
__global__ __launch_bounds__(kGsBlock) void k_gather(const T* __restrict__ in, T* __restrict__ out,
                                                     const int32_t* __restrict__ idx, int64_t ngrp, uint32_t from,
                                                     uint32_t span, uint32_t nall, uint8_t* __restrict__ done,
                                                     int phase, int ufs, uint32_t nchunk) {
    __shared__ T win[kGsWin];
    __shared__ uint32_t wlo[kGsBlock / kWave], whi[kGsBlock / kWave];
    // a gather's later passes use the first pass's vectors per lane: one flag
    // per workgroup
    (void)ufs;
    (void)nchunk;
    if (phase == kGsLater && done[blockIdx.x]) return;  // the same in the whole workgroup
    const int64_t g0 = (int64_t)blockIdx.x * kGsBlock * U + threadIdx.x;
    uint32_t pos[U][4];
    #pragma unroll
    for (int u = 0; u < U; ++u) {
        int64_t g = g0 + (int64_t)u * kGsBlock;
        v4u32 x = {nall, nall, nall, nall};  // out of range
        if (g < ngrp) x = ld_stream(reinterpret_cast<const v4u32*>(idx) + g);
        #pragma unroll
        for (int k = 0; k < 4; ++k) pos[u][k] = x[k];
    }
    T val[U][4];
    bool window = false;
    if (phase != kGsLater) {
        uint32_t lo, hi;
        chunk_window<U>(pos, nall, wlo, whi, lo, hi);
        window = lo <= hi && hi - lo < (uint32_t)kGsWin;  // the same in the whole workgroup
        if (phase == kGsFirst && threadIdx.x == 0) done[blockIdx.x] = window ? 1 : 0;
        if (window) {
            for (uint32_t o = threadIdx.x; o <= hi - lo; o += kGsBlock) win[o] = in[lo + o];
            __syncthreads();
            #pragma unroll
            for (int u = 0; u < U; ++u)
                #pragma unroll
                for (int k = 0; k < 4; ++k) val[u][k] = pos[u][k] < nall ? win[pos[u][k] - lo] : T(0);
            // every index of the chunk is done: the slice of this pass is all of [0, n)
            from = 0;
            span = nall;
        }
    }
    if (!window) {
        #pragma unroll
        for (int u = 0; u < U; ++u)
            #pragma unroll
            for (int k = 0; k < 4; ++k)
                val[u][k] = (pos[u][k] - from < span) ? (RNT ? ld_stream(in + pos[u][k]) : in[pos[u][k]]) : T(0);
    }
    #pragma unroll
    for (int u = 0; u < U; ++u) {
        int64_t g = g0 + (int64_t)u * kGsBlock;
        if (g < ngrp) {
            bool all = true;
            #pragma unroll
            for (int k = 0; k < 4; ++k) all = all && (pos[u][k] - from < span);
            if (all) {
                store4(out + 4 * g, val[u]);
            } else {
                #pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (pos[u][k] - from < span) out[4 * g + k] = val[u][k];
            }
        }
    }
}

// out[idx[i]] = in[i] for idx[i] in [from, from+span)
template <class T, int U, bool RNT>
__global__ __launch_bounds__(kGsBlock) void k_scatter(const T* __restrict__ in, T* __restrict__ out,
                                                      const int32_t* __restrict__ idx, int64_t ngrp, uint32_t from,
                                                      uint32_t span, uint32_t nall, uint8_t* __restrict__ done,
                                                      int phase, int ufs, uint32_t nchunk) {
    __shared__ T win[kGsWin];
    __shared__ uint32_t cov[kGsWin / 32];  // positions of the window this chunk writes
    __shared__ uint32_t wlo[kGsBlock / kWave], whi[kGsBlock / kWave];
    bool skip[U] = {};
    if (phase == kGsLater && later_skips<U>(done, ngrp, ufs, nchunk, skip)) return;
    for (int i = threadIdx.x; i < kGsWin / 32; i += kGsBlock) cov[i] = 0;
    const int64_t g0 = (int64_t)blockIdx.x * kGsBlock * U + threadIdx.x;
    uint32_t pos[U][4];
    T val[U][4];
    #pragma unroll
    for (int u = 0; u < U; ++u) {
        int64_t g = g0 + (int64_t)u * kGsBlock;
        if (g < ngrp && !skip[u]) {
            v4u32 x = ld_stream(reinterpret_cast<const v4u32*>(idx) + g);
            #pragma unroll
            for (int k = 0; k < 4; ++k) pos[u][k] = x[k];
            load4(in + 4 * g, val[u]);
        } else {
            #pragma unroll
            for (int k = 0; k < 4; ++k) { pos[u][k] = nall; val[u][k] = 0; }
        }
    }
    bool window = false;
    if (phase != kGsLater) {
        uint32_t lo, hi;
        chunk_window<U>(pos, nall, wlo, whi, lo, hi);  // its barrier also orders the cov clearing
        window = lo <= hi && hi - lo < (uint32_t)kGsWin;  // the same in the whole workgroup
        if (phase == kGsFirst && threadIdx.x == 0) done[blockIdx.x] = window ? 1 : 0;
        if (window) {
            #pragma unroll
            for (int u = 0; u < U; ++u)
                #pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (pos[u][k] < nall) {
                        const uint32_t o = pos[u][k] - lo;
                        win[o] = val[u][k];
                        atomicOr(cov + (o >> 5), 1u << (o & 31));
                    }
            __syncthreads();
            for (uint32_t o = threadIdx.x; o <= hi - lo; o += kGsBlock)
                if ((cov[o >> 5] >> (o & 31)) & 1u) out[lo + o] = win[o];
        }
    }
    if (!window) {
        #pragma unroll
        for (int u = 0; u < U; ++u)
            #pragma unroll
            for (int k = 0; k < 4; ++k)
                if (pos[u][k] - from < span) {
                    if (RNT) st_stream(val[u][k], out + pos[u][k]);
                    else out[pos[u][k]] = val[u][k];
                }
    }
}

// elements [lo, n): scalar tail and the path for unaligned pointers
template <class T>
__global__ void k_gs_scalar(const T* in, T* out, const int32_t* idx, int64_t lo, int64_t n, uint32_t from,
                            uint32_t span, bool scatter) {
    for (int64_t i = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t p = (uint32_t)idx[i];
        if (p - from < span) {
            if (scatter) out[p] = in[i];
            else out[i] = in[p];
        }
    }
}

struct GsPass {
    uint32_t from, span, nall;
    uint8_t* done;  // one flag per workgroup chunk, written by the first of several passes
    int phase;
    int ufs;          // log2 of the first pass's index vectors per lane (one done flag per kGsBlock << ufs)
    uint32_t nchunk;  // done flags (workgroups of the first pass)
};

template <class T, int U, bool RNT>
static void launch_gs_kernel(edpp_ctx* c, const T* in, T* out, const int32_t* idx, int64_t ngrp, const GsPass& q,
                             bool scatter) {
    const unsigned grid = (unsigned)ceil_div(ngrp, (int64_t)kGsBlock * U);
    if (scatter)
        hipLaunchKernelGGL((k_scatter<T, U, RNT>), dim3(grid), dim3(kGsBlock), 0, c->stream, in, out, idx, ngrp,
                           q.from, q.span, q.nall, q.done, q.phase, q.ufs, q.nchunk);
    else
        hipLaunchKernelGGL((k_gather<T, U, RNT>), dim3(grid), dim3(kGsBlock), 0, c->stream, in, out, idx, ngrp,
                           q.from, q.span, q.nall, q.done, q.phase, q.ufs, q.nchunk);
}

template <class T, bool RNT>
static void dispatch_unroll(edpp_ctx* c, int u, const T* in, T* out, const int32_t* idx, int64_t ngrp,
                            const GsPass& q, bool scatter) {
    switch (u) {
        case 1: launch_gs_kernel<T, 1, RNT>(c, in, out, idx, ngrp, q, scatter); break;
        case 2: launch_gs_kernel<T, 2, RNT>(c, in, out, idx, ngrp, q, scatter); break;
        case 8: launch_gs_kernel<T, 8, RNT>(c, in, out, idx, ngrp, q, scatter); break;
        default: launch_gs_kernel<T, 4, RNT>(c, in, out, idx, ngrp, q, scatter); break;
    }
}

template <class T>
int launch_gather(edpp_ctx* c, const T* in, T* out, const int32_t* idx, int64_t n, int passes, bool scatter) {
    if (n == 0) return EDPP_OK;
    if (passes == 0) passes = default_passes(c, n, (int)sizeof(T), scatter);
    if (passes < 1 || passes > n) passes = passes < 1 ? 1 : (int)n;
    const bool vec = aligned16(in) && aligned16(out) && aligned16(idx);
    const int64_t ngrp = vec ? n / 4 : 0;
    const int64_t len = ceil_div(n, passes);  // gatherImpl.cpp:37
    uint8_t* done = nullptr;
    if (ngrp && passes > 1) {
        const int64_t grid = ceil_div(ngrp, (int64_t)kGsBlock * c->tune.gs_unroll);
        void* w;
        int rc = ws_reserve(c, (size_t)grid, &w);
        if (rc) return rc;
        done = static_cast<uint8_t*>(w);
    }
    for (int p = 0; p < passes; ++p) {
        const int64_t lo = (int64_t)p * len;
        if (lo >= n) break;
        GsPass q;
        q.from = (uint32_t)lo;
        q.span = (uint32_t)(lo + len <= n ? len : n - lo);
        q.nall = (uint32_t)n;
        q.done = done;
        q.phase = passes == 1 ? kGsOnly : (p == 0 ? kGsFirst : kGsLater);
        const int uf = c->tune.gs_unroll;  // 1, 2, 4 or 8
        q.ufs = uf == 1 ? 0 : uf == 2 ? 1 : uf == 4 ? 2 : 3;
        q.nchunk = (uint32_t)ceil_div(ngrp, (int64_t)kGsBlock * uf);
        const uint32_t from = q.from, span = q.span;
        // scatter passes after the first skip the window test, so for 4-byte
        // values they take 8 index vectors per lane: half the workgroups
        // re-reading the index and values (2^28, random, 4 passes: 4.51
        // against 4.95 ms; gathers measured no steady gain)
        const int u = (q.phase == kGsLater && scatter && sizeof(T) == 4) ? std::max(8, uf) : uf;
        if (ngrp) {
            if (c->tune.gs_random_nt > 0)
                dispatch_unroll<T, true>(c, u, in, out, idx, ngrp, q, scatter);
            else
                dispatch_unroll<T, false>(c, u, in, out, idx, ngrp, q, scatter);
        }
        if (ngrp * 4 < n) {
            int64_t rest = n - ngrp * 4;
            unsigned g = (unsigned)(rest < 256 * 4096 ? ceil_div(rest, 256) : 4096);
            hipLaunchKernelGGL((k_gs_scalar<T>), dim3(g), dim3(256), 0, c->stream, in, out, idx, ngrp * 4, n, from,
                               span, scatter);
        }
    }
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

template int launch_gather<uint32_t>(edpp_ctx*, const uint32_t*, uint32_t*, const int32_t*, int64_t, int, bool);
template int launch_gather<uint64_t>(edpp_ctx*, const uint64_t*, uint64_t*, const int32_t*, int64_t, int, bool);

}  // namespace edpp
