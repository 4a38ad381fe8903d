// scan_local.hip — single-workgroup tile scans ("matrix scans") and the
// intra-workgroup networks that scan the per-lane partials.
//
// Reference: opencl/kernels/scan_local_kernel.cl:478-695 (scan_local and
// matrix_scan_lm / _reg / _lm_reg / _lm_serial), networks :42-474 (serial,
// serial on 128 lanes, Blelloch, conflict-free Blelloch, Sklansky, warp scan,
// and the warp / Sklansky / Blelloch networks run on a global-memory buffer),
// driver opencl/test/test_scan_local.cpp:6-106
// (one workgroup of 512 lanes scanning 512*TILE ints, exclusive).  Each lane
// owns a row of `tile` consecutive elements: reduce the row, scan the row
// sums across the workgroup with the chosen network, rescan the row.  The
// warp network of the reference assumes 32-lane lockstep without barriers;
// the wave network here uses DPP on 64-lane waves and one LDS exchange.
#include "ctx.hpp"
#include "device.hpp"

namespace edpp {

constexpr int kMaxRow = 64;    // largest tile (elements per lane)
constexpr int kMaxLocal = 1024;

// Bank-conflict-free index of the padded Blelloch network: one pad word per
// 32 entries, so the power-of-two strides of the sweeps spread over the LDS
// banks (the reference's CON_OFFSET, scan_local_kernel.cl:140-178).
__device__ __forceinline__ int cf_index(int i) { return i + (i >> 5); }

// Work-efficient up/down sweep over x[0..L) addressed through idx(), in LDS
// or global memory; the total goes to *tot.
template <class Idx>
__device__ __forceinline__ void blelloch_sweeps(uint32_t* x, int L, Idx idx, uint32_t* tot) {
    const int t = threadIdx.x;
    int stride = 1;
    for (; stride < L; stride <<= 1) {
        int i = (t + 1) * stride * 2 - 1;
        if (i < L) x[idx(i)] += x[idx(i - stride)];
        __syncthreads();
    }
    if (t == 0) {
        *tot = x[idx(L - 1)];
        x[idx(L - 1)] = 0;
    }
    __syncthreads();
    for (stride = L >> 1; stride >= 1; stride >>= 1) {
        int i = (t + 1) * stride * 2 - 1;
        if (i < L) {
            uint32_t v = x[idx(i - stride)];
            x[idx(i - stride)] = x[idx(i)];
            x[idx(i)] += v;
        }
        __syncthreads();
    }
}

// Exclusive Sklansky network over x[0..L) (LDS or global memory); returns
// the total in every lane.
__device__ __forceinline__ uint32_t sklansky_network(uint32_t* x, int L) {
    const int t = threadIdx.x;
    uint32_t own = t < L ? x[t] : 0u;
    for (int d = 1; d < L; d <<= 1) {
        uint32_t add = 0;
        if (t < L && (t & d)) add = x[(t & ~(d - 1)) - 1];  // last element of the left block
        __syncthreads();
        if (t < L && (t & d)) x[t] += add;
        __syncthreads();
    }
    const uint32_t total = x[L - 1];
    __syncthreads();
    if (t < L) x[t] -= own;
    __syncthreads();
    return total;
}

// exclusive scan of part[0..L) in LDS, in place; returns the total (all lanes).
// For the conflict-free Blelloch network part has L + L/32 words (its padded
// copy); pad: 128 LDS words for the multi-lane serial network; gbuf: L words
// of global memory for the global-memory networks (the call's output,
// overwritten later).
__device__ uint32_t scan_partials_network(uint32_t* part, int L, int network, uint32_t* wtmp, uint32_t* pad,
                                          uint32_t* gbuf) {
    const int t = threadIdx.x;
    uint32_t total = 0;
    if (network >= EDPP_NET_GLOBAL_BLELLOCH) {
        // the network runs on a global-memory copy of the partials, as the
        // reference's global_* networks do; __syncthreads orders the
        // workgroup's global accesses between steps
        if (t < L) gbuf[t] = part[t];
        __syncthreads();
        if (network == EDPP_NET_GLOBAL_BLELLOCH) {
            blelloch_sweeps(gbuf, L, [](int i) { return i; }, wtmp);
            total = wtmp[0];
        } else if (network == EDPP_NET_GLOBAL_SKLANSKY) {
            total = sklansky_network(gbuf, L);
        } else {  // EDPP_NET_GLOBAL_WAVE: DPP wave scans, wave totals through global memory
            const int w = t / kWave, lane = lane_id();
            const uint32_t x = t < L ? gbuf[t] : 0u;
            const uint32_t inc = wave_inclusive_add(x);
            const int nw = (L + kWave - 1) / kWave;
            __syncthreads();  // every lane read its element before wave totals overwrite the buffer head
            if (lane == kWave - 1 && w < nw) gbuf[w] = inc;
            __syncthreads();
            uint32_t before = 0, all = 0;
            for (int q = 0; q < nw; ++q) {
                const uint32_t y = gbuf[q];
                before += q < w ? y : 0u;
                all += y;
            }
            __syncthreads();
            if (t < L) gbuf[t] = before + inc - x;
            __syncthreads();
            total = all;
        }
        if (t < L) part[t] = gbuf[t];
        __syncthreads();
        return total;
    }
    if (network == EDPP_NET_SERIAL) {
        if (t == 0) {
            uint32_t acc = 0;
            for (int i = 0; i < L; ++i) {
                uint32_t x = part[i];
                part[i] = acc;
                acc += x;
            }
            wtmp[0] = acc;
        }
        __syncthreads();
        total = wtmp[0];
    } else if (network == EDPP_NET_SERIAL_MT) {
        // the serial scan spread over T = min(128, L) lanes: each reduces a
        // run of L/T partials, lane 0 scans the T sums, each rescans its run
        // (local_serial_scan_multithread, scan_local_kernel.cl:57-100)
        const int T = L < 128 ? L : 128, E = L / T;
        uint32_t* sums = pad;  // [T]
        if (t < T) {
            uint32_t a = 0;
            for (int i = 0; i < E; ++i) a += part[t * E + i];
            sums[t] = a;
        }
        __syncthreads();
        if (t == 0) {
            uint32_t acc = 0;
            for (int i = 0; i < T; ++i) {
                const uint32_t x = sums[i];
                sums[i] = acc;
                acc += x;
            }
            wtmp[0] = acc;
        }
        __syncthreads();
        if (t < T) {
            uint32_t acc = sums[t];
            for (int i = 0; i < E; ++i) {
                const uint32_t x = part[t * E + i];
                part[t * E + i] = acc;
                acc += x;
            }
        }
        __syncthreads();
        total = wtmp[0];
    } else if (network == EDPP_NET_BLELLOCH) {
        blelloch_sweeps(part, L, [](int i) { return i; }, wtmp);
        total = wtmp[0];
    } else if (network == EDPP_NET_BLELLOCH_CF) {
        // spread the partials to their padded slots in place (every lane holds
        // its value across the barrier), sweep, and gather them back
        uint32_t x = t < L ? part[t] : 0u;
        __syncthreads();
        if (t < L) part[cf_index(t)] = x;
        __syncthreads();
        blelloch_sweeps(part, L, [](int i) { return cf_index(i); }, wtmp);
        x = t < L ? part[cf_index(t)] : 0u;
        __syncthreads();
        if (t < L) part[t] = x;
        __syncthreads();
        total = wtmp[0];
    } else if (network == EDPP_NET_SKLANSKY) {
        total = sklansky_network(part, L);
    } else {  // EDPP_NET_WAVE
        const int w = t / kWave, lane = lane_id();
        uint32_t x = t < L ? part[t] : 0u;
        uint32_t inc = wave_inclusive_add(x);
        if (lane == kWave - 1) wtmp[w] = inc;
        __syncthreads();
        uint32_t before = 0, all = 0;
        const int nw = (L + kWave - 1) / kWave;
        for (int q = 0; q < nw; ++q) {
            uint32_t y = wtmp[q];
            before += q < w ? y : 0u;
            all += y;
        }
        if (t < L) part[t] = before + inc - x;
        __syncthreads();
        total = all;
    }
    return total;
}

// MATRIX: EDPP_MATRIX_* ; the whole tile is n = L*tile elements
// Coalesced copy of n = L*tile words between global memory and LDS: 16-byte
// vectors (at most kMaxRow/4 per lane, four loads in flight before their
// stores) when both sides are 16-byte aligned, else words.
__device__ __forceinline__ void copy_tile(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, int n,
                                          bool vec) {
    const int t = threadIdx.x, L = blockDim.x;
    if (vec) {
        const int nv = n / 4;  // L is a power of two >= 64, so 4 divides n
        const uint4* s4 = reinterpret_cast<const uint4*>(src);
        uint4* d4 = reinterpret_cast<uint4*>(dst);
        // four rounds of four vectors: 16 registers of loads in flight
        #pragma unroll
        for (int r = 0; r < kMaxRow / 16; ++r) {
            uint4 v[4];
            #pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = t + (4 * r + q) * L;
                v[q] = i < nv ? s4[i] : make_uint4(0u, 0u, 0u, 0u);
            }
            #pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = t + (4 * r + q) * L;
                if (i < nv) d4[i] = v[q];
            }
        }
    } else {
        for (int i = t; i < n; i += L) dst[i] = src[i];
    }
}

// Row-wise copy for a padded LDS layout (row stride S = tile + 1, odd, so the
// lanes' rows fall in different banks): wave w moves rows w, w+W, ..., lane l
// element l of the row (tile <= 64).  to_lds: dst[r*S + l] = src[r*tile + l];
// else dst[r*tile + l] = src[r*S + l].
__device__ __forceinline__ void copy_rows(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, int tile,
                                          int S, bool to_lds) {
    const int W = blockDim.x / kWave, w = threadIdx.x / kWave, l = lane_id();
    const int L = blockDim.x;
    if (l >= tile) return;
    const int ss = to_lds ? tile : S, ds = to_lds ? S : tile;
    for (int r = w; r < L; r += W) dst[r * ds + l] = src[r * ss + l];
}

template <int MATRIX>
__global__ __launch_bounds__(kMaxLocal) void k_matrix_scan(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                           int tile, int network, bool vec, int S, int part_words,
                                                           int pad_words) {
    extern __shared__ __align__(16) uint32_t lds[];
    const int L = blockDim.x, t = threadIdx.x;
    const int n = L * tile;
    uint32_t* part = lds;                // [part_words]: L, or L + L/32 for the padded network
    uint32_t* wtmp = part + part_words;  // [16]
    uint32_t* pad = wtmp + 16;           // [pad_words]: network scratch (multiple of 4 words)
    uint32_t* data = pad + pad_words;  // [n] for the LDS-staged variants (16-byte aligned: L >= 64)
    if (MATRIX == EDPP_MATRIX_LM_SERIAL) {
        if (t == 0) {
            uint32_t acc = 0;
            for (int i = 0; i < n; ++i) {
                uint32_t x = in[i];
                out[i] = acc;
                acc += x;
            }
        }
        return;
    }
    uint32_t reg[kMaxRow];
    uint32_t s = 0;
    if (MATRIX == EDPP_MATRIX_REG) {
        // row straight from global memory (lane-strided, uncoalesced)
        #pragma unroll
        for (int k = 0; k < kMaxRow; ++k)
            if (k < tile) { reg[k] = in[t * tile + k]; s += reg[k]; }
    } else {
        // coalesced copy into LDS (rows of stride S), then rows out of LDS
        if (S == tile) copy_tile(in, data, n, vec);
        else copy_rows(in, data, tile, S, true);
        __syncthreads();
        #pragma unroll
        for (int k = 0; k < kMaxRow; ++k)
            if (k < tile) { reg[k] = data[t * S + k]; s += reg[k]; }
    }
    part[t] = s;
    __syncthreads();
    scan_partials_network(part, L, network, wtmp, pad, out);
    uint32_t acc = part[t];
    if (MATRIX == EDPP_MATRIX_LM_REG) {
        __syncthreads();
        #pragma unroll
        for (int k = 0; k < kMaxRow; ++k)
            if (k < tile) { data[t * S + k] = acc; acc += reg[k]; }
        __syncthreads();
        if (S == tile) copy_tile(data, out, n, vec);
        else copy_rows(data, out, tile, S, false);
    } else {
        // rows written from registers (the reference's LM and REG final scans)
        #pragma unroll
        for (int k = 0; k < kMaxRow; ++k)
            if (k < tile) { out[t * tile + k] = acc; acc += reg[k]; }
    }
}

int launch_scan_local(edpp_ctx* c, const int32_t* in, int32_t* out, int tile, int local_size, int matrix_type,
                      int network) {
    if (tile < 1 || tile > kMaxRow)
        return fail(c, EDPP_EINVAL, "tile_size %d outside [1,%d]", tile, kMaxRow);
    if (local_size < kWave || local_size > kMaxLocal || (local_size & (local_size - 1)))
        return fail(c, EDPP_EINVAL, "local_size %d must be a power of two in [64,1024]", local_size);
    if (network < EDPP_NET_SERIAL || network > EDPP_NET_GLOBAL_WAVE)
        return fail(c, EDPP_EINVAL, "unknown network %d", network);
    const size_t n = (size_t)local_size * tile;
    // network scratch: the padded partials of the conflict-free Blelloch
    // network, the 128 run sums of the multi-lane serial network
    int part_words = local_size;
    if (network == EDPP_NET_BLELLOCH_CF) part_words = (local_size + local_size / 32 + 3) & ~3;
    const int pad_words = network == EDPP_NET_SERIAL_MT ? 128 : 0;
    size_t lds = (size_t)(part_words + 16 + pad_words) * 4;
    // LDS rows of an even tile are padded to an odd stride where that still fits
    int S = tile;
    if (matrix_type == EDPP_MATRIX_LM || matrix_type == EDPP_MATRIX_LM_REG) {
        if ((tile & 1) == 0 && lds + (size_t)local_size * (tile + 1) * 4 <= 160 * 1024) S = tile + 1;
        lds += (size_t)local_size * S * 4;
    }
    if (lds > 160 * 1024) return fail(c, EDPP_EINVAL, "tile of %zu elements does not fit in LDS", n);
    const uint32_t* a = (const uint32_t*)in;
    uint32_t* b = (uint32_t*)out;
    const bool vec = aligned16(a) && aligned16(b);
    switch (matrix_type) {
        case EDPP_MATRIX_LM:
            EDPP_HIP(c, hipFuncSetAttribute((const void*)k_matrix_scan<EDPP_MATRIX_LM>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL(k_matrix_scan<EDPP_MATRIX_LM>, dim3(1), dim3(local_size), lds, c->stream, a, b, tile,
                               network, vec, S, part_words, pad_words);
            break;
        case EDPP_MATRIX_REG:
            hipLaunchKernelGGL(k_matrix_scan<EDPP_MATRIX_REG>, dim3(1), dim3(local_size), lds, c->stream, a, b, tile,
                               network, vec, S, part_words, pad_words);
            break;
        case EDPP_MATRIX_LM_REG:
            EDPP_HIP(c, hipFuncSetAttribute((const void*)k_matrix_scan<EDPP_MATRIX_LM_REG>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL(k_matrix_scan<EDPP_MATRIX_LM_REG>, dim3(1), dim3(local_size), lds, c->stream, a, b,
                               tile, network, vec, S, part_words, pad_words);
            break;
        case EDPP_MATRIX_LM_SERIAL:
            hipLaunchKernelGGL(k_matrix_scan<EDPP_MATRIX_LM_SERIAL>, dim3(1), dim3(local_size), lds, c->stream, a, b,
                               tile, network, vec, S, part_words, pad_words);
            break;
        default:
            return fail(c, EDPP_EINVAL, "unknown matrix scan type %d", matrix_type);
    }
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

}  // namespace edpp
