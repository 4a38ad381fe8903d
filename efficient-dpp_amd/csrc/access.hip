// access.hip — memory access-pattern and workgroup-order probes.
//
// Reference: opencl/kernels/mem_kernel.cl:33-86 (scale_column, scale_row,
// scale_mixed, wg_access, cache_heat) driven by opencl/test/test_access.cpp
// and opencl/test/test_access_wg.cpp.  They measure how the lane-to-address
// mapping and the order in which workgroups take their chunks change the
// achieved bandwidth; here they run with 64-lane waves, and the chunk probe
// returns a deterministic result (the count of odd elements) instead of the
// reference's racy single output word.
#include <algorithm>
#include "ctx.hpp"
#include "device.hpp"

namespace edpp {

namespace {

constexpr int kAccessBlock = 512;  // the reference's local size (test_access.cpp: localSize = 512)

// lanes = n / repeat; lane t of the launch
//   COLUMN: elements t + k*lanes          (each step of the wave is one coalesced line run)
//   ROW:    elements t*repeat + k         (each lane walks its own run: strided across the wave)
//   MIXED:  wave w, lane l: w*64*repeat + k*64 + l  (each wave walks its own coalesced run)
template <int PATTERN>
__global__ __launch_bounds__(kAccessBlock) void k_access_scale(const int32_t* __restrict__ in,
                                                               int32_t* __restrict__ out, int64_t lanes, int repeat) {
    const int64_t t = (int64_t)blockIdx.x * kAccessBlock + threadIdx.x;
    if (t >= lanes) return;
    int64_t idx, step;
    if (PATTERN == EDPP_ACCESS_COLUMN) {
        idx = t;
        step = lanes;
    } else if (PATTERN == EDPP_ACCESS_ROW) {
        idx = t * repeat;
        step = 1;
    } else {
        idx = (t / kWave) * kWave * repeat + (t % kWave);
        step = kWave;
    }
    for (int k = 0; k < repeat; ++k, idx += step) out[idx] = in[idx] * 3;
}

__global__ void k_access_zero(unsigned long long* p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0ull;
}

// Workgroup ticket t takes chunk order[t] of `chunk` elements and counts its
// odd elements; the counts add into *count.  Tickets come from a device
// counter, so chunks are taken in ticket order as workgroups start.
__global__ __launch_bounds__(256) void k_access_chunks(const int32_t* __restrict__ in, int64_t n, int64_t chunk,
                                                       const int32_t* __restrict__ order, int64_t chunks,
                                                       uint32_t* ticket, unsigned long long* count) {
    __shared__ int64_t s_chunk;
    __shared__ uint32_t s_part[256 / kWave];
    if (threadIdx.x == 0) {
        const uint32_t t = atomicAdd(ticket, 1u);
        s_chunk = (int64_t)t < chunks ? order[t] : -1;
    }
    __syncthreads();
    const int64_t c = s_chunk;
    uint32_t odd = 0;
    if (c >= 0 && c < chunks) {
        const int64_t b = c * chunk, e = std::min<int64_t>(b + chunk, n);
        for (int64_t i = b + threadIdx.x; i < e; i += 256) odd += (uint32_t)(in[i] & 1);
    }
    odd = wave_sum(odd);
    if (lane_id() == 0) s_part[threadIdx.x / kWave] = odd;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int q = 0; q < 256 / kWave; ++q) s += s_part[q];
        if (s) atomicAdd(count, s);
    }
}

__global__ void k_access_reset_ticket(uint32_t* ticket) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *ticket = 0u;
}

// MurmurHash3's 32-bit finaliser (Appleby): a bijection of 32-bit words
// whose output bits all depend on every input bit.
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

// Random-access probe: access i touches table[fmix32(i + seed) & mask].
// Workgroup g takes accesses [g*S, g*S + S) of each round of gridDim.x*S
// (S = 256*U), lane t the accesses g*S + t + u*256, u < U: every lane keeps U
// independent accesses in flight and the positions come from registers only.
// Reads add the values up (64-bit, per wave, then one atomic); writes store
// the position itself, so colliding accesses write the same word.
constexpr int kRandomBlock = 256;
template <int U, bool WRITE>
__global__ __launch_bounds__(kRandomBlock) void k_access_random(int32_t* __restrict__ table, uint32_t mask,
                                                                int64_t accesses, uint32_t seed,
                                                                unsigned long long* __restrict__ sum) {
    constexpr int64_t S = (int64_t)kRandomBlock * U;
    const int64_t stride = (int64_t)gridDim.x * S;
    uint64_t acc = 0;
    for (int64_t base = (int64_t)blockIdx.x * S + threadIdx.x; base < accesses; base += stride) {
        uint32_t p[U];
        #pragma unroll
        for (int u = 0; u < U; ++u) p[u] = fmix32((uint32_t)(base + (int64_t)u * kRandomBlock) + seed) & mask;
        if (WRITE) {
            #pragma unroll
            for (int u = 0; u < U; ++u)
                if (base + (int64_t)u * kRandomBlock < accesses) table[p[u]] = (int32_t)p[u];
        } else {
            uint32_t v[U];
            #pragma unroll
            for (int u = 0; u < U; ++u) v[u] = base + (int64_t)u * kRandomBlock < accesses ? (uint32_t)table[p[u]] : 0u;
            #pragma unroll
            for (int u = 0; u < U; ++u) acc += v[u];
        }
    }
    if (!WRITE) {
        const uint64_t w = wave_sum(acc);
        if (lane_id() == 0 && w) atomicAdd(sum, (unsigned long long)w);
    }
}

}  // namespace

int launch_access_scale(edpp_ctx* c, int pattern, const int32_t* in, int32_t* out, int64_t n, int repeat) {
    const int64_t lanes = n / repeat;
    if (lanes == 0) return EDPP_OK;
    const unsigned grid = (unsigned)ceil_div(lanes, (int64_t)kAccessBlock);
    switch (pattern) {
        case EDPP_ACCESS_COLUMN:
            hipLaunchKernelGGL(k_access_scale<EDPP_ACCESS_COLUMN>, dim3(grid), dim3(kAccessBlock), 0, c->stream, in,
                               out, lanes, repeat);
            break;
        case EDPP_ACCESS_ROW:
            hipLaunchKernelGGL(k_access_scale<EDPP_ACCESS_ROW>, dim3(grid), dim3(kAccessBlock), 0, c->stream, in, out,
                               lanes, repeat);
            break;
        default:
            hipLaunchKernelGGL(k_access_scale<EDPP_ACCESS_MIXED>, dim3(grid), dim3(kAccessBlock), 0, c->stream, in,
                               out, lanes, repeat);
    }
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

int launch_access_chunks(edpp_ctx* c, const int32_t* in, int64_t n, int64_t chunk, const int32_t* order,
                         unsigned long long* d_count) {
    const int64_t chunks = ceil_div(n, chunk);
    void* w;
    int rc = ws_reserve(c, 256, &w);
    if (rc) return rc;
    uint32_t* ticket = static_cast<uint32_t*>(w);
    hipLaunchKernelGGL(k_access_zero, dim3(1), dim3(64), 0, c->stream, d_count);
    hipLaunchKernelGGL(k_access_reset_ticket, dim3(1), dim3(64), 0, c->stream, ticket);
    if (chunks > 0)
        hipLaunchKernelGGL(k_access_chunks, dim3((unsigned)chunks), dim3(256), 0, c->stream, in, n, chunk, order,
                           chunks, ticket, d_count);
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

}  // namespace edpp

namespace edpp {

int launch_access_random(edpp_ctx* c, int32_t* table, int64_t table_n, int64_t accesses, uint32_t seed, bool write,
                         unsigned long long* d_sum) {
    constexpr int U = 8;
    if (!write) hipLaunchKernelGGL(k_access_zero, dim3(1), dim3(64), 0, c->stream, d_sum);
    if (accesses > 0) {
        const uint32_t mask = (uint32_t)(table_n - 1);
        const int64_t groups = ceil_div(accesses, (int64_t)kRandomBlock * U);
        const unsigned grid = (unsigned)std::min<int64_t>(groups, (int64_t)c->cus * 16);
        if (write)
            hipLaunchKernelGGL((k_access_random<U, true>), dim3(grid), dim3(kRandomBlock), 0, c->stream, table, mask,
                               accesses, seed, d_sum);
        else
            hipLaunchKernelGGL((k_access_random<U, false>), dim3(grid), dim3(kRandomBlock), 0, c->stream, table, mask,
                               accesses, seed, d_sum);
    }
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

}  // namespace edpp
