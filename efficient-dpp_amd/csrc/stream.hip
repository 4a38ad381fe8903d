// stream.hip — STREAM bandwidth probe over int32 (copy, scale, add, triad).
//
// Reference: opencl/kernels/mem_kernel.cl:9-31 (one int per work-item, scalar
// 3), cuda/test_bandwidth.cu:14-44, driver opencl/test/test_bandwidth.cpp:8-131.
// Here every lane moves 16-byte vectors (dwordx4) in a grid-stride loop with
// streaming (non-temporal) hints, which is what the HBM roofline measurement
// needs; the arithmetic is the reference's.
#include "ctx.hpp"
#include "device.hpp"

namespace edpp {

constexpr int kStreamScalar = 3;

template <int OP>
__device__ __forceinline__ v4i32 stream_op(v4i32 a, v4i32 b) {
    if (OP == EDPP_STREAM_COPY) return a;
    if (OP == EDPP_STREAM_SCALE) return a * kStreamScalar;
    if (OP == EDPP_STREAM_ADD) return a + b;
    return a + b * kStreamScalar;
}
template <int OP>
__device__ __forceinline__ int32_t stream_op1(int32_t a, int32_t b) {
    if (OP == EDPP_STREAM_COPY) return a;
    if (OP == EDPP_STREAM_SCALE) return a * kStreamScalar;
    if (OP == EDPP_STREAM_ADD) return a + b;
    return a + b * kStreamScalar;
}

// each lane handles U vectors per grid stride, issued back to back
template <int OP, int U>
__global__ __launch_bounds__(256) void k_stream(const v4i32* __restrict__ a, const v4i32* __restrict__ b,
                                                v4i32* __restrict__ c, int64_t nvec) {
    const int64_t stride = (int64_t)gridDim.x * 256 * U;
    for (int64_t base = (int64_t)blockIdx.x * 256 * U + threadIdx.x; base < nvec; base += stride) {
        v4i32 x[U], y[U];
        #pragma unroll
        for (int u = 0; u < U; ++u) {
            int64_t i = base + (int64_t)u * 256;
            if (i < nvec) {
                x[u] = ld_stream(a + i);
                if (OP >= EDPP_STREAM_ADD) y[u] = ld_stream(b + i);
            }
        }
        #pragma unroll
        for (int u = 0; u < U; ++u) {
            int64_t i = base + (int64_t)u * 256;
            if (i < nvec) st_stream(stream_op<OP>(x[u], y[u]), c + i);
        }
    }
}

template <int OP>
__global__ void k_stream_scalar(const int32_t* a, const int32_t* b, int32_t* c, int64_t lo, int64_t n) {
    for (int64_t i = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        c[i] = stream_op1<OP>(a[i], OP >= EDPP_STREAM_ADD ? b[i] : 0);
}

template <int OP>
static int run_stream(edpp_ctx* c, const int32_t* a, const int32_t* b, int32_t* out, int64_t n) {
    constexpr int U = 4;
    const bool vec = aligned16(a) && aligned16(out) && (OP < EDPP_STREAM_ADD || aligned16(b));
    int64_t nvec = vec ? n / 4 : 0;
    if (nvec) {
        // enough resident waves to cover HBM latency: 8 workgroups of 4 waves per CU
        int64_t want = ceil_div(nvec, 256 * U);
        int64_t cap = (int64_t)c->cus * 8;
        unsigned grid = (unsigned)(want < cap ? want : cap);
        hipLaunchKernelGGL((k_stream<OP, U>), dim3(grid), dim3(256), 0, c->stream, (const v4i32*)a,
                           (const v4i32*)b, (v4i32*)out, nvec);
    }
    if (nvec * 4 < n) {
        int64_t rest = n - nvec * 4;
        unsigned grid = (unsigned)(rest < 256 * 1024 ? ceil_div(rest, 256) : 1024);
        hipLaunchKernelGGL((k_stream_scalar<OP>), dim3(grid), dim3(256), 0, c->stream, a, b, out, nvec * 4, n);
    }
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

int launch_stream_i32(edpp_ctx* c, int op, const int32_t* a, const int32_t* b, int32_t* out, int64_t n) {
    if (n == 0) return EDPP_OK;
    switch (op) {
        case EDPP_STREAM_COPY: return run_stream<EDPP_STREAM_COPY>(c, a, b, out, n);
        case EDPP_STREAM_SCALE: return run_stream<EDPP_STREAM_SCALE>(c, a, b, out, n);
        case EDPP_STREAM_ADD: return run_stream<EDPP_STREAM_ADD>(c, a, b, out, n);
        case EDPP_STREAM_TRIAD: return run_stream<EDPP_STREAM_TRIAD>(c, a, b, out, n);
        default: return fail(c, EDPP_EINVAL, "unknown stream op %d", op);
    }
}

}  // namespace edpp
