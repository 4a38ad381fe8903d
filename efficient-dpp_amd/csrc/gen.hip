// gen.hip — reproducible synthetic inputs on the device.
//
// The reference draws inputs with rand_r per OpenMP thread under a dynamic
// schedule and with a time-seeded swap shuffle (opencl/util/utility.cpp:214-241),
// so its inputs cannot be reproduced.  Every generator here is a pure function
// of (seed, i) built on SplitMix64, so the CPU oracle and the GPU produce the
// same arrays under any launch shape.
#include <math.h>
#include <vector>
#include "ctx.hpp"
#include "device.hpp"

namespace edpp {

__global__ void k_gen_u32(uint32_t* out, int64_t n, uint64_t seed, uint32_t mask) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (uint32_t)(splitmix_at(seed, (uint64_t)i) >> 32) & mask;
}
__global__ void k_gen_u64(uint64_t* out, int64_t n, uint64_t seed, uint64_t mask) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = splitmix_at(seed, (uint64_t)i) & mask;
}

constexpr uint64_t kPermKeySalt = 0x243F6A8885A308D3ull;   // pi digits: key for the whole-array / block order
constexpr uint64_t kBlockKeySalt = 0x13198A2E03707344ull;  // key salt for the order inside blocks
constexpr int kPermBlockBits = 12;                         // 4096-element blocks

__global__ void k_gen_perm_random(int32_t* idx, int64_t n, int half, uint64_t key) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        idx[i] = (int32_t)perm_at((uint64_t)i, (uint64_t)n, half, key);
}

// position i of block b reads source block P(b) at offset Q_b(i mod 4096); the
// ragged tail (n mod 4096 positions) is permuted within itself
__global__ void k_gen_perm_block(int32_t* idx, int64_t n, uint64_t seed, int64_t nblocks, int half_b, int half_t) {
    const uint64_t full = (uint64_t)nblocks << kPermBlockBits;
    const uint64_t kb = mix64(seed ^ kPermKeySalt);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t u = (uint64_t)i, r;
        if (u < full) {
            uint64_t b = u >> kPermBlockBits, j = u & ((1u << kPermBlockBits) - 1);
            uint64_t pb = perm_at(b, (uint64_t)nblocks, half_b, kb);
            uint64_t kj = mix64(seed ^ kBlockKeySalt ^ (b * 0x9E3779B97F4A7C15ull));
            r = (pb << kPermBlockBits) | feistel(j, kPermBlockBits / 2, kj);
        } else {
            uint64_t t = (uint64_t)n - full;
            uint64_t kt = mix64(seed ^ kBlockKeySalt ^ 0xA4093822299F31D0ull);
            r = full + perm_at(u - full, t, half_t, kt);
        }
        idx[i] = (int32_t)r;
    }
}

__global__ void k_gen_iota(int32_t* idx, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        idx[i] = (int32_t)i;
}

// bucket from u = high 32 bits of x_i: uniform -> top log2(B) bits;
// Zipf -> smallest k with u < thr[k] (thr[B-1] = 2^32)
template <class K>
__global__ void k_gen_split_keys(K* keys, int32_t* vals, int64_t n, uint64_t seed, int lg, const uint64_t* thr,
                                 int zipf) {
    const int B = 1 << lg;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t r = splitmix_at(seed, (uint64_t)i);
        uint32_t u = (uint32_t)(r >> 32);
        uint32_t b;
        if (!zipf) {
            b = lg ? (u >> (32 - lg)) : 0u;
        } else {
            int lo = 0, hi = B - 1;  // answer in [lo, hi]
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if ((uint64_t)u < thr[mid]) hi = mid;
                else lo = mid + 1;
            }
            b = (uint32_t)lo;
        }
        if (sizeof(K) == 4) keys[i] = (K)((((uint32_t)r) << lg) | b);
        else keys[i] = (K)((mix64(r) << lg) | b);
        if (vals) vals[i] = (int32_t)i;
    }
}

static unsigned gen_grid(edpp_ctx* c, int64_t n) {
    int64_t want = ceil_div(n, 256);
    int64_t cap = (int64_t)c->cus * 16;
    return (unsigned)(want < cap ? (want < 1 ? 1 : want) : cap);
}

int launch_gen_u32(edpp_ctx* c, uint32_t* out, int64_t n, uint64_t seed, uint32_t mask) {
    if (!n) return EDPP_OK;
    hipLaunchKernelGGL(k_gen_u32, dim3(gen_grid(c, n)), dim3(256), 0, c->stream, out, n, seed, mask);
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}
int launch_gen_u64(edpp_ctx* c, uint64_t* out, int64_t n, uint64_t seed, uint64_t mask) {
    if (!n) return EDPP_OK;
    hipLaunchKernelGGL(k_gen_u64, dim3(gen_grid(c, n)), dim3(256), 0, c->stream, out, n, seed, mask);
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

int launch_gen_perm(edpp_ctx* c, int32_t* idx, int64_t n, uint64_t seed, int kind) {
    if (!n) return EDPP_OK;
    if (kind == EDPP_PERM_RANDOM) {
        hipLaunchKernelGGL(k_gen_perm_random, dim3(gen_grid(c, n)), dim3(256), 0, c->stream, idx, n,
                           feistel_half((uint64_t)n), mix64(seed ^ kPermKeySalt));
    } else if (kind == EDPP_PERM_BLOCK_LOCAL) {
        int64_t nb = n >> kPermBlockBits;
        int64_t tail = n - (nb << kPermBlockBits);
        hipLaunchKernelGGL(k_gen_perm_block, dim3(gen_grid(c, n)), dim3(256), 0, c->stream, idx, n, seed, nb,
                           feistel_half(nb ? (uint64_t)nb : 1), feistel_half(tail ? (uint64_t)tail : 1));
    } else if (kind == EDPP_PERM_IDENTITY) {
        hipLaunchKernelGGL(k_gen_iota, dim3(gen_grid(c, n)), dim3(256), 0, c->stream, idx, n);
    } else {
        return fail(c, EDPP_EINVAL, "unknown permutation kind %d", kind);
    }
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

// Zipf(s=1) thresholds over ranks 1..B scaled to 2^32: thr[k] = floor(F(k+1) * 2^32)
// with F the CDF accumulated in double in rank order; thr[B-1] = 2^32.
void zipf_thresholds(int B, std::vector<uint64_t>& thr) {
    double h = 0.0;
    for (int r = 1; r <= B; ++r) h += 1.0 / (double)r;
    thr.resize(B);
    double cum = 0.0;
    for (int k = 0; k < B; ++k) {
        cum += 1.0 / (double)(k + 1);
        thr[k] = (k == B - 1) ? (1ull << 32) : (uint64_t)floor(cum / h * 4294967296.0);
    }
}

template <class K>
int launch_gen_split_keys(edpp_ctx* c, K* keys, int32_t* vals, int64_t n, int buckets, int dist, uint64_t seed) {
    if (!n) return EDPP_OK;
    int lg = 0;
    while ((1 << lg) < buckets) ++lg;
    const uint64_t* d_thr = nullptr;
    if (dist == EDPP_KEYS_ZIPF) {
        std::vector<uint64_t> thr;
        zipf_thresholds(buckets, thr);
        void* w;
        int rc = ws_reserve(c, thr.size() * sizeof(uint64_t), &w);
        if (rc) return rc;
        EDPP_HIP(c, hipMemcpyAsync(w, thr.data(), thr.size() * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
        EDPP_HIP(c, hipStreamSynchronize(c->stream));  // thr is a host temporary
        d_thr = (const uint64_t*)w;
    } else if (dist != EDPP_KEYS_UNIFORM) {
        return fail(c, EDPP_EINVAL, "unknown key distribution %d", dist);
    }
    hipLaunchKernelGGL((k_gen_split_keys<K>), dim3(gen_grid(c, n)), dim3(256), 0, c->stream, keys, vals, n, seed, lg,
                       d_thr, dist == EDPP_KEYS_ZIPF ? 1 : 0);
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

template int launch_gen_split_keys<uint32_t>(edpp_ctx*, uint32_t*, int32_t*, int64_t, int, int, uint64_t);
template int launch_gen_split_keys<uint64_t>(edpp_ctx*, uint64_t*, int32_t*, int64_t, int, int, uint64_t);

}  // namespace edpp
