// learn.hip — schedule candidates, exhaustive autotuning, device-side call
// features and the FP32 MLP cost model (forward, backward, Adam) that picks a
// candidate per call.  See include/edpp_learn.h for the feature definitions
// and the parameter layout.
#include <math.h>
#include <string.h>
#include <algorithm>
#include <vector>
#include "ctx.hpp"
#include "device.hpp"
#include "edpp_learn.h"

// One remembered model decision: the call signature, the parameter version
// it was computed under (0: empty slot) and the chosen candidate.
struct ChoiceMemo {
    int op = 0, elem_bytes = 0, buckets = 0, has_vals = 0;
    int64_t n = 0;
    const void* in = nullptr;
    const void* idx = nullptr;
    uint64_t version = 0;
    uint64_t used = 0;
    int cand = -1;
};
constexpr int kMemoSlots = 64;

struct edpp_mlp {
    edpp_ctx* ctx = nullptr;
    uint64_t version = 1;      // bumped whenever the parameters change
    bool memo_on = true;
    uint64_t memo_clock = 0;
    ChoiceMemo memo[kMemoSlots];
    float* params = nullptr;   // device, P floats
    float* partial = nullptr;  // device, kGradGroups * (P + 1) floats
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* loss = nullptr;     // device, 1 float
    float* losses = nullptr;   // device, the per-epoch losses of one training call
    int32_t losses_cap = 0;
    int32_t* pinned_choice = nullptr;  // host-pinned, written by the selection kernel
    int64_t adam_step = 0;
};

namespace edpp {

namespace {

constexpr int F = EDPP_NUM_FEATURES, H = EDPP_HIDDEN, K = EDPP_MAX_CANDIDATES;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H * F, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + K * H, NPARAM = OFF_B3 + K;
constexpr int kGradGroups = 256;  // most workgroups (and partials) of one gradient

struct Cand {
    int family;
    const char* name;
    int scheme;     // scan scheme / gather passes / split strategy
    int knob_a;     // scan: block; gather: unroll; split: ipt
    int knob_b;     // scan: vpt;   split: tile_mult
};

const Cand kCands[K] = {
    {0, "scan look-back 1024x8", EDPP_SCAN_LOOKBACK, 1024, 8},
    {0, "scan look-back 1024x4", EDPP_SCAN_LOOKBACK, 1024, 4},
    {0, "scan look-back 1024x16", EDPP_SCAN_LOOKBACK, 1024, 16},
    {0, "scan reduce-then-scan", EDPP_SCAN_RTS, 0, 0},
    {0, "scan persistent look-back 1024x8", EDPP_SCAN_PERSIST, 1024, 8},
    {1, "gather/scatter 1 pass, unroll 4", 1, 4, 0},
    {1, "gather/scatter 1 pass, unroll 8", 1, 8, 0},
    {1, "gather/scatter 2 passes", 2, 4, 0},
    {1, "gather/scatter 4 passes", 4, 4, 0},
    {1, "gather/scatter 8 passes", 8, 4, 0},
    {2, "split WG 16/lane, tile 16B", EDPP_SPLIT_WG, 16, 16},
    {2, "split WG 16/lane, tile 64B", EDPP_SPLIT_WG, 16, 64},
    {2, "split WG 8/lane, tile 16B", EDPP_SPLIT_WG, 8, 16},
    {2, "split WI columns 64B", EDPP_SPLIT_WI, 16, 64},
    {2, "split WG 16/lane, tile 256B", EDPP_SPLIT_WG, 16, 256},
};

int family_of(int op) { return op == EDPP_OP_SCAN ? 0 : (op == EDPP_OP_SPLIT ? 2 : 1); }

// ---------------------------------------------------------------- features
struct StaticFeat {
    float f[F];
};

// Fills feats[0..F) from the static part plus device samples of the inputs.
__device__ void compute_features(int op, const void* in, const int32_t* idx, int64_t n, int lg, int wide,
                                 const StaticFeat& sf, float* feats, uint32_t* lds) {
    const int t = threadIdx.x;
    if (t < F) feats[t] = sf.f[t];
    __syncthreads();
    if (op == EDPP_OP_SPLIT && n > 0) {
        const int B = 1 << lg;
        const int S = (int)std::min<int64_t>(n, 4096);
        for (int b = t; b < B; b += blockDim.x) lds[b] = 0;
        __syncthreads();
        for (int k = t; k < S; k += blockDim.x) {
            int64_t i = (int64_t)((double)k * (double)n / (double)S);
            uint32_t key = wide ? (uint32_t) static_cast<const uint64_t*>(in)[i] : static_cast<const uint32_t*>(in)[i];
            atomicAdd(&lds[key & (uint32_t)(B - 1)], 1u);
        }
        __syncthreads();
        uint32_t m = 0;
        for (int b = t; b < B; b += blockDim.x) m = max(m, lds[b]);
        // block maximum: wave maxima through LDS
        uint32_t wm = m;
        for (int o = 32; o > 0; o >>= 1) wm = max(wm, (uint32_t)__shfl_xor((int)wm, o));
        __syncthreads();
        if (lane_id() == 0) lds[t / kWave] = wm;
        __syncthreads();
        if (t == 0) {
            uint32_t mm = 0;
            for (int q = 0; q < (int)(blockDim.x / kWave); ++q) mm = max(mm, lds[q]);
            float skew = (float)mm * (float)B / (float)S;
            feats[6] = log2f(fmaxf(skew, 1.0f)) / 12.0f;
        }
        __syncthreads();
    } else if ((op == EDPP_OP_GATHER || op == EDPP_OP_SCATTER) && n > 1) {
        const int S = (int)std::min<int64_t>(n - 1, 2048);
        uint32_t local = 0;
        for (int k = t; k < S; k += blockDim.x) {
            // hashed positions: a regular stride could keep landing on block edges
            int64_t i = (int64_t)(mix64((uint64_t)k + 0x51ED27ull) % (uint64_t)(n - 1));
            int64_t d = (int64_t)idx[i + 1] - (int64_t)idx[i];
            local += (d < 4096 && d > -4096) ? 1u : 0u;
        }
        for (int o = 32; o > 0; o >>= 1) local += (uint32_t)__shfl_xor((int)local, o);
        __syncthreads();
        if (lane_id() == 0) lds[t / kWave] = local;
        __syncthreads();
        if (t == 0) {
            uint32_t s = 0;
            for (int q = 0; q < (int)(blockDim.x / kWave); ++q) s += lds[q];
            feats[7] = (float)s / (float)S;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_features(int op, const void* in, const int32_t* idx, int64_t n, int lg,
                                                  int wide, StaticFeat sf, float* out) {
    __shared__ uint32_t lds[4096];
    __shared__ float feats[F];
    compute_features(op, in, idx, n, lg, wide, sf, feats, lds);
    if (threadIdx.x < F) out[threadIdx.x] = feats[threadIdx.x];
}

// ---------------------------------------------------------------- MLP pieces
// One row through the network by 64 lanes (lane j = hidden unit j); x, h1,
// h2, z in LDS.  Returns nothing; leaves probabilities of valid candidates
// in p[] (LDS) and the masked argmax in *choice.
__device__ void mlp_row_forward(const float* __restrict__ P, const float* x, const float* mask, float* h1, float* h2,
                                float* z, float* p, int* choice) {
    const int j = threadIdx.x;
    if (j < H) {
        float a = P[OFF_B1 + j];
        #pragma unroll
        for (int k = 0; k < F; ++k) a += P[OFF_W1 + j * F + k] * x[k];
        h1[j] = fmaxf(a, 0.0f);
    }
    __syncthreads();
    if (j < H) {
        float a = P[OFF_B2 + j];
        for (int k = 0; k < H; ++k) a += P[OFF_W2 + j * H + k] * h1[k];
        h2[j] = fmaxf(a, 0.0f);
    }
    __syncthreads();
    if (j < K) {
        float a = P[OFF_B3 + j];
        for (int k = 0; k < H; ++k) a += P[OFF_W3 + j * H + k] * h2[k];
        z[j] = a;
    }
    __syncthreads();
    if (j < K) {
        float mx = -INFINITY;
        int best = -1;
        for (int c = 0; c < K; ++c)
            if (mask[c] > 0.0f && z[c] > mx) {
                mx = z[c];
                best = c;
            }
        float s = 0.0f;
        for (int c = 0; c < K; ++c)
            if (mask[c] > 0.0f) s += expf(z[c] - mx);
        p[j] = mask[j] > 0.0f ? expf(z[j] - mx) / s : 0.0f;
        if (j == 0) *choice = best;
    }
    __syncthreads();
}

__global__ __launch_bounds__(64) void k_mlp_forward(const float* __restrict__ P, const float* __restrict__ X,
                                                    const float* __restrict__ M, int32_t rows, float* probs,
                                                    int32_t* choice) {
    __shared__ float x[F], mask[K], h1[H], h2[H], z[K], p[K];
    __shared__ int ch;
    for (int r = blockIdx.x; r < rows; r += gridDim.x) {
        if (threadIdx.x < F) x[threadIdx.x] = X[(int64_t)r * F + threadIdx.x];
        if (threadIdx.x < K) mask[threadIdx.x] = M[(int64_t)r * K + threadIdx.x];
        __syncthreads();
        mlp_row_forward(P, x, mask, h1, h2, z, p, &ch);
        if (probs && threadIdx.x < K) probs[(int64_t)r * K + threadIdx.x] = p[threadIdx.x];
        if (choice && threadIdx.x == 0) choice[r] = ch;
        __syncthreads();
    }
}

// Gradient of the mean cross-entropy, one workgroup of kGradBlock lanes per
// group of chunks of kGradRows consecutive rows (chunks g, g+G, ... for
// workgroup g of G).  The weights are staged once per workgroup in LDS,
// transposed with odd row strides (W1T[k][j], W2T[k][j], W3T[k][c]), so that
// every layer reads them conflict-free both forward (lanes over the output
// unit) and backward (lanes over the input unit).  A chunk goes through the
// network as small matrices: wave w takes rows w, w+4, ... of each layer with
// one lane per unit; the logits take one lane per (row, candidate).  Each lane
// keeps its share of the weight gradients in registers over all the
// workgroup's chunks and writes them once as the workgroup's partial sums;
// k_mlp_reduce and k_mlp_reduce_adam add the partials in a fixed order, so
// the result is deterministic.
constexpr int kGradBlock = 256;
constexpr int kGradRows = 16;
constexpr int kGradWaves = kGradBlock / kWave;  // 4
constexpr int SF = F + 1, SH = H + 1, SK = 17;   // LDS row strides (odd: no bank conflicts)
static_assert(H == kWave && K < 16 && F % kGradWaves == 0 && H % kGradWaves == 0, "lane mapping of k_mlp_grad");
static_assert(kGradRows * 16 == kGradBlock && kGradRows % kGradWaves == 0, "lane mapping of k_mlp_grad");

__global__ __launch_bounds__(kGradBlock) void k_mlp_grad(const float* __restrict__ P, const float* __restrict__ X,
                                                         const float* __restrict__ T, const float* __restrict__ M,
                                                         int32_t rows, float* __restrict__ partial) {
    constexpr int R = kGradRows, RW = R / kGradWaves;
    __shared__ float sW1T[F * SH], sW2T[H * SH], sW3T[H * SK], sb1[H], sb2[H], sb3[16];
    __shared__ float sX[R * SF], sM[R * 16], sT[R * 16], sH1[R * SH], sH2[R * SH], sdH2[R * SH], sdH1[R * SH],
        sdZ[R * SK], sLoss[kGradBlock / 16];
    const int t = threadIdx.x, w = t / kWave, l = t % kWave;
    for (int i = t; i < H * F; i += kGradBlock) sW1T[(i % F) * SH + i / F] = P[OFF_W1 + i];
    for (int i = t; i < H * H; i += kGradBlock) sW2T[(i % H) * SH + i / H] = P[OFF_W2 + i];
    for (int i = t; i < K * H; i += kGradBlock) sW3T[(i % H) * SK + i / H] = P[OFF_W3 + i];
    if (t < H) {
        sb1[t] = P[OFF_B1 + t];
        sb2[t] = P[OFF_B2 + t];
    }
    if (t < K) sb3[t] = P[OFF_B3 + t];
    constexpr int QK = (K + kGradWaves - 1) / kGradWaves;
    float gW2[H / kGradWaves], gW1[F / kGradWaves], gW3[QK], gb = 0.f, loss = 0.f;
    #pragma unroll
    for (int q = 0; q < H / kGradWaves; ++q) gW2[q] = 0.f;
    #pragma unroll
    for (int q = 0; q < QK; ++q) gW3[q] = 0.f;
    #pragma unroll
    for (int q = 0; q < F / kGradWaves; ++q) gW1[q] = 0.f;
    const float inv = 1.0f / (float)rows;
    const int64_t chunks = ((int64_t)rows + R - 1) / R;
    for (int64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
        const int64_t r0 = ch * R;
        __syncthreads();  // weights staged; the previous chunk's reads are done
        for (int i = t; i < R * F; i += kGradBlock) {
            const int r = i / F, k = i % F;
            sX[r * SF + k] = r0 + r < rows ? X[(r0 + r) * F + k] : 0.f;
        }
        for (int i = t; i < R * K; i += kGradBlock) {
            const int r = i / K, c = i % K;
            const bool ok = r0 + r < rows;
            sM[r * 16 + c] = ok ? M[(r0 + r) * K + c] : 0.f;  // rows past the end: no candidate, no gradient
            sT[r * 16 + c] = ok ? T[(r0 + r) * K + c] : 0.f;
        }
        __syncthreads();
        #pragma unroll 1
        for (int q = 0; q < RW; ++q) {  // hidden layer 1: lane l = unit j
            const int r = w + kGradWaves * q;
            float a = sb1[l];
            #pragma unroll
            for (int k = 0; k < F; ++k) a += sW1T[k * SH + l] * sX[r * SF + k];
            sH1[r * SH + l] = fmaxf(a, 0.0f);
        }
        __syncthreads();
        #pragma unroll 1
        for (int q = 0; q < RW; ++q) {  // hidden layer 2
            const int r = w + kGradWaves * q;
            float a = sb2[l];
            #pragma unroll 8
            for (int k = 0; k < H; ++k) a += sW2T[k * SH + l] * sH1[r * SH + k];
            sH2[r * SH + l] = fmaxf(a, 0.0f);
        }
        __syncthreads();
        {  // logits, masked softmax and its gradient: lane t = (row t/16, candidate t%16)
            const int r = t / 16, c = t % 16;
            const bool valid = c < K && sM[r * 16 + c] > 0.0f;
            float z = 0.f;
            if (c < K) {
                z = sb3[c];
                #pragma unroll 8
                for (int k = 0; k < H; ++k) z += sW3T[k * SK + c] * sH2[r * SH + k];
            }
            float mx = valid ? z : -INFINITY;
            #pragma unroll
            for (int o = 8; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
            const float e = valid ? expf(z - mx) : 0.f;
            float s = e;
            #pragma unroll
            for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o);
            const float p = s > 0.f ? e / s : 0.f;
            const float tg = c < K ? sT[r * 16 + c] : 0.f;
            sdZ[r * SK + c] = valid ? (p - tg) * inv : 0.f;
            float lt = valid && tg > 0.f ? -tg * logf(fmaxf(p, 1e-30f)) : 0.f;
            #pragma unroll
            for (int o = 8; o > 0; o >>= 1) lt += __shfl_xor(lt, o);
            if (c == 0) loss += lt * inv;
        }
        __syncthreads();
        #pragma unroll 1
        for (int q = 0; q < RW; ++q) {  // back through W3 and the ReLU of layer 2
            const int r = w + kGradWaves * q;
            float d = 0.f;
            #pragma unroll
            for (int c = 0; c < K; ++c) d += sdZ[r * SK + c] * sW3T[l * SK + c];
            sdH2[r * SH + l] = sH2[r * SH + l] > 0.0f ? d : 0.0f;
        }
        __syncthreads();
        #pragma unroll 1
        for (int q = 0; q < RW; ++q) {  // back through W2 and the ReLU of layer 1: lane l = input unit k
            const int r = w + kGradWaves * q;
            float d = 0.f;
            #pragma unroll 8
            for (int j = 0; j < H; ++j) d += sdH2[r * SH + j] * sW2T[l * SH + j];
            sdH1[r * SH + l] = sH1[r * SH + l] > 0.0f ? d : 0.0f;
        }
        __syncthreads();
        // weight and bias gradients of this chunk, one row at a time in order:
        // W2[j][k] (j = w + 4q, k = l), W1[j][k] (j = l, k = w + 4q),
        // W3[c][j] (c = w + 4q, j = l); biases: wave 0 b1, wave 1 b2, wave 2 b3
        #pragma unroll 1
        for (int r = 0; r < R; ++r) {
            const float h1 = sH1[r * SH + l], h2 = sH2[r * SH + l], d1 = sdH1[r * SH + l];
            #pragma unroll
            for (int q = 0; q < H / kGradWaves; ++q) gW2[q] += sdH2[r * SH + w + kGradWaves * q] * h1;
            #pragma unroll
            for (int q = 0; q < F / kGradWaves; ++q) gW1[q] += d1 * sX[r * SF + w + kGradWaves * q];
            #pragma unroll
            for (int q = 0; q < (K + kGradWaves - 1) / kGradWaves; ++q)
                if (w + kGradWaves * q < K) gW3[q] += sdZ[r * SK + w + kGradWaves * q] * h2;
            gb += w == 0 ? d1 : (w == 1 ? sdH2[r * SH + l] : (w == 2 && l < K ? sdZ[r * SK + l] : 0.f));
        }
    }
    float* out = partial + (int64_t)blockIdx.x * (NPARAM + 1);
    #pragma unroll
    for (int q = 0; q < H / kGradWaves; ++q) out[OFF_W2 + (w + kGradWaves * q) * H + l] = gW2[q];
    #pragma unroll
    for (int q = 0; q < F / kGradWaves; ++q) out[OFF_W1 + l * F + w + kGradWaves * q] = gW1[q];
    #pragma unroll
    for (int q = 0; q < QK; ++q)
        if (w + kGradWaves * q < K) out[OFF_W3 + (w + kGradWaves * q) * H + l] = gW3[q];
    if (w == 0) out[OFF_B1 + l] = gb;
    if (w == 1) out[OFF_B2 + l] = gb;
    if (w == 2 && l < K) out[OFF_B3 + l] = gb;
    // the loss: lanes t = 16*row hold it, added in row order
    if (t % 16 == 0) sLoss[t / 16] = loss;
    __syncthreads();
    if (t == 0) {
        float s = 0.f;
        for (int q = 0; q < kGradBlock / 16; ++q) s += sLoss[q];
        out[NPARAM] = s;
    }
}

// Sum of entry i over the first `groups` partials.  A block of 256 lanes
// takes 64 entries; the four waves each add a quarter of the partials
// (partials q, q+4, q+8, ... in wave q, as two interleaved chains) and the
// quarters are combined in a fixed order, so every call returns the same
// float.  Valid in wave 0; every lane of the block must call it.
constexpr int kReduceEntries = 64;
__device__ __forceinline__ float sum_partials(const float* __restrict__ partial, int groups, int i, float* lds) {
    const int q = threadIdx.x / kWave;
    float a0 = 0.f, a1 = 0.f;
    if (i <= NPARAM) {
        int g = q;
        for (; g + 4 < groups; g += 8) {
            a0 += partial[(int64_t)g * (NPARAM + 1) + i];
            a1 += partial[(int64_t)(g + 4) * (NPARAM + 1) + i];
        }
        if (g < groups) a0 += partial[(int64_t)g * (NPARAM + 1) + i];
    }
    lds[threadIdx.x] = a0 + a1;
    __syncthreads();
    const int e = threadIdx.x % kWave;
    return (lds[e] + lds[kWave + e]) + (lds[2 * kWave + e] + lds[3 * kWave + e]);
}

__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ partial, int groups,
                                                    float* __restrict__ grad, float* loss) {
    __shared__ float lds[256];
    const int i = blockIdx.x * kReduceEntries + (int)(threadIdx.x % kWave);
    const float s = sum_partials(partial, groups, i, lds);
    if (threadIdx.x < kWave) {
        if (i < NPARAM) grad[i] = s;
        else if (i == NPARAM && loss) *loss = s;
    }
}

// The reduction fused with the Adam step of training: parameter i is updated
// from its summed gradient; the loss goes to loss_at (the epoch's slot).
__global__ __launch_bounds__(256) void k_mlp_reduce_adam(const float* __restrict__ partial, int groups,
                                                         float* __restrict__ P, float* __restrict__ m,
                                                         float* __restrict__ v, float lr, float c1, float c2,
                                                         float* loss_at) {
    __shared__ float lds[256];
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int i = blockIdx.x * kReduceEntries + (int)(threadIdx.x % kWave);
    const float g = sum_partials(partial, groups, i, lds);
    if (threadIdx.x >= kWave || i > NPARAM) return;
    if (i == NPARAM) {
        *loss_at = g;
        return;
    }
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// features + one forward row + masked argmax, result to host-pinned memory
__global__ __launch_bounds__(256) void k_select(const float* __restrict__ P, int op, const void* in,
                                                const int32_t* idx, int64_t n, int lg, int wide, StaticFeat sf,
                                                int32_t* out_choice) {
    __shared__ uint32_t lds[4096];
    __shared__ float feats[F], mask[K], h1[H], h2[H], z[K], p[K];
    __shared__ int ch;
    compute_features(op, in, idx, n, lg, wide, sf, feats, lds);
    const int fam = op == EDPP_OP_SCAN ? 0 : (op == EDPP_OP_SPLIT ? 2 : 1);
    if (threadIdx.x < K) {
        int lo = fam == 0 ? 0 : (fam == 1 ? 5 : 10);
        mask[threadIdx.x] = (threadIdx.x >= lo && threadIdx.x < lo + 5) ? 1.0f : 0.0f;
    }
    __syncthreads();
    mlp_row_forward(P, feats, mask, h1, h2, z, p, &ch);
    if (threadIdx.x == 0) *out_choice = ch;
}

StaticFeat static_features(edpp_ctx* c, const edpp_call& q) {
    StaticFeat s;
    memset(&s, 0, sizeof s);
    const int fam = family_of(q.op);
    s.f[fam] = 1.0f;
    const double n = (double)std::max<int64_t>(q.n, 1);
    s.f[3] = (float)(log2(n) / 28.0);
    s.f[4] = (float)q.elem_bytes / 8.0f;
    int lg = 0;
    if (q.op == EDPP_OP_SPLIT)
        while ((1 << lg) < q.buckets) ++lg;
    s.f[5] = (float)lg / 12.0f;
    // bytes the call moves under the per-element models
    double per = q.op == EDPP_OP_SCAN ? 2.0 * q.elem_bytes
                                      : (q.op == EDPP_OP_SPLIT ? 2.0 * (q.elem_bytes + (q.vals_in ? 4 : 0))
                                                              : 4.0 + 2.0 * q.elem_bytes);
    double bytes = std::max(1.0, n * per);
    s.f[8] = (float)(log2(bytes / (double)c->mall_bytes) / 8.0);
    s.f[9] = (float)(log2(bytes / (double)std::max<int64_t>(c->l2_bytes, 1)) / 16.0);
    s.f[10] = (float)(log2(std::max(n / 4096.0 / std::max(c->cus, 1), 1e-6)) / 16.0);
    s.f[11] = q.op == EDPP_OP_SCATTER ? 1.0f : 0.0f;
    return s;
}

int lg_of(int b) {
    int lg = 0;
    while ((1 << lg) < b) ++lg;
    return lg;
}

}  // namespace

int validate_call(edpp_ctx* c, const edpp_call& q) {
    if (q.op < EDPP_OP_SCAN || q.op > EDPP_OP_SPLIT) return fail(c, EDPP_EINVAL, "call: unknown op %d", q.op);
    if (q.elem_bytes != 4 && q.elem_bytes != 8) return fail(c, EDPP_EINVAL, "call: elem_bytes must be 4 or 8");
    if (q.n < 0 || q.n >= (1ll << 31)) return fail(c, EDPP_EINVAL, "call: n outside [0, 2^31)");
    if (q.n > 0 && (!q.in || !q.out)) return fail(c, EDPP_EINVAL, "call: null buffer");
    if ((q.op == EDPP_OP_GATHER || q.op == EDPP_OP_SCATTER) && q.n > 0 && !q.idx)
        return fail(c, EDPP_EINVAL, "call: gather/scatter need idx");
    if (q.op == EDPP_OP_SPLIT && (q.buckets < 1 || q.buckets > 4096 || (q.buckets & (q.buckets - 1))))
        return fail(c, EDPP_EINVAL, "call: bad bucket count");
    if (q.op == EDPP_OP_SPLIT && ((q.vals_in == nullptr) != (q.vals_out == nullptr)))
        return fail(c, EDPP_EINVAL, "call: values need both vals_in and vals_out");
    return EDPP_OK;
}

// Runs the call with candidate `cand` (its knobs replace the context's for
// this call only).
int run_candidate(edpp_ctx* c, const edpp_call& q, int cand, void* aux) {
    const Cand& k = kCands[cand];
    if (k.family != family_of(q.op)) return fail(c, EDPP_EINVAL, "candidate %d does not fit op %d", cand, q.op);
    const edpp_tuning saved = c->tune;
    int rc;
    const bool w = q.elem_bytes == 8;
    if (k.family == 0) {
        if (k.scheme == EDPP_SCAN_LOOKBACK || k.scheme == EDPP_SCAN_PERSIST) {
            c->tune.scan_block = k.knob_a;
            c->tune.scan_vpt = k.knob_b;
        }
        rc = w ? launch_scan<uint64_t>(c, (const uint64_t*)q.in, (uint64_t*)q.out, q.n, k.scheme, (uint64_t*)aux)
               : launch_scan<uint32_t>(c, (const uint32_t*)q.in, (uint32_t*)q.out, q.n, k.scheme, (uint32_t*)aux);
    } else if (k.family == 1) {
        c->tune.gs_unroll = k.knob_a;
        const bool sc = q.op == EDPP_OP_SCATTER;
        rc = w ? launch_gather<uint64_t>(c, (const uint64_t*)q.in, (uint64_t*)q.out, q.idx, q.n, k.scheme, sc)
               : launch_gather<uint32_t>(c, (const uint32_t*)q.in, (uint32_t*)q.out, q.idx, q.n, k.scheme, sc);
    } else {
        c->tune.split_ipt = k.knob_a;
        c->tune.split_tile_mult = k.knob_b;
        rc = w ? launch_split<uint64_t>(c, (const uint64_t*)q.in, (uint64_t*)q.out, q.vals_in, q.vals_out,
                                        (int32_t*)aux, q.n, q.buckets, k.scheme)
               : launch_split<uint32_t>(c, (const uint32_t*)q.in, (uint32_t*)q.out, q.vals_in, q.vals_out,
                                        (int32_t*)aux, q.n, q.buckets, k.scheme);
    }
    c->tune = saved;
    return rc;
}

// The model's candidate for a call: one kernel computes the features from
// device samples, evaluates the network and writes the masked argmax to
// host-pinned memory.
//
// Decisions are remembered per call signature (op, width, n, buckets, values,
// input and index pointers) until the parameters change: a repeated call skips
// the selection kernel and its synchronisation (about 20 us), which otherwise
// outweighs the schedule gain on calls below a few milliseconds.  Every
// candidate is bit-exact, so a remembered choice for a buffer whose contents
// have since changed can cost time, never correctness.
int model_choose(edpp_ctx* c, const edpp_call& q, int* cand) {
    edpp_mlp* m = c->model;
    const int has_vals = q.vals_in ? 1 : 0;
    auto same = [&](const ChoiceMemo& e) {
        return e.version == m->version && e.op == q.op && e.elem_bytes == q.elem_bytes && e.buckets == q.buckets &&
               e.has_vals == has_vals && e.n == q.n && e.in == q.in && e.idx == (const void*)q.idx;
    };
    if (m->memo_on) {
        for (ChoiceMemo& e : m->memo)
            if (same(e)) {
                e.used = ++m->memo_clock;
                *cand = e.cand;
                return EDPP_OK;
            }
    }
    StaticFeat sf = static_features(c, q);
    *m->pinned_choice = -1;
    hipLaunchKernelGGL(k_select, dim3(1), dim3(256), 0, c->stream, (const float*)m->params, q.op, q.in, q.idx, q.n,
                       lg_of(q.buckets), q.elem_bytes == 8 ? 1 : 0, sf, m->pinned_choice);
    EDPP_HIP(c, hipGetLastError());
    EDPP_HIP(c, hipStreamSynchronize(c->stream));
    int k = *m->pinned_choice;
    if (k < 0 || k >= K || kCands[k].family != family_of(q.op)) return fail(c, EDPP_EHIP, "model gave candidate %d", k);
    *cand = k;
    if (m->memo_on) {
        ChoiceMemo* slot = &m->memo[0];  // an empty or stale slot, else the least recently used
        for (ChoiceMemo& e : m->memo) {
            if (e.version != m->version) {
                slot = &e;
                break;
            }
            if (e.used < slot->used) slot = &e;
        }
        slot->op = q.op;
        slot->elem_bytes = q.elem_bytes;
        slot->buckets = q.buckets;
        slot->has_vals = has_vals;
        slot->n = q.n;
        slot->in = q.in;
        slot->idx = q.idx;
        slot->version = m->version;
        slot->used = ++m->memo_clock;
        slot->cand = k;
    }
    return EDPP_OK;
}

}  // namespace edpp

using namespace edpp;

extern "C" {

int edpp_num_candidates(void) { return K; }

int edpp_candidate_info(int32_t cand, int32_t* family, char* name, int32_t name_len) {
    if (cand < 0 || cand >= K) return EDPP_EINVAL;
    if (family) *family = kCands[cand].family;
    if (name && name_len > 0) {
        strncpy(name, kCands[cand].name, (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    return EDPP_OK;
}

int edpp_call_run(edpp_ctx* c, const edpp_call* q, int32_t cand, double* ms) {
    if (!c || !q) return EDPP_EINVAL;
    int rc = validate_call(c, *q);
    if (rc) return rc;
    if (cand >= K) return fail(c, EDPP_EINVAL, "candidate %d out of range", cand);
    rc = timing_begin(c, ms);
    if (rc) return rc;
    if (q->n > 0) {
        if (cand < 0) {
            if (c->model) {
                int k;
                rc = model_choose(c, *q, &k);
                if (!rc) rc = run_candidate(c, *q, k, nullptr);
            } else {
                // built-in defaults
                const edpp_call& a = *q;
                const bool w = a.elem_bytes == 8;
                if (a.op == EDPP_OP_SCAN)
                    rc = w ? launch_scan<uint64_t>(c, (const uint64_t*)a.in, (uint64_t*)a.out, a.n, EDPP_SCAN_AUTO, nullptr)
                           : launch_scan<uint32_t>(c, (const uint32_t*)a.in, (uint32_t*)a.out, a.n, EDPP_SCAN_AUTO, nullptr);
                else if (a.op == EDPP_OP_SPLIT)
                    rc = w ? launch_split<uint64_t>(c, (const uint64_t*)a.in, (uint64_t*)a.out, a.vals_in, a.vals_out,
                                                    nullptr, a.n, a.buckets, EDPP_SPLIT_AUTO)
                           : launch_split<uint32_t>(c, (const uint32_t*)a.in, (uint32_t*)a.out, a.vals_in, a.vals_out,
                                                    nullptr, a.n, a.buckets, EDPP_SPLIT_AUTO);
                else
                    rc = w ? launch_gather<uint64_t>(c, (const uint64_t*)a.in, (uint64_t*)a.out, a.idx, a.n, 0,
                                                     a.op == EDPP_OP_SCATTER)
                           : launch_gather<uint32_t>(c, (const uint32_t*)a.in, (uint32_t*)a.out, a.idx, a.n, 0,
                                                     a.op == EDPP_OP_SCATTER);
            }
        } else {
            rc = run_candidate(c, *q, cand, nullptr);
        }
    }
    if (rc) return rc;
    return timing_end(c, ms);
}

int edpp_autotune(edpp_ctx* c, const edpp_call* q, int32_t reps, double* times) {
    if (!c || !q || !times || reps < 1) return EDPP_EINVAL;
    int rc = validate_call(c, *q);
    if (rc) return rc;
    const int fam = family_of(q->op);
    std::vector<double> t(reps);
    for (int k = 0; k < K; ++k) {
        times[k] = -1.0;
        if (kCands[k].family != fam) continue;
        double ms;
        rc = edpp_call_run(c, q, k, &ms);  // warm-up (also grows any workspace)
        if (rc) return rc;
        for (int r = 0; r < reps; ++r) {
            rc = edpp_call_run(c, q, k, &t[r]);
            if (rc) return rc;
        }
        std::sort(t.begin(), t.end());
        times[k] = t[reps / 2];
    }
    return EDPP_OK;
}

int edpp_default_candidate(edpp_ctx* c, const edpp_call* q, int32_t* cand) {
    if (!c || !q || !cand) return EDPP_EINVAL;
    int rc = validate_call(c, *q);
    if (rc) return rc;
    const int fam = family_of(q->op);
    int scheme = 0, a = 0, b = 0;
    if (fam == 0) {
        scheme = default_scan_scheme(c, q->n, q->elem_bytes);
        if (scheme == EDPP_SCAN_LOOKBACK) default_scan_shape(c, q->n, &a, &b);
        if (scheme == EDPP_SCAN_PERSIST) a = 1024, b = 8;
    } else if (fam == 1) {
        scheme = default_passes(c, q->n, q->elem_bytes, q->op == EDPP_OP_SCATTER);
        a = c->tune.gs_unroll;
    } else {
        scheme = default_split_strategy(c, q->n, q->buckets);
        default_split_shape(c, q->n, q->buckets, q->elem_bytes, &a, &b);
    }
    *cand = -1;
    for (int k = 0; k < K; ++k) {
        const Cand& d = kCands[k];
        if (d.family != fam || d.scheme != scheme) continue;
        // knobs a scheme does not read match anything
        const bool lb = fam == 0 && (scheme == EDPP_SCAN_LOOKBACK || scheme == EDPP_SCAN_PERSIST);
        const bool ka = fam == 0 ? (!lb || d.knob_a == a) : d.knob_a == a;
        const bool kb = fam == 1 ? true : (fam == 0 ? (!lb || d.knob_b == b) : d.knob_b == b);
        if (ka && kb) {
            *cand = k;
            break;
        }
    }
    return EDPP_OK;
}

int edpp_call_features(edpp_ctx* c, const edpp_call* q, float* d_features) {
    if (!c || !q || !d_features) return EDPP_EINVAL;
    int rc = validate_call(c, *q);
    if (rc) return rc;
    StaticFeat sf = static_features(c, *q);
    hipLaunchKernelGGL(k_features, dim3(1), dim3(256), 0, c->stream, q->op, q->in, q->idx, q->n, lg_of(q->buckets),
                       q->elem_bytes == 8 ? 1 : 0, sf, d_features);
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

int edpp_mlp_num_params(void) { return NPARAM; }

int edpp_mlp_create(edpp_ctx* c, uint64_t seed, edpp_mlp** out) {
    if (!c || !out) return EDPP_EINVAL;
    edpp_mlp* m = new edpp_mlp();
    m->ctx = c;
    const size_t pb = sizeof(float) * NPARAM;
    if (hipMalloc(&m->params, pb) != hipSuccess ||
        hipMalloc(&m->adam_m, pb) != hipSuccess || hipMalloc(&m->adam_v, pb) != hipSuccess ||
        hipMalloc(&m->loss, sizeof(float)) != hipSuccess ||
        hipMalloc(&m->partial, sizeof(float) * (size_t)kGradGroups * (NPARAM + 1)) != hipSuccess ||
        hipHostMalloc((void**)&m->pinned_choice, sizeof(int32_t), hipHostMallocDefault) != hipSuccess) {
        edpp_mlp_destroy(m);
        return fail(c, EDPP_ENOMEM, "mlp allocation");
    }
    // He-uniform initial weights from SplitMix64, zero biases
    std::vector<float> p(NPARAM, 0.0f);
    uint64_t i = 0;
    auto uni = [&](float bound) {
        double u = (double)(splitmix_at(seed, i++) >> 11) * (1.0 / 9007199254740992.0);
        return (float)((2.0 * u - 1.0) * bound);
    };
    for (int k = 0; k < H * F; ++k) p[OFF_W1 + k] = uni(sqrtf(6.0f / F));
    for (int k = 0; k < H * H; ++k) p[OFF_W2 + k] = uni(sqrtf(6.0f / H));
    for (int k = 0; k < K * H; ++k) p[OFF_W3 + k] = uni(sqrtf(6.0f / H));
    int rc = edpp_mlp_set_params(m, p.data(), NPARAM);
    if (rc) {
        edpp_mlp_destroy(m);
        return rc;
    }
    *out = m;
    return EDPP_OK;
}

int edpp_mlp_destroy(edpp_mlp* m) {
    if (!m) return EDPP_OK;
    if (m->ctx && m->ctx->model == m) m->ctx->model = nullptr;
    if (m->ctx) hipStreamSynchronize(m->ctx->stream);
    hipFree(m->params);
    hipFree(m->adam_m);
    hipFree(m->adam_v);
    hipFree(m->loss);
    hipFree(m->losses);
    hipFree(m->partial);
    if (m->pinned_choice) hipHostFree(m->pinned_choice);
    delete m;
    return EDPP_OK;
}

int edpp_mlp_get_params(edpp_mlp* m, float* host, int64_t count) {
    if (!m || !host || count != NPARAM) return EDPP_EINVAL;
    edpp_ctx* c = m->ctx;
    EDPP_HIP(c, hipMemcpyAsync(host, m->params, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, c->stream));
    EDPP_HIP(c, hipStreamSynchronize(c->stream));
    return EDPP_OK;
}

int edpp_mlp_set_params(edpp_mlp* m, const float* host, int64_t count) {
    if (!m || !host || count != NPARAM) return EDPP_EINVAL;
    edpp_ctx* c = m->ctx;
    EDPP_HIP(c, hipMemcpyAsync(m->params, host, sizeof(float) * NPARAM, hipMemcpyHostToDevice, c->stream));
    int zrc = zero_fill(c, m->adam_m, sizeof(float) * NPARAM);
    if (!zrc) zrc = zero_fill(c, m->adam_v, sizeof(float) * NPARAM);
    if (zrc) return zrc;
    EDPP_HIP(c, hipStreamSynchronize(c->stream));
    m->adam_step = 0;
    ++m->version;
    return EDPP_OK;
}

int edpp_mlp_forward(edpp_mlp* m, const float* x, const float* mask, int32_t rows, float* probs, int32_t* choice) {
    if (!m || rows < 0 || (rows > 0 && (!x || !mask))) return EDPP_EINVAL;
    if (rows == 0) return EDPP_OK;
    edpp_ctx* c = m->ctx;
    const unsigned grid = (unsigned)std::min<int64_t>(rows, (int64_t)c->cus * 8);
    hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(64), 0, c->stream, (const float*)m->params, x, mask, rows,
                       probs, choice);
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

// workgroups of k_mlp_grad for `rows` rows (one partial each)
static int grad_groups(int32_t rows) {
    return (int)std::min<int64_t>(((int64_t)rows + kGradRows - 1) / kGradRows, kGradGroups);
}

constexpr int kReduceBlocks = (NPARAM + 1 + kReduceEntries - 1) / kReduceEntries;

static int mlp_grad(edpp_mlp* m, const float* x, const float* t, const float* mask, int32_t rows, float* loss,
                    float* grad) {
    edpp_ctx* c = m->ctx;
    const int groups = grad_groups(rows);
    hipLaunchKernelGGL(k_mlp_grad, dim3(groups), dim3(kGradBlock), 0, c->stream, (const float*)m->params, x, t, mask,
                       rows, m->partial);
    hipLaunchKernelGGL(k_mlp_reduce, dim3(kReduceBlocks), dim3(256), 0, c->stream, (const float*)m->partial, groups,
                       grad, loss);
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

int edpp_mlp_loss_grad(edpp_mlp* m, const float* x, const float* t, const float* mask, int32_t rows, float* loss,
                       float* grad) {
    if (!m || rows <= 0 || !x || !t || !mask || !grad) return EDPP_EINVAL;
    return mlp_grad(m, x, t, mask, rows, loss, grad);
}

int edpp_mlp_train(edpp_mlp* m, const float* x, const float* t, const float* mask, int32_t rows, int32_t epochs,
                   float lr, float* loss_host) {
    if (!m || rows <= 0 || epochs < 0 || !x || !t || !mask) return EDPP_EINVAL;
    edpp_ctx* c = m->ctx;
    ++m->version;
    if (epochs > m->losses_cap) {  // per-epoch losses stay on the device until the end
        EDPP_HIP(c, hipStreamSynchronize(c->stream));
        hipFree(m->losses);
        m->losses = nullptr;
        m->losses_cap = 0;
        if (hipMalloc(&m->losses, sizeof(float) * (size_t)epochs) != hipSuccess) {
            m->losses = nullptr;
            return fail(c, EDPP_ENOMEM, "mlp loss buffer of %d epochs", epochs);
        }
        m->losses_cap = epochs;
    }
    const int groups = grad_groups(rows);
    for (int e = 0; e < epochs; ++e) {
        hipLaunchKernelGGL(k_mlp_grad, dim3(groups), dim3(kGradBlock), 0, c->stream, (const float*)m->params, x, t,
                           mask, rows, m->partial);
        ++m->adam_step;
        const float c1 = 1.0f - powf(0.9f, (float)m->adam_step), c2 = 1.0f - powf(0.999f, (float)m->adam_step);
        hipLaunchKernelGGL(k_mlp_reduce_adam, dim3(kReduceBlocks), dim3(256), 0, c->stream,
                           (const float*)m->partial, groups, m->params, m->adam_m, m->adam_v, lr, c1, c2,
                           m->losses + e);
    }
    EDPP_HIP(c, hipGetLastError());
    if (loss_host && epochs > 0)
        EDPP_HIP(c, hipMemcpyAsync(loss_host, m->losses, sizeof(float) * (size_t)epochs, hipMemcpyDeviceToHost,
                                   c->stream));
    EDPP_HIP(c, hipStreamSynchronize(c->stream));
    return EDPP_OK;
}

int edpp_model_choice(edpp_ctx* c, const edpp_call* q, int32_t* cand) {
    if (!c || !q || !cand) return EDPP_EINVAL;
    if (!c->model) return fail(c, EDPP_EINVAL, "no model set on this context");
    int rc = validate_call(c, *q);
    if (rc) return rc;
    int k = -1;
    rc = model_choose(c, *q, &k);
    if (rc) return rc;
    *cand = k;
    return EDPP_OK;
}

int edpp_mlp_memo(edpp_mlp* m, int32_t enabled) {
    if (!m || enabled < 0 || enabled > 1) return EDPP_EINVAL;
    m->memo_on = enabled != 0;
    for (ChoiceMemo& e : m->memo) e.version = 0;
    return EDPP_OK;
}

int edpp_ctx_set_model(edpp_ctx* c, edpp_mlp* m) {
    if (!c) return EDPP_EINVAL;
    if (m && m->ctx != c) return fail(c, EDPP_EINVAL, "model belongs to another context");
    c->model = m;
    return EDPP_OK;
}

}  // extern "C"
