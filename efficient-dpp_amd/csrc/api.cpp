// api.cpp — context management and the C ABI entry points of include/edpp.h.
//
// Argument checking follows the reference's conventions where it has them
// (opencl/primitives/scanImpl.cpp:19-22 rejects R==L==0;
// opencl/primitives/splitImpl.cpp:29-45 rejects missing SOA values) but every
// failure is a status code: nothing here exits the process.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <new>
#include "ctx.hpp"

namespace edpp {

int fail(edpp_ctx* c, int code, const char* fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        c->last_error = buf;
    }
    return code;
}

int hip_fail(edpp_ctx* c, hipError_t e, const char* what) {
    return fail(c, EDPP_EHIP, "%s: %s", what, hipGetErrorString(e));
}

static int grow(edpp_ctx* c, void** ws, size_t* ws_bytes, size_t bytes, void** out) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes > *ws_bytes) {
        if (*ws) {
            EDPP_HIP(c, hipStreamSynchronize(c->stream));  // earlier kernels may still use it
            EDPP_HIP(c, hipFree(*ws));
            *ws = nullptr;
            *ws_bytes = 0;
        }
        size_t want = std::max(bytes, (size_t)1 << 20);
        if (hipMalloc(ws, want) != hipSuccess) {
            *ws = nullptr;
            return fail(c, EDPP_ENOMEM, "workspace of %zu bytes", want);
        }
        *ws_bytes = want;
    }
    *out = *ws;
    return EDPP_OK;
}
int ws_reserve(edpp_ctx* c, size_t bytes, void** out) { return grow(c, &c->ws, &c->ws_bytes, bytes, out); }
int ws_outer_reserve(edpp_ctx* c, size_t bytes, void** out) {
    return grow(c, &c->ws_outer, &c->ws_outer_bytes, bytes, out);
}

// Zeroes `bytes` (a multiple of 4) of device memory with a kernel on the
// context's stream: a kernel node, unlike a memset node, is safe to capture
// into a graph and replay.
__global__ void k_zero_words(uint32_t* p, int64_t words) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}
int zero_fill(edpp_ctx* c, void* p, size_t bytes) {
    const int64_t words = (int64_t)(bytes / 4);
    if (words == 0) return EDPP_OK;
    const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(words, 256), (int64_t)std::max(c->cus, 1) * 8);
    hipLaunchKernelGGL(k_zero_words, dim3(grid), dim3(256), 0, c->stream, static_cast<uint32_t*>(p), words);
    EDPP_HIP(c, hipGetLastError());
    return EDPP_OK;
}

int scan_state_acquire(edpp_ctx* c, int64_t tiles, uint32_t* epoch_out) {
    if (tiles > c->scan_cap_tiles) {
        int64_t cap = std::max<int64_t>(tiles, 1 << 16);
        if (c->scan_status) {
            EDPP_HIP(c, hipStreamSynchronize(c->stream));
            hipFree(c->scan_status);
            hipFree(c->scan_status64);
            c->scan_status = nullptr;
            c->scan_status64 = nullptr;
        }
        if (hipMalloc(&c->scan_status, cap * sizeof(uint64_t)) != hipSuccess ||
            hipMalloc(&c->scan_status64, cap * 2 * sizeof(uint64_t)) != hipSuccess) {
            c->scan_cap_tiles = 0;
            return fail(c, EDPP_ENOMEM, "look-back state for %lld tiles", (long long)cap);
        }
        int rc = zero_fill(c, c->scan_status, cap * sizeof(uint64_t));
        if (!rc) rc = zero_fill(c, c->scan_status64, cap * 2 * sizeof(uint64_t));
        if (!rc) rc = zero_fill(c, c->tickets, 2 * sizeof(uint32_t));
        if (rc) return rc;
        c->scan_cap_tiles = cap;
        c->epoch = 1;
    }
    if (c->epoch >= (1u << 30) - 1) {
        // epochs are 30-bit: clear the state once per 2^30 calls
        int rc = zero_fill(c, c->scan_status, c->scan_cap_tiles * sizeof(uint64_t));
        if (!rc) rc = zero_fill(c, c->scan_status64, c->scan_cap_tiles * 2 * sizeof(uint64_t));
        if (!rc) rc = zero_fill(c, c->tickets, 2 * sizeof(uint32_t));
        if (rc) return rc;
        c->epoch = 1;
    }
    *epoch_out = c->epoch++;
    return EDPP_OK;
}

int chain_state_acquire(edpp_ctx* c, int64_t words, int epochs, uint32_t* epoch_out) {
    if (words > c->chain_cap_words) {
        const int64_t cap = std::max<int64_t>(words, 1 << 20);
        if (c->chain_status) {
            EDPP_HIP(c, hipStreamSynchronize(c->stream));
            hipFree(c->chain_status);
            c->chain_status = nullptr;
            c->chain_cap_words = 0;
        }
        if (hipMalloc(&c->chain_status, cap * sizeof(uint64_t)) != hipSuccess) {
            c->chain_status = nullptr;
            return fail(c, EDPP_ENOMEM, "chained-split state of %lld words", (long long)cap);
        }
        int rc = zero_fill(c, c->chain_status, cap * sizeof(uint64_t));
        if (rc) return rc;
        c->chain_cap_words = cap;
        c->chain_epoch = 1;
    }
    if (c->chain_epoch + (uint32_t)epochs >= (1u << 30) - 1) {
        // 30-bit epochs of its own (the scan's alternate its ticket counters)
        int rc = zero_fill(c, c->chain_status, c->chain_cap_words * sizeof(uint64_t));
        if (rc) return rc;
        c->chain_epoch = 1;
    }
    *epoch_out = c->chain_epoch;
    c->chain_epoch += (uint32_t)epochs;
    return EDPP_OK;
}

int timing_begin(edpp_ctx* c, double* ms) {
    if (ms) EDPP_HIP(c, hipEventRecord(c->ev_start, c->stream));
    return EDPP_OK;
}

int check_device_error(edpp_ctx* c) {
    int32_t err = 0;
    EDPP_HIP(c, hipMemcpy(&err, c->d_err, sizeof err, hipMemcpyDeviceToHost));
    if (err) {
        int32_t zero = 0;
        EDPP_HIP(c, hipMemcpy(c->d_err, &zero, sizeof zero, hipMemcpyHostToDevice));
        return fail(c, err, "device watchdog reported status %d", err);
    }
    return EDPP_OK;
}

int timing_end(edpp_ctx* c, double* ms) {
    if (!ms) return EDPP_OK;
    EDPP_HIP(c, hipEventRecord(c->ev_stop, c->stream));
    EDPP_HIP(c, hipEventSynchronize(c->ev_stop));
    float f = 0.f;
    EDPP_HIP(c, hipEventElapsedTime(&f, c->ev_start, c->ev_stop));
    *ms = (double)f;
    return check_device_error(c);
}

// ---------------------------------------------------------------- schedules
// Built-in schedules for AUTO calls when no cost model is set, read off the
// autotuner's database on MI355X (2,000 calls, n in [2^16, 2^28]).
int default_passes(edpp_ctx* c, int64_t n, int elem_bytes, bool scatter) {
    (void)c;
    if (!scatter) return 1;  // random reads gain nothing from re-reading the index
    // random writes: a pass per quarter of the Infinity Cache keeps the
    // written slice resident, so lines leave it whole
    const int64_t bytes = n * (int64_t)elem_bytes;
    if (bytes >= (512ll << 20)) return 4;
    if (bytes >= (64ll << 20)) return 2;
    return 1;
}
// The look-back scan is the fastest scheme at every size for both widths
// (2^14..2^28 measured); reduce-then-scan ties it only near 2^21-2^22.
// Above 2^25 elements the persistent look-back wins (3 * 2^24 int32: 78
// against 86 us; 2^26: 101 against 110 us; 2^28: 375 against 421 us, a copy
// takes 370 us); up to 2^25 one workgroup per tile (2^25: 55 against 57 us;
// 2^24: 30.0 against 34.6 us).  The tile shape is default_scan_shape's.
int default_scan_scheme(edpp_ctx* c, int64_t n, int elem_bytes) {
    (void)c; (void)elem_bytes;
    return n > (1ll << 25) ? EDPP_SCAN_PERSIST : EDPP_SCAN_LOOKBACK;
}
// Look-back tile of the scan from a sweep of 2^22..2^28 int32 (1024 lanes;
// larger tiles mean fewer look-backs, which the single workgroup per CU of
// 8 and 16 vectors per lane cannot hide): 4 vectors per lane below 2^21, 8
// below 2^23, 16 above (2^24: 29.6 against 33.4 us; 2^28: 410 against 423
// us).  64-bit scans take 8 where 16 is asked.
void default_scan_shape(edpp_ctx* c, int64_t n, int* block, int* vpt) {
    (void)c;
    *block = 1024;
    *vpt = n < (1ll << 21) ? 4 : (n < (1ll << 23) ? 8 : 16);
}
int default_split_strategy(edpp_ctx* c, int64_t n, int buckets) {
    (void)c; (void)n; (void)buckets;
    return EDPP_SPLIT_WG;
}
// Keys per lane of the WG split, from an autotuner sweep of 2^16..2^28
// pairs: 8 below 2^22 (more workgroups while the call is latency-bound),
// and for up to 16 buckets up to 2^24 and for two-pass bucket counts up to
// 2^26; 16 otherwise (from 128 buckets the 512-lane scatter needs it: B=256
// at 2^28 is 1.33 ms at 16 against 2.31 ms at 8).
void default_split_shape(edpp_ctx* c, int64_t n, int buckets, int key_bytes, int* ipt, int* mult) {
    (void)c;
    (void)key_bytes;
    int lg = 0;
    while ((1 << lg) < buckets) ++lg;
    const bool small = n < (1ll << 22) || (lg <= 4 && n <= (1ll << 24)) || (lg > 8 && n <= (1ll << 26));
    *ipt = small ? 8 : 16;  // wg_ipt() also falls back to 8 when the LDS reorder buffer would not fit
    *mult = 16;
}

}  // namespace edpp

using namespace edpp;

#define EDPP_CHECK_CTX(c) \
    do { if (!(c)) return EDPP_EINVAL; } while (0)

extern "C" {

int edpp_abi_version(void) { return EDPP_ABI_VERSION; }

const char* edpp_strerror(int s) {
    switch (s) {
        case EDPP_OK: return "ok";
        case EDPP_EINVAL: return "invalid argument";
        case EDPP_ENOMEM: return "out of device memory";
        case EDPP_ETIMEOUT: return "device wait timed out";
        case EDPP_EHIP: return "HIP runtime error";
        case EDPP_ENODEV: return "no HIP device";
        case EDPP_EUNSUP: return "unsupported combination";
        default: return "unknown status";
    }
}

const char* edpp_last_error(edpp_ctx* c) { return c ? c->last_error.c_str() : "no context"; }

int edpp_ctx_create(int device, void* stream, edpp_ctx** out) {
    if (!out) return EDPP_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return EDPP_ENODEV;
    if (device < 0 || device >= count) return EDPP_EINVAL;
    edpp_ctx* c = new (std::nothrow) edpp_ctx();
    if (!c) return EDPP_ENOMEM;
    c->device = device;
    int rc = EDPP_OK;
    do {
        if (hipSetDevice(device) != hipSuccess) { rc = EDPP_EHIP; break; }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) != hipSuccess) { rc = EDPP_EHIP; break; }
        c->cus = prop.multiProcessorCount;
        c->l2_bytes = prop.l2CacheSize;
        int xcc = 0;
        if (hipDeviceGetAttribute(&xcc, hipDeviceAttributeNumberOfXccs, device) == hipSuccess && xcc > 0)
            c->xcds = xcc;
        // Infinity Cache of MI355X (not reported by the runtime)
        c->mall_bytes = (int64_t)256 << 20;
        // NULL selects the device's default stream, as HIP and torch do
        c->stream = (hipStream_t)stream;
        if (hipEventCreate(&c->ev_start) != hipSuccess || hipEventCreate(&c->ev_stop) != hipSuccess) {
            rc = EDPP_EHIP;
            break;
        }
        if (hipMalloc(&c->tickets, 64) != hipSuccess || hipMalloc(&c->d_err, 64) != hipSuccess) {
            rc = EDPP_ENOMEM;
            break;
        }
        if (hipMemset(c->tickets, 0, 64) != hipSuccess || hipMemset(c->d_err, 0, 64) != hipSuccess) {
            rc = EDPP_EHIP;
            break;
        }
        rc = noalloc_setup(c);
        if (rc) break;
    } while (0);
    if (rc) {
        edpp_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return EDPP_OK;
}

int edpp_ctx_destroy(edpp_ctx* c) {
    if (!c) return EDPP_OK;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    hipFree(c->ws);
    hipFree(c->ws_outer);
    hipFree(c->scan_status);
    hipFree(c->scan_status64);
    hipFree(c->chain_status);
    hipFree(c->tickets);
    hipFree(c->d_err);
    if (c->ev_start) hipEventDestroy(c->ev_start);
    if (c->ev_stop) hipEventDestroy(c->ev_stop);
    delete c;
    return EDPP_OK;
}

int edpp_set_tuning(edpp_ctx* c, const edpp_tuning* t) {
    EDPP_CHECK_CTX(c);
    edpp_tuning d = edpp_ctx::default_tuning();
    if (!t) {
        c->tune = d;
        return EDPP_OK;
    }
    edpp_tuning v = *t;
    auto pick = [](int32_t x, int32_t dflt) { return x == 0 ? dflt : x; };
    v.gs_unroll = pick(v.gs_unroll, d.gs_unroll);
    v.gs_random_nt = v.gs_random_nt < 0 ? d.gs_random_nt : v.gs_random_nt;
    v.scan_vpt = pick(v.scan_vpt, d.scan_vpt);
    v.split_ipt = pick(v.split_ipt, d.split_ipt);
    v.split_tile_mult = pick(v.split_tile_mult, d.split_tile_mult);
    v.scan_block = pick(v.scan_block, d.scan_block);
    if ((v.gs_unroll != 1 && v.gs_unroll != 2 && v.gs_unroll != 4 && v.gs_unroll != 8) || v.gs_random_nt > 1 ||
        (v.scan_vpt != 0 && v.scan_vpt != 2 && v.scan_vpt != 4 && v.scan_vpt != 8 && v.scan_vpt != 16) ||
        (v.split_ipt != 8 && v.split_ipt != 16) ||
        v.split_tile_mult < 16 || v.split_tile_mult > 1024 ||
        (v.scan_block != 256 && v.scan_block != 512 && v.scan_block != 1024) ||
        (v.split_block != 0 && v.split_block != 256 && v.split_block != 512) ||
        v.split_reorder < EDPP_REORDER_DEFAULT || v.split_reorder > EDPP_REORDER_VARIED ||
        (v.sort_strategy != 0 && v.sort_strategy != EDPP_SPLIT_WG && v.sort_strategy != EDPP_SPLIT_CHAINED))
        return fail(c, EDPP_EINVAL, "tuning value out of range");
    c->tune = v;
    return EDPP_OK;
}

int edpp_get_tuning(edpp_ctx* c, edpp_tuning* t) {
    EDPP_CHECK_CTX(c);
    if (!t) return EDPP_EINVAL;
    *t = c->tune;
    return EDPP_OK;
}

int edpp_ctx_sync(edpp_ctx* c) {
    EDPP_CHECK_CTX(c);
    EDPP_HIP(c, hipStreamSynchronize(c->stream));
    return check_device_error(c);
}

int edpp_ctx_set_watchdog(edpp_ctx* c, uint32_t polls) {
    EDPP_CHECK_CTX(c);
    c->spin_limit = polls;
    return EDPP_OK;
}

void* edpp_ctx_stream(edpp_ctx* c) { return c ? (void*)c->stream : nullptr; }

int edpp_ctx_device_info(edpp_ctx* c, int* cus, int* xcds, int64_t* l2, int64_t* mall) {
    EDPP_CHECK_CTX(c);
    if (cus) *cus = c->cus;
    if (xcds) *xcds = c->xcds;
    if (l2) *l2 = c->l2_bytes;
    if (mall) *mall = c->mall_bytes;
    return EDPP_OK;
}

// ---------------------------------------------------------------- STREAM
int edpp_stream_i32(edpp_ctx* c, int op, const int32_t* a, const int32_t* b, int32_t* out, int64_t n, double* ms) {
    EDPP_CHECK_CTX(c);
    if (n < 0 || (n > 0 && (!a || !out))) return fail(c, EDPP_EINVAL, "stream: null pointer or n < 0");
    if ((op == EDPP_STREAM_ADD || op == EDPP_STREAM_TRIAD) && n > 0 && !b)
        return fail(c, EDPP_EINVAL, "stream add/triad need b");
    if (op < EDPP_STREAM_COPY || op > EDPP_STREAM_TRIAD) return fail(c, EDPP_EINVAL, "unknown stream op %d", op);
    int rc = timing_begin(c, ms);
    if (!rc) rc = launch_stream_i32(c, op, a, b, out, n);
    if (!rc) rc = timing_end(c, ms);
    return rc;
}

// ---------------------------------------------------------------- gather / scatter
}  // extern "C"

template <class T>
static int gs_common(edpp_ctx* c, const T* in, T* out, const int32_t* idx, int64_t n, int passes, double* ms,
                     bool scatter) {
    EDPP_CHECK_CTX(c);
    if (n < 0 || n >= (1ll << 31)) return fail(c, EDPP_EINVAL, "n=%lld outside [0, 2^31)", (long long)n);
    if (n > 0 && (!in || !out || !idx)) return fail(c, EDPP_EINVAL, "null buffer");
    if (passes < 0) return fail(c, EDPP_EINVAL, "pass count %d < 0", passes);
    if (n > 0 && (const void*)in == (const void*)out) return fail(c, EDPP_EINVAL, "in and out must differ");
    int rc = timing_begin(c, ms);
    if (!rc) {
        if (passes == 0 && c->model && n > 0) {
            edpp_call q = {};
            q.op = scatter ? EDPP_OP_SCATTER : EDPP_OP_GATHER;
            q.elem_bytes = (int32_t)sizeof(T);
            q.n = n;
            q.in = in;
            q.out = out;
            q.idx = idx;
            int k = 0;
            rc = model_choose(c, q, &k);
            if (!rc) rc = run_candidate(c, q, k, nullptr);
        } else {
            rc = launch_gather<T>(c, in, out, idx, n, passes, scatter);
        }
    }
    if (!rc) rc = timing_end(c, ms);
    return rc;
}

extern "C" {

int edpp_gather_i32(edpp_ctx* c, const int32_t* in, int32_t* out, const int32_t* idx, int64_t n, int passes,
                    double* ms) {
    return gs_common<uint32_t>(c, (const uint32_t*)in, (uint32_t*)out, idx, n, passes, ms, false);
}
int edpp_scatter_i32(edpp_ctx* c, const int32_t* in, int32_t* out, const int32_t* idx, int64_t n, int passes,
                     double* ms) {
    return gs_common<uint32_t>(c, (const uint32_t*)in, (uint32_t*)out, idx, n, passes, ms, true);
}
int edpp_gather_i64(edpp_ctx* c, const int64_t* in, int64_t* out, const int32_t* idx, int64_t n, int passes,
                    double* ms) {
    return gs_common<uint64_t>(c, (const uint64_t*)in, (uint64_t*)out, idx, n, passes, ms, false);
}
int edpp_scatter_i64(edpp_ctx* c, const int64_t* in, int64_t* out, const int32_t* idx, int64_t n, int passes,
                     double* ms) {
    return gs_common<uint64_t>(c, (const uint64_t*)in, (uint64_t*)out, idx, n, passes, ms, true);
}
int edpp_default_passes(edpp_ctx* c, int64_t n, int elem_bytes, int scatter, int* passes) {
    EDPP_CHECK_CTX(c);
    if (!passes || n < 0 || (elem_bytes != 4 && elem_bytes != 8)) return fail(c, EDPP_EINVAL, "edpp_default_passes: bad argument");
    *passes = default_passes(c, n, elem_bytes, scatter != 0);
    return EDPP_OK;
}

// ---------------------------------------------------------------- scans
}  // extern "C"

template <class T>
static int scan_common(edpp_ctx* c, const T* in, T* out, int64_t n, int scheme, T* d_total, double* ms) {
    EDPP_CHECK_CTX(c);
    if (n < 0 || n >= (1ll << 31)) return fail(c, EDPP_EINVAL, "n=%lld outside [0, 2^31)", (long long)n);
    if (n > 0 && (!in || !out)) return fail(c, EDPP_EINVAL, "null buffer");
    if (scheme < EDPP_SCAN_AUTO || scheme > EDPP_SCAN_PERSIST) return fail(c, EDPP_EINVAL, "unknown scheme %d", scheme);
    int rc = timing_begin(c, ms);
    if (!rc) {
        if (scheme == EDPP_SCAN_AUTO && c->model && n > 0) {
            edpp_call q = {};
            q.op = EDPP_OP_SCAN;
            q.elem_bytes = (int32_t)sizeof(T);
            q.n = n;
            q.in = in;
            q.out = out;
            int k = 0;
            rc = model_choose(c, q, &k);
            if (!rc) rc = run_candidate(c, q, k, d_total);
        } else {
            rc = launch_scan<T>(c, in, out, n, scheme, d_total);
        }
    }
    if (!rc) rc = timing_end(c, ms);
    return rc;
}

extern "C" {

int edpp_scan_exclusive_i32(edpp_ctx* c, const int32_t* in, int32_t* out, int64_t n, int scheme, int32_t* d_total,
                            double* ms) {
    return scan_common<uint32_t>(c, (const uint32_t*)in, (uint32_t*)out, n, scheme, (uint32_t*)d_total, ms);
}
int edpp_scan_exclusive_i64(edpp_ctx* c, const int64_t* in, int64_t* out, int64_t n, int scheme, int64_t* d_total,
                            double* ms) {
    return scan_common<uint64_t>(c, (const uint64_t*)in, (uint64_t*)out, n, scheme, (uint64_t*)d_total, ms);
}

int edpp_scan_local_i32(edpp_ctx* c, const int32_t* in, int32_t* out, int tile_size, int local_size, int matrix_type,
                        int network, double* ms) {
    EDPP_CHECK_CTX(c);
    if (!in || !out) return fail(c, EDPP_EINVAL, "null buffer");
    int rc = timing_begin(c, ms);
    if (!rc) rc = launch_scan_local(c, in, out, tile_size, local_size, matrix_type, network);
    if (!rc) rc = timing_end(c, ms);
    return rc;
}

// ---------------------------------------------------------------- split
}  // extern "C"

template <class K>
static int split_common(edpp_ctx* c, const K* kin, K* kout, const int32_t* vin, int32_t* vout, int32_t* d_start,
                        int64_t n, int buckets, int strategy, double* ms) {
    EDPP_CHECK_CTX(c);
    if (n < 0 || n >= (1ll << 31)) return fail(c, EDPP_EINVAL, "n=%lld outside [0, 2^31)", (long long)n);
    if (buckets < 1 || buckets > 4096 || (buckets & (buckets - 1)))
        return fail(c, EDPP_EINVAL, "buckets=%d must be a power of two in [1,4096]", buckets);
    if (n > 0 && (!kin || !kout)) return fail(c, EDPP_EINVAL, "null key buffer");
    if ((vin == nullptr) != (vout == nullptr))
        return fail(c, EDPP_EINVAL, "values: both or neither of vals_in/vals_out");  // splitImpl.cpp:29-34
    if (n > 0 && (const void*)kin == (const void*)kout) return fail(c, EDPP_EINVAL, "keys in and out must differ");
    if (n > 0 && vin && (const void*)vin == (const void*)vout) return fail(c, EDPP_EINVAL, "vals in and out must differ");
    if (strategy < EDPP_SPLIT_AUTO || strategy > EDPP_SPLIT_CHAINED)
        return fail(c, EDPP_EINVAL, "unknown split strategy %d", strategy);
    int rc = timing_begin(c, ms);
    if (!rc) {
        if (strategy == EDPP_SPLIT_AUTO && c->model && n > 0) {
            edpp_call q = {};
            q.op = EDPP_OP_SPLIT;
            q.elem_bytes = (int32_t)sizeof(K);
            q.n = n;
            q.buckets = buckets;
            q.in = kin;
            q.out = kout;
            q.vals_in = vin;
            q.vals_out = vout;
            int k = 0;
            rc = model_choose(c, q, &k);
            if (!rc) rc = run_candidate(c, q, k, d_start);
        } else {
            rc = launch_split<K>(c, kin, kout, vin, vout, d_start, n, buckets, strategy);
        }
    }
    if (!rc) rc = timing_end(c, ms);
    return rc;
}

extern "C" {

int edpp_split_i32(edpp_ctx* c, const int32_t* kin, int32_t* kout, const int32_t* vin, int32_t* vout,
                   int32_t* d_start, int64_t n, int buckets, int strategy, double* ms) {
    return split_common<uint32_t>(c, (const uint32_t*)kin, (uint32_t*)kout, vin, vout, d_start, n, buckets, strategy,
                                  ms);
}
int edpp_split_i64(edpp_ctx* c, const int64_t* kin, int64_t* kout, const int32_t* vin, int32_t* vout,
                   int32_t* d_start, int64_t n, int buckets, int strategy, double* ms) {
    return split_common<uint64_t>(c, (const uint64_t*)kin, (uint64_t*)kout, vin, vout, d_start, n, buckets, strategy,
                                  ms);
}

// ---------------------------------------------------------------- radix sort
static int sort_common(edpp_ctx* c, const uint32_t* kin, uint32_t* kout, const int32_t* vin, int32_t* vout,
                       int64_t n, int begin_bit, int end_bit, bool is_signed, double* ms) {
    EDPP_CHECK_CTX(c);
    if (n < 0 || n >= (1ll << 31)) return fail(c, EDPP_EINVAL, "n=%lld outside [0, 2^31)", (long long)n);
    if (n > 0 && (!kin || !kout)) return fail(c, EDPP_EINVAL, "null key buffer");
    if ((vin == nullptr) != (vout == nullptr))
        return fail(c, EDPP_EINVAL, "values: both or neither of vals_in/vals_out");
    if (begin_bit < 0 || end_bit > 32 || begin_bit > end_bit)
        return fail(c, EDPP_EINVAL, "key bits [%d, %d) outside [0, 32)", begin_bit, end_bit);
    int rc = timing_begin(c, ms);
    if (!rc) rc = launch_sort_u32(c, kin, kout, vin, vout, n, begin_bit, end_bit, is_signed);
    if (!rc) rc = timing_end(c, ms);
    return rc;
}
int edpp_sort_u32(edpp_ctx* c, const uint32_t* keys_in, uint32_t* keys_out, const int32_t* vals_in, int32_t* vals_out,
                  int64_t n, int begin_bit, int end_bit, double* ms) {
    return sort_common(c, keys_in, keys_out, vals_in, vals_out, n, begin_bit, end_bit, false, ms);
}
int edpp_sort_i32(edpp_ctx* c, const int32_t* keys_in, int32_t* keys_out, const int32_t* vals_in, int32_t* vals_out,
                  int64_t n, int begin_bit, int end_bit, double* ms) {
    return sort_common(c, (const uint32_t*)keys_in, (uint32_t*)keys_out, vals_in, vals_out, n, begin_bit, end_bit, true,
                       ms);
}

// ---------------------------------------------------------------- access probes
int edpp_access_scale_i32(edpp_ctx* c, int pattern, const int32_t* in, int32_t* out, int64_t n, int repeat,
                          double* ms) {
    EDPP_CHECK_CTX(c);
    if (pattern < EDPP_ACCESS_COLUMN || pattern > EDPP_ACCESS_MIXED)
        return fail(c, EDPP_EINVAL, "unknown access pattern %d", pattern);
    if (repeat < 1 || n < 0 || n % ((int64_t)repeat * 64) != 0)
        return fail(c, EDPP_EINVAL, "n=%lld must be a multiple of 64*repeat (repeat=%d)", (long long)n, repeat);
    if (n > 0 && (!in || !out)) return fail(c, EDPP_EINVAL, "null buffer");
    int rc = timing_begin(c, ms);
    if (!rc) rc = launch_access_scale(c, pattern, in, out, n, repeat);
    if (!rc) rc = timing_end(c, ms);
    return rc;
}
int edpp_access_chunks_i32(edpp_ctx* c, const int32_t* in, int64_t n, int64_t chunk, const int32_t* order,
                           int64_t* d_count, double* ms) {
    EDPP_CHECK_CTX(c);
    if (n < 0 || chunk < 1 || !d_count || (n > 0 && (!in || !order)))
        return fail(c, EDPP_EINVAL, "access chunks: bad argument");
    if (ceil_div(n, chunk) >= (1ll << 31)) return fail(c, EDPP_EINVAL, "too many chunks");
    int rc = timing_begin(c, ms);
    if (!rc) rc = launch_access_chunks(c, in, n, chunk, order, (unsigned long long*)d_count);
    if (!rc) rc = timing_end(c, ms);
    return rc;
}

int edpp_access_random_i32(edpp_ctx* c, int32_t* table, int64_t table_n, int64_t accesses, uint64_t seed, int write,
                           int64_t* d_sum, double* ms) {
    EDPP_CHECK_CTX(c);
    if (table_n < 1 || table_n > (1ll << 31) || (table_n & (table_n - 1)))
        return fail(c, EDPP_EINVAL, "table_n=%lld must be a power of two in [1, 2^31]", (long long)table_n);
    if (accesses < 0 || accesses >= (1ll << 32))
        return fail(c, EDPP_EINVAL, "accesses=%lld outside [0, 2^32)", (long long)accesses);
    if (!table || (!write && !d_sum)) return fail(c, EDPP_EINVAL, "access random: null buffer");
    int rc = timing_begin(c, ms);
    if (!rc)
        rc = launch_access_random(c, table, table_n, accesses, (uint32_t)seed, write != 0,
                                  (unsigned long long*)d_sum);
    if (!rc) rc = timing_end(c, ms);
    return rc;
}

// ---------------------------------------------------------------- hash join
int edpp_hash_join_count(edpp_ctx* c, const int32_t* r_keys, int64_t r_n, const int32_t* s_keys, int64_t s_n,
                         uint64_t* d_count, double* ms) {
    EDPP_CHECK_CTX(c);
    if (r_n < 0 || s_n < 0 || r_n >= (1ll << 31) || s_n >= (1ll << 31))
        return fail(c, EDPP_EINVAL, "relation sizes outside [0, 2^31)");
    if (!d_count || (r_n > 0 && !r_keys) || (s_n > 0 && !s_keys)) return fail(c, EDPP_EINVAL, "null buffer");
    int rc = timing_begin(c, ms);
    if (!rc)
        rc = launch_join_count(c, (const uint32_t*)r_keys, r_n, (const uint32_t*)s_keys, s_n,
                               (unsigned long long*)d_count);
    if (!rc) rc = timing_end(c, ms);
    return rc;
}

// ---------------------------------------------------------------- generators
int edpp_gen_i32(edpp_ctx* c, int32_t* out, int64_t n, uint64_t seed, int32_t mask) {
    EDPP_CHECK_CTX(c);
    if (n < 0 || (n > 0 && !out)) return fail(c, EDPP_EINVAL, "gen: bad buffer");
    return launch_gen_u32(c, (uint32_t*)out, n, seed, (uint32_t)mask);
}
int edpp_gen_i64(edpp_ctx* c, int64_t* out, int64_t n, uint64_t seed, int64_t mask) {
    EDPP_CHECK_CTX(c);
    if (n < 0 || (n > 0 && !out)) return fail(c, EDPP_EINVAL, "gen: bad buffer");
    return launch_gen_u64(c, (uint64_t*)out, n, seed, (uint64_t)mask);
}
int edpp_gen_perm(edpp_ctx* c, int32_t* idx, int64_t n, uint64_t seed, int kind) {
    EDPP_CHECK_CTX(c);
    if (n < 0 || n >= (1ll << 31) || (n > 0 && !idx)) return fail(c, EDPP_EINVAL, "gen_perm: bad buffer or n");
    return launch_gen_perm(c, idx, n, seed, kind);
}
int edpp_gen_split_keys_i32(edpp_ctx* c, int32_t* keys, int32_t* vals, int64_t n, int buckets, int dist,
                            uint64_t seed) {
    EDPP_CHECK_CTX(c);
    if (n < 0 || n >= (1ll << 31) || (n > 0 && !keys)) return fail(c, EDPP_EINVAL, "gen_split_keys: bad buffer");
    if (buckets < 1 || buckets > 4096 || (buckets & (buckets - 1))) return fail(c, EDPP_EINVAL, "bad bucket count");
    return launch_gen_split_keys<uint32_t>(c, (uint32_t*)keys, vals, n, buckets, dist, seed);
}
int edpp_gen_split_keys_i64(edpp_ctx* c, int64_t* keys, int32_t* vals, int64_t n, int buckets, int dist,
                            uint64_t seed) {
    EDPP_CHECK_CTX(c);
    if (n < 0 || n >= (1ll << 31) || (n > 0 && !keys)) return fail(c, EDPP_EINVAL, "gen_split_keys: bad buffer");
    if (buckets < 1 || buckets > 4096 || (buckets & (buckets - 1))) return fail(c, EDPP_EINVAL, "bad bucket count");
    return launch_gen_split_keys<uint64_t>(c, (uint64_t*)keys, vals, n, buckets, dist, seed);
}

}  // extern "C"
