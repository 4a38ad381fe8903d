// noalloc.hip — the exclusive scan for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  No counterpart in the reference.
//
// The look-back scan's state (two ticket counters and one status word per
// tile) lives in `scratch`, a fixed device array of 512 KiB + 256 B: one copy
// per GPU, held until the process ends.  In host code the array's name is a
// host-side stand-in, so its device address on each GPU is looked up once,
// with that GPU current, when the first context on it is created, and only
// that address is handed to kernels.  Every call on a GPU shares its copy and
// clears the words it uses before the scan: calls that overlap on one GPU
// race.
#include <mutex>
#include "ctx.hpp"

namespace edpp {

namespace {

constexpr int64_t kScratchHeadWords = 32;       // two ticket counters, padded to 256 bytes
constexpr int64_t kScratchStatusWords = 65536;  // look-back words: ceil(n / tile) of them
constexpr int64_t kScratchWords = kScratchHeadWords + kScratchStatusWords;

static __device__ __align__(256) uint64_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint64_t* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

// Called by edpp_ctx_create with the context's device current: looks up the
// device address of this GPU's copy of `scratch` once per device.
int noalloc_setup(edpp_ctx* c) {
    if (c->device < 0 || c->device >= kMaxDevices) return fail(c, EDPP_EUNSUP, "device %d: no scratch slot", c->device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[c->device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(c, e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[c->device] = static_cast<uint64_t*>(p);
    }
    c->noalloc_scratch = g_scratch_ptr[c->device];
    return EDPP_OK;
}

}  // namespace edpp

using namespace edpp;

extern "C" int edpp_scan_exclusive_i32_noalloc(edpp_ctx* c, const int32_t* in, int32_t* out, int64_t n,
                                               int32_t* d_total, double* ms) {
    if (!c) return EDPP_EINVAL;
    if (n < 0 || n >= (1ll << 31)) return fail(c, EDPP_EINVAL, "n=%lld outside [0, 2^31)", (long long)n);
    if (n > 0 && (!in || !out)) return fail(c, EDPP_EINVAL, "null buffer");
    if (!c->noalloc_scratch) return fail(c, EDPP_EHIP, "no-allocation scratch not set up");
    uint64_t* base = c->noalloc_scratch;
    int rc = timing_begin(c, ms);
    if (!rc)
        rc = launch_scan_lookback_u32_on(c, (const uint32_t*)in, (uint32_t*)out, n, (uint32_t*)d_total,
                                         base + kScratchHeadWords, kScratchStatusWords, (uint32_t*)base);
    if (!rc) rc = timing_end(c, ms);
    return rc;
}
