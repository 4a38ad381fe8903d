// ctx.hpp — host-side context shared by all primitives (internal, not ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "edpp.h"
#include "edpp_learn.h"
#include "edpp_tune.h"

struct edpp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    int cus = 0, xcds = 8;
    int64_t l2_bytes = 0, mall_bytes = 0;

    // grow-only scratch for per-call temporaries (histograms, partials)
    void* ws = nullptr;
    size_t ws_bytes = 0;
    // scratch of composite calls (hash join) that hold buffers across the
    // primitives they call, which use ws
    void* ws_outer = nullptr;
    size_t ws_outer_bytes = 0;

    // persistent look-back state of the single-pass scans: tile status words
    // tagged with a per-call epoch, plus two ticket counters used alternately
    uint64_t* scan_status = nullptr;    // u32 scans: packed {epoch, flag, value}
    uint64_t* scan_status64 = nullptr;  // u64 scans: two tagged words per tile
    int64_t scan_cap_tiles = 0;
    uint32_t* tickets = nullptr;        // [2]
    int32_t* d_err = nullptr;           // device error word written by watchdogs
    uint32_t epoch = 1;                 // 30-bit, 0 never used
    // chained split: per-tile, per-bucket status words tagged like the scan's,
    // with epochs of their own
    uint64_t* chain_status = nullptr;
    int64_t chain_cap_words = 0;
    uint32_t chain_epoch = 1;
    uint32_t spin_limit = 1u << 22;     // look-back watchdog (polls before giving up, seconds)

    // this device's copy of the no-allocation scan's fixed scratch array
    uint64_t* noalloc_scratch = nullptr;

    edpp_tuning tune = default_tuning();
    edpp_mlp* model = nullptr;          // when set, AUTO calls ask the cost model
    std::string last_error;

    static edpp_tuning default_tuning() {
        edpp_tuning t = {};
        t.gs_unroll = 4;
        t.gs_random_nt = 0;
        t.scan_vpt = 0;  // by size (default_scan_shape)
        t.scan_block = 1024;
        t.split_ipt = 16;
        t.split_tile_mult = 16;
        return t;
    }
};

namespace edpp {

int fail(edpp_ctx* c, int code, const char* fmt, ...);
int hip_fail(edpp_ctx* c, hipError_t e, const char* what);
#define EDPP_HIP(c, expr)                                                      \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if (_e != hipSuccess) return ::edpp::hip_fail((c), _e, #expr);          \
    } while (0)

// workspace of at least `bytes` (256-byte aligned); may synchronise to grow
int ws_reserve(edpp_ctx* c, size_t bytes, void** out);
int ws_outer_reserve(edpp_ctx* c, size_t bytes, void** out);
int zero_fill(edpp_ctx* c, void* p, size_t bytes);  // kernel, graph-capturable
// look-back state for at least `tiles` tiles; returns the epoch for this call
int scan_state_acquire(edpp_ctx* c, int64_t tiles, uint32_t* epoch_out);
// chained-split status of at least `words` words; returns the first of
// `epochs` consecutive epochs reserved for this call
int chain_state_acquire(edpp_ctx* c, int64_t words, int epochs, uint32_t* epoch_out);

// timing helpers: begin() records the start event when ms != nullptr,
// end() records the stop event, synchronises and writes the elapsed ms
int timing_begin(edpp_ctx* c, double* ms);
int timing_end(edpp_ctx* c, double* ms);
int check_device_error(edpp_ctx* c);

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// launchers implemented by the kernel translation units
int launch_stream_i32(edpp_ctx* c, int op, const int32_t* a, const int32_t* b, int32_t* out, int64_t n);
template <class T>
int launch_scan(edpp_ctx* c, const T* in, T* out, int64_t n, int scheme, T* d_total);
template <class T>
int launch_gather(edpp_ctx* c, const T* in, T* out, const int32_t* idx, int64_t n, int passes, bool scatter);
int noalloc_setup(edpp_ctx* c);
int launch_scan_lookback_u32_on(edpp_ctx* c, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* d_total,
                                uint64_t* status, int64_t status_cap, uint32_t* tickets);
int launch_scan_local(edpp_ctx* c, const int32_t* in, int32_t* out, int tile, int local_size,
                      int matrix_type, int network);
template <class K>
int launch_split(edpp_ctx* c, const K* kin, K* kout, const int32_t* vin, int32_t* vout,
                 int32_t* d_start, int64_t n, int buckets, int strategy);
int launch_sort_u32(edpp_ctx* c, const uint32_t* kin, uint32_t* kout, const int32_t* vin, int32_t* vout, int64_t n,
                    int begin_bit, int end_bit, bool is_signed);
int launch_access_scale(edpp_ctx* c, int pattern, const int32_t* in, int32_t* out, int64_t n, int repeat);
int launch_access_chunks(edpp_ctx* c, const int32_t* in, int64_t n, int64_t chunk, const int32_t* order,
                         unsigned long long* d_count);
int launch_access_random(edpp_ctx* c, int32_t* table, int64_t table_n, int64_t accesses, uint32_t seed, bool write,
                         unsigned long long* d_sum);
int launch_join_count(edpp_ctx* c, const uint32_t* rk, int64_t rn, const uint32_t* sk, int64_t sn,
                      unsigned long long* d_count);
int launch_gen_u32(edpp_ctx* c, uint32_t* out, int64_t n, uint64_t seed, uint32_t mask);
int launch_gen_u64(edpp_ctx* c, uint64_t* out, int64_t n, uint64_t seed, uint64_t mask);
int launch_gen_perm(edpp_ctx* c, int32_t* idx, int64_t n, uint64_t seed, int kind);
template <class K>
int launch_gen_split_keys(edpp_ctx* c, K* keys, int32_t* vals, int64_t n, int buckets, int dist, uint64_t seed);

// schedule candidates and the cost model (learn.hip)
int validate_call(edpp_ctx* c, const edpp_call& q);
int model_choose(edpp_ctx* c, const edpp_call& q, int* cand);
int run_candidate(edpp_ctx* c, const edpp_call& q, int cand, void* aux);

// schedule defaults (overridden by the tuner / cost model)
int default_passes(edpp_ctx* c, int64_t n, int elem_bytes, bool scatter);
int default_scan_scheme(edpp_ctx* c, int64_t n, int elem_bytes);
void default_scan_shape(edpp_ctx* c, int64_t n, int* block, int* vpt);
int default_split_strategy(edpp_ctx* c, int64_t n, int buckets);
void default_split_shape(edpp_ctx* c, int64_t n, int buckets, int key_bytes, int* ipt, int* mult);

}  // namespace edpp
