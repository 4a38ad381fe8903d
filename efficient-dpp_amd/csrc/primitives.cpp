// primitives.cpp — the reference's C++ host API (include/edpp_primitives.hpp)
// over the C ABI.  Every function validates its arguments the way the
// reference does, runs synchronously on the process-wide context and returns
// device milliseconds.
#include <hip/hip_runtime.h>
#include <mutex>
#include "edpp_primitives.hpp"
#include "edpp_tune.h"

namespace edpp {

namespace {
std::once_flag g_once;
edpp_ctx* g_ctx = nullptr;

double ms_or_fail(int rc, double ms) { return rc == EDPP_OK ? ms : -1.0; }
}  // namespace

edpp_ctx* context() {
    std::call_once(g_once, [] {
        if (edpp_ctx_create(0, nullptr, &g_ctx) != EDPP_OK) g_ctx = nullptr;
    });
    return g_ctx;
}

double gather(const int* d_in, int* d_out, int length, const int* d_loc, int localSize, int gridSize, int pass) {
    (void)localSize;
    (void)gridSize;
    edpp_ctx* c = context();
    if (!c) return -1.0;
    double ms = 0.0;
    return ms_or_fail(edpp_gather_i32(c, d_in, d_out, d_loc, length, pass < 1 ? 1 : pass, &ms), ms);
}

double scatter(const int* d_in, int* d_out, int length, const int* d_loc, int localSize, int gridSize, int pass) {
    (void)localSize;
    (void)gridSize;
    edpp_ctx* c = context();
    if (!c) return -1.0;
    double ms = 0.0;
    return ms_or_fail(edpp_scatter_i32(c, d_in, d_out, d_loc, length, pass < 1 ? 1 : pass, &ms), ms);
}

double scan_chained(const int* d_in, int* d_out, int length, int localSize, int gridSize, int R, int L) {
    (void)localSize;
    (void)gridSize;
    if (R == 0 && L == 0) return 1;  // opencl/primitives/scanImpl.cpp:19-22
    edpp_ctx* c = context();
    if (!c) return -1.0;
    double ms = 0.0;
    return ms_or_fail(edpp_scan_exclusive_i32(c, d_in, d_out, length, EDPP_SCAN_LOOKBACK, nullptr, &ms), ms);
}

double scan_RSS(const int* d_in, int* d_out, unsigned length, int local_size, int grid_size) {
    (void)local_size;
    (void)grid_size;
    edpp_ctx* c = context();
    if (!c) return -1.0;
    double ms = 0.0;
    return ms_or_fail(edpp_scan_exclusive_i32(c, d_in, d_out, length, EDPP_SCAN_RTS, nullptr, &ms), ms);
}

double scan_RSS_single(const int* d_in, int* d_out, unsigned length) {
    edpp_ctx* c = context();
    if (!c) return -1.0;
    double ms = 0.0;
    return ms_or_fail(edpp_scan_exclusive_i32(c, d_in, d_out, length, EDPP_SCAN_SSA, nullptr, &ms), ms);
}

double matrix_scan(const int* d_in, int* d_out, int tile_size, MatrixScanType type, int local_size) {
    edpp_ctx* c = context();
    if (!c) return -1.0;
    double ms = 0.0;
    return ms_or_fail(edpp_scan_local_i32(c, d_in, d_out, tile_size, local_size, (int)type, EDPP_NET_WAVE, &ms), ms);
}

namespace {

// KO and KVS_SOA map to the key/value split; KVS_AOS holds (key, value) int
// pairs, which split as 64-bit elements whose low word is the key.
double split_common(const int* d_in, int* d_out, int* d_start, int length, int buckets, DataStruc structure,
                    const int* d_in_values, int* d_out_values, int strategy) {
    if (structure == KVS_SOA && (!d_in_values || !d_out_values)) return -1;  // splitImpl.cpp:29-34
    if (structure != KVS_SOA && (d_in_values || d_out_values)) return -1;
    edpp_ctx* c = context();
    if (!c) return -1.0;
    double ms = 0.0;
    int rc;
    if (structure == KVS_AOS)
        rc = edpp_split_i64(c, reinterpret_cast<const int64_t*>(d_in), reinterpret_cast<int64_t*>(d_out), nullptr,
                            nullptr, d_start, length, buckets, strategy, &ms);
    else
        rc = edpp_split_i32(c, d_in, d_out, structure == KVS_SOA ? d_in_values : nullptr,
                            structure == KVS_SOA ? d_out_values : nullptr, d_start, length, buckets, strategy, &ms);
    return ms_or_fail(rc, ms);
}

}  // namespace

double WI_split(const int* d_in, int* d_out, int* d_start, int length, int buckets, DataStruc structure,
                const int* d_in_values, int* d_out_values, int local_size, int grid_size) {
    (void)local_size;
    (void)grid_size;
    return split_common(d_in, d_out, d_start, length, buckets, structure, d_in_values, d_out_values, EDPP_SPLIT_WI);
}

double WG_split(const int* d_in, int* d_out, int* d_start, int length, int buckets, ReorderType reorder_type,
                DataStruc structure, const int* d_in_values, int* d_out_values, int local_size, int grid_size) {
    (void)local_size;
    (void)grid_size;
    // the reorder type picks how the WG split writes its tiles; every variant
    // gives the same stable output
    edpp_ctx* c = context();
    if (!c) return -1.0;
    if (reorder_type != NO_REORDER && reorder_type != FIXED_REORDER && reorder_type != VARIED_REORDER) return -1;
    edpp_tuning saved, t;
    edpp_get_tuning(c, &saved);
    t = saved;
    t.split_reorder = reorder_type == NO_REORDER ? EDPP_REORDER_NONE
                      : reorder_type == FIXED_REORDER ? EDPP_REORDER_FIXED : EDPP_REORDER_VARIED;
    if (edpp_set_tuning(c, &t) != EDPP_OK) return -1.0;
    const double ms =
        split_common(d_in, d_out, d_start, length, buckets, structure, d_in_values, d_out_values, EDPP_SPLIT_WG);
    edpp_set_tuning(c, &saved);
    return ms;
}

double single_split(const int* d_in, int* d_out, int length, int buckets, bool reorder, DataStruc structure) {
    (void)reorder;
    if (structure == KVS_SOA) return -1;  // splitImpl.cpp:376-379: SOA not supported
    return split_common(d_in, d_out, nullptr, length, buckets, structure, nullptr, nullptr, EDPP_SPLIT_SINGLE);
}

double radixSort(int* d_source_keys, int* d_source_values, int len, bool isRecord) {
    edpp_ctx* c = context();
    if (!c) return -1.0;
    double ms = 0.0;
    if (isRecord)
        return ms_or_fail(edpp_sort_i32(c, d_source_keys, d_source_keys, d_source_values, d_source_values, len, 0, 32,
                                        &ms),
                          ms);
    return ms_or_fail(edpp_sort_i32(c, d_source_values, d_source_values, nullptr, nullptr, len, 0, 32, &ms), ms);
}

double hashjoin(const int* d_R_keys, const int* d_R_values, int rLen, const int* d_S_keys, const int* d_S_values,
                int sLen, int& res_len) {
    (void)d_R_values;
    (void)d_S_values;
    edpp_ctx* c = context();
    if (!c) return -1.0;
    static std::once_flag once;
    static uint64_t* d_count = nullptr;  // the process-wide result word, like the context
    std::call_once(once, [] {
        if (hipMalloc(&d_count, sizeof(uint64_t)) != hipSuccess) d_count = nullptr;
    });
    if (!d_count) return -1.0;
    double ms = 0.0;
    const int rc = edpp_hash_join_count(c, d_R_keys, rLen, d_S_keys, sLen, d_count, &ms);
    uint64_t h = 0;
    if (rc == EDPP_OK && hipMemcpy(&h, d_count, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return -1.0;
    res_len = (int)h;  // the reference's int result
    return ms_or_fail(rc, ms);
}

}  // namespace edpp
