/*
 * edpp.h — C ABI of the MI355X (gfx950) data-parallel primitives library.
 *
 * Every entry point takes device pointers that the caller allocated, plain
 * sizes and enums, and returns an int status (EDPP_OK == 0).  No C++ types,
 * no exceptions and no torch types cross this boundary.  Work is enqueued on
 * the context's HIP stream; when `ms` is non-NULL the call synchronises and
 * stores the summed device time of its kernels in milliseconds, which is the
 * value the reference's primitives return (opencl/util/utility.cpp:99-104,
 * opencl/primitives/gatherImpl.cpp:52).
 *
 * Reference interfaces replaced (file:line in RapidsAtHKUST/Efficient-DPP):
 *   gather            opencl/primitives.h:12-14   -> edpp_gather_i32
 *   scatter           opencl/primitives.h:17-19   -> edpp_scatter_i32
 *   scan_chained      opencl/primitives.h:22-24   -> edpp_scan_exclusive_i32 (EDPP_SCAN_LOOKBACK)
 *   scan_RSS          opencl/primitives.h:26-27   -> edpp_scan_exclusive_i32 (EDPP_SCAN_RTS)
 *   scan_RSS_single   opencl/primitives.h:29      -> edpp_scan_exclusive_i32 (EDPP_SCAN_SSA)
 *   matrix scans      opencl/kernels/scan_local_kernel.cl:530-695,
 *                     opencl/test/test_scan_local.cpp:6-106 -> edpp_scan_local_i32
 *   WI_split          opencl/primitives.h:32-37   -> edpp_split_i32 (EDPP_SPLIT_WI)
 *   WG_split          opencl/primitives.h:39-44   -> edpp_split_i32 (EDPP_SPLIT_WG*)
 *   single_split      opencl/primitives.h:46-49   -> edpp_split_i32 (EDPP_SPLIT_SINGLE)
 *   radixSort         obsolete/CUDA/inc/kernels.h:37-45 -> edpp_sort_u32 / edpp_sort_i32
 *   hashjoin          obsolete/OpenCL/phjImpl.cpp:14-66 -> edpp_hash_join_count
 *   STREAM kernels    opencl/kernels/mem_kernel.cl:9-31 -> edpp_stream_i32
 *   access probes     opencl/kernels/mem_kernel.cl:33-86 -> edpp_access_scale_i32 / edpp_access_chunks_i32
 *   data generators   opencl/util/utility.cpp:214-241 -> edpp_gen_* (reproducible)
 *
 * All results are bit-exact with the sequential definitions:
 *   scan:    out[i] = sum_{j<i} in[j]  (two's-complement wrap-around)
 *   gather:  out[i] = in[idx[i]]       scatter: out[idx[i]] = in[i]
 *   split:   stable counting sort by (key & (buckets-1)); start[b] = #keys in buckets < b
 * and they do not depend on the schedule (scheme, pass count, tile) chosen.
 */
#ifndef EDPP_H
#define EDPP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EDPP_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
enum edpp_status {
    EDPP_OK       = 0,
    EDPP_EINVAL   = -1,  /* bad argument (null pointer, n<0, n>=2^31, B not a power of two, ...) */
    EDPP_ENOMEM   = -2,  /* workspace allocation failed                                            */
    EDPP_ETIMEOUT = -3,  /* a bounded device wait ran out (look-back watchdog)                     */
    EDPP_EHIP     = -4,  /* a HIP runtime call failed                                              */
    EDPP_ENODEV   = -5,  /* no HIP device                                                           */
    EDPP_EUNSUP   = -6   /* combination not supported                                               */
};

/* ---- enums mirroring opencl/types.h:8-46 -------------------------------- */
enum edpp_scan_scheme {          /* SCAN_GLOBAL_TYPE, opencl/types.h:8-10 */
    EDPP_SCAN_AUTO     = 0,
    EDPP_SCAN_LOOKBACK = 1,      /* single pass, decoupled look-back (replaces CHAINED)            */
    EDPP_SCAN_RTS      = 2,      /* reduce -> scan partials -> scan tiles with offsets (RSS)       */
    EDPP_SCAN_SSA      = 3,      /* scan tiles -> scan partials -> add offsets                     */
    EDPP_SCAN_PERSIST  = 4       /* the look-back on persistent workgroups, each loading its next
                                    tile while the current one looks back (large calls)             */
};

enum edpp_matrix_scan {          /* MatrixScanType, opencl/types.h:12-14 */
    EDPP_MATRIX_LM        = 0,   /* tile staged in LDS, per-thread rows                           */
    EDPP_MATRIX_REG       = 1,   /* rows loaded straight into registers (strided)                 */
    EDPP_MATRIX_LM_REG    = 2,   /* coalesced load to LDS, rows to registers, back through LDS    */
    EDPP_MATRIX_LM_SERIAL = 3    /* one lane scans the whole tile                                 */
};

enum edpp_local_network {        /* intra-workgroup scan of the per-thread partials,
                                    opencl/kernels/scan_local_kernel.cl:42-474 */
    EDPP_NET_SERIAL          = 0, /* lane 0 walks the partials (the reference's active choice, :504)   */
    EDPP_NET_BLELLOCH        = 1, /* work-efficient up/down sweep in LDS (:103-137)                      */
    EDPP_NET_SKLANSKY        = 2, /* log-depth Sklansky in LDS (:181-211)                                */
    EDPP_NET_WAVE            = 3, /* 64-lane DPP wave scan + LDS across waves (the warp scan, :214-259)  */
    EDPP_NET_SERIAL_MT       = 4, /* serial runs on 128 lanes, lane 0 scans their sums (:57-100)         */
    EDPP_NET_BLELLOCH_CF     = 5, /* Blelloch on a padded LDS copy, one pad word per 32 (:140-178)       */
    EDPP_NET_GLOBAL_BLELLOCH = 6, /* Blelloch on a global-memory buffer, the output's head (:437-474)    */
    EDPP_NET_GLOBAL_SKLANSKY = 7, /* Sklansky on a global-memory buffer (:403-435)                       */
    EDPP_NET_GLOBAL_WAVE     = 8  /* wave scans, wave totals through global memory (:309-401)            */
};

enum edpp_split_strategy {       /* SPLIT_ALGO / ReorderType, opencl/types.h:32-42 */
    EDPP_SPLIT_AUTO     = 0,
    EDPP_SPLIT_WG       = 1,     /* per-tile histogram -> scan of [B][tiles] -> stable scatter, LDS reorder */
    EDPP_SPLIT_WI       = 2,     /* per-wave histograms [B][waves] -> scan -> stable scatter, no reorder     */
    EDPP_SPLIT_SINGLE   = 3,     /* one lane per contiguous chunk (the reference's single_split)            */
    EDPP_SPLIT_CHAINED  = 4      /* bucket counts per XCD range of tiles, then a scatter whose tiles chain
                                    their bucket offsets by a per-bucket look-back (no [B][tiles] matrix) */
};

enum edpp_stream_op {            /* opencl/kernels/mem_kernel.cl:9-31 */
    EDPP_STREAM_COPY  = 0,       /* c[i] = a[i]          */
    EDPP_STREAM_SCALE = 1,       /* c[i] = 3*a[i]        */
    EDPP_STREAM_ADD   = 2,       /* c[i] = a[i] + b[i]   */
    EDPP_STREAM_TRIAD = 3        /* c[i] = a[i] + 3*b[i] */
};

enum edpp_perm_kind {
    EDPP_PERM_RANDOM      = 0,   /* pseudo-random permutation of [0,n) (keyed Feistel, cycle walking) */
    EDPP_PERM_BLOCK_LOCAL = 1,   /* 4096-element blocks permuted, order inside each block permuted    */
    EDPP_PERM_IDENTITY    = 2
};

enum edpp_key_dist {
    EDPP_KEYS_UNIFORM = 0,       /* bucket uniform on [0,B)                   */
    EDPP_KEYS_ZIPF    = 1        /* bucket ~ Zipf(s=1) over ranks 1..B        */
};

typedef struct edpp_ctx edpp_ctx;

/* ---- context ------------------------------------------------------------
 * One context per host thread.  `stream` is the hipStream_t every call
 * enqueues on; NULL means the device's default (null) stream, which is also
 * what torch reports for its default stream.  The context owns a
 * device workspace that grows on demand (hipMalloc outside the timed
 * kernels, as the reference allocates its temporaries per call,
 * opencl/primitives/scanImpl.cpp:39, splitImpl.cpp:76).                    */
int  edpp_ctx_create(int device, void* stream, edpp_ctx** out);
int  edpp_ctx_destroy(edpp_ctx* ctx);
int  edpp_ctx_sync(edpp_ctx* ctx);
void* edpp_ctx_stream(edpp_ctx* ctx);
/* Number of compute units and XCDs the context's device reports. */
int  edpp_ctx_device_info(edpp_ctx* ctx, int* cus, int* xcds, int64_t* l2_bytes, int64_t* mall_bytes);
/* Polls a device-side wait makes (a look-back over a predecessor tile's
 * status, a batch phase's counter) before it gives up: the kernel then
 * finishes with an invalid result and flags EDPP_ETIMEOUT, which the next
 * edpp_ctx_sync or timed call returns.  Default 2^22 polls (seconds); 0 gives
 * up at the first wait.  No counterpart in the reference, whose chained scan
 * spins without a bound (opencl/kernels/scan_global_chain_kernel.cl:51).     */
int  edpp_ctx_set_watchdog(edpp_ctx* ctx, uint32_t polls);
const char* edpp_strerror(int status);
const char* edpp_last_error(edpp_ctx* ctx);   /* message of the last failing call on ctx */
int  edpp_abi_version(void);

/* ---- STREAM probe: opencl/kernels/mem_kernel.cl:9-31, cuda/test_bandwidth.cu:14-44 */
int edpp_stream_i32(edpp_ctx* ctx, int op, const int32_t* a, const int32_t* b,
                    int32_t* c, int64_t n, double* ms);

/* ---- multi-pass gather / scatter ----------------------------------------
 * opencl/primitives/gatherImpl.cpp:7-55, scatterImpl.cpp:9-54.  Pass k of
 * `passes` touches only indices in [k*ceil(n/passes), (k+1)*ceil(n/passes)).
 * passes == 0 picks the count from the context's schedule (tuner or model).
 * idx must hold values in [0,n); for scatter it must be a permutation.      */
int edpp_gather_i32 (edpp_ctx* ctx, const int32_t* in, int32_t* out, const int32_t* idx,
                     int64_t n, int passes, double* ms);
int edpp_scatter_i32(edpp_ctx* ctx, const int32_t* in, int32_t* out, const int32_t* idx,
                     int64_t n, int passes, double* ms);
int edpp_gather_i64 (edpp_ctx* ctx, const int64_t* in, int64_t* out, const int32_t* idx,
                     int64_t n, int passes, double* ms);
int edpp_scatter_i64(edpp_ctx* ctx, const int64_t* in, int64_t* out, const int32_t* idx,
                     int64_t n, int passes, double* ms);
/* The pass count a passes == 0 call on n elements of elem_bytes (4 or 8) uses
 * while no cost model is set (the built-in schedule).  No reference
 * counterpart: the reference always takes the pass count from its caller.   */
int edpp_default_passes(edpp_ctx* ctx, int64_t n, int elem_bytes, int scatter, int* passes);

/* ---- global exclusive scan ----------------------------------------------
 * opencl/primitives/scanImpl.cpp:15-236.  in == out is allowed (the
 * reference scans split histograms in place, splitImpl.cpp:253).
 * d_total (device pointer, may be NULL) receives sum(in) mod 2^32 / 2^64.   */
int edpp_scan_exclusive_i32(edpp_ctx* ctx, const int32_t* in, int32_t* out, int64_t n,
                            int scheme, int32_t* d_total, double* ms);
int edpp_scan_exclusive_i64(edpp_ctx* ctx, const int64_t* in, int64_t* out, int64_t n,
                            int scheme, int64_t* d_total, double* ms);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the look-back scan of edpp_scan_exclusive_i32 (the same
 * output as every scheme of it) whose look-back state lives in a fixed device
 * array of 512 KiB + 256 B per GPU, held until the process ends, instead of
 * the context's workspace.  The call allocates and frees nothing.  Limits:
 * n < 2^31 when in and out are 16-byte aligned, n <= 2^26 otherwise
 * (EDPP_EINVAL beyond).  Every call on one GPU shares that array and clears
 * the part it uses first: calls that overlap on one GPU (other streams,
 * contexts or threads) race and must not be issued.                        */
int edpp_scan_exclusive_i32_noalloc(edpp_ctx* ctx, const int32_t* in, int32_t* out, int64_t n,
                                    int32_t* d_total, double* ms);

/* ---- single-workgroup tile scans ----------------------------------------
 * opencl/kernels/scan_local_kernel.cl:478-695, driver
 * opencl/test/test_scan_local.cpp:6-106: one workgroup of `local_size`
 * lanes scans n = local_size*tile_size values, exclusive.                    */
int edpp_scan_local_i32(edpp_ctx* ctx, const int32_t* in, int32_t* out, int tile_size,
                        int local_size, int matrix_type, int network, double* ms);

/* ---- stable multi-split --------------------------------------------------
 * opencl/primitives/splitImpl.cpp:21-498, cuda/test_split.cu:18-88.
 * bucket = key & (buckets-1); buckets is a power of two in [1, 4096].
 * vals_in/vals_out may both be NULL (key-only, KO) or both non-NULL (SOA).
 * d_bucket_start (may be NULL) receives int32[buckets] start offsets
 * (the reference's gatherStartPos, opencl/kernels/split_kernel.cl:320-332).  */
int edpp_split_i32(edpp_ctx* ctx, const int32_t* keys_in, int32_t* keys_out,
                   const int32_t* vals_in, int32_t* vals_out, int32_t* d_bucket_start,
                   int64_t n, int buckets, int strategy, double* ms);
int edpp_split_i64(edpp_ctx* ctx, const int64_t* keys_in, int64_t* keys_out,
                   const int32_t* vals_in, int32_t* vals_out, int32_t* d_bucket_start,
                   int64_t n, int buckets, int strategy, double* ms);

/* ---- LSD radix sort on the stable split -------------------------------------
 * Replaces the reference's radixSort (obsolete/CUDA/kernels/radixSort.cu:29-200,
 * declared at obsolete/CUDA/inc/kernels.h:37-45) and its CUB SortKeys /
 * SortPairs calls (obsolete/CUDA/radixSortCUB.cu:25,93): stable sort of n
 * 32-bit keys by key bits [begin_bit, end_bit) (0 <= begin <= end <= 32),
 * values (both NULL or both non-NULL) following their keys.  _u32 orders keys
 * as unsigned, _i32 as signed (bit 31 inverted when it is among the sorted
 * bits).  keys_in may equal keys_out (and vals_in vals_out); otherwise the
 * buffers must not overlap.                                                    */
int edpp_sort_u32(edpp_ctx* ctx, const uint32_t* keys_in, uint32_t* keys_out,
                  const int32_t* vals_in, int32_t* vals_out, int64_t n,
                  int begin_bit, int end_bit, double* ms);
int edpp_sort_i32(edpp_ctx* ctx, const int32_t* keys_in, int32_t* keys_out,
                  const int32_t* vals_in, int32_t* vals_out, int64_t n,
                  int begin_bit, int end_bit, double* ms);

/* ---- access-pattern and workgroup-order probes ------------------------------
 * opencl/kernels/mem_kernel.cl:33-86 with opencl/test/test_access.cpp and
 * opencl/test/test_access_wg.cpp.
 * edpp_access_scale_i32: out[i] = 3 * in[i] for i < n, touched by
 * lanes = n / repeat lanes in the pattern's mapping (lane t, step k):
 *   EDPP_ACCESS_COLUMN  t + k*lanes               (scale_column, coalesced)
 *   EDPP_ACCESS_ROW     t*repeat + k              (scale_row, a run per lane)
 *   EDPP_ACCESS_MIXED   (t/64)*64*repeat + k*64 + t%64   (scale_mixed, a run per wave)
 * n must be a multiple of 64*repeat, repeat >= 1.
 * edpp_access_chunks_i32: workgroups take chunks of `chunk` elements in the
 * order order[0], order[1], ... (a permutation of the ceil(n/chunk) chunk ids,
 * device memory) and *d_count (device) = number of odd elements of in
 * (wg_access; its output is racy in the reference, here it is exact).      */
enum edpp_access_pattern {
    EDPP_ACCESS_COLUMN = 0,
    EDPP_ACCESS_ROW    = 1,
    EDPP_ACCESS_MIXED  = 2
};
int edpp_access_scale_i32(edpp_ctx* ctx, int pattern, const int32_t* in, int32_t* out, int64_t n,
                          int repeat, double* ms);
int edpp_access_chunks_i32(edpp_ctx* ctx, const int32_t* in, int64_t n, int64_t chunk,
                           const int32_t* order, int64_t* d_count, double* ms);
/* edpp_access_random_i32: the random-access ceiling that gather and scatter
 * are graded against.  `accesses` 4-byte accesses to table positions
 * p_i = fmix32(i + seed) & (table_n - 1), i < accesses (fmix32: the MurmurHash3
 * finaliser, a bijection of 32-bit words), computed in registers, so no index
 * stream is read.  write = 0: *d_sum (device, int64) = sum over i of
 * (uint32)table[p_i], modulo 2^64.  write = 1: table[p_i] = p_i for every i
 * (d_sum may be NULL).  table_n is a power of two in [1, 2^31], accesses in
 * [0, 2^32).  The probe is the sequential-index counterpart of the
 * reference's cold/warm cache probe (opencl/test/test_access_wg.cpp:36-46):
 * a table of 1 GiB prices random reads from HBM, 128 MiB from the Infinity
 * Cache, 2 MiB from each XCD's L2. */
int edpp_access_random_i32(edpp_ctx* ctx, int32_t* table, int64_t table_n, int64_t accesses,
                           uint64_t seed, int write, int64_t* d_sum, double* ms);

/* ---- partitioned hash join ----------------------------------------------
 * Replaces the reference's hashjoin (obsolete/OpenCL/phjImpl.cpp:14-66 with
 * obsolete/OpenCL/hj_partitioned_kernel.cl:7-115), which partitions both
 * relations with WG_split and counts the joined pairs: *d_count (device
 * memory, overwritten) = number of pairs (i, j) with r_keys[i] == s_keys[j].
 * Every int32 key value takes part (the reference reserves -1 as "empty").
 * The values of the reference's relations do not change the count and are
 * not taken.                                                                 */
int edpp_hash_join_count(edpp_ctx* ctx, const int32_t* r_keys, int64_t r_n,
                         const int32_t* s_keys, int64_t s_n, uint64_t* d_count, double* ms);

/* ---- reproducible synthetic inputs ---------------------------------------
 * Replace opencl/util/utility.cpp:214-241 (time-seeded, schedule-dependent)
 * with counter-based SplitMix64: x_i = i-th output of SplitMix64(seed).      */
/* out[i] = (int32)(x_i >> 32) & mask  (mask = -1 for full-range values)     */
int edpp_gen_i32(edpp_ctx* ctx, int32_t* out, int64_t n, uint64_t seed, int32_t mask);
/* out[i] = (int64)x_i & mask */
int edpp_gen_i64(edpp_ctx* ctx, int64_t* out, int64_t n, uint64_t seed, int64_t mask);
int edpp_gen_perm(edpp_ctx* ctx, int32_t* idx, int64_t n, uint64_t seed, int kind);
/* keys with bucket (key & (B-1)) drawn from `dist`; vals (may be NULL) = i   */
int edpp_gen_split_keys_i32(edpp_ctx* ctx, int32_t* keys, int32_t* vals, int64_t n,
                            int buckets, int dist, uint64_t seed);
int edpp_gen_split_keys_i64(edpp_ctx* ctx, int64_t* keys, int32_t* vals, int64_t n,
                            int buckets, int dist, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif /* EDPP_H */
