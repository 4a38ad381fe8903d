// edpp_primitives.hpp — the reference's primitives-style C++ host API
// (opencl/primitives.h:12-49, enums of opencl/types.h:8-46) over libedpp.so.
//
// Same names, argument order and meaning; cl_mem becomes a device pointer.
// Each call runs on a process-wide context (device 0, default stream), like
// the reference's Plat singleton (opencl/util/Plat.h:37-62), is synchronous,
// and returns the summed device time of its kernels in milliseconds.
// Launch-shape arguments (localSize, gridSize, R, L, local_size, grid_size)
// are schedule hints only: results are the same for every value.
// Error returns follow the reference: scan_chained returns 1 when R and L
// are both 0 (opencl/primitives/scanImpl.cpp:19-22); the splits return -1
// on bad value / layout arguments (opencl/primitives/splitImpl.cpp:29-45);
// any other failure returns -1 (the reference exits the process instead,
// opencl/util/utility.cpp:85-90) and edpp_last_error(edpp::context())
// holds the message.
#pragma once

#include "edpp.h"

namespace edpp {

// opencl/types.h:8-10
enum SCAN_GLOBAL_TYPE { RSS, RSS_SINGLE_THREAD, CHAINED };
// opencl/types.h:12-14
enum MatrixScanType { LM, REG, LM_REG, LM_SERIAL };
// opencl/types.h:16-20
struct scan_arg {
    int R;
    int L;
    SCAN_GLOBAL_TYPE algo;
};
// opencl/types.h:29-31: key-only, key-value array of structures, structure of arrays
enum DataStruc { KO, KVS_AOS, KVS_SOA };
// opencl/types.h:39-41
enum SPLIT_ALGO { WI, WG, WG_fixed_reorder, WG_varied_reorder, Single, Single_reorder };
// opencl/types.h:43-45
enum ReorderType { NO_REORDER, FIXED_REORDER, VARIED_REORDER };

// the process-wide context used by the functions below (created on first use)
edpp_ctx* context();

// opencl/primitives.h:12-14
double gather(const int* d_source_values, int* d_dest_values, int length, const int* d_loc, int localSize,
              int gridSize, int pass);
// opencl/primitives.h:17-19
double scatter(const int* d_source_values, int* d_dest_values, int length, const int* d_loc, int localSize,
               int gridSize, int pass);
// opencl/primitives.h:22-24 (single pass, decoupled look-back)
double scan_chained(const int* d_in, int* d_out, int length, int localSize, int gridSize, int R, int L);
// opencl/primitives.h:26-27 (reduce, scan of partials, scan with offsets)
double scan_RSS(const int* d_in, int* d_out, unsigned length, int local_size, int grid_size);
// opencl/primitives.h:29 (scan-scan-add)
double scan_RSS_single(const int* d_in, int* d_out, unsigned length);
// opencl/test/test_scan_local.cpp:6-106: one workgroup scans local_size*tile_size values
double matrix_scan(const int* d_in, int* d_out, int tile_size, MatrixScanType type, int local_size = 512);

// opencl/primitives.h:32-37.  For KVS_AOS d_in/d_out hold `length` (key, value)
// int pairs and the value arguments must be null.
double WI_split(const int* d_in, int* d_out, int* d_start, int length, int buckets, DataStruc structure,
                const int* d_in_values = 0, int* d_out_values = 0, int local_size = 256, int grid_size = 32768);
// opencl/primitives.h:39-44.  reorder_type selects how the stable WG split
// (EDPP_SPLIT_WG) writes a tile, as the reference's kernels do
// (opencl/kernels/split_kernel.cl:455-665): NO_REORDER stores each element
// straight to its output position, FIXED_REORDER stages each bucket in a
// fixed-length LDS buffer written whenever it fills, VARIED_REORDER stages
// each bucket's whole run of a sub-tile (the library's default).  All three
// give the same stable output; any other value returns -1.
double WG_split(const int* d_in, int* d_out, int* d_start, int length, int buckets, ReorderType reorder_type,
                DataStruc structure, const int* d_in_values = 0, int* d_out_values = 0, int local_size = 256,
                int grid_size = 32768);
// opencl/primitives.h:46-49 (KO and KVS_AOS only, as in the reference)
double single_split(const int* d_in, int* d_out, int length, int buckets, bool reorder, DataStruc structure);

// obsolete/CUDA/inc/kernels.h:37-45 (the reference's radixSort): sorts in
// place, ascending as int32 like its std::sort check
// (obsolete/CUDA/src/testRadixSort.cpp:57-60).  isRecord: d_source_keys are
// the keys and d_source_values follow them; otherwise d_source_values are the
// keys and d_source_keys is unused.
double radixSort(int* d_source_keys, int* d_source_values, int len, bool isRecord);

// obsolete/OpenCL/phjImpl.cpp:14 (the reference's partitioned hash join):
// res_len = number of pairs with equal keys (the reference counts only); the
// value arrays are accepted and do not change the result.  Returns device ms.
double hashjoin(const int* d_R_keys, const int* d_R_values, int rLen, const int* d_S_keys, const int* d_S_values,
                int sLen, int& res_len);

}  // namespace edpp
