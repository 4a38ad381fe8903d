// ref_shim.cpp — C entry points over functions of the reference's own OpenMP
// sources, which oracle/build.py compiles unchanged from /root/reference into
// oracle/_ref/libedpp_ref_openmp.so.  Only the checker and the CPU baseline
// use that library.
//
// Declarations restate the signatures found in the reference:
//   openmp/test_gather_scatter_CPU.cpp:24,33   gather / scatter
//   openmp/test_bandwidth_CPU.cpp:39,49,59,71  copy_omp / scale_omp / *_ss
//   openmp/util/utility.cpp:23,27              compute_bandwidth / average_Hampel
#include <stdint.h>

double gather(int* input, int* output, int* idx, uint64_t len);
double scatter(int* input, int* output, int* idx, uint64_t len);
double copy_omp(int* input, int* output, uint64_t len);
double scale_omp(int* input, int* output, uint64_t len);
double copy_omp_ss(int* input, int* output, uint64_t len);
double scale_omp_ss(int* input, int* output, uint64_t len);
double compute_bandwidth(uint64_t num, int wordSize, double kernel_time);
double average_Hampel(double* input, int num);

extern "C" {
double edpp_ref_gather(int* in, int* out, int* idx, uint64_t n) { return gather(in, out, idx, n); }
double edpp_ref_scatter(int* in, int* out, int* idx, uint64_t n) { return scatter(in, out, idx, n); }
double edpp_ref_copy(int* in, int* out, uint64_t n) { return copy_omp(in, out, n); }
double edpp_ref_scale(int* in, int* out, uint64_t n) { return scale_omp(in, out, n); }
double edpp_ref_copy_ss(int* in, int* out, uint64_t n) { return copy_omp_ss(in, out, n); }
double edpp_ref_scale_ss(int* in, int* out, uint64_t n) { return scale_omp_ss(in, out, n); }
double edpp_ref_compute_bandwidth(uint64_t num, int word, double ms) { return compute_bandwidth(num, word, ms); }
double edpp_ref_average_hampel(double* x, int n) { return average_Hampel(x, n); }
}
