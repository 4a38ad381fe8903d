/*
 * baseline_omp.c — OpenMP CPU baseline: restatement of the reference's host
 * primitives, compiled like the reference (gcc -O3 -fopenmp -mavx2,
 * openmp/CMakeLists.txt:5).
 *
 * TEST / BASELINE INFRASTRUCTURE ONLY: bench.py times these on the host
 * cores beside the GPU numbers; the product never calls them.  Each function
 * returns the elapsed wall time of its parallel region in milliseconds, as
 * the reference's Timer does (openmp/util/timer.h:7-24).
 */
#include <omp.h>
#include <stdint.h>

#define BASELINE_MAX_THREADS 256 /* openmp/test_scan_CPU.cpp:36 */

int baseline_threads(void) { return omp_get_max_threads(); }

/* scan-scan-add, 4n accesses: openmp/test_scan_CPU.cpp:84-122 */
double baseline_scan_ssa(const uint32_t* input, uint32_t* output, int64_t len) {
    uint32_t partial[BASELINE_MAX_THREADS] = {0};
    double t0 = omp_get_wtime();
#pragma omp parallel num_threads(omp_get_max_threads() < BASELINE_MAX_THREADS ? omp_get_max_threads() : BASELINE_MAX_THREADS)
    {
        int nth = omp_get_num_threads(), tid = omp_get_thread_num();
        uint32_t local = 0;
#pragma omp for schedule(static) nowait
        for (int64_t i = 0; i < len; i++) {
            output[i] = local;
            local += input[i];
        }
        partial[tid] = local;
#pragma omp barrier
#pragma omp single
        {
            uint32_t acc = 0;
            for (int k = 0; k < nth; k++) {
                uint32_t x = partial[k];
                partial[k] = acc;
                acc += x;
            }
        }
        uint32_t add = partial[tid];
#pragma omp for schedule(static) nowait
        for (int64_t i = 0; i < len; i++) output[i] += add;
    }
    return (omp_get_wtime() - t0) * 1000.0;
}

/* reduce-then-scan, 3n accesses: openmp/test_scan_CPU.cpp:124-163 */
double baseline_scan_rts(const uint32_t* input, uint32_t* output, int64_t len) {
    uint32_t partial[BASELINE_MAX_THREADS] = {0};
    double t0 = omp_get_wtime();
#pragma omp parallel num_threads(omp_get_max_threads() < BASELINE_MAX_THREADS ? omp_get_max_threads() : BASELINE_MAX_THREADS)
    {
        int nth = omp_get_num_threads(), tid = omp_get_thread_num();
        uint32_t local = 0;
#pragma omp for schedule(static) nowait
        for (int64_t i = 0; i < len; i++) local += input[i];
        partial[tid] = local;
#pragma omp barrier
#pragma omp single
        {
            uint32_t acc = 0;
            for (int k = 0; k < nth; k++) {
                uint32_t x = partial[k];
                partial[k] = acc;
                acc += x;
            }
        }
        local = partial[tid];
#pragma omp for schedule(static) nowait
        for (int64_t i = 0; i < len; i++) {
            output[i] = local;
            local += input[i];
        }
    }
    return (omp_get_wtime() - t0) * 1000.0;
}

/* openmp/test_gather_scatter_CPU.cpp:24-41 */
double baseline_gather(const int32_t* input, int32_t* output, const int32_t* idx, int64_t len) {
    double t0 = omp_get_wtime();
#pragma omp parallel for schedule(auto)
    for (int64_t i = 0; i < len; i++) output[i] = input[idx[i]];
    return (omp_get_wtime() - t0) * 1000.0;
}

double baseline_scatter(const int32_t* input, int32_t* output, const int32_t* idx, int64_t len) {
    double t0 = omp_get_wtime();
#pragma omp parallel for schedule(auto)
    for (int64_t i = 0; i < len; i++) output[idx[i]] = input[i];
    return (omp_get_wtime() - t0) * 1000.0;
}

/* openmp/test_bandwidth_CPU.cpp:39-56 */
double baseline_copy(const int32_t* input, int32_t* output, int64_t len) {
    double t0 = omp_get_wtime();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < len; i++) output[i] = input[i];
    return (omp_get_wtime() - t0) * 1000.0;
}

double baseline_scale(const int32_t* input, int32_t* output, int64_t len) {
    double t0 = omp_get_wtime();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < len; i++) output[i] = (int32_t)((uint32_t)input[i] * 3u);
    return (omp_get_wtime() - t0) * 1000.0;
}

/* first-touch initialisation in parallel, so pages land where the workers run */
void baseline_fill(int32_t* a, int64_t len, int32_t v) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < len; i++) a[i] = v;
}
