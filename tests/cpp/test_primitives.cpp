// test_primitives.cpp — checks of the reference-shaped C++ API in the style of
// the reference's own drivers: opencl/test/test_gather.cpp:53-70 (out[i] ==
// in[loc[i]] with in[i] = i), test_scatter.cpp:63-69, test_scan_global.cpp:
// 62-73 (serial exclusive prefix of rand() & 0xf), test_scan_local.cpp:84-94,
// test_split.cpp:164-203 (bucket ids non-decreasing, sums kept; here also
// stability, through values = original index) and cuda/test_split.cu:196-204.
// Exit status 0 when every check passes.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <algorithm>
#include <vector>
#include "edpp_primitives.hpp"

static int failures = 0;
#define CHECK(cond, ...)                        \
    do {                                        \
        if (!(cond)) {                          \
            printf("FAILED: " __VA_ARGS__);     \
            printf("\n");                       \
            ++failures;                         \
        }                                       \
    } while (0)

template <class T>
static T* dev(const std::vector<T>& h) {
    T* d = nullptr;
    hipMalloc(&d, sizeof(T) * (h.size() ? h.size() : 1));
    if (!h.empty()) hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
    return d;
}
template <class T>
static std::vector<T> host(const T* d, size_t n) {
    std::vector<T> h(n);
    if (n) hipMemcpy(h.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost);
    return h;
}

static void test_gather_scatter(int len) {
    std::vector<int> h_in(len);
    for (int i = 0; i < len; ++i) h_in[i] = i;
    int *d_in = dev(h_in), *d_out = nullptr, *d_loc = nullptr;
    hipMalloc(&d_out, sizeof(int) * len);
    hipMalloc(&d_loc, sizeof(int) * len);
    edpp_gen_perm(edpp::context(), d_loc, len, 1234, EDPP_PERM_RANDOM);
    std::vector<int> loc = host(d_loc, len);
    for (int pass = 1; pass <= 16; pass *= 2) {
        double ms = edpp::gather(d_in, d_out, len, d_loc, 1024, len / 1024 / 16 + 1, pass);
        std::vector<int> out = host(d_out, len);
        bool ok = ms >= 0;
        for (int i = 0; i < len && ok; ++i) ok = out[i] == h_in[loc[i]];
        CHECK(ok, "gather len=%d pass=%d", len, pass);
        ms = edpp::scatter(d_in, d_out, len, d_loc, 1024, len / 1024 / 16 + 1, pass);
        out = host(d_out, len);
        ok = ms >= 0;
        for (int i = 0; i < len && ok; ++i) ok = h_in[i] == out[loc[i]];
        CHECK(ok, "scatter len=%d pass=%d", len, pass);
    }
    hipFree(d_in);
    hipFree(d_out);
    hipFree(d_loc);
}

static void test_scans(int len) {
    int *d_in = nullptr, *d_out = nullptr;
    hipMalloc(&d_in, sizeof(int) * len);
    hipMalloc(&d_out, sizeof(int) * len);
    edpp_gen_i32(edpp::context(), d_in, len, 99, 0xf);
    std::vector<int> in = host(d_in, len);
    for (int which = 0; which < 3; ++which) {
        double ms = which == 0 ? edpp::scan_chained(d_in, d_out, len, 1024, 80, 0, 11)
                  : which == 1 ? edpp::scan_RSS(d_in, d_out, (unsigned)len, 1024, 80)
                               : edpp::scan_RSS_single(d_in, d_out, (unsigned)len);
        std::vector<int> out = host(d_out, len);
        unsigned acc = 0;
        bool ok = ms >= 0;
        for (int i = 0; i < len && ok; ++i) {
            ok = (unsigned)out[i] == acc;
            acc += (unsigned)in[i];
        }
        CHECK(ok, "scan variant %d len=%d", which, len);
    }
    CHECK(edpp::scan_chained(d_in, d_out, len, 1024, 80, 0, 0) == 1, "scan_chained R=L=0 must return 1");
    hipFree(d_in);
    hipFree(d_out);
}

static void test_matrix_scans() {
    for (int tile = 1; tile < 50; tile += 12) {
        int len = 512 * tile;
        int *d_in = nullptr, *d_out = nullptr;
        hipMalloc(&d_in, sizeof(int) * len);
        hipMalloc(&d_out, sizeof(int) * len);
        edpp_gen_i32(edpp::context(), d_in, len, tile, 0xf);
        std::vector<int> in = host(d_in, len);
        for (int t = edpp::LM; t <= edpp::LM_SERIAL; ++t) {
            double ms = edpp::matrix_scan(d_in, d_out, tile, (edpp::MatrixScanType)t);
            std::vector<int> out = host(d_out, len);
            int acc = 0;
            bool ok = ms >= 0;
            for (int i = 0; i < len && ok; ++i) {
                ok = out[i] == acc;
                acc += in[i];
            }
            CHECK(ok, "matrix scan type %d tile %d", t, tile);
        }
        hipFree(d_in);
        hipFree(d_out);
    }
}

static void test_splits(int len, int buckets) {
    int *d_keys = nullptr, *d_vals = nullptr, *d_ko = nullptr, *d_vo = nullptr, *d_start = nullptr;
    hipMalloc(&d_keys, sizeof(int) * len);
    hipMalloc(&d_vals, sizeof(int) * len);
    hipMalloc(&d_ko, sizeof(int) * len);
    hipMalloc(&d_vo, sizeof(int) * len);
    hipMalloc(&d_start, sizeof(int) * buckets);
    edpp_gen_split_keys_i32(edpp::context(), d_keys, d_vals, len, buckets, EDPP_KEYS_ZIPF, 1234);
    std::vector<int> keys = host(d_keys, len);
    const unsigned mask = buckets - 1;
    // algo 0 WI, 1..3 WG with VARIED / NO / FIXED reorder, 4 single
    for (int algo = 0; algo < 5; ++algo) {
        const edpp::ReorderType ro = algo == 2 ? edpp::NO_REORDER : algo == 3 ? edpp::FIXED_REORDER
                                                                              : edpp::VARIED_REORDER;
        double ms = algo == 0 ? edpp::WI_split(d_keys, d_ko, d_start, len, buckets, edpp::KVS_SOA, d_vals, d_vo)
                  : algo <= 3 ? edpp::WG_split(d_keys, d_ko, d_start, len, buckets, ro, edpp::KVS_SOA, d_vals, d_vo)
                              : edpp::single_split(d_keys, d_ko, len, buckets, false, edpp::KO);
        CHECK(ms >= 0, "split algo %d returned %f", algo, ms);
        std::vector<int> ko = host(d_ko, len), vo = host(d_vo, len), st = host(d_start, buckets);
        unsigned long in_sum = 0, out_sum = 0;
        bool ok = true;
        for (int i = 0; i < len; ++i) {
            in_sum += keys[i] & mask;
            out_sum += ko[i] & mask;
            if (i && (unsigned)(ko[i] & mask) < (unsigned)(ko[i - 1] & mask)) ok = false;
            if (algo < 4) {
                if (ko[i] != keys[vo[i]]) ok = false;                                    // keys follow values
                if (i && (ko[i] & mask) == (ko[i - 1] & mask) && vo[i] <= vo[i - 1]) ok = false;  // stable
            }
        }
        CHECK(ok && in_sum == out_sum, "split algo %d len %d buckets %d", algo, len, buckets);
        if (algo < 4) {
            std::vector<int> cnt(buckets, 0);
            for (int i = 0; i < len; ++i) cnt[keys[i] & mask]++;
            int acc = 0;
            bool sok = true;
            for (int b = 0; b < buckets; ++b) {
                sok = sok && st[b] == acc;
                acc += cnt[b];
            }
            CHECK(sok, "bucket starts algo %d buckets %d", algo, buckets);
        }
    }
    CHECK(edpp::WG_split(d_keys, d_ko, 0, len, buckets, edpp::NO_REORDER, edpp::KVS_SOA) == -1,
          "SOA without values must return -1");
    // AOS: (key, value) pairs
    std::vector<int> aos(2 * (size_t)len);
    for (int i = 0; i < len; ++i) {
        aos[2 * i] = keys[i];
        aos[2 * i + 1] = i;
    }
    int* d_aos = dev(aos);
    int* d_aos_out = nullptr;
    hipMalloc(&d_aos_out, sizeof(int) * 2 * (size_t)len);
    double ms = edpp::WG_split(d_aos, d_aos_out, 0, len, buckets, edpp::VARIED_REORDER, edpp::KVS_AOS);
    std::vector<int> out = host(d_aos_out, 2 * (size_t)len);
    bool ok = ms >= 0;
    for (int i = 1; i < len && ok; ++i) {
        unsigned b0 = out[2 * i - 2] & mask, b1 = out[2 * i] & mask;
        ok = b1 > b0 || (b1 == b0 && out[2 * i + 1] > out[2 * i - 1]);
        ok = ok && out[2 * i] == keys[out[2 * i + 1]];
    }
    CHECK(ok, "AOS split buckets %d", buckets);
    hipFree(d_aos);
    hipFree(d_aos_out);
    hipFree(d_keys);
    hipFree(d_vals);
    hipFree(d_ko);
    hipFree(d_vo);
    hipFree(d_start);
}

// radixSort (obsolete/CUDA/inc/kernels.h:37-45) checked against std::sort /
// std::stable_sort, as the reference's own test does
static void test_radix_sort(int len) {
    std::vector<int> keys(len), vals(len);
    unsigned s = 12345u;
    for (int i = 0; i < len; ++i) {
        s = s * 1664525u + 1013904223u;
        keys[i] = (int)s;
        if (i % 7 == 0 && i > 0) keys[i] = keys[i / 2];  // duplicates
        vals[i] = i;
    }
    int* d_keys = dev(keys);
    int* d_vals = dev(vals);
    double ms = edpp::radixSort(d_keys, d_vals, len, true);
    std::vector<int> ko = host(d_keys, len), vo = host(d_vals, len);
    std::vector<int> order(len);
    for (int i = 0; i < len; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return keys[a] < keys[b]; });
    bool ok = ms >= 0;
    for (int i = 0; i < len && ok; ++i) ok = vo[i] == order[i] && ko[i] == keys[order[i]];
    CHECK(ok, "radixSort records len %d", len);
    // keys only: d_source_values are the keys
    hipMemcpy(d_vals, keys.data(), sizeof(int) * (size_t)len, hipMemcpyHostToDevice);
    ms = edpp::radixSort(nullptr, d_vals, len, false);
    std::vector<int> sorted = keys;
    std::sort(sorted.begin(), sorted.end());
    CHECK(ms >= 0 && host(d_vals, len) == sorted, "radixSort keys len %d", len);
    hipFree(d_keys);
    hipFree(d_vals);
}

// hashjoin (obsolete/OpenCL/phjImpl.cpp:14): count of equal-key pairs
static void test_hash_join(int len) {
    std::vector<int> r(len), s(len / 2 + 3), rv(len, 0), sv(len / 2 + 3, 0);
    unsigned x = 777u;
    for (auto& k : r) { x = x * 1664525u + 1013904223u; k = (int)(x >> 12) % 50000; }
    for (auto& k : s) { x = x * 1664525u + 1013904223u; k = (int)(x >> 12) % 50000; }
    std::vector<long long> cr(50000, 0), cs(50000, 0);
    for (int k : r) cr[k]++;
    for (int k : s) cs[k]++;
    long long expect = 0;
    for (int k = 0; k < 50000; ++k) expect += cr[k] * cs[k];
    int *dr = dev(r), *ds = dev(s), *drv = dev(rv), *dsv = dev(sv);
    int res = -1;
    double ms = edpp::hashjoin(dr, drv, len, ds, dsv, (int)s.size(), res);
    CHECK(ms >= 0 && res == (int)expect, "hashjoin len %d: %d vs %lld", len, res, expect);
    hipFree(dr);
    hipFree(ds);
    hipFree(drv);
    hipFree(dsv);
}

int main(int argc, char** argv) {
    int len = argc > 1 ? atoi(argv[1]) : (1 << 20);
    if (!edpp::context()) {
        printf("no device\n");
        return 2;
    }
    test_gather_scatter(len);
    test_gather_scatter(len + 7);
    test_scans(len);
    test_scans(3 * len + 5);
    test_matrix_scans();
    for (int b = 2; b <= 4096; b *= 8) test_splits(len, b);
    test_splits(len + 13, 4096);
    test_radix_sort(len);
    test_radix_sort(len / 3 + 1);
    test_hash_join(len);
    printf("%s: %d failure(s)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
