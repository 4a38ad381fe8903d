/*
 * oracle.c — sequential CPU restatement of the reference's primitives.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the GPU library
 * is compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  The product (efficient-dpp_amd/libedpp.so)
 * never links, calls or falls back to it.
 *
 * Every function states the reference code whose semantics it restates
 * (RapidsAtHKUST/Efficient-DPP, file:line).  Arithmetic is unsigned with
 * wrap-around, which is the two's-complement result the reference's int
 * accumulators produce whenever they do not overflow (for example the
 * all-ones scan of openmp/test_scan_CPU.cpp:174-177) and the only defined
 * result when they do.
 *
 * Parity pins (tests/test_oracle.py): SplitMix64 known-answer vector,
 * numpy cumsum / fancy indexing / stable argsort / bincount as independent
 * library routines, the reference's own known answers (all-ones scan,
 * iota gather/scatter, split ordering and sum checks), brute force on tiny
 * inputs, and the Hampel mean on hand-computed vectors.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ generators
 * SplitMix64 (Steele, Lea, Flood 2014): state advances by the golden gamma,
 * output = mix(state).  x_i = output number i+1 of the generator seeded with
 * `seed`.  Replaces the non-reproducible generators of
 * opencl/util/utility.cpp:214-241.                                          */
static uint64_t o_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t oracle_splitmix_at(uint64_t seed, uint64_t i) {
    uint64_t state = seed;
    state += (i + 1) * 0x9E3779B97F4A7C15ull;
    return o_mix(state);
}

/* sequential form, for the known-answer test of the generator */
void oracle_splitmix_seq(uint64_t seed, uint64_t* out, int64_t n) {
    uint64_t state = seed;
    for (int64_t i = 0; i < n; ++i) {
        state += 0x9E3779B97F4A7C15ull;
        out[i] = o_mix(state);
    }
}

void oracle_gen_i32(int32_t* out, int64_t n, uint64_t seed, int32_t mask) {
    for (int64_t i = 0; i < n; ++i) out[i] = (int32_t)((uint32_t)(oracle_splitmix_at(seed, (uint64_t)i) >> 32) & (uint32_t)mask);
}

void oracle_gen_i64(int64_t* out, int64_t n, uint64_t seed, int64_t mask) {
    for (int64_t i = 0; i < n; ++i) out[i] = (int64_t)(oracle_splitmix_at(seed, (uint64_t)i) & (uint64_t)mask);
}

/* 4-round balanced Feistel on 2*h bits, round function = low h bits of
 * mix(right ^ (key + C*(round+1))) */
static uint64_t o_feistel(uint64_t x, int h, uint64_t key) {
    uint64_t m = (1ull << h) - 1, left = x >> h, right = x & m;
    for (int k = 1; k <= 4; ++k) {
        uint64_t f = o_mix(right ^ (key + 0xD1B54A32D192ED03ull * (uint64_t)k)) & m;
        uint64_t t = left ^ f;
        left = right;
        right = t;
    }
    return (left << h) | right;
}

static int o_half(uint64_t n) {
    int h = 1;
    while ((1ull << (2 * h)) < n) h++;
    return h;
}

static uint64_t o_perm_at(uint64_t i, uint64_t n, int h, uint64_t key) {
    uint64_t x = o_feistel(i, h, key);
    while (x >= n) x = o_feistel(x, h, key);
    return x;
}

#define O_PERM_SALT 0x243F6A8885A308D3ull
#define O_BLOCK_SALT 0x13198A2E03707344ull
#define O_TAIL_SALT 0xA4093822299F31D0ull

/* kind 0: random permutation of [0,n); 1: 4096-element blocks permuted and
 * each block's order permuted; 2: identity */
void oracle_gen_perm(int32_t* idx, int64_t n, uint64_t seed, int kind) {
    if (kind == 2) {
        for (int64_t i = 0; i < n; ++i) idx[i] = (int32_t)i;
        return;
    }
    if (kind == 0) {
        int h = o_half((uint64_t)n);
        uint64_t key = o_mix(seed ^ O_PERM_SALT);
        for (int64_t i = 0; i < n; ++i) idx[i] = (int32_t)o_perm_at((uint64_t)i, (uint64_t)n, h, key);
        return;
    }
    int64_t nb = n / 4096;
    int64_t full = nb * 4096;
    int hb = o_half(nb ? (uint64_t)nb : 1);
    int ht = o_half((n - full) ? (uint64_t)(n - full) : 1);
    uint64_t kb = o_mix(seed ^ O_PERM_SALT);
    uint64_t kt = o_mix(seed ^ O_BLOCK_SALT ^ O_TAIL_SALT);
    for (int64_t b = 0; b < nb; ++b) {
        uint64_t src = o_perm_at((uint64_t)b, (uint64_t)nb, hb, kb);
        uint64_t kj = o_mix(seed ^ O_BLOCK_SALT ^ ((uint64_t)b * 0x9E3779B97F4A7C15ull));
        for (int64_t j = 0; j < 4096; ++j) idx[b * 4096 + j] = (int32_t)(src * 4096 + o_feistel((uint64_t)j, 6, kj));
    }
    for (int64_t i = full; i < n; ++i)
        idx[i] = (int32_t)(full + (int64_t)o_perm_at((uint64_t)(i - full), (uint64_t)(n - full), ht, kt));
}

/* Zipf(s=1) over ranks 1..B, CDF scaled to 2^32 */
void oracle_zipf_thresholds(int B, uint64_t* thr) {
    double h = 0.0, cum = 0.0;
    for (int r = 1; r <= B; ++r) h += 1.0 / (double)r;
    for (int k = 0; k < B; ++k) {
        cum += 1.0 / (double)(k + 1);
        thr[k] = (k == B - 1) ? (1ull << 32) : (uint64_t)floor(cum / h * 4294967296.0);
    }
}

static int o_log2(int B) {
    int lg = 0;
    while ((1 << lg) < B) lg++;
    return lg;
}

static uint32_t o_bucket(uint32_t u, int lg, int zipf, const uint64_t* thr) {
    if (!zipf) return lg ? (u >> (32 - lg)) : 0u;
    int k = 0;
    while ((uint64_t)u >= thr[k]) k++;  /* linear search: independent of the device's bisection */
    return (uint32_t)k;
}

void oracle_gen_split_keys_i32(int32_t* keys, int32_t* vals, int64_t n, int B, int dist, uint64_t seed) {
    int lg = o_log2(B);
    uint64_t* thr = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)B);
    oracle_zipf_thresholds(B, thr);
    for (int64_t i = 0; i < n; ++i) {
        uint64_t r = oracle_splitmix_at(seed, (uint64_t)i);
        uint32_t b = o_bucket((uint32_t)(r >> 32), lg, dist == 1, thr);
        keys[i] = (int32_t)((((uint32_t)r) << lg) | b);
        if (vals) vals[i] = (int32_t)i;
    }
    free(thr);
}

void oracle_gen_split_keys_i64(int64_t* keys, int32_t* vals, int64_t n, int B, int dist, uint64_t seed) {
    int lg = o_log2(B);
    uint64_t* thr = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)B);
    oracle_zipf_thresholds(B, thr);
    for (int64_t i = 0; i < n; ++i) {
        uint64_t r = oracle_splitmix_at(seed, (uint64_t)i);
        uint32_t b = o_bucket((uint32_t)(r >> 32), lg, dist == 1, thr);
        keys[i] = (int64_t)((o_mix(r) << lg) | b);
        if (vals) vals[i] = (int32_t)i;
    }
    free(thr);
}

/* ------------------------------------------------------------------ scan
 * Exclusive prefix sum: the serial check loops of
 * opencl/test/test_scan_global.cpp:62-73 and openmp/test_scan_CPU.cpp:39-49,
 * with unsigned accumulators.  Returns the total.                          */
uint32_t oracle_scan_exclusive_u32(const uint32_t* in, uint32_t* out, int64_t n) {
    uint32_t acc = 0;
    for (int64_t i = 0; i < n; ++i) {
        uint32_t x = in[i];
        out[i] = acc;
        acc += x;
    }
    return acc;
}

uint64_t oracle_scan_exclusive_u64(const uint64_t* in, uint64_t* out, int64_t n) {
    uint64_t acc = 0;
    for (int64_t i = 0; i < n; ++i) {
        uint64_t x = in[i];
        out[i] = acc;
        acc += x;
    }
    return acc;
}

/* ------------------------------------------------------------------ gather / scatter
 * The multi-pass loop of opencl/primitives/gatherImpl.cpp:37-53 with the
 * kernel body of opencl/kernels/gather_kernel.cl:12-17: len = ceil(n/pass),
 * pass k copies the elements whose index lies in [k*len, (k+1)*len).
 * elem = 4 or 8 bytes.                                                     */
void oracle_gather(const void* in, void* out, const int32_t* idx, int64_t n, int passes, int elem) {
    if (passes < 1) passes = 1;
    int64_t len = (n + passes - 1) / passes;
    for (int p = 0; p < passes; ++p) {
        int64_t from = (int64_t)p * len, to = from + len;
        for (int64_t i = 0; i < n; ++i) {
            int64_t pos = idx[i];
            if (pos >= from && pos < to && pos < n)
                memcpy((char*)out + i * elem, (const char*)in + pos * elem, (size_t)elem);
        }
    }
}

/* opencl/primitives/scatterImpl.cpp:9-54, opencl/kernels/scatter_kernel.cl:14-20 */
void oracle_scatter(const void* in, void* out, const int32_t* idx, int64_t n, int passes, int elem) {
    if (passes < 1) passes = 1;
    int64_t len = (n + passes - 1) / passes;
    for (int p = 0; p < passes; ++p) {
        int64_t from = (int64_t)p * len, to = from + len;
        for (int64_t i = 0; i < n; ++i) {
            int64_t pos = idx[i];
            if (pos >= from && pos < to && pos < n)
                memcpy((char*)out + pos * elem, (const char*)in + i * elem, (size_t)elem);
        }
    }
}

/* ------------------------------------------------------------------ stable split
 * Sequential stable partition by bucket = key & (B-1): the single-work-item
 * histogram/scan/shuffle of opencl/kernels/split_kernel.cl:675-736 with one
 * chunk, i.e. a counting sort.  start[b] = number of keys in buckets < b
 * (gatherStartPos, split_kernel.cl:320-332).  key_bytes = 4 or 8; vals may
 * be NULL.                                                                  */
void oracle_split(const void* keys_in, void* keys_out, const int32_t* vals_in, int32_t* vals_out, int32_t* start,
                  int64_t n, int B, int key_bytes) {
    int64_t* count = (int64_t*)calloc((size_t)B + 1, sizeof(int64_t));
    uint32_t mask = (uint32_t)B - 1u;
    for (int64_t i = 0; i < n; ++i) {
        uint32_t lowword;
        memcpy(&lowword, (const char*)keys_in + i * key_bytes, 4); /* little endian: low 32 bits */
        count[(lowword & mask) + 1]++;
    }
    for (int b = 0; b < B; ++b) count[b + 1] += count[b];
    if (start)
        for (int b = 0; b < B; ++b) start[b] = (int32_t)count[b];
    for (int64_t i = 0; i < n; ++i) {
        uint32_t lowword;
        memcpy(&lowword, (const char*)keys_in + i * key_bytes, 4);
        int64_t p = count[lowword & mask]++;
        memcpy((char*)keys_out + p * key_bytes, (const char*)keys_in + i * key_bytes, (size_t)key_bytes);
        if (vals_in) vals_out[p] = vals_in[i];
    }
    free(count);
}

/* ------------------------------------------------------------------ hash join
 * Count of equal-key pairs of R and S, the result of the reference's
 * partitioned hash join (obsolete/OpenCL/phjImpl.cpp:13-66: partition both
 * relations by key bits, then per partition pair build a hash table of one
 * side and probe it with the other, counting matches,
 * obsolete/OpenCL/hj_partitioned_kernel.cl:60-115), restated sequentially:
 * partition both by the low 8 key bits with a counting sort, then per
 * partition count matches through a chained hash table of the S side.  Every
 * key value takes part (the reference's kernel reserves -1 as "empty").      */
static uint32_t o_mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x85EBCA6Bu;
    x ^= x >> 13;
    x *= 0xC2B2AE35u;
    x ^= x >> 16;
    return x;
}

static void o_partition(const uint32_t* k, int64_t n, uint32_t* out, int64_t* start) {
    int64_t cnt[257] = {0};
    for (int64_t i = 0; i < n; ++i) cnt[(k[i] & 255u) + 1]++;
    for (int b = 0; b < 256; ++b) cnt[b + 1] += cnt[b];
    for (int b = 0; b <= 256; ++b) start[b] = cnt[b];
    for (int64_t i = 0; i < n; ++i) out[cnt[k[i] & 255u]++] = k[i];
}

uint64_t oracle_join_count(const int32_t* r, int64_t rn, const int32_t* s, int64_t sn) {
    uint32_t* rp = (uint32_t*)malloc((size_t)(rn > 0 ? rn : 1) * 4);
    uint32_t* sp = (uint32_t*)malloc((size_t)(sn > 0 ? sn : 1) * 4);
    int64_t rst[257], sst[257];
    o_partition((const uint32_t*)r, rn, rp, rst);
    o_partition((const uint32_t*)s, sn, sp, sst);
    uint64_t total = 0;
    for (int b = 0; b < 256; ++b) {
        const int64_t s0 = sst[b], s1 = sst[b + 1], r0 = rst[b], r1 = rst[b + 1];
        if (s1 == s0 || r1 == r0) continue;
        /* chained table of the S partition: heads by hash, next links */
        int64_t len = s1 - s0, nb = 1;
        while (nb < 2 * len) nb <<= 1;
        int64_t* head = (int64_t*)malloc((size_t)nb * sizeof(int64_t));
        int64_t* next = (int64_t*)malloc((size_t)len * sizeof(int64_t));
        for (int64_t h = 0; h < nb; ++h) head[h] = -1;
        for (int64_t j = 0; j < len; ++j) {
            const int64_t h = (int64_t)(o_mix32(sp[s0 + j]) & (uint32_t)(nb - 1));
            next[j] = head[h];
            head[h] = j;
        }
        for (int64_t i = r0; i < r1; ++i) {
            const int64_t h = (int64_t)(o_mix32(rp[i]) & (uint32_t)(nb - 1));
            for (int64_t j = head[h]; j >= 0; j = next[j]) total += sp[s0 + j] == rp[i];
        }
        free(head);
        free(next);
    }
    free(rp);
    free(sp);
    return total;
}

/* ------------------------------------------------------------------ radix sort
 * LSD radix sort as the reference's radixSort runs it
 * (obsolete/CUDA/kernels/radixSort.cu:20-26: count the digits, exclusive-scan
 * the counts, scatter; SORT_BITS = 4 digits per pass,
 * obsolete/CUDA/inc/dataDef.h:16-17), one digit of key bits [begin, end) per
 * pass, lowest first; every pass is a stable counting sort, so the result is
 * the stable sort by those bits.  is_signed orders the keys as int32 (bit 31
 * inverted when it is among the sorted bits).  vals may be NULL.             */
void oracle_sort_u32(const uint32_t* keys, const int32_t* vals, uint32_t* keys_out, int32_t* vals_out, int64_t n,
                     int begin_bit, int end_bit, int is_signed) {
    uint32_t* ka = (uint32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(uint32_t));
    int32_t* va = (int32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    uint32_t* kb = (uint32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(uint32_t));
    int32_t* vb = (int32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    const uint32_t twiddle = is_signed ? 0x80000000u : 0u;
    for (int64_t i = 0; i < n; ++i) {
        ka[i] = keys[i] ^ twiddle;
        va[i] = vals ? vals[i] : 0;
    }
    for (int lo = begin_bit; lo < end_bit; lo += 4) {
        const int bits = end_bit - lo < 4 ? end_bit - lo : 4;
        const uint32_t mask = (1u << bits) - 1u;
        int64_t count[16] = {0}, pos[16];
        for (int64_t i = 0; i < n; ++i) count[(ka[i] >> lo) & mask]++;
        int64_t acc = 0;
        for (int d = 0; d < 16; ++d) {
            pos[d] = acc;
            acc += count[d];
        }
        for (int64_t i = 0; i < n; ++i) {
            const int64_t p = pos[(ka[i] >> lo) & mask]++;
            kb[p] = ka[i];
            vb[p] = va[i];
        }
        uint32_t* tk = ka; ka = kb; kb = tk;
        int32_t* tv = va; va = vb; vb = tv;
    }
    for (int64_t i = 0; i < n; ++i) {
        keys_out[i] = ka[i] ^ twiddle;
        if (vals_out) vals_out[i] = va[i];
    }
    free(ka);
    free(va);
    free(kb);
    free(vb);
}

/* ------------------------------------------------------------------ STREAM
 * opencl/kernels/mem_kernel.cl:9-31 with SCALAR = 3 (mem_kernel.cl:6):
 * 0 copy b=a, 1 scale b=3a, 2 add c=a+b, 3 triad c=a+3b (int32 wrap)       */
void oracle_stream(int op, const int32_t* a, const int32_t* b, int32_t* c, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        uint32_t x = (uint32_t)a[i], y = b ? (uint32_t)b[i] : 0u, r;
        switch (op) {
            case 0: r = x; break;
            case 1: r = x * 3u; break;
            case 2: r = x + y; break;
            default: r = x + 3u * y; break;
        }
        c[i] = (int32_t)r;
    }
}

/* ------------------------------------------------------------------ statistics
 * average_Hampel of opencl/util/utility.cpp:38-83 (= openmp/util/utility.cpp):
 * median m, MAD = median |x-m| / 0.6745; keep samples with |x-m|/MAD <= 3.5
 * (only |x-m| == 0 when MAD == 0); return their mean.                       */
static int o_cmp(const void* a, const void* b) {
    double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

static double o_median_sorted(const double* v, int n) {
    return (n % 2 == 0) ? 0.5 * (v[n / 2 - 1] + v[n / 2]) : v[(n - 1) / 2];
}

double oracle_hampel_mean(const double* x, int n) {
    if (n <= 0) return 0.0;
    double* s = (double*)malloc(sizeof(double) * (size_t)n);
    double* d = (double*)malloc(sizeof(double) * (size_t)n);
    memcpy(s, x, sizeof(double) * (size_t)n);
    qsort(s, (size_t)n, sizeof(double), o_cmp);
    double med = o_median_sorted(s, n);
    for (int i = 0; i < n; ++i) d[i] = fabs(x[i] - med);
    double* ds = (double*)malloc(sizeof(double) * (size_t)n);
    memcpy(ds, d, sizeof(double) * (size_t)n);
    qsort(ds, (size_t)n, sizeof(double), o_cmp);
    double mad = o_median_sorted(ds, n) / 0.6745;
    double total = 0.0;
    int valid = 0;
    for (int i = 0; i < n; ++i) {
        int keep = (mad == 0.0) ? (d[i] == 0.0) : (d[i] / mad <= 3.5);
        if (keep) {
            total += x[i];
            valid++;
        }
    }
    free(s);
    free(d);
    free(ds);
    return total / valid;
}
