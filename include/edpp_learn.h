/*
 * edpp_learn.h — schedule candidates, exhaustive autotuning and the learned
 * FP32 cost model that picks a schedule per call.
 *
 * No counterpart in the reference beyond its exhaustive launch-parameter
 * sweep (opencl/test/test_split.cpp:273-358) and hand-picked scan settings
 * (opencl/test/test_scan_global.cpp:87-94).  Every candidate is a complete
 * schedule of one primitive; all candidates of a primitive give bit-identical
 * results, so a wrong prediction costs time, never correctness.
 *
 * Model: 12 features -> 64 -> 64 -> EDPP_MAX_CANDIDATES logits (ReLU), softmax
 * restricted to the candidates of the call's primitive.  Features (in order):
 *   0-2  one-hot primitive family (scan, gather/scatter, split)
 *   3    log2(n) / 28
 *   4    element (key) bytes / 8
 *   5    log2(buckets) / 12            (split; 0 otherwise)
 *   6    log2(sampled bucket skew) / 12 (split: max bucket share x B on a
 *        4096-key device sample; 0 otherwise)
 *   7    sampled index locality         (gather/scatter: share of sampled
 *        adjacent positions whose indices are within 4096; 0 otherwise)
 *   8    log2(bytes moved / 256 MiB Infinity Cache) / 8
 *   9    log2(bytes moved / 4 MiB L2) / 16
 *   10   log2(4096-element tiles per CU) / 16
 *   11   scatter flag (1 for scatter)
 * Parameters (float32, this order): W1[64][12] b1[64] W2[64][64] b2[64]
 * W3[K][64] b3[K] with K = EDPP_MAX_CANDIDATES.
 */
#ifndef EDPP_LEARN_H
#define EDPP_LEARN_H

#include <stdint.h>
#include "edpp.h"

#ifdef __cplusplus
extern "C" {
#endif

#define EDPP_NUM_FEATURES 12
#define EDPP_HIDDEN 64
#define EDPP_MAX_CANDIDATES 15

enum edpp_op { EDPP_OP_SCAN = 0, EDPP_OP_GATHER = 1, EDPP_OP_SCATTER = 2, EDPP_OP_SPLIT = 3 };

typedef struct edpp_call {
    int32_t op;              /* edpp_op                                                      */
    int32_t elem_bytes;      /* 4 or 8: scan values, gather/scatter values, split keys       */
    int64_t n;
    int32_t buckets;         /* split only                                                   */
    int32_t reserved;
    const void* in;          /* scan in / gather-scatter source / split keys in              */
    void* out;               /* scan out / gather-scatter destination / split keys out       */
    const int32_t* idx;      /* gather / scatter                                             */
    const int32_t* vals_in;  /* split, may be NULL                                           */
    int32_t* vals_out;
} edpp_call;

int edpp_num_candidates(void);
/* family: 0 scan, 1 gather/scatter, 2 split; name: short description */
int edpp_candidate_info(int32_t cand, int32_t* family, char* name, int32_t name_len);
/* Run one call with candidate `cand`, or with the context's choice when cand < 0. */
int edpp_call_run(edpp_ctx* ctx, const edpp_call* call, int32_t cand, double* ms);
/* Exhaustive autotune: time every candidate of the call's family (`reps`
 * timed runs after one warm-up, median); times[k] = ms, or -1 for candidates
 * of other families.  times has EDPP_MAX_CANDIDATES entries. */
int edpp_autotune(edpp_ctx* ctx, const edpp_call* call, int32_t reps, double* times);
/* The candidate the built-in schedule (no model) gives this call, or -1 when
 * that schedule is none of the candidates. */
int edpp_default_candidate(edpp_ctx* ctx, const edpp_call* call, int32_t* cand);
/* Features of a call, computed on the device from samples of its inputs. */
int edpp_call_features(edpp_ctx* ctx, const edpp_call* call, float* d_features);

typedef struct edpp_mlp edpp_mlp;
int edpp_mlp_num_params(void);
int edpp_mlp_create(edpp_ctx* ctx, uint64_t seed, edpp_mlp** out);
int edpp_mlp_destroy(edpp_mlp* m);
int edpp_mlp_get_params(edpp_mlp* m, float* host_params, int64_t count);
int edpp_mlp_set_params(edpp_mlp* m, const float* host_params, int64_t count);
/* Row-major device arrays: x[rows][12], mask/target/probs[rows][K]; mask is
 * 1 for candidates allowed in that row, 0 otherwise. */
int edpp_mlp_forward(edpp_mlp* m, const float* d_x, const float* d_mask, int32_t rows, float* d_probs,
                     int32_t* d_choice);
/* Mean cross-entropy against soft targets and its gradient (device, num_params floats). */
int edpp_mlp_loss_grad(edpp_mlp* m, const float* d_x, const float* d_target, const float* d_mask, int32_t rows,
                       float* d_loss, float* d_grad);
/* Full-batch Adam for `epochs` steps on the device; loss_host (may be NULL)
 * receives the loss of every step. */
int edpp_mlp_train(edpp_mlp* m, const float* d_x, const float* d_target, const float* d_mask, int32_t rows,
                   int32_t epochs, float lr, float* loss_host);
/* Let AUTO calls on ctx pick their schedule with the model (NULL: built-in defaults). */
int edpp_ctx_set_model(edpp_ctx* ctx, edpp_mlp* m);
/* Decision memo of AUTO calls (no counterpart in the reference, like the rest
 * of this header).  On (the default), the model remembers its choice per call
 * signature: op, element width, n, buckets, whether values are given, and the
 * input and index pointers.  A repeated call then skips the selection kernel
 * and its synchronisation.  Every remembered choice is dropped when the
 * parameters change (edpp_mlp_train, edpp_mlp_set_params).  A buffer whose
 * contents change under the same signature keeps its first choice; that can
 * cost time, never correctness, since every candidate is bit-exact.
 * enabled = 0 turns the memo off (every AUTO call runs the selection kernel);
 * either setting forgets the remembered choices. */
int edpp_mlp_memo(edpp_mlp* m, int32_t enabled);
/* The candidate the context's model picks for the call (through the memo), as
 * an AUTO call would; EDPP_EINVAL when no model is set. */
int edpp_model_choice(edpp_ctx* ctx, const edpp_call* call, int32_t* cand);

#ifdef __cplusplus
}
#endif
#endif /* EDPP_LEARN_H */
