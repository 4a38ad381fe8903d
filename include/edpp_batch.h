/*
 * edpp_batch.h — many primitives of uneven length in one persistent launch.
 *
 * No counterpart in the reference, whose primitives are one synchronous call
 * each (opencl/primitives/*.cpp: every launch followed by clFinish).  A batch
 * is planned once: every job is cut into tiles, the tiles form one device
 * work queue (split histograms first, then the scans of those histograms,
 * then scan, gather and split-scatter tiles; a split job of one tile has
 * only its scatter tile), and one persistent kernel
 * drains it.  A tile only ever waits for tiles with smaller queue tickets,
 * which belong to workgroups that are already running, so the launch always
 * completes.  Results are bit-exact per job with the single-call API.
 */
#ifndef EDPP_BATCH_H
#define EDPP_BATCH_H

#include <stdint.h>
#include "edpp.h"

#ifdef __cplusplus
extern "C" {
#endif

enum edpp_job_kind {
    EDPP_JOB_SCAN   = 0,   /* exclusive scan: in -> out, aux = total (may be NULL)                       */
    EDPP_JOB_SPLIT  = 1,   /* stable split: in = keys, out = keys out, vals_in/vals_out optional,
                              aux = int32[buckets] bucket starts (may be NULL); buckets <= 256         */
    EDPP_JOB_GATHER = 2    /* out[i] = in[idx[i]]                                                        */
};

typedef struct edpp_job {
    int32_t kind;           /* edpp_job_kind                                   */
    int32_t elem_bytes;     /* 4 or 8: scan values, split keys, gather values  */
    int64_t n;              /* elements, < 2^31                                 */
    int32_t buckets;        /* split only: power of two in [1, 256]            */
    int32_t reserved;
    const void* in;
    void* out;
    const int32_t* idx;     /* gather only                                     */
    const int32_t* vals_in; /* split only, may be NULL                         */
    int32_t* vals_out;
    void* aux;
} edpp_job;

typedef struct edpp_batch edpp_batch;

/* Plan a batch (validates every job, builds and uploads the work queue and
 * the per-job state).  The job buffers must stay valid while the batch is
 * used; the descriptors are copied. */
int edpp_batch_create(edpp_ctx* ctx, const edpp_job* jobs, int32_t njobs, edpp_batch** out);
/* One state reset kernel plus one persistent launch on the context's stream. */
int edpp_batch_run(edpp_batch* b, double* ms);
/* Capture the two launches into a hipGraph once; later calls replay it. */
int edpp_batch_run_graph(edpp_batch* b, double* ms);
int edpp_batch_destroy(edpp_batch* b);
/* Workgroups of the persistent launch per compute unit (tuning; default 4). */
int edpp_batch_set_occupancy(edpp_batch* b, int32_t wgs_per_cu);
/* Tile shapes of the persistent launch.  Shape k fixes the lanes (waves) per
 * workgroup and the elements per tile of each kind: 32- and 64-bit scan
 * tiles, split tiles (lanes x keys per lane) and gather tiles.  Shape 5 is
 * the default.  Setting a shape re-plans the batch's work queue; results do
 * not depend on the shape or the occupancy. */
int edpp_batch_num_shapes(void);
int edpp_batch_shape_info(int32_t shape, int32_t* block, int32_t* scan_tile32, int32_t* scan_tile64,
                          int32_t* split_tile, int32_t* gather_tile);
int edpp_batch_set_shape(edpp_batch* b, int32_t shape);
/* Queue order of the launch: every split's scatter tiles after all scan and
 * gather tiles (EDPP_BATCH_ORDER_PHASES), or each job's main tiles in job
 * order, split scatters included (EDPP_BATCH_ORDER_JOBS), so that streaming
 * tiles run beside random-gather tiles, or the jobs dealt to one queue per
 * XCD, each queue in job order (EDPP_BATCH_ORDER_XCD, the default), so that
 * all tiles of a job run on one XCD and share its L2; a workgroup whose
 * queue is used up takes tiles from the next.  Re-plans the queue; results
 * do not depend on it. */
enum edpp_batch_order { EDPP_BATCH_ORDER_PHASES = 0, EDPP_BATCH_ORDER_JOBS = 1, EDPP_BATCH_ORDER_XCD = 2 };
int edpp_batch_set_order(edpp_batch* b, int32_t order);
int edpp_batch_get_shape(edpp_batch* b, int32_t* shape, int32_t* wgs_per_cu, int32_t* order);
/* Exhaustive autotune of one batch, as the reference's launch-parameter
 * search tries every local and grid size of a split and keeps the fastest
 * (opencl/test/test_split.cpp:273-358): the three queue orders, every
 * shape, at 2, 4 and 8 workgroups per compute unit, one warm-up run and
 * `reps` timed runs each (median); the fastest setting stays.  times (may be
 * NULL) receives 3 * edpp_batch_num_shapes() * 3 medians in ms, indexed
 * [order][shape][occupancy].  Runs the batch repeatedly, so jobs writing
 * their own input are rejected (EDPP_EINVAL). */
int edpp_batch_autotune(edpp_batch* b, int32_t reps, double* times);
/* Queue statistics: total tiles, split histogram tiles, histogram-scan tiles
 * (split jobs of one tile have neither). */
int edpp_batch_info(edpp_batch* b, int64_t* tiles, int64_t* hist_tiles, int64_t* hscan_tiles);

/* The baseline the batch replaces: one ordinary library call per job. */
int edpp_batch_run_plain(edpp_ctx* ctx, const edpp_job* jobs, int32_t njobs, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* EDPP_BATCH_H */
