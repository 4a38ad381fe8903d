/*
 * edpp_tune.h — schedule knobs of the primitives, for the autotuner and the
 * learned cost model.  Every knob only changes how the work is scheduled;
 * results stay bit-exact for every setting.
 *
 * The reference tunes launch shapes by exhaustive search over local and
 * grid sizes (opencl/test/test_split.cpp:273-358) and hand-picked scan
 * settings (opencl/test/test_scan_global.cpp:87-94); these knobs are the
 * MI355X counterparts (tile lengths, vectors per lane, cache hints).
 */
#ifndef EDPP_TUNE_H
#define EDPP_TUNE_H

#include <stdint.h>
#include "edpp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct edpp_tuning {
    int32_t gs_unroll;        /* index vectors (4 elements each) per lane: 1, 2, 4, 8; 0 = default      */
    int32_t gs_random_nt;     /* random side of gather/scatter with non-temporal hint: 0 / 1; -1 default */
    int32_t scan_vpt;         /* 16-byte vectors per lane in a look-back tile: 2, 4, 8, 16; 0 = by size
                                 (4 below 2^21 elements, 8 below 2^23, 16 above; 64-bit scans stop at 8) */
    int32_t split_ipt;        /* keys per lane per split sub-tile: 8 or 16; 0 = default                  */
    int32_t split_tile_mult;  /* split tile >= mult * buckets elements: 16..1024; 0 = default (16)       */
    int32_t scan_block;       /* lanes per look-back workgroup: 256, 512, 1024; 0 = default              */
    int32_t split_block;      /* lanes per WG-split scatter workgroup: 256, or 512 (4-byte keys, values
                                 staged after keys, 16 keys per lane); 0 = default: at the large-input
                                 shape 512 lanes from 128 buckets, and 512 lanes with 8 keys per lane
                                 for 8-byte keys from 128 buckets and 4-byte keys at 64               */
    int32_t split_reorder;    /* how the WG split writes a tile (the reference's ReorderType,
                                 opencl/types.h:43-45): EDPP_REORDER_VARIED (the default), _NONE or
                                 _FIXED; 0 = default                                                  */
    int32_t sort_strategy;    /* digit passes of a radix sort: EDPP_SPLIT_WG (per-tile histogram, scan,
                                 scatter) or EDPP_SPLIT_CHAINED (per-range counts, chained scatter);
                                 0 = default                                                          */
    int32_t reserved[7];
} edpp_tuning;

enum edpp_reorder {
    EDPP_REORDER_DEFAULT = 0,
    EDPP_REORDER_NONE    = 1,  /* each element straight from registers to its output position (NO_REORDER) */
    EDPP_REORDER_FIXED   = 2,  /* per-bucket LDS staging of 16 elements, written whenever full (FIXED_REORDER) */
    EDPP_REORDER_VARIED  = 3   /* each bucket's run of a sub-tile staged whole in LDS (VARIED_REORDER)       */
};

/* Applies to every later call on ctx; NULL restores the defaults. */
int edpp_set_tuning(edpp_ctx* ctx, const edpp_tuning* t);
int edpp_get_tuning(edpp_ctx* ctx, edpp_tuning* t);

#ifdef __cplusplus
}
#endif
#endif /* EDPP_TUNE_H */
