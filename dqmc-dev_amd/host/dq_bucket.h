/* dq_bucket.h -- files of uneven size in one engine (size buckets).
 *
 * No counterpart in the reference, which runs one file per process
 * (src/main_stack.c).  Files whose N falls in one bucket (N <= 64, 121, 196,
 * 289, 400, 512) and that share n_matmul, period_eqlt, period_uneqlt, the
 * element type and the default measurement sets (no optional groups) run as
 * one engine: each file is padded to the group's largest N and L.  The
 * padding sites are inert (identity blocks in every exponential,
 * exp_lambda = 1, del = 0, map entries -1: no measurement bin) and the
 * padding slices are whole identity groups (every file's L is a multiple of
 * n_matmul); the engine skips both (dqmc_tuning.chain_L / chain_N).  Every
 * file keeps its own random stream, recalculation slices and measurement
 * slices, so it runs as it would alone.  Unequal-time accumulators are laid
 * out [L][bins]: the file's slices and bins are the first ones of the padded
 * [L_pad][bins_pad] arrays. */
#ifndef DQ_BUCKET_H
#define DQ_BUCKET_H

#include "dq_h5.h"

/* the bucket of a file (its padded site-count bound), or 0 when the file
 * cannot share an engine with files of other sizes */
int dq_bucket_of(const dqmc_params *p);

/* two files may share a bucketed engine */
int dq_bucket_compatible(const dqmc_params *a, const dqmc_params *b);

typedef struct dq_pad {
    dqmc_params p;     /* padded parameters; arrays in own storage */
    int32_t *hs;       /* padded HS field [L][N] */
    dqmc_meas meq;     /* padded equal-time accumulators */
    dqmc_meas mue;     /* padded unequal-time accumulators (period_uneqlt > 0) */
    int N0, L0;        /* the file's own sizes */
    void *own[96];
    int n_own;
} dq_pad;

/* padded copy of a file's parameters for a group of N_pad sites, L_pad slices
 * and num_i_pad / num_ij_pad bins; 0 or a negative error */
int dq_pad_build(const dq_sim *s, int N_pad, int L_pad, int num_i_pad, int num_ij_pad,
                 dq_pad *d);
/* the file's field and accumulators into the padded buffers, and back */
void dq_pad_in(const dq_sim *s, dq_pad *d);
void dq_pad_out(const dq_pad *d, dq_sim *s);
void dq_pad_free(dq_pad *d);

#endif
