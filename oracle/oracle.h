/*
 * oracle.h -- CPU FP64 restatement of the CTBB_ICD reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the ctbb_icd_amd
 * library, its C ABI or its driver) links, loads or calls this code; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * as the checker.  Every function cites the reference lines it restates.
 *
 * Parity status of each function (what pins it) is listed in oracle.c.
 */
#ifndef CTBB_ORACLE_H
#define CTBB_ORACLE_H

#include <stdint.h>
#include "../include/ctbb_icd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A3/A4: stored system matrix, ray-driven, in the reference's loop order.
 * Pass 1 (col_cnt != NULL): col_cnt[nx*ny] = entries per column; returns the
 * total.  Pass 2 (col_ptr given): fills idx/val in ascending ray order. */
int64_t orc_ssm_count(const ctbb_geom *g, int64_t *col_cnt);
int orc_ssm_fill(const ctbb_geom *g, const int64_t *col_ptr, int32_t *idx,
                 float *val);

/* A6: yhat = A x over every slice (FP64, sequential sums). */
void orc_forward_project(const ctbb_geom *g, const int64_t *col_ptr,
                         const int32_t *idx, const float *val,
                         const int64_t *slice_index, const double *x,
                         double *yhat);

/* A8: penalty steps for voxel q of x (FP64), 8-neighbour flat stencil. */
double orc_quadratic(const ctbb_geom *g, int64_t q, double alpha, double beta,
                     double lambda, const double *x);
double orc_edge_preserving(const ctbb_geom *g, int64_t q, double alpha,
                           double beta, double lambda, double delta,
                           const double *x);

/* A7: ICD iterations.  x (FP64 volume) and yhat (FP64 A x) are updated in
 * place.  weights may be NULL.  spacing = 0 replays the reference order
 * (columns in raster order, slices ascending, every update applied at
 * once); spacing >= 2 replays the batched z-group order of the GPU sweep:
 * batches of columns at one local position of every tile x tile block
 * (raster over local positions), per batch the slice groups
 * {j, j+s, j+2s, ...}, j = 0..s-1, every voxel of a step computed from the
 * state at the start of the step.  tile >= max(nx, ny) gives single-column
 * batches.  Returns the number of voxel updates. */
int64_t orc_icd(const ctbb_geom *g, const int64_t *col_ptr,
                const int32_t *idx, const float *val, const float *raw,
                const float *weights, const int64_t *slice_index,
                int penalty, double lambda, double delta, int64_t spacing,
                int64_t tile, int n_iter, double *x, double *yhat);
/* orc_icd with the batch surrogate option (surrogate != 0, batched order
 * only): the curvature of voxel (c, k) is sum_i w a_ic * (sum of the batch's
 * values on SSM ray i) instead of sum_i w a_ic^2. */
int64_t orc_icd_ex(const ctbb_geom *g, const int64_t *col_ptr,
                   const int32_t *idx, const float *val, const float *raw,
                   const float *weights, const int64_t *slice_index,
                   int penalty, double lambda, double delta, int64_t spacing,
                   int64_t tile, int surrogate, int n_iter, double *x,
                   double *yhat);

/* A9: global cost of x (computes a fresh A x in FP64). */
double orc_cost(const ctbb_geom *g, const int64_t *col_ptr,
                const int32_t *idx, const float *val, const float *raw,
                const float *weights, const int64_t *slice_index,
                int penalty, double lambda, double delta, const double *x);

/* A10: rotate every slice in place (direction +1 rotating->fixed,
 * -1 fixed->rotating). */
void orc_rotate_slices(const ctbb_geom *g, const double *tube_angles,
                       const int64_t *slice_index, float *volume,
                       int direction);

/* Exact footprint test: 1 if column c's entries, shifted by dv views, meet
 * the unshifted entries (S and S + dv*n_channels*n_rows intersect). */
int orc_column_overlaps(const ctbb_geom *g, const int64_t *col_ptr,
                        const int32_t *idx, int64_t c, int64_t dv);

#ifdef __cplusplus
}
#endif
#endif
