/*
 * oracle.h -- CPU FP64 restatement of the CTBB_ICD reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the ctbb_icd_amd
 * library, its C ABI or its driver) links, loads or calls this code; only
 * tests/, __graft_entry__.smoke() and bench.py's CPU baseline legs (ICD and
 * matrix build) use it, as the checker or the timed CPU baseline.  Every
 * function cites the reference lines it restates.
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

/* The ray-driven build restricted to SSM views [v0, v1): the same rays,
 * crossings and entries as orc_ssm_count/fill over those views (their q
 * keep the full-matrix numbering), so the matrix is the concatenation, view
 * range by view range, of these pieces per column.  Returns -1 for a range
 * outside [0, n_views_sm).  Used to time the reference's single-threaded
 * build on a sample of views. */
int64_t orc_ssm_count_views(const ctbb_geom *g, int64_t v0, int64_t v1,
                            int64_t *col_cnt);
int orc_ssm_fill_views(const ctbb_geom *g, int64_t v0, int64_t v1,
                       const int64_t *col_ptr, int32_t *idx, float *val);

/* A3/A4 for a subset of columns, voxel-driven (the same per-crossing
 * expressions as the ray loop, every ray of every view screened): pass 1
 * writes cnt[t] = entries of column cols[t] and returns the total; pass 2
 * fills idx/val at ptr[t] (exclusive scan of cnt) in ascending ray order.
 * Equal to the matching columns of orc_ssm_fill; cheap for a few columns of
 * a clinical geometry, where the full ray-driven build is out of reach. */
int64_t orc_ssm_cols_count(const ctbb_geom *g, int64_t ncols, const int64_t *cols,
                           int64_t *cnt);
void orc_ssm_cols_fill(const ctbb_geom *g, int64_t ncols, const int64_t *cols,
                       const int64_t *ptr, int32_t *idx, float *val);

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
/* orc_icd with the batch surrogate option (surrogate != 0, batched order,
 * multi-column batches; 2 also for single-column batches): the curvature of
 * voxel v of a step is sum_r w_r a_vr S_r, S_r = sum over every voxel of the
 * step (all its columns and slices) of their values on data sample r,
 * instead of sum_r w_r a_vr^2. */
int64_t orc_icd_ex(const ctbb_geom *g, const int64_t *col_ptr,
                   const int32_t *idx, const float *val, const float *raw,
                   const float *weights, const int64_t *slice_index,
                   int penalty, double lambda, double delta, int64_t spacing,
                   int64_t tile, int surrogate, int n_iter, double *x,
                   double *yhat);

/* orc_icd_ex that also records step_cost[0] = the cost before the first
 * step and step_cost[t] = the cost after step t (t < max_costs; batched
 * order only), the number of steps in *n_steps, and in shared[3] the data
 * samples reached by two voxels of one step: [0] of different columns and
 * different slices, [1] of one column and different slices (footprint
 * conflicts inside a column), [2] of different columns and one slice.
 * step_cost, n_steps and shared may be NULL. */
int64_t orc_icd_trace(const ctbb_geom *g, const int64_t *col_ptr,
                      const int32_t *idx, const float *val, const float *raw,
                      const float *weights, const int64_t *slice_index,
                      int penalty, double lambda, double delta, int64_t spacing,
                      int64_t tile, int surrogate, int n_iter, double *x,
                      double *yhat, double *step_cost, int64_t max_costs,
                      int64_t *n_steps, int64_t *shared);

/* orc_icd_trace with the penalty neighbourhood: 8 (the reference's flat
 * in-plane stencil) or 26 (8 in-plane neighbours without wrap plus 9 in each
 * adjacent slice through the slice rotation, see oracle.c). */
int64_t orc_icd_nb(const ctbb_geom *g, const int64_t *col_ptr,
                   const int32_t *idx, const float *val, const float *raw,
                   const float *weights, const int64_t *slice_index,
                   int penalty, double lambda, double delta, int neighbours,
                   int64_t spacing, int64_t tile, int surrogate, int n_iter,
                   double *x, double *yhat, double *step_cost, int64_t max_costs,
                   int64_t *n_steps, int64_t *shared);
/* Penalty step of voxel q (8 or 26 neighbours), before the clamp. */
double orc_penalty_step(const ctbb_geom *g, const int64_t *slice_index,
                        int neighbours, int penalty, int64_t q, double alpha,
                        double beta, double lambda, double delta, const double *x);
/* Cost with the penalty neighbourhood (8: orc_cost). */
double orc_cost_nb(const ctbb_geom *g, const int64_t *col_ptr,
                   const int32_t *idx, const float *val, const float *raw,
                   const float *weights, const int64_t *slice_index,
                   int penalty, double lambda, double delta, int neighbours,
                   const double *x);
/* The slice-to-slice rotation pi_k of the 26-neighbour penalty (inverse != 0:
 * its inverse) applied to grid point (i, j). */
void orc_rot26(const ctbb_geom *g, const int64_t *slice_index, int64_t k,
               int inverse, int64_t i, int64_t j, int64_t *io, int64_t *jo);

/* CPU baseline: the reference's OpenMP structure (oracle_cpu_ref.c). */
int64_t orc_icd_ref_omp(const ctbb_geom *g, const int64_t *col_ptr,
                        const int32_t *idx, const float *val, const float *raw,
                        const float *weights, const int64_t *slice_index,
                        int penalty, double lambda, double delta, int n_iter,
                        int threads, int64_t row_begin, int64_t row_end,
                        double *x, double *yhat);

/* The same sweep with the matrix read from a matrix.bin-format file on every
 * iteration, as src/icd_iteration.cpp:164-179,270-273 reads it: the file
 * holds the columns of rows [row_begin, row_end) (x fastest), each a uint64
 * nnz and nnz {int32 index, float32 value} pairs; every column is read into
 * a fresh array, inside the FOV or not, and the file is rewound after each
 * iteration.  Returns the update count, or -1 if the file cannot be opened
 * or ends early. */
int64_t orc_icd_ref_omp_file(const ctbb_geom *g, const char *path, const float *raw,
                             const float *weights, const int64_t *slice_index,
                             int penalty, double lambda, double delta, int n_iter,
                             int threads, int64_t row_begin, int64_t row_end,
                             double *x, double *yhat);

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
