/*
 * oracle_cpu_ref.c -- the reference's OpenMP ICD structure, restated, for the
 * CPU baseline timed beside the GPU path.
 *
 * TEST / BASELINE INFRASTRUCTURE ONLY.  Same arithmetic as orc_icd (FP64
 * image and sinogram estimate, loop order y -> x -> z), organised like
 * src/icd_iteration.cpp:157-255: per voxel update one
 * `omp parallel for reduction(+:alpha,beta)` over the column's entries and
 * one `omp parallel for` scatter, with a team of `threads` threads
 * (the reference hard-codes 11, src/icd_iteration.cpp:40).  The matrix is
 * held in RAM (the reference re-reads it from disk each iteration,
 * src/icd_iteration.cpp:164-179; that I/O is not timed here).
 * Only the quadratic and edge-preserving 8-neighbour penalties are used.
 * Pin: equals the sequential orc_icd in the reference order (spacing 0)
 * within 1e-12 on tiny and tiny_ffs (tests/test_oracle_clinical.py).
 */
#include "oracle.h"

#include <omp.h>
#include <stdlib.h>

int64_t orc_icd_ref_omp(const ctbb_geom *g, const int64_t *col_ptr,
                        const int32_t *idx, const float *val, const float *raw,
                        const float *weights, const int64_t *slice_index,
                        int penalty, double lambda, double delta, int n_iter,
                        int threads, int64_t row_begin, int64_t row_end,
                        double *x, double *yhat)
{
    int64_t updates = 0;
    int64_t data_size = g->readings * g->n_channels * g->n_rows;
    double fov_limit = (g->acquisition_fov / 2.0) * (g->acquisition_fov / 2.0);
    if (row_end <= 0 || row_end > g->ny)
        row_end = g->ny;
    for (int n = 0; n < n_iter; n++) {
        for (int64_t j = row_begin; j < row_end; j++) {
            double y = (j - g->center_voxel_y) * g->voxel_size_y;
            for (int64_t i = 0; i < g->nx; i++) {
                double xx = (i - g->center_voxel_x) * g->voxel_size_x;
                if (!((xx * xx + y * y) < fov_limit))
                    continue;
                int64_t c = i + g->nx * j;
                int64_t p0 = col_ptr[c];
                int nnz = (int)(col_ptr[c + 1] - p0);
                for (int64_t k = 0; k < g->nz; k++) {
                    int64_t q = c + g->nx * g->ny * k;
                    int64_t off = (slice_index[k] - g->n_views_sm / 2)
                                  * g->n_channels * g->n_rows;
                    double alpha = 0.0, beta = 0.0;
#pragma omp parallel for num_threads(threads) reduction(+:alpha,beta)
                    for (int m = 0; m < nnz; m++) {
                        int64_t r = idx[p0 + m] + off;
                        if (r > -1 && r < data_size) {
                            /* FP32 square, FP64 sums (src/icd_iteration.cpp:208-209) */
                            float af = val[p0 + m];
                            double a = af;
                            double w = weights ? (double)weights[r] : 1.0;
                            alpha += weights ? w * (double)(af * af) : (double)(af * af);
                            beta += w * a * ((double)raw[r] - yhat[r]);
                        }
                    }
                    double du;
                    if (penalty == CTBB_PENALTY_QUADRATIC)
                        du = orc_quadratic(g, q, alpha, beta, lambda, x);
                    else
                        du = orc_edge_preserving(g, q, alpha, beta, lambda,
                                                 delta, x);
                    if (du + x[q] < 0)
                        du = -x[q];
                    x[q] += du;
#pragma omp parallel for num_threads(threads)
                    for (int m = 0; m < nnz; m++) {
                        int64_t r = idx[p0 + m] + off;
                        if (r > -1 && r < data_size)
                            yhat[r] += du * val[p0 + m];
                    }
                    updates++;
                }
            }
        }
    }
    return updates;
}
