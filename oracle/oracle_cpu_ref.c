/*
 * oracle_cpu_ref.c -- the reference's OpenMP ICD structure, restated, for the
 * CPU baseline timed beside the GPU path.
 *
 * TEST / BASELINE INFRASTRUCTURE ONLY.  Same arithmetic as orc_icd (FP64
 * image and sinogram estimate, loop order y -> x -> z), organised like
 * src/icd_iteration.cpp:157-255: per voxel update one
 * `omp parallel for reduction(+:alpha,beta)` over the column's entries and
 * one `omp parallel for` scatter, with a team of `threads` threads
 * (the reference hard-codes 11, src/icd_iteration.cpp:40).  Two ways to get
 * the matrix: held in RAM (orc_icd_ref_omp), or streamed from a matrix.bin
 * file on every iteration as the reference reads it
 * (orc_icd_ref_omp_file, src/icd_iteration.cpp:164-179,270-273).
 * Only the quadratic and edge-preserving 8-neighbour penalties are used.
 * Pins: orc_icd_ref_omp equals the sequential orc_icd in the reference order
 * (spacing 0) within 1e-12 on tiny and tiny_ffs; the streamed form equals
 * the in-RAM one on the same rows (tests/test_oracle_clinical.py).
 */
#include "oracle.h"

#include <omp.h>
#include <stdio.h>
#include <stdlib.h>

/* Every slice of column c = i + nx*j, from its nnz entries
 * (idx[m*stride], val[m*stride]): the body of src/icd_iteration.cpp:183-250.
 * stride 1: separate index and value arrays; stride 2: the file's
 * interleaved {int32 index, float32 value} pairs. */
static void update_column(const ctbb_geom *g, int64_t c, const int32_t *idx,
                          const float *val, int stride, int nnz, const float *raw,
                          const float *weights, const int64_t *slice_index,
                          int penalty, double lambda, double delta, int threads,
                          double *x, double *yhat)
{
    int64_t data_size = g->readings * g->n_channels * g->n_rows;
    for (int64_t k = 0; k < g->nz; k++) {
        int64_t q = c + g->nx * g->ny * k;
        int64_t off = (slice_index[k] - g->n_views_sm / 2) * g->n_channels * g->n_rows;
        double alpha = 0.0, beta = 0.0;
#pragma omp parallel for num_threads(threads) reduction(+:alpha,beta)
        for (int m = 0; m < nnz; m++) {
            int64_t r = idx[(int64_t)m * stride] + off;
            if (r > -1 && r < data_size) {
                /* FP32 square, FP64 sums (src/icd_iteration.cpp:208-209) */
                float af = val[(int64_t)m * stride];
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
            du = orc_edge_preserving(g, q, alpha, beta, lambda, delta, x);
        if (du + x[q] < 0)
            du = -x[q];
        x[q] += du;
#pragma omp parallel for num_threads(threads)
        for (int m = 0; m < nnz; m++) {
            int64_t r = idx[(int64_t)m * stride] + off;
            if (r > -1 && r < data_size)
                yhat[r] += du * val[(int64_t)m * stride];
        }
    }
}

static int inside_fov(const ctbb_geom *g, int64_t i, int64_t j)
{
    double fov_limit = (g->acquisition_fov / 2.0) * (g->acquisition_fov / 2.0);
    double y = (j - g->center_voxel_y) * g->voxel_size_y;
    double xx = (i - g->center_voxel_x) * g->voxel_size_x;
    return (xx * xx + y * y) < fov_limit;
}

int64_t orc_icd_ref_omp(const ctbb_geom *g, const int64_t *col_ptr,
                        const int32_t *idx, const float *val, const float *raw,
                        const float *weights, const int64_t *slice_index,
                        int penalty, double lambda, double delta, int n_iter,
                        int threads, int64_t row_begin, int64_t row_end,
                        double *x, double *yhat)
{
    int64_t updates = 0;
    if (row_end <= 0 || row_end > g->ny)
        row_end = g->ny;
    for (int n = 0; n < n_iter; n++) {
        for (int64_t j = row_begin; j < row_end; j++) {
            for (int64_t i = 0; i < g->nx; i++) {
                if (!inside_fov(g, i, j))
                    continue;
                int64_t c = i + g->nx * j;
                int64_t p0 = col_ptr[c];
                update_column(g, c, idx + p0, val + p0, 1, (int)(col_ptr[c + 1] - p0),
                              raw, weights, slice_index, penalty, lambda, delta,
                              threads, x, yhat);
                updates += g->nz;
            }
        }
    }
    return updates;
}

/* One {index, value} pair of the matrix file (8 bytes, as written by
 * src/generate_system_matrix.cpp:212-246). */
typedef struct {
    int32_t index;
    float value;
} file_pair;

int64_t orc_icd_ref_omp_file(const ctbb_geom *g, const char *path, const float *raw,
                             const float *weights, const int64_t *slice_index,
                             int penalty, double lambda, double delta, int n_iter,
                             int threads, int64_t row_begin, int64_t row_end,
                             double *x, double *yhat)
{
    if (row_end <= 0 || row_end > g->ny)
        row_end = g->ny;
    FILE *f = fopen(path, "rb");
    if (!f)
        return -1;
    int64_t updates = 0;
    for (int n = 0; n < n_iter; n++) {
        for (int64_t j = row_begin; j < row_end; j++) {
            for (int64_t i = 0; i < g->nx; i++) {
                uint64_t nnz = 0;
                if (fread(&nnz, sizeof(nnz), 1, f) != 1 || nnz > (uint64_t)INT32_MAX) {
                    fclose(f);
                    return -1;
                }
                /* a fresh array per column, read whether or not the column
                 * is inside the FOV (src/icd_iteration.cpp:164-179) */
                file_pair *p = (file_pair *)malloc(sizeof(file_pair) * (nnz ? nnz : 1));
                if (!p || (nnz && fread(p, sizeof(file_pair), nnz, f) != nnz)) {
                    free(p);
                    fclose(f);
                    return -1;
                }
                if (inside_fov(g, i, j)) {
                    update_column(g, i + g->nx * j, (const int32_t *)(const void *)p,
                                  (const float *)(const void *)p + 1, 2, (int)nnz,
                                  raw, weights, slice_index, penalty, lambda, delta,
                                  threads, x, yhat);
                    updates += g->nz;
                }
                free(p);
            }
        }
        /* rewind for the next iteration (src/icd_iteration.cpp:270-273) */
        if (fseek(f, 0, SEEK_SET) != 0) {
            fclose(f);
            return -1;
        }
    }
    fclose(f);
    return updates;
}
