/*
 * oracle.c -- CPU FP64 restatement of the CTBB_ICD reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled with -ffp-contract=off
 * so that every FP64 expression rounds exactly as written, like the
 * reference binary (x86-64 baseline, -O0, no FMA: src/makefile:12).
 *
 * What pins each function (tests/test_oracle_*.py):
 *   orc_ssm_count/fill   closed forms (partition of unity per ray step,
 *                        A*1 along a ray), brute-force single-ray re-evaluation,
 *                        analytic line integral of a smooth object,
 *                        symmetric central-view cases.  uBLAS expression
 *                        semantics (element-wise order, norm_2 as the plain
 *                        sqrt of the sum of squares) are restated, not run:
 *                        parity with uBLAS itself is unpinned (Boost absent).
 *                        Flying focal spots (in-plane, z, both): footprint
 *                        centroids and water-cylinder chords on the FFS scans
 *                        (a flipped deflection or swapped view parity fails);
 *                        the deflection constants themselves
 *                        (src/generate_system_matrix_ffs.cpp:66-67,242-283)
 *                        are restated, parity unpinned.
 *   orc_ssm_*_views      the ray-driven build over a view range: view range
 *                        by view range, exactly the full build's entries of
 *                        those views (it times the single-threaded CPU build
 *                        on a sample of views).
 *   orc_forward_project  dense brute-force A x (numpy) on tiny inputs.
 *   orc_quadratic        reference src/penalties.cpp compiled from source
 *   orc_edge_preserving  (oracle/_ref), bit-exact; closed forms.
 *   orc_icd              dense Gauss-Seidel (numpy) with lambda=0; NNLS limit;
 *                        single-voxel du = beta/alpha; cost non-increasing.
 *   orc_icd_ex           batch surrogate: no batched sweep raises the cost
 *                        where plain Jacobi curvature does; single-column
 *                        batches give the reference's curvature bit for bit.
 *   orc_column_overlaps  brute force: np.intersect1d of a column's sample set
 *                        with the same set dv views later, every column and
 *                        shift; it checks the device's group spacing.
 *   orc_ssm_cols         equals the ray-driven build on its columns; in-plane
 *                        flying focal spot: even/odd-view footprint centroids
 *                        of central voxels 0.50 +- 0.02 channel apart (the
 *                        interlace the deflection da exists for).
 *   orc_icd_trace        every step of a default-schedule sweep on the
 *                        clinical cone lowers the cost, with shared samples
 *                        between members counted (> 0).
 *   orc_*26              three-shear rotation a bijection near the true
 *                        rotation; steps are exact coordinate minimisers of
 *                        the 26-neighbour cost; constant image -> zero step;
 *                        sweeps never raise the cost.
 *   orc_cost             gradient of the cost equals the update optimality
 *                        residual; finite differences.
 *   orc_rotate_slices    reference src/rotate_slices.cpp + src/interp.h
 *                        compiled from source (oracle/_ref), bit-exact.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_SQRT2 1.41421356237309504880168872420969807
#define ORC_PI_ROT 3.141592653589793238   /* src/rotate_slices.cpp:29 */
#define ORC_PI_FFS 3.14159265359          /* src/generate_system_matrix_ffs.cpp:41 */

/* ------------------------------------------------------------------------ */
/* Ray geometry of one SSM view (src/generate_system_matrix.cpp:85-100,      */
/* src/generate_system_matrix_ffs.cpp:112-134, 242-261).                     */
typedef struct {
    double ew[3], eu[3], ez[3];
    double s[3];      /* source position used by the ray loops (FFS-shifted) */
    double off[3];    /* focal-spot offset (zero without FFS)                */
} view_geom;

/* Focal-spot deflection of SSM view i: generate_ffs_offset,
 * src/generate_system_matrix_ffs.cpp:242-283 (radial = e_w, anti-radial =
 * e_u; the z term takes tan() of the prm anode angle as given, in degrees,
 * while dr at :66 converts it to radians -- reproduced as written). */
static void orc_ffs_offset(const ctbb_geom *g, int i, double da, double dr,
                           const view_geom *v, double off[3])
{
    double t = tan(g->anode_angle);
    double cr, ca, cz;             /* multiples of dr e_w, da e_u, dr tan e_z */
    if (g->phiffs && !g->zffs) {
        cr = 0.0; ca = (i % 2 == 0) ? 1.0 : -1.0; cz = 0.0;
    } else if (!g->phiffs && g->zffs) {
        cr = (i % 2 == 0) ? -1.0 : 1.0; ca = 0.0; cz = cr;
    } else {
        int rho = i % 4;
        cr = rho < 2 ? -1.0 : 1.0;
        ca = (rho % 2 == 0) ? -1.0 : 1.0;
        cz = cr;
    }
    for (int e = 0; e < 3; e++) {
        double o = 0.0;
        if (g->zffs)
            o = (cr * dr) * v->ew[e];
        if (g->phiffs)
            o = g->zffs ? o + (ca * da) * v->eu[e] : (ca * da) * v->eu[e];
        if (g->zffs)
            o = o + ((cz * dr) * t) * v->ez[e];
        off[e] = o;
    }
}

static void orc_view(const ctbb_geom *g, int i, int ffs, double da,
                     view_geom *v)
{
    double n_sm = (double)g->n_views_sm;
    double tube_angle = (i - 0.5 * n_sm) * g->tube_angle_increment;
    double cw = cos(tube_angle), sw = sin(tube_angle);
    v->ew[0] = cw;   v->ew[1] = sw;  v->ew[2] = 0.0;
    v->eu[0] = -sw;  v->eu[1] = cw;  v->eu[2] = 0.0;
    v->ez[0] = 0.0;  v->ez[1] = 0.0; v->ez[2] = 1.0;
    double zs = (double)g->table_direction * ((double)i - 0.5 * n_sm)
                * g->tube_z_increment;
    double base[3];
    for (int e = 0; e < 3; e++)
        base[e] = g->focal_spot_radius * v->ew[e] + zs * v->ez[e];
    double dr = g->source_detector_distance * g->coll_slicewidth
                / (4.0 * (g->source_detector_distance - g->focal_spot_radius)
                   * tan(g->anode_angle * ORC_PI_FFS / 180.0));
    if (ffs)
        orc_ffs_offset(g, i, da, dr, v, v->off);
    else
        v->off[0] = v->off[1] = v->off[2] = 0.0;
    for (int e = 0; e < 3; e++)
        v->s[e] = ffs ? base[e] + v->off[e] : base[e];
}

/* Direction of ray (channel j, row k) for view geometry v. */
static void orc_dir(const ctbb_geom *g, const view_geom *v, int ffs, int j,
                    int k, double d[3])
{
    double D = g->source_detector_distance;
    double gamma;
    if (ffs)
        gamma = (j - g->center_channel_non_ffs) * g->fan_angle_increment;
    else
        gamma = ((j - g->center_channel_non_ffs)
                 * g->transaxial_detector_spacing) / D;
    double cg = cos(gamma), sg = sin(gamma);
    double ap = (k - g->center_row) * g->axial_detector_spacing;
    double A = -D * cg, B = D * sg;
    double t[3];
    for (int e = 0; e < 3; e++) {
        t[e] = (A * v->ew[e] - B * v->eu[e]) - ap * v->ez[e];
        if (ffs)
            t[e] = t[e] - v->off[e];
    }
    double nrm;
    if (ffs) {
        double acc = 0.0;
        for (int e = 0; e < 3; e++) {
            double u = fabs(t[e]);
            acc += u * u;
        }
        nrm = sqrt(acc);
    } else {
        nrm = sqrt(ap * ap + D * D);
    }
    for (int e = 0; e < 3; e++)
        d[e] = t[e] / nrm;
}

typedef void (*emit_fn)(void *ctx, int64_t col, int32_t q, float val);

/* One crossing of ray d (origin v->s) with the voxel line of index m along
 * its major axis (ix = m for x-major rays, iy = m for y-major rays):
 * src/generate_system_matrix.cpp:116-184 for one loop step.  Returns 0 when
 * the reference pushes nothing; else 1 with the crossed minor index *lo (the
 * two entries go to lo and lo + 1) and their FP32 values.  Shared by the
 * ray-driven build and the column-subset build, so both evaluate the same
 * expressions in the same order. */
static int orc_crossing(const ctbb_geom *g, const view_geom *v,
                        const double d[3], int xmajor, int m, int *lo,
                        float *v_lo, float *v_hi)
{
    int nx = (int)g->nx, ny = (int)g->ny;
    double aax = fabs(d[0]), aay = fabs(d[1]);
    if (xmajor) {
        double x = (m - g->center_voxel_x) * g->voxel_size_x;
        double y_hat = v->s[1] + d[1] / d[0] * (x - v->s[0]);
        double j_hat = y_hat / g->voxel_size_y + g->center_voxel_y;
        int jy = (int)floor(j_hat);
        double z_hat = v->s[2] + d[2] / d[0] * (x - v->s[0]);
        double k_hat = z_hat / g->voxel_size_z + g->center_voxel_z;
        int kz = (int)floor(k_hat);
        if (!(jy >= 0 && jy <= ny - 2))
            return 0;
        double scale = g->voxel_size_x / aax;
        double tri;
        if (kz == -1)
            tri = scale * (k_hat - kz);
        else if (kz == 0)
            tri = scale * (1 - (k_hat - kz));
        else
            return 0;
        *lo = jy;
        *v_lo = (float)(tri * (1 - (j_hat - jy)));
        *v_hi = (float)(tri * (j_hat - jy));
        return 1;
    } else {
        double y = (m - g->center_voxel_y) * g->voxel_size_y;
        double x_hat = v->s[0] + d[0] / d[1] * (y - v->s[1]);
        double i_hat = x_hat / g->voxel_size_x + g->center_voxel_x;
        int ix = (int)floor(i_hat);
        double z_hat = v->s[2] + d[2] / d[1] * (y - v->s[1]);
        double k_hat = z_hat / g->voxel_size_z + g->center_voxel_z;
        int kz = (int)floor(k_hat);
        if (!(ix >= 0 && ix <= nx - 2))
            return 0;
        double scale = g->voxel_size_y / aay;
        double tri;
        if (kz == -1)
            tri = scale * (k_hat - kz);
        else if (kz == 0)
            tri = scale * (1 - (k_hat - kz));
        else
            return 0;
        *lo = ix;
        *v_lo = (float)(tri * (1 - (i_hat - ix)));
        *v_hi = (float)(tri * (i_hat - ix));
        return 1;
    }
}

static double orc_da(const ctbb_geom *g)
{
    /* in-plane deflection, src/generate_system_matrix_ffs.cpp:67 */
    double D = g->source_detector_distance, R = g->focal_spot_radius;
    return D * R * sin(g->fan_angle_increment) / (4.0 * (D - R));
}

/* The ray loop of src/generate_system_matrix.cpp:80-186 (and its FFS twin
 * src/generate_system_matrix_ffs.cpp:104-222): emits (column, q, value) in
 * ascending q. */
static int orc_ssm_rays(const ctbb_geom *g, int v0, int v1, emit_fn emit,
                        void *ctx)
{
    int ffs = (g->zffs != 0 || g->phiffs != 0);
    double da = orc_da(g);
    int nx = (int)g->nx, ny = (int)g->ny;
    for (int i = v0; i < v1; i++) {
        view_geom v;
        orc_view(g, i, ffs, da, &v);
        for (int j = 0; j < (int)g->n_channels; j++) {
            for (int k = 0; k < (int)g->n_rows; k++) {
                double d[3];
                orc_dir(g, &v, ffs, j, k, d);
                int32_t q = (int32_t)(k + g->n_rows * j
                                      + g->n_rows * g->n_channels * i);
                int xmajor = fabs(d[0]) > fabs(d[1]);
                int nmaj = xmajor ? nx : ny;
                for (int m = 0; m < nmaj; m++) {
                    int lo;
                    float a0, a1;
                    if (!orc_crossing(g, &v, d, xmajor, m, &lo, &a0, &a1))
                        continue;
                    if (xmajor) {
                        emit(ctx, m + (int64_t)nx * lo, q, a0);
                        emit(ctx, m + (int64_t)nx * (lo + 1), q, a1);
                    } else {
                        emit(ctx, lo + (int64_t)nx * m, q, a0);
                        emit(ctx, (lo + 1) + (int64_t)nx * m, q, a1);
                    }
                }
            }
        }
    }
    return 0;
}

/* Entries of one column (ix0, iy0), voxel-driven: every ray of every view
 * is tested, in ascending q, with the crossing evaluation of the ray loop
 * above, so the column comes out exactly as orc_ssm_fill writes it.  A ray
 * crosses the column's line once (x-major: at ix0; y-major: at iy0) and
 * pushes into it iff the crossed minor index is iy0 or iy0 - 1 (x-major)
 * resp. ix0 or ix0 - 1.  Channels are screened first with the ray of the
 * central row, keeping every channel whose crossing lies within two voxels
 * of the column (the in-plane crossing of a ray does not depend on its row
 * except through rounding of the normalised direction, far below a voxel);
 * all rows of the kept channels are then evaluated exactly.
 * emit == NULL: count only.  Returns the entry count. */
static int64_t orc_ssm_column(const ctbb_geom *g, int64_t col, int32_t *idx,
                              float *val)
{
    int ffs = (g->zffs != 0 || g->phiffs != 0);
    double da = orc_da(g);
    int ix0 = (int)(col % g->nx), iy0 = (int)(col / g->nx);
    int64_t n = 0;
    int kc = (int)(g->n_rows / 2);
    for (int i = 0; i < (int)g->n_views_sm; i++) {
        view_geom v;
        orc_view(g, i, ffs, da, &v);
        for (int j = 0; j < (int)g->n_channels; j++) {
            double d[3];
            orc_dir(g, &v, ffs, j, kc, d);
            int xmajor = fabs(d[0]) > fabs(d[1]);
            double cross;
            if (xmajor) {
                double x = (ix0 - g->center_voxel_x) * g->voxel_size_x;
                cross = (v.s[1] + d[1] / d[0] * (x - v.s[0])) / g->voxel_size_y
                        + g->center_voxel_y - iy0;
            } else {
                double y = (iy0 - g->center_voxel_y) * g->voxel_size_y;
                cross = (v.s[0] + d[0] / d[1] * (y - v.s[1])) / g->voxel_size_x
                        + g->center_voxel_x - ix0;
            }
            if (!(cross > -3.0 && cross < 2.0))
                continue;
            for (int k = 0; k < (int)g->n_rows; k++) {
                orc_dir(g, &v, ffs, j, k, d);
                xmajor = fabs(d[0]) > fabs(d[1]);
                int lo;
                float a0, a1;
                if (!orc_crossing(g, &v, d, xmajor, xmajor ? ix0 : iy0, &lo, &a0, &a1))
                    continue;
                int mine = xmajor ? iy0 : ix0;
                if (lo != mine && lo + 1 != mine)
                    continue;
                if (idx) {
                    idx[n] = (int32_t)(k + g->n_rows * j + g->n_rows * g->n_channels * i);
                    val[n] = (lo == mine) ? a0 : a1;
                }
                n++;
            }
        }
    }
    return n;
}

int64_t orc_ssm_cols_count(const ctbb_geom *g, int64_t ncols, const int64_t *cols,
                           int64_t *cnt)
{
    int64_t tot = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+:tot)
    for (int64_t t = 0; t < ncols; t++) {
        cnt[t] = orc_ssm_column(g, cols[t], NULL, NULL);
        tot += cnt[t];
    }
    return tot;
}

void orc_ssm_cols_fill(const ctbb_geom *g, int64_t ncols, const int64_t *cols,
                       const int64_t *ptr, int32_t *idx, float *val)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t t = 0; t < ncols; t++)
        (void)orc_ssm_column(g, cols[t], idx + ptr[t], val + ptr[t]);
}

static void emit_count(void *ctx, int64_t col, int32_t q, float val)
{
    (void)q; (void)val;
    ((int64_t *)ctx)[col] += 1;
}

typedef struct {
    int64_t *cursor;
    int32_t *idx;
    float *val;
} fill_ctx;

static void emit_fill(void *ctx, int64_t col, int32_t q, float val)
{
    fill_ctx *f = (fill_ctx *)ctx;
    int64_t p = f->cursor[col]++;
    f->idx[p] = q;
    f->val[p] = val;
}

int64_t orc_ssm_count(const ctbb_geom *g, int64_t *col_cnt)
{
    int64_t ncol = g->nx * g->ny;
    memset(col_cnt, 0, sizeof(int64_t) * ncol);
    if (orc_ssm_rays(g, 0, (int)g->n_views_sm, emit_count, col_cnt) != 0)
        return -1;
    int64_t tot = 0;
    for (int64_t c = 0; c < ncol; c++)
        tot += col_cnt[c];
    return tot;
}

int orc_ssm_fill(const ctbb_geom *g, const int64_t *col_ptr, int32_t *idx,
                 float *val)
{
    int64_t ncol = g->nx * g->ny;
    fill_ctx f;
    f.cursor = (int64_t *)malloc(sizeof(int64_t) * ncol);
    if (!f.cursor)
        return -1;
    memcpy(f.cursor, col_ptr, sizeof(int64_t) * ncol);
    f.idx = idx;
    f.val = val;
    int rc = orc_ssm_rays(g, 0, (int)g->n_views_sm, emit_fill, &f);
    free(f.cursor);
    return rc;
}

int64_t orc_ssm_count_views(const ctbb_geom *g, int64_t v0, int64_t v1,
                            int64_t *col_cnt)
{
    int64_t ncol = g->nx * g->ny;
    if (v0 < 0 || v1 > g->n_views_sm || v0 > v1)
        return -1;
    memset(col_cnt, 0, sizeof(int64_t) * ncol);
    orc_ssm_rays(g, (int)v0, (int)v1, emit_count, col_cnt);
    int64_t tot = 0;
    for (int64_t c = 0; c < ncol; c++)
        tot += col_cnt[c];
    return tot;
}

int orc_ssm_fill_views(const ctbb_geom *g, int64_t v0, int64_t v1,
                       const int64_t *col_ptr, int32_t *idx, float *val)
{
    int64_t ncol = g->nx * g->ny;
    if (v0 < 0 || v1 > g->n_views_sm || v0 > v1)
        return -1;
    fill_ctx f;
    f.cursor = (int64_t *)malloc(sizeof(int64_t) * ncol);
    if (!f.cursor)
        return -1;
    memcpy(f.cursor, col_ptr, sizeof(int64_t) * ncol);
    f.idx = idx;
    f.val = val;
    int rc = orc_ssm_rays(g, (int)v0, (int)v1, emit_fill, &f);
    free(f.cursor);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Slice offset in samples: src/icd_iteration.cpp:88-96 (forward projection) */
/* and :189-196 (sweep); the n_ffs alignment is a no-op for slice tables     */
/* built by src/setup.cpp:394-406 (central views step by n_ffs).             */
static int64_t orc_slice_offset(const ctbb_geom *g, int64_t central_idx)
{
    int64_t nffs = ((int64_t)1 << g->zffs) * ((int64_t)1 << g->phiffs);
    if (g->zffs || g->phiffs)
        central_idx -= central_idx % nffs;
    return (central_idx - g->n_views_sm / 2) * g->n_channels * g->n_rows;
}

void orc_forward_project(const ctbb_geom *g, const int64_t *col_ptr,
                         const int32_t *idx, const float *val,
                         const int64_t *slice_index, const double *x,
                         double *yhat)
{
    int64_t data_size = g->readings * g->n_channels * g->n_rows;
    int64_t nxy = g->nx * g->ny;
    memset(yhat, 0, sizeof(double) * data_size);
    for (int64_t c = 0; c < nxy; c++) {
        for (int64_t k = 0; k < g->nz; k++) {
            int64_t off = orc_slice_offset(g, slice_index[k]);
            double xv = x[c + nxy * k];
            for (int64_t p = col_ptr[c]; p < col_ptr[c + 1]; p++) {
                int64_t r = idx[p] + off;
                if (r > -1 && r < data_size)
                    yhat[r] = yhat[r] + xv * val[p];
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Penalties: src/penalties.cpp:39-154.                                       */
static const double *orc_weights8(double *W)
{
    static double w[8];
    w[0] = 1.0 / ORC_SQRT2; w[1] = 1.0; w[2] = 1.0 / ORC_SQRT2; w[3] = 1.0;
    w[4] = 1.0; w[5] = 1.0 / ORC_SQRT2; w[6] = 1.0; w[7] = 1.0 / ORC_SQRT2;
    double s = 0.0;
    for (int i = 0; i < 8; i++)
        s += w[i];
    *W = s;
    return w;
}

static void orc_neighbours8(const ctbb_geom *g, int64_t q, int64_t nb[8])
{
    int64_t nx = g->nx;
    nb[0] = q - nx - 1; nb[1] = q - nx; nb[2] = q - nx + 1; nb[3] = q - 1;
    nb[4] = q + 1; nb[5] = q + nx - 1; nb[6] = q + nx; nb[7] = q + nx + 1;
}

double orc_quadratic(const ctbb_geom *g, int64_t q, double alpha, double beta,
                     double lambda, const double *x)
{
    double W;
    const double *w = orc_weights8(&W);
    int64_t nb[8], n = g->nx * g->ny * g->nz;
    orc_neighbours8(g, q, nb);
    double sum = 0.0;
    for (int i = 0; i < 8; i++)
        if (nb[i] >= 0 && nb[i] < n)
            sum += w[i] * (x[nb[i]] - x[q]);
    return (beta + lambda * sum) / (alpha + lambda * W);
}

static double orc_ep_func(double du, double alpha, double beta, double delta,
                          double lambda, const double *x, const double *w,
                          const int64_t *nb, int64_t q, int64_t n)
{
    double sum = 0;
    for (int i = 0; i < 8; i++) {
        if (nb[i] >= 0 && nb[i] < n) {
            double t = (du - (x[nb[i]] - x[q]));
            if (t >= 0)
                sum = sum + w[i] * (1.0 - 1.0 / (1.0 + fabs(t / delta)));
            else
                sum = sum - w[i] * (1.0 - 1.0 / (1.0 + fabs(t / delta)));
        }
    }
    return 2 * alpha * du - 2 * beta + lambda * sum;
}

static int orc_sign(double v) { return v < 0 ? -1 : 1; }

double orc_edge_preserving(const ctbb_geom *g, int64_t q, double alpha,
                           double beta, double lambda, double delta,
                           const double *x)
{
    double W;
    const double *w = orc_weights8(&W);
    int64_t nb[8], n = g->nx * g->ny * g->nz;
    orc_neighbours8(g, q, nb);
    if (alpha == 0.0)
        return 0.0;
    double a = beta / alpha, b = beta / alpha;
    for (int i = 0; i < 8; i++) {
        if (nb[i] >= 0 && nb[i] < n) {
            double diff = (x[nb[i]] - x[q]);
            if (diff < a) a = diff;
            if (diff > b) b = diff;
        }
    }
    const int max_iterations = 1000000;
    const double tol = 0.00000001;
    double c = (a + b) / 2.0;
    for (int i = 0; i < max_iterations; i++) {
        c = (a + b) / 2.0;
        double f_c = orc_ep_func(c, alpha, beta, delta, lambda, x, w, nb, q, n);
        double f_a = orc_ep_func(a, alpha, beta, delta, lambda, x, w, nb, q, n);
        if ((f_c == 0) || (b - a) / 2 < tol)
            return c;
        if (orc_sign(f_c) == orc_sign(f_a))
            a = c;
        else
            b = c;
    }
    /* The reference falls off the end here (no return); never reached in
     * FP64 because the bracket halves every step. */
    return c;
}

/* ------------------------------------------------------------------------ */
/* 26-neighbour penalty (no reference counterpart: the reference stencil is  */
/* 8 in-plane neighbours, "limited to 2D", src/icd_iteration.cpp:140).       */
/* The slices are rotated: voxel p of slice k sits at R(theta_k) p (theta_k  */
/* the tube angle of its central view), so the point of slice k+1 above it   */
/* is R(theta_k - theta_{k+1}) p.  That rotation is taken as three shears    */
/* with integer-rounded shifts (Paeth): an exact bijection of the grid with  */
/* an exact inverse.  Pairs: in-plane the 8 neighbours without wrap, and     */
/* (p in slice k, pi_k(p) + d in slice k+1) for the 9 offsets d of a 3x3     */
/* block; weights 1/|offset| in voxel units (1, 1/sqrt2, 1/sqrt3).            */
static double orc_rnd(double v) { return floor(v + 0.5); }

/* shear parameters of pi_k (slices k -> k+1): a = -tan(phi/2), b = sin(phi),
 * phi = theta_k - theta_{k+1} = -(c_{k+1} - c_k) * tube_angle_increment */
static void orc_shears(const ctbb_geom *g, const int64_t *slice_index,
                       double *sh)
{
    for (int64_t k = 0; k + 1 < g->nz; k++) {
        double phi = -(double)(slice_index[k + 1] - slice_index[k])
                     * g->tube_angle_increment;
        sh[2 * k] = -tan(phi / 2.0);
        sh[2 * k + 1] = sin(phi);
    }
}

static void orc_rot_fwd(const ctbb_geom *g, const double *sh, int64_t k,
                        int64_t i, int64_t j, int64_t *io, int64_t *jo)
{
    double a = sh[2 * k], b = sh[2 * k + 1];
    double x = (double)i - g->center_voxel_x, y = (double)j - g->center_voxel_y;
    double x1 = x + orc_rnd(a * y);
    double y2 = y + orc_rnd(b * x1);
    double x3 = x1 + orc_rnd(a * y2);
    *io = (int64_t)(x3 + g->center_voxel_x);
    *jo = (int64_t)(y2 + g->center_voxel_y);
}

static void orc_rot_inv(const ctbb_geom *g, const double *sh, int64_t k,
                        int64_t i, int64_t j, int64_t *io, int64_t *jo)
{
    double a = sh[2 * k], b = sh[2 * k + 1];
    double x3 = (double)i - g->center_voxel_x, y2 = (double)j - g->center_voxel_y;
    double x1 = x3 - orc_rnd(a * y2);
    double y1 = y2 - orc_rnd(b * x1);
    double x = x1 - orc_rnd(a * y1);
    *io = (int64_t)(x + g->center_voxel_x);
    *jo = (int64_t)(y1 + g->center_voxel_y);
}

void orc_rot26(const ctbb_geom *g, const int64_t *slice_index, int64_t k,
               int inverse, int64_t i, int64_t j, int64_t *io, int64_t *jo)
{
    double *sh = (double *)malloc(sizeof(double) * 2 * (g->nz > 1 ? g->nz - 1 : 1));
    orc_shears(g, slice_index, sh);
    if (inverse)
        orc_rot_inv(g, sh, k, i, j, io, jo);
    else
        orc_rot_fwd(g, sh, k, i, j, io, jo);
    free(sh);
}

/* The 26-neighbourhood of voxel q: up to 26 indices and weights. */
static int orc_nbrs26(const ctbb_geom *g, const double *sh, int64_t q,
                      int64_t *nb, double *w)
{
    int64_t nx = g->nx, ny = g->ny, nxy = nx * ny;
    int64_t k = q / nxy, c = q - k * nxy, j = c / nx, i = c - j * nx;
    int n = 0;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            if (!dx && !dy)
                continue;
            int64_t ii = i + dx, jj = j + dy;
            if (ii >= 0 && ii < nx && jj >= 0 && jj < ny) {
                nb[n] = ii + nx * jj + nxy * k;
                w[n++] = (dx && dy) ? 1.0 / ORC_SQRT2 : 1.0;
            }
        }
    if (k + 1 < g->nz) {
        int64_t iu, ju;
        orc_rot_fwd(g, sh, k, i, j, &iu, &ju);
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                int64_t ii = iu + dx, jj = ju + dy;
                if (ii >= 0 && ii < nx && jj >= 0 && jj < ny) {
                    nb[n] = ii + nx * jj + nxy * (k + 1);
                    w[n++] = 1.0 / sqrt((double)(dx * dx + dy * dy + 1));
                }
            }
    }
    if (k >= 1) {
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                int64_t ii, jj;
                orc_rot_inv(g, sh, k - 1, i - dx, j - dy, &ii, &jj);
                if (ii >= 0 && ii < nx && jj >= 0 && jj < ny) {
                    nb[n] = ii + nx * jj + nxy * (k - 1);
                    w[n++] = 1.0 / sqrt((double)(dx * dx + dy * dy + 1));
                }
            }
    }
    return n;
}

/* Quadratic step over the neighbours actually present (the exact
 * coordinate minimiser; the 8-neighbour reference form divides by the full
 * weight sum, src/penalties.cpp:76-84). */
static double orc_quadratic26(const ctbb_geom *g, const double *sh, int64_t q,
                              double alpha, double beta, double lambda,
                              const double *x)
{
    int64_t nb[26];
    double w[26];
    int n = orc_nbrs26(g, sh, q, nb, w);
    double sum = 0.0, W = 0.0;
    for (int t = 0; t < n; t++) {
        sum += w[t] * (x[nb[t]] - x[q]);
        W += w[t];
    }
    return (beta + lambda * sum) / (alpha + lambda * W);
}

/* Edge-preserving step: the bisection of src/penalties.cpp:88-139 over the
 * 26-neighbourhood. */
static double orc_edge_preserving26(const ctbb_geom *g, const double *sh,
                                    int64_t q, double alpha, double beta,
                                    double lambda, double delta, const double *x)
{
    if (alpha == 0.0)
        return 0.0;
    int64_t nb[26];
    double w[26], diff[26];
    int n = orc_nbrs26(g, sh, q, nb, w);
    double a = beta / alpha, b = beta / alpha;
    for (int t = 0; t < n; t++) {
        diff[t] = x[nb[t]] - x[q];
        if (diff[t] < a) a = diff[t];
        if (diff[t] > b) b = diff[t];
    }
    const double tol = 0.00000001;
    double c = (a + b) / 2.0;
    for (int it = 0; it < 1000000; it++) {
        c = (a + b) / 2.0;
        double fc = 0.0, fa = 0.0;
        for (int t = 0; t < n; t++) {
            double u = c - diff[t], v = a - diff[t];
            fc += (u >= 0 ? 1.0 : -1.0) * w[t] * (1.0 - 1.0 / (1.0 + fabs(u / delta)));
            fa += (v >= 0 ? 1.0 : -1.0) * w[t] * (1.0 - 1.0 / (1.0 + fabs(v / delta)));
        }
        fc = 2 * alpha * c - 2 * beta + lambda * fc;
        fa = 2 * alpha * a - 2 * beta + lambda * fa;
        if ((fc == 0) || (b - a) / 2 < tol)
            return c;
        if (orc_sign(fc) == orc_sign(fa))
            a = c;
        else
            b = c;
    }
    return c;
}

double orc_penalty_step(const ctbb_geom *g, const int64_t *slice_index,
                        int neighbours, int penalty, int64_t q, double alpha,
                        double beta, double lambda, double delta, const double *x)
{
    if (neighbours != 26)
        return penalty == CTBB_PENALTY_QUADRATIC
                   ? orc_quadratic(g, q, alpha, beta, lambda, x)
                   : orc_edge_preserving(g, q, alpha, beta, lambda, delta, x);
    double *sh = (double *)malloc(sizeof(double) * 2 * (g->nz > 1 ? g->nz - 1 : 1));
    orc_shears(g, slice_index, sh);
    double du = penalty == CTBB_PENALTY_QUADRATIC
                    ? orc_quadratic26(g, sh, q, alpha, beta, lambda, x)
                    : orc_edge_preserving26(g, sh, q, alpha, beta, lambda, delta, x);
    free(sh);
    return du;
}

/* Penalty sum over the 26-neighbour pairs, each pair once: in-plane
 * (i+1, j), (i-1, j+1), (i, j+1), (i+1, j+1) inside the slice; across
 * slices (p, pi_k(p) + d). */
static double orc_penalty26(const ctbb_geom *g, const double *sh, int penalty,
                            double delta, const double *x)
{
    int64_t nx = g->nx, ny = g->ny, nxy = nx * ny;
    double pen = 0.0;
    for (int64_t k = 0; k < g->nz; k++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t i = 0; i < nx; i++) {
                int64_t q = i + nx * j + nxy * k;
                int64_t pr[13];
                double pw[13];
                int n = 0;
                const int ox[4] = {1, -1, 0, 1}, oy[4] = {0, 1, 1, 1};
                for (int t = 0; t < 4; t++) {
                    int64_t ii = i + ox[t], jj = j + oy[t];
                    if (ii >= 0 && ii < nx && jj >= 0 && jj < ny) {
                        pr[n] = ii + nx * jj + nxy * k;
                        pw[n++] = (ox[t] && oy[t]) ? 1.0 / ORC_SQRT2 : 1.0;
                    }
                }
                if (k + 1 < g->nz) {
                    int64_t iu, ju;
                    orc_rot_fwd(g, sh, k, i, j, &iu, &ju);
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            int64_t ii = iu + dx, jj = ju + dy;
                            if (ii >= 0 && ii < nx && jj >= 0 && jj < ny) {
                                pr[n] = ii + nx * jj + nxy * (k + 1);
                                pw[n++] = 1.0 / sqrt((double)(dx * dx + dy * dy + 1));
                            }
                        }
                }
                for (int t = 0; t < n; t++) {
                    double diff = x[q] - x[pr[t]];
                    if (penalty == CTBB_PENALTY_QUADRATIC) {
                        pen += pw[t] * diff * diff;
                    } else {
                        double a = fabs(diff);
                        pen += pw[t] * (a - delta * log1p(a / delta));
                    }
                }
            }
    return pen;
}

/* ------------------------------------------------------------------------ */
/* ICD: src/icd_iteration.cpp:149-274 (plus optional weights and the        */
/* batched z-group order).                                                   */
/* theta sums and penalty step of voxel (c, k) against the current x, yhat.
 * alpha squares the FP32 value in FP32 and accumulates in FP64, as
 * `alpha += nonzeros[m].value * nonzeros[m].value` does
 * (src/icd_iteration.cpp:208, float * float); beta multiplies in FP64
 * (:209).  With S (batch surrogate, no reference counterpart) the curvature
 * is sum_r w_r a_r S[r] over the voxel's data samples r instead. */
static double orc_voxel_step(const ctbb_geom *g, const int64_t *col_ptr,
                             const int32_t *idx, const float *val,
                             const float *raw, const float *weights, int64_t c,
                             int64_t k, int64_t off, int penalty,
                             double lambda, double delta, const double *x,
                             const double *yhat, const double *S,
                             const double *sh26)
{
    int64_t data_size = g->readings * g->n_channels * g->n_rows;
    int64_t q = c + g->nx * g->ny * k;
    double alpha = 0.0, beta = 0.0;
    for (int64_t p = col_ptr[c]; p < col_ptr[c + 1]; p++) {
        int64_t r = idx[p] + off;
        if (r > -1 && r < data_size) {
            float af = val[p];
            double a = af;
            double w = weights ? (double)weights[r] : 1.0;
            if (S)
                alpha += w * a * S[r];
            else if (weights)
                alpha += w * (double)(af * af);
            else
                alpha += (double)(af * af);
            beta += w * a * ((double)raw[r] - yhat[r]);
        }
    }
    double du;
    if (sh26)
        du = penalty == CTBB_PENALTY_QUADRATIC
                 ? orc_quadratic26(g, sh26, q, alpha, beta, lambda, x)
                 : orc_edge_preserving26(g, sh26, q, alpha, beta, lambda, delta, x);
    else if (penalty == CTBB_PENALTY_QUADRATIC)
        du = orc_quadratic(g, q, alpha, beta, lambda, x);
    else
        du = orc_edge_preserving(g, q, alpha, beta, lambda, delta, x);
    if (du + x[q] < 0)
        du = -x[q];
    return du;
}

static void orc_apply(const ctbb_geom *g, const int64_t *col_ptr,
                      const int32_t *idx, const float *val, int64_t c,
                      int64_t k, int64_t off, double du, double *x,
                      double *yhat)
{
    int64_t data_size = g->readings * g->n_channels * g->n_rows;
    x[c + g->nx * g->ny * k] += du;
    for (int64_t p = col_ptr[c]; p < col_ptr[c + 1]; p++) {
        int64_t r = idx[p] + off;
        if (r > -1 && r < data_size)
            yhat[r] += du * val[p];
    }
}

/* Cost from the maintained yhat (no projection): the orc_cost definition.
 * par != 0 sums the data term with an OpenMP reduction (the per-step trace,
 * where the order of the additions does not matter); par == 0 sums in
 * sample order. */
static double orc_cost_state(const ctbb_geom *g, const float *raw,
                             const float *weights, int penalty, double lambda,
                             double delta, const double *x, const double *yhat,
                             int par, const double *sh26)
{
    int64_t data_size = g->readings * g->n_channels * g->n_rows;
    double data = 0.0;
#pragma omp parallel for reduction(+:data) if (par)
    for (int64_t r = 0; r < data_size; r++) {
        double e = (double)raw[r] - yhat[r];
        double w = weights ? (double)weights[r] : 1.0;
        data += w * e * e;
    }
    int64_t n = g->nx * g->ny * g->nz, nx = g->nx;
    int64_t dd[4] = {1, nx - 1, nx, nx + 1};
    double wd[4] = {1.0, 1.0 / ORC_SQRT2, 1.0, 1.0 / ORC_SQRT2};
    double pen = 0.0;
    if (sh26)
        n = 0;   /* the 26-neighbour pairs instead of the flat stencil */
    for (int64_t q = 0; q < n; q++) {
        for (int t = 0; t < 4; t++) {
            if (q + dd[t] >= n)
                continue;
            double diff = x[q] - x[q + dd[t]];
            if (penalty == CTBB_PENALTY_QUADRATIC) {
                pen += wd[t] * diff * diff;
            } else {
                double a = fabs(diff);
                pen += wd[t] * (a - delta * log1p(a / delta));
            }
        }
    }
    if (sh26)
        pen = orc_penalty26(g, sh26, penalty, delta, x);
    if (penalty == CTBB_PENALTY_QUADRATIC)
        return 0.5 * data + 0.5 * lambda * pen;
    return data + lambda * pen;
}

int64_t orc_icd(const ctbb_geom *g, const int64_t *col_ptr,
                const int32_t *idx, const float *val, const float *raw,
                const float *weights, const int64_t *slice_index,
                int penalty, double lambda, double delta, int64_t spacing,
                int64_t tile, int n_iter, double *x, double *yhat)
{
    return orc_icd_trace(g, col_ptr, idx, val, raw, weights, slice_index, penalty,
                         lambda, delta, spacing, tile, 0, n_iter, x, yhat, NULL, 0,
                         NULL, NULL);
}

int64_t orc_icd_ex(const ctbb_geom *g, const int64_t *col_ptr,
                   const int32_t *idx, const float *val, const float *raw,
                   const float *weights, const int64_t *slice_index,
                   int penalty, double lambda, double delta, int64_t spacing,
                   int64_t tile, int surrogate, int n_iter, double *x,
                   double *yhat)
{
    return orc_icd_trace(g, col_ptr, idx, val, raw, weights, slice_index, penalty,
                         lambda, delta, spacing, tile, surrogate, n_iter, x, yhat,
                         NULL, 0, NULL, NULL);
}

int64_t orc_icd_trace(const ctbb_geom *g, const int64_t *col_ptr,
                      const int32_t *idx, const float *val, const float *raw,
                      const float *weights, const int64_t *slice_index,
                      int penalty, double lambda, double delta, int64_t spacing,
                      int64_t tile, int surrogate, int n_iter, double *x,
                      double *yhat, double *step_cost, int64_t max_costs,
                      int64_t *n_steps, int64_t *shared)
{
    return orc_icd_nb(g, col_ptr, idx, val, raw, weights, slice_index, penalty,
                      lambda, delta, 8, spacing, tile, surrogate, n_iter, x, yhat,
                      step_cost, max_costs, n_steps, shared);
}

int64_t orc_icd_nb(const ctbb_geom *g, const int64_t *col_ptr,
                   const int32_t *idx, const float *val, const float *raw,
                   const float *weights, const int64_t *slice_index,
                   int penalty, double lambda, double delta, int neighbours,
                   int64_t spacing, int64_t tile, int surrogate, int n_iter,
                   double *x, double *yhat, double *step_cost, int64_t max_costs,
                   int64_t *n_steps, int64_t *shared)
{
    int64_t updates = 0;
    double *sh26 = NULL;
    if (neighbours == 26) {
        sh26 = (double *)malloc(sizeof(double) * 2 * (g->nz > 1 ? g->nz - 1 : 1));
        orc_shears(g, slice_index, sh26);
    }
    int64_t data_size = g->readings * g->n_channels * g->n_rows;
    double fov_limit = (g->acquisition_fov / 2.0) * (g->acquisition_fov / 2.0);
    int64_t *offs = (int64_t *)malloc(sizeof(int64_t) * g->nz);
    for (int64_t k = 0; k < g->nz; k++)
        offs[k] = (slice_index[k] - g->n_views_sm / 2)
                  * g->n_channels * g->n_rows;
    if (n_steps)
        *n_steps = 0;
    if (shared)
        shared[0] = shared[1] = shared[2] = 0;
    if (spacing <= 1) {
        /* the reference's order: y -> x -> z, each update applied at once */
        for (int n = 0; n < n_iter; n++)
            for (int64_t j = 0; j < g->ny; j++) {
                double y = (j - g->center_voxel_y) * g->voxel_size_y;
                for (int64_t i = 0; i < g->nx; i++) {
                    double xx = (i - g->center_voxel_x) * g->voxel_size_x;
                    if (!((xx * xx + y * y) < fov_limit))
                        continue;
                    int64_t c = i + g->nx * j;
                    for (int64_t k = 0; k < g->nz; k++) {
                        double du = orc_voxel_step(g, col_ptr, idx, val, raw,
                                                   weights, c, k, offs[k],
                                                   penalty, lambda, delta, x,
                                                   yhat, NULL, sh26);
                        orc_apply(g, col_ptr, idx, val, c, k, offs[k], du, x,
                                  yhat);
                        updates++;
                    }
                }
            }
        free(offs);
        free(sh26);
        return updates;
    }
    /* batched z-group order: batches of columns at one local position of
     * every tile x tile block, raster over local positions; per batch, slice
     * groups j = 0..s-1; all voxels of a step see the state at its start. */
    int64_t ntx = (g->nx + tile - 1) / tile, nty = (g->ny + tile - 1) / tile;
    int64_t *bcols = (int64_t *)malloc(sizeof(int64_t) * ntx * nty);
    double *dus = (double *)malloc(sizeof(double) * ntx * nty * g->nz);
    /* batch surrogate: S[r] = sum over every voxel v of the step of its
     * matrix value on data sample r.  By Cauchy-Schwarz,
     * (sum_v a_vr d_v)^2 <= S_r sum_v a_vr d_v^2, so the separable quadratic
     * with curvatures sum_r w_r a_vr S_r majorises the weighted data term of
     * the whole step -- whatever the columns and slices of the step share --
     * and the step's voxels are never penalty neighbours (columns >= tile
     * apart, slices >= s apart), so every step lowers the cost.  Single-column
     * batches keep the plain curvature unless surrogate == 2 (the step's
     * samples are disjoint there for a conflict-free spacing, S_r = a_vr). */
    int single = tile >= g->nx && tile >= g->ny;
    int use_s = surrogate == 2 || (surrogate && !single);
    double *S = use_s ? (double *)calloc((size_t)data_size, sizeof(double)) : NULL;
    int32_t *mk_k = NULL, *mk_c = NULL;
    if (shared) {
        mk_k = (int32_t *)malloc(sizeof(int32_t) * data_size);
        mk_c = (int32_t *)malloc(sizeof(int32_t) * data_size);
        for (int64_t r = 0; r < data_size; r++)
            mk_k[r] = -1;
    }
    int64_t nst = 0;
    if (step_cost && max_costs > 0)
        step_cost[0] = orc_cost_state(g, raw, weights, penalty, lambda, delta, x, yhat, 1,
                                      sh26);
    for (int n = 0; n < n_iter; n++) {
        for (int64_t ly = 0; ly < tile; ly++)
            for (int64_t lx = 0; lx < tile; lx++) {
                int64_t nb = 0;
                for (int64_t ty = 0; ty < nty; ty++)
                    for (int64_t tx = 0; tx < ntx; tx++) {
                        int64_t i = tx * tile + lx, j = ty * tile + ly;
                        if (i >= g->nx || j >= g->ny)
                            continue;
                        double xx = (i - g->center_voxel_x) * g->voxel_size_x;
                        double y = (j - g->center_voxel_y) * g->voxel_size_y;
                        if ((xx * xx + y * y) < fov_limit)
                            bcols[nb++] = i + g->nx * j;
                    }
                if (nb == 0)
                    continue;
                for (int64_t jj = 0; jj < spacing; jj++) {
                    if (S || shared)
                        for (int64_t b = 0; b < nb; b++)
                            for (int64_t k = jj; k < g->nz; k += spacing)
                                for (int64_t p = col_ptr[bcols[b]]; p < col_ptr[bcols[b] + 1]; p++) {
                                    int64_t r = idx[p] + offs[k];
                                    if (!(r > -1 && r < data_size))
                                        continue;
                                    if (S)
                                        S[r] += val[p];
                                    if (shared) {
                                        if (mk_k[r] < 0) {
                                            mk_k[r] = (int32_t)k;
                                            mk_c[r] = (int32_t)bcols[b];
                                        } else if (mk_k[r] != k && mk_c[r] != bcols[b]) {
                                            shared[0]++;
                                        } else if (mk_k[r] != k) {
                                            shared[1]++;
                                        } else if (mk_c[r] != bcols[b]) {
                                            shared[2]++;
                                        }
                                    }
                                }
                    int64_t t = 0;
                    for (int64_t b = 0; b < nb; b++)
                        for (int64_t k = jj; k < g->nz; k += spacing)
                            dus[t++] = orc_voxel_step(g, col_ptr, idx, val, raw,
                                                      weights, bcols[b], k,
                                                      offs[k], penalty, lambda,
                                                      delta, x, yhat, S, sh26);
                    t = 0;
                    for (int64_t b = 0; b < nb; b++)
                        for (int64_t k = jj; k < g->nz; k += spacing) {
                            orc_apply(g, col_ptr, idx, val, bcols[b], k, offs[k],
                                      dus[t++], x, yhat);
                            updates++;
                        }
                    if (S || shared)
                        for (int64_t b = 0; b < nb; b++)
                            for (int64_t k = jj; k < g->nz; k += spacing)
                                for (int64_t p = col_ptr[bcols[b]]; p < col_ptr[bcols[b] + 1]; p++) {
                                    int64_t r = idx[p] + offs[k];
                                    if (!(r > -1 && r < data_size))
                                        continue;
                                    if (S)
                                        S[r] = 0.0;
                                    if (shared)
                                        mk_k[r] = -1;
                                }
                    nst++;
                    if (step_cost && nst < max_costs)
                        step_cost[nst] = orc_cost_state(g, raw, weights, penalty,
                                                        lambda, delta, x, yhat, 1, sh26);
                }
            }
    }
    if (n_steps)
        *n_steps = nst;
    free(bcols);
    free(dus);
    free(S);
    free(mk_k);
    free(mk_c);
    free(offs);
    free(sh26);
    return updates;
}

/* ------------------------------------------------------------------------ */
/* Cost (no reference counterpart; consistent with the update rules of      */
/* src/penalties.cpp:84 (quadratic) and :141-154 (edge-preserving)).        */
double orc_cost(const ctbb_geom *g, const int64_t *col_ptr,
                const int32_t *idx, const float *val, const float *raw,
                const float *weights, const int64_t *slice_index,
                int penalty, double lambda, double delta, const double *x)
{
    int64_t data_size = g->readings * g->n_channels * g->n_rows;
    double *yhat = (double *)malloc(sizeof(double) * data_size);
    orc_forward_project(g, col_ptr, idx, val, slice_index, x, yhat);
    double c = orc_cost_state(g, raw, weights, penalty, lambda, delta, x, yhat, 0, NULL);
    free(yhat);
    return c;
}

double orc_cost_nb(const ctbb_geom *g, const int64_t *col_ptr,
                   const int32_t *idx, const float *val, const float *raw,
                   const float *weights, const int64_t *slice_index,
                   int penalty, double lambda, double delta, int neighbours,
                   const double *x)
{
    if (neighbours != 26)
        return orc_cost(g, col_ptr, idx, val, raw, weights, slice_index, penalty,
                        lambda, delta, x);
    int64_t data_size = g->readings * g->n_channels * g->n_rows;
    double *yhat = (double *)malloc(sizeof(double) * data_size);
    double *sh = (double *)malloc(sizeof(double) * 2 * (g->nz > 1 ? g->nz - 1 : 1));
    orc_shears(g, slice_index, sh);
    orc_forward_project(g, col_ptr, idx, val, slice_index, x, yhat);
    double c = orc_cost_state(g, raw, weights, penalty, lambda, delta, x, yhat, 0, sh);
    free(yhat);
    free(sh);
    return c;
}

/* ------------------------------------------------------------------------ */
/* Slice rotation: src/rotate_slices.cpp:48-154, interp2 src/interp.h:41-76. */
static float orc_interp2(const float *a, int d1, int d2, float i1, float i2)
{
    float val;
    if (i1 > (float)(d1 - 2) || (i1 < 0) || i2 > (float)(d2 - 2) || (i2 < 0)) {
        val = 0.0f;
    } else {
        float v = i1 - floorf(i1);
        float w = i2 - floorf(i2);
        int a1 = (int)floorf(i2) * d1 + (int)floorf(i1);
        int a2 = (int)floorf(i2) * d1 + (int)floorf(i1) + 1;
        int a3 = ((int)floorf(i2) + 1) * d1 + (int)floorf(i1);
        int a4 = ((int)floorf(i2) + 1) * d1 + (int)floorf(i1) + 1;
        val = a[a1] * (1.0f - v) * (1.0f - w) + a[a2] * v * (1.0f - w)
              + a[a3] * (1.0f - v) * w + a[a4] * v * w;
    }
    return val;
}

static void orc_rotate_slice(const ctbb_geom *g, const float *in, float *out,
                             double angle)
{
    int nx = (int)g->nx, ny = (int)g->ny;
    for (int xi = 0; xi < nx; xi++) {
        for (int yi = 0; yi < ny; yi++) {
            int64_t out_idx = xi + (int64_t)yi * ny;
            double x = xi - g->center_voxel_x;
            double y = yi - g->center_voxel_y;
            float xp = (float)(cos(angle) * x - sin(angle) * y);
            float yp = (float)(sin(angle) * x + cos(angle) * y);
            out[out_idx] = orc_interp2(in, nx, ny,
                                       (float)(xp + g->center_voxel_x),
                                       (float)(yp + g->center_voxel_y));
        }
    }
}

void orc_rotate_slices(const ctbb_geom *g, const double *tube_angles,
                       const int64_t *slice_index, float *volume,
                       int direction)
{
    int64_t npix = g->nx * g->ny;
    float *tmp = (float *)malloc(sizeof(float) * npix);
    for (int64_t k = 0; k < g->nz; k++) {
        double angle = -(tube_angles[slice_index[k]]) * 2.0 * ORC_PI_ROT / 360.0;
        if (direction < 0)
            angle = -angle;
        memcpy(tmp, volume + k * npix, sizeof(float) * npix);
        orc_rotate_slice(g, tmp, volume + k * npix, angle);
    }
    free(tmp);
}

/* ------------------------------------------------------------------------ */
int orc_column_overlaps(const ctbb_geom *g, const int64_t *col_ptr,
                        const int32_t *idx, int64_t c, int64_t dv)
{
    int64_t shift = dv * g->n_channels * g->n_rows;
    int64_t a = col_ptr[c], b = col_ptr[c], end = col_ptr[c + 1];
    /* S intersect (S + shift): walk idx[a] against idx[b] + shift */
    while (a < end && b < end) {
        int64_t va = idx[a], vb = (int64_t)idx[b] + shift;
        if (va == vb)
            return 1;
        if (va < vb)
            a++;
        else
            b++;
    }
    return 0;
}
