// fbp.hip -- filtered backprojection start image on gfx950.
//
// Stands in for the external wFBP program that initialize_recon_volume runs
// when `wfbp_initialize` is set (src/initialize_recon_volume.cpp:82-178),
// whose fixed-frame image the reference then rotates into the rotating
// slices.  Here every rotated slice is reconstructed directly in its own
// frame: an equiangular fan-beam FBP (Kak & Slaney, Principles of
// Computerized Tomographic Imaging, sec. 3.4.2) with an FDK cone weight,
// backprojected through the stored system matrix.
//
//   R(view, ch, row) = y * R_src * cos(gamma_ch) * D / sqrt(D^2 + ap_row^2)
//   Q(view, ch, row) = alpha * sum_k R(view, k, row) g(ch - k),
//       g(0) = 1/(8 alpha^2), g(n even) = 0, g(n odd) = -1/(2 pi^2 sin^2(n alpha))
//   x(c, k) = 2 pi * sum_v w_v P_v / sum_v w_v,
//   P_v = [sum_e a_e Q(ray_e + off_k)] / [sum_e a_e] / L_v(c)^2
//
// with e the entries of column c in view v whose sample lies inside the
// sinogram (P_v: the footprint-weighted value of Q where the voxel projects)
// and L_v(c) the in-plane distance of the column from the view's nominal
// source.  The SSM window spans more than a turn, so a column can see one
// source angle from two views N_turn apart: w_v = 1 / n_v, n_v the number of
// views of the column's view range [first, last] at v's source angle, counts
// every angle once (first / last: the views of the column's first / last
// entry, clipped to the views whose samples lie in the sinogram).
#include "common.hpp"

#include <hip/hip_fp16.h>

#include <cmath>
#include <string>
#include <vector>

using namespace ctbb;

namespace {

constexpr int kFbpBlock = 256;

// One workgroup per (view, row): the channel line, pre-weighted, in LDS;
// thread j forms Q[j] (FP32 products and sums: FP64 would cost 2.4e11 FP64
// FMAs at the 6-turn clinical scan; the start image does not need it).
__global__ __launch_bounds__(kFbpBlock) void fbp_filter_kernel(
    const float *raw, int64_t nch, int64_t nr, const double *pre_ch,
    const double *pre_row, const float *godd, double g0, double alpha, float *q)
{
    extern __shared__ float line[];
    const int64_t view = blockIdx.x / nr, row = blockIdx.x - view * nr;
    const int64_t base = row + nr * nch * view;
    for (int64_t j = threadIdx.x; j < nch; j += kFbpBlock)
        line[j] = (float)((double)raw[base + nr * j] * pre_ch[j] * pre_row[row]);
    __syncthreads();
    for (int64_t j = threadIdx.x; j < nch; j += kFbpBlock) {
        float acc = (float)g0 * line[j];
        // odd distances only: k = j - 1, j - 3, ... and j + 1, j + 3, ...
        for (int64_t k = j - 1; k >= 0; k -= 2)
            acc += line[k] * godd[j - k];
        for (int64_t k = j + 1; k < nch; k += 2)
            acc += line[k] * godd[k - j];
        q[base + nr * j] = (float)(alpha * (double)acc);
    }
}

constexpr int kFbpUnroll = 8;

// Squared in-plane distance of a column at (px, py) from the nominal source
// of SSM view v.
__device__ __forceinline__ double fbp_l2(int64_t v, int64_t nsm, double dtheta, double R,
                                         double px, double py)
{
    const double th = ((double)v - 0.5 * (double)nsm) * dtheta;
    const double dx = R * cos(th) - px, dy = R * sin(th) - py;
    return dx * dx + dy * dy;
}

// One wavefront per FOV column; lane l takes slices l, l + 64, ...  The
// column's entries are read once per 64 slices (every lane reads the same
// entry); each lane gathers Q at its slice's sample.  Sums in FP64.
// nvr = n_channels * n_rows (samples per view), nturn = views per turn.
template <bool FP16>
__global__ __launch_bounds__(kFbpBlock) void fbp_backproject_kernel(
    const int64_t *col_ptr, const int32_t *idx, const float *val, const __half *val16,
    const float *col_scale, const int32_t *view_off, const int32_t *cols, int64_t ncols,
    int64_t nx, int64_t nxy, int64_t nz, int64_t nsm, double cx, double cy, double vx,
    double vy, double R, double dtheta, const int64_t *slice_off,
    int64_t ds, int64_t nvr, int64_t nturn, const float *q, float *x)
{
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = (int64_t)gridDim.x * (kFbpBlock / 64);
    for (int64_t i = ((int64_t)blockIdx.x * kFbpBlock + threadIdx.x) / 64; i < ncols;
         i += nwaves) {
        const int64_t col = cols[i];
        const int64_t ix = col % nx, iy = col / nx;
        const double px = (ix - cx) * vx, py = (iy - cy) * vy;
        const int64_t b = col_ptr[col], len = col_ptr[col + 1] - b;
        const int32_t *vo = view_off + col * nsm;
        const float sc = FP16 ? col_scale[col] : 1.0f;
        // views of the column's first and last entry (entries run view by view)
        const int64_t fv = len > 0 ? (int64_t)idx[b] / nvr : 0;
        const int64_t lv = len > 0 ? (int64_t)idx[b + len - 1] / nvr : -1;
        for (int64_t k0 = 0; k0 < nz; k0 += 64) {
            const int64_t k = k0 + lane;
            const bool own = k < nz;
            const int64_t off = own ? slice_off[k] : 0;
            // the slice's views whose samples lie in the sinogram: off is a
            // whole number of views, so the in-range views are a range
            const int64_t offv = off / nvr;
            const int64_t first = fv > -offv ? fv : -offv;
            const int64_t last = lv < ds / nvr - 1 - offv ? lv : ds / nvr - 1 - offv;
            // the column's entries run view by view; kFbpUnroll of them are
            // loaded and gathered at once, then summed in entry order, a
            // view's footprint sum closing where its entries end
            double acc = 0.0, sa = 0.0, sq = 0.0, wsum = 0.0;
            int64_t v = 0;
            // closes view v: its weight 1 / n_v (sa > 0 puts v in [first, last])
            auto close_view = [&]() {
                const double wv =
                    1.0 / (double)((last - v) / nturn + (v - first) / nturn + 1);
                acc += wv * (sq / sa / fbp_l2(v, nsm, dtheta, R, px, py));
                wsum += wv;
            };
            int64_t e1 = nsm > 1 ? (int64_t)vo[1] : len;
            for (int64_t e0 = 0; e0 < len; e0 += kFbpUnroll) {
                double av[kFbpUnroll], qv[kFbpUnroll];
                bool ok[kFbpUnroll];
#pragma unroll
                for (int u = 0; u < kFbpUnroll; u++) {
                    const int64_t e = e0 + u;
                    const bool in = e < len;
                    const int64_t r = in ? (int64_t)idx[b + e] + off : -1;
                    ok[u] = own && in && r >= 0 && r < ds;
                    av[u] = in ? (FP16 ? (double)__half2float(val16[b + e]) * (double)sc
                                       : (double)val[b + e])
                               : 0.0;
                    qv[u] = ok[u] ? (double)q[r] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < kFbpUnroll; u++) {
                    const int64_t e = e0 + u;
                    if (e >= len)
                        break;
                    while (e >= e1) {     // close view v (and any empty ones)
                        if (sa > 0.0)
                            close_view();
                        sa = 0.0;
                        sq = 0.0;
                        v++;
                        e1 = v + 1 < nsm ? (int64_t)vo[v + 1] : len;
                    }
                    if (ok[u]) {
                        sa += av[u];
                        sq += av[u] * qv[u];
                    }
                }
            }
            if (sa > 0.0)
                close_view();
            if (own)
                x[col + nxy * k] = wsum > 0.0 ? (float)(2.0 * M_PI * acc / wsum) : 0.0f;
        }
    }
}

}  // namespace

extern "C" int ctbb_fbp(const ctbb_ssm *m, const float *raw, const int64_t *slice_index,
                        float *volume)
{
    if (!m || !raw || !slice_index || !volume)
        return fail(CTBB_E_ARG, "null argument");
    const ctbb_geom *g = &m->g;
    if (!m->view_off)
        return fail(CTBB_E_STATE, "matrix without view offsets");
    const int64_t nch = g->n_channels, nr = g->n_rows, nsm = g->n_views_sm;
    const int64_t nxy = g->nx * g->ny, nvox = nxy * g->nz, ds = data_size_of(g);
    if (nch * (int64_t)sizeof(float) > 65536)
        return fail(CTBB_E_ARG, "detector line of " + std::to_string(nch) +
                                    " channels exceeds the 64 KiB filter tile");
    const bool ffs = g->zffs || g->phiffs;
    // filter tables (host libm, FP64)
    const double alpha = ffs ? g->fan_angle_increment
                             : g->transaxial_detector_spacing / g->source_detector_distance;
    const double D = g->source_detector_distance, Rs = g->focal_spot_radius;
    std::vector<double> pre_ch(nch), pre_row(nr), godd(nch, 0.0);
    for (int64_t j = 0; j < nch; j++)
        pre_ch[j] = Rs * cos(((double)j - g->center_channel_non_ffs) * alpha);
    for (int64_t k = 0; k < nr; k++) {
        const double ap = ((double)k - g->center_row) * g->axial_detector_spacing;
        pre_row[k] = D / sqrt(D * D + ap * ap);
    }
    for (int64_t n = 1; n < nch; n += 2) {
        const double sn = sin((double)n * alpha);
        godd[n] = -1.0 / (2.0 * M_PI * M_PI * sn * sn);
    }
    const double g0 = 1.0 / (8.0 * alpha * alpha);
    const int64_t nturn = (int64_t)std::llround(2.0 * M_PI / g->tube_angle_increment);
    if (nturn < 1)
        return fail(CTBB_E_ARG, "tube angle increment above a turn");
    // FOV columns and slice offsets
    const double fov = (g->acquisition_fov / 2.0) * (g->acquisition_fov / 2.0);
    std::vector<int32_t> cols;
    for (int64_t j = 0; j < g->ny; j++)
        for (int64_t i = 0; i < g->nx; i++) {
            const double x = (i - g->center_voxel_x) * g->voxel_size_x;
            const double y = (j - g->center_voxel_y) * g->voxel_size_y;
            if (x * x + y * y < fov)
                cols.push_back((int32_t)(i + g->nx * j));
        }
    std::vector<int64_t> off(g->nz);
    for (int64_t k = 0; k < g->nz; k++)
        off[k] = slice_offset(g, slice_index[k]);
    float *d_raw = nullptr, *d_q = nullptr, *d_x = nullptr;
    double *d_tab = nullptr;
    int32_t *d_cols = nullptr;
    int64_t *d_off = nullptr;
    float *d_godd = nullptr;
    auto cleanup = [&]() {
        if (d_godd) (void)hipFree(d_godd);
        if (d_raw) (void)hipFree(d_raw);
        if (d_q) (void)hipFree(d_q);
        if (d_x) (void)hipFree(d_x);
        if (d_tab) (void)hipFree(d_tab);
        if (d_cols) (void)hipFree(d_cols);
        if (d_off) (void)hipFree(d_off);
    };
    int rc = CTBB_OK;
    if ((rc = dev_alloc_n(&d_raw, ds, "raw data")) ||
        (rc = dev_alloc_n(&d_q, ds, "filtered sinogram")) ||
        (rc = dev_alloc_n(&d_x, nvox, "volume")) ||
        (rc = dev_alloc_n(&d_tab, nch + nr, "pre-weights")) ||
        (rc = dev_alloc_n(&d_cols, std::max<size_t>(1, cols.size()), "FOV columns")) ||
        (rc = dev_alloc_n(&d_off, g->nz, "slice offsets")) ||
        (rc = dev_alloc_n(&d_godd, nch, "ramp kernel"))) {
        cleanup();
        return rc;
    }
    hipError_t e = hipMemcpy(d_raw, raw, ds * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_tab, pre_ch.data(), nch * 8, hipMemcpyHostToDevice);
    std::vector<float> goddf(godd.begin(), godd.end());
    if (e == hipSuccess) e = hipMemcpy(d_godd, goddf.data(), nch * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_tab + nch, pre_row.data(), nr * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess && !cols.empty())
        e = hipMemcpy(d_cols, cols.data(), cols.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_off, off.data(), g->nz * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(d_x, 0, nvox * 4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fbp_filter_kernel, dim3((unsigned)(g->readings * nr)),
                           dim3(kFbpBlock), nch * sizeof(float), 0, d_raw, nch, nr, d_tab,
                           d_tab + nch, d_godd, g0, alpha, d_q);
        e = hipGetLastError();
    }
    if (e == hipSuccess && !cols.empty()) {
        const int64_t ncols = (int64_t)cols.size();
        const unsigned grid = (unsigned)std::min<int64_t>((ncols + 3) / 4, 65536);
        if (m->val16)
            hipLaunchKernelGGL(fbp_backproject_kernel<true>, dim3(grid), dim3(kFbpBlock), 0, 0,
                               m->col_ptr, m->idx, m->val, m->val16, m->col_scale,
                               m->view_off, d_cols, ncols, g->nx, nxy, g->nz, nsm,
                               g->center_voxel_x, g->center_voxel_y, g->voxel_size_x,
                               g->voxel_size_y, Rs, g->tube_angle_increment, d_off,
                               ds, nch * nr, nturn, d_q, d_x);
        else
            hipLaunchKernelGGL(fbp_backproject_kernel<false>, dim3(grid), dim3(kFbpBlock), 0, 0,
                               m->col_ptr, m->idx, m->val, m->val16, m->col_scale,
                               m->view_off, d_cols, ncols, g->nx, nxy, g->nz, nsm,
                               g->center_voxel_x, g->center_voxel_y, g->voxel_size_x,
                               g->voxel_size_y, Rs, g->tube_angle_increment, d_off,
                               ds, nch * nr, nturn, d_q, d_x);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpy(volume, d_x, nvox * 4, hipMemcpyDeviceToHost);
    cleanup();
    CTBB_HIP(e);
    return CTBB_OK;
}
