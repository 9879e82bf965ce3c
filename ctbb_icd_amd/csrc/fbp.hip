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

// floor(a / d) for a, d >= 1 below 2^31 with a quotient below 2^20: the
// FP32 estimate a * (1/d) is within 2^20 * 4 * 2^-24 < 1 of a / d, so one
// correction step makes it exact.
__device__ __forceinline__ uint32_t fbp_div(uint32_t a, uint32_t d, float inv)
{
    uint32_t q = (uint32_t)((float)a * inv);
    if (q * d > a)
        q--;
    else if ((q + 1) * d <= a)
        q++;
    return q;
}

__device__ __forceinline__ double fbp_readlane(double v, int l)
{
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// One wavefront per FOV column.  Lanes take 64 consecutive entries of the
// column (entries run view by view, so a view's two or three entries sit in
// neighbouring lanes and their samples in one or two cache lines); for each
// of kFbpSlices slices at a time every lane gathers Q at its entry's sample,
// runs of one view are summed with DPP lane moves (a view cut by the chunk
// boundary is carried into the next chunk), and the lane that ends a view
// adds the view's weighted, footprint-normalised value to its slice sums;
// the slice sums are added over the lanes at the end.  The column's entries
// are read once per kFbpSlices slices.  Per view (once per chunk, not per
// slice): 1 / (L^2 * footprint sum) from the source table src (R cos th_v,
// R sin th_v per SSM view) and the redundancy weight over the column's whole
// view range, which every slice whose range is not clipped by the sinogram's
// ends uses.  Sums in FP64.  nvr = n_channels * n_rows (samples per view),
// nturn = views per turn; inv_* = their FP32 reciprocals.
constexpr int kFbpSlices = 8;

template <bool FP16>
__global__ __launch_bounds__(kFbpBlock) void fbp_backproject_kernel(
    const int64_t *col_ptr, const int32_t *idx, const float *val, const __half *val16,
    const float *col_scale, const int32_t *cols, int64_t ncols, int64_t nx, int64_t nxy,
    int64_t nz, double cx, double cy, double vx, double vy, const double *src, int64_t nsm,
    const int64_t *slice_off, int64_t ds, uint32_t nvr, float inv_nvr, uint32_t nturn,
    float inv_nturn, const float *q, float *x)
{
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = (int64_t)gridDim.x * (kFbpBlock / 64);
    const int ndv = (int)(ds / nvr);   // data views
    for (int64_t i = __builtin_amdgcn_readfirstlane(
             (int)(((int64_t)blockIdx.x * kFbpBlock + threadIdx.x) / 64));
         i < ncols; i += nwaves) {
        const int64_t col = cols[i];
        const int64_t ix = col % nx, iy = col / nx;
        const double px = (ix - cx) * vx, py = (iy - cy) * vy;
        const int64_t b = col_ptr[col], len = col_ptr[col + 1] - b;
        const double sc = FP16 ? (double)col_scale[col] : 1.0;
        // views of the column's first and last entry
        const int fv = len > 0 ? (int)fbp_div((uint32_t)idx[b], nvr, inv_nvr) : 0;
        const int lv = len > 0 ? (int)fbp_div((uint32_t)idx[b + len - 1], nvr, inv_nvr) : -1;
        for (int64_t k0 = 0; k0 < nz; k0 += kFbpSlices) {
            int64_t off[kFbpSlices];
            int first[kFbpSlices], last[kFbpSlices];
            bool slice[kFbpSlices], clip[kFbpSlices];
#pragma unroll
            for (int g = 0; g < kFbpSlices; g++) {
                slice[g] = k0 + g < nz;
                off[g] = slice[g] ? slice_off[k0 + g] : 0;
                // the slice's views whose samples lie in the sinogram: off is
                // a whole number of views, so they are a range
                const int offv = (int)(off[g] / (int64_t)nvr);
                first[g] = fv > -offv ? fv : -offv;
                last[g] = lv < ndv - 1 - offv ? lv : ndv - 1 - offv;
                if (!slice[g]) {   // no view is in an absent slice's range
                    first[g] = 1;
                    last[g] = 0;
                }
                clip[g] = first[g] != fv || last[g] != lv;
            }
            double acc[kFbpSlices], wsum[kFbpSlices];
#pragma unroll
            for (int g = 0; g < kFbpSlices; g++) {
                acc[g] = 0.0;
                wsum[g] = 0.0;
            }
            // the last view of the previous chunk when it continues here
            double carry_a = 0.0, carry_q[kFbpSlices];
#pragma unroll
            for (int g = 0; g < kFbpSlices; g++)
                carry_q[g] = 0.0;
            for (int64_t c0 = 0; c0 < len; c0 += 64) {
                const int64_t e = c0 + lane;
                const bool in = e < len;
                const int32_t r0 = in ? idx[b + e] : 0;
                const double a = in ? (FP16 ? (double)__half2float(val16[b + e]) * sc
                                            : (double)val[b + e])
                                    : 0.0;
                const int v = in ? (int)fbp_div((uint32_t)r0, nvr, inv_nvr) : -1;
                // does this chunk's last view go on in the next chunk?
                const bool cont =
                    c0 + 64 < len && fbp_div((uint32_t)idx[b + c0 + 64], nvr, inv_nvr) ==
                                         fbp_div((uint32_t)idx[b + c0 + 63], nvr, inv_nvr);
                const uint64_t head = run_heads(v, lane);
                const bool end = in && run_end(head, lane) && !(cont && lane == 63);
                // lane 0 takes the carried part of its view
                const double sa = run_sum(lane == 0 ? a + carry_a : a, head, lane);
                // per view: f = 1 / (L^2 sa), and the redundancy weight over
                // the unclipped range; both only where a view ends
                const bool use = end && sa > 0.0 && v < nsm;
                double f = 0.0, w0 = 0.0;
                if (use) {
                    const double dx = src[2 * v] - px, dy = src[2 * v + 1] - py;
                    f = 1.0 / ((dx * dx + dy * dy) * sa);
                    w0 = 1.0 / (double)(fbp_div((uint32_t)(lv - v), nturn, inv_nturn) +
                                        fbp_div((uint32_t)(v - fv), nturn, inv_nturn) + 1);
                }
                double sq[kFbpSlices];
#pragma unroll
                for (int g = 0; g < kFbpSlices; g++) {
                    // in the slice's view range <=> the sample is in the sinogram
                    const int64_t rr = (int64_t)r0 + off[g];
                    const bool ok = in && v >= first[g] && v <= last[g] && rr >= 0 && rr < ds;
                    const double t = ok ? a * (double)q[rr] : 0.0;
                    sq[g] = lane == 0 ? t + carry_q[g] : t;
                }
#pragma unroll
                for (int g = 0; g < kFbpSlices; g++) {
                    const double s = run_sum(sq[g], head, lane);
                    if (use && v >= first[g] && v <= last[g]) {
                        // weight 1 / n_v: every source angle of the range counts once
                        const double wv =
                            clip[g] ? 1.0 / (double)(fbp_div((uint32_t)(last[g] - v), nturn,
                                                             inv_nturn) +
                                                     fbp_div((uint32_t)(v - first[g]), nturn,
                                                             inv_nturn) +
                                                     1)
                                    : w0;
                        acc[g] += wv * (s * f);
                        wsum[g] += wv;
                    }
                    carry_q[g] = cont ? fbp_readlane(s, 63) : 0.0;
                }
                carry_a = cont ? fbp_readlane(sa, 63) : 0.0;
            }
#pragma unroll
            for (int g = 0; g < kFbpSlices; g++) {
                double s1 = acc[g], s2 = wsum[g];
                for (int o = 32; o > 0; o >>= 1) {
                    s1 += __shfl_xor(s1, o, 64);
                    s2 += __shfl_xor(s2, o, 64);
                }
                if (lane == 0 && slice[g])
                    x[col + nxy * (k0 + g)] = s2 > 0.0 ? (float)(2.0 * M_PI * s1 / s2) : 0.0f;
            }
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
    const int64_t nch = g->n_channels, nr = g->n_rows, nsm = g->n_views_sm;
    // the backprojection's FP32-estimated integer quotients (views) are exact
    // below 2^20
    if (nsm >= (1 << 20))
        return fail(CTBB_E_ARG, "more than 2^20 SSM views");
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
    // nominal source of each SSM view, (R cos th_v, R sin th_v)
    std::vector<double> src(2 * nsm);
    for (int64_t v = 0; v < nsm; v++) {
        const double th = ((double)v - 0.5 * (double)nsm) * g->tube_angle_increment;
        src[2 * v] = Rs * cos(th);
        src[2 * v + 1] = Rs * sin(th);
    }
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
        (rc = dev_alloc_n(&d_tab, nch + nr + 2 * nsm, "pre-weights, source table")) ||
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
    if (e == hipSuccess)
        e = hipMemcpy(d_tab + nch + nr, src.data(), 2 * nsm * 8, hipMemcpyHostToDevice);
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
        const uint32_t nvr = (uint32_t)(nch * nr), nturn32 = (uint32_t)nturn;
        const float inv_nvr = (float)(1.0 / (double)nvr), inv_nturn = (float)(1.0 / (double)nturn);
        const unsigned grid = (unsigned)std::min<int64_t>((ncols + 3) / 4, 65536);
        if (m->val16)
            hipLaunchKernelGGL(fbp_backproject_kernel<true>, dim3(grid), dim3(kFbpBlock), 0, 0,
                               m->col_ptr, m->idx, m->val, m->val16, m->col_scale, d_cols,
                               ncols, g->nx, nxy, g->nz, g->center_voxel_x,
                               g->center_voxel_y, g->voxel_size_x, g->voxel_size_y,
                               d_tab + nch + nr, nsm, d_off, ds, nvr, inv_nvr, nturn32,
                               inv_nturn, d_q, d_x);
        else
            hipLaunchKernelGGL(fbp_backproject_kernel<false>, dim3(grid), dim3(kFbpBlock), 0, 0,
                               m->col_ptr, m->idx, m->val, m->val16, m->col_scale, d_cols,
                               ncols, g->nx, nxy, g->nz, g->center_voxel_x,
                               g->center_voxel_y, g->voxel_size_x, g->voxel_size_y,
                               d_tab + nch + nr, nsm, d_off, ds, nvr, inv_nvr, nturn32,
                               inv_nturn, d_q, d_x);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpy(volume, d_x, nvox * 4, hipMemcpyDeviceToHost);
    cleanup();
    CTBB_HIP(e);
    return CTBB_OK;
}
