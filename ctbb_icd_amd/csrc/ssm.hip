// ssm.hip -- stored system matrix (SSM) build on gfx950.
//
// Replaces generate_system_matrix (src/generate_system_matrix.cpp:54-209)
// and the in-plane flying-focal-spot variant generate_system_matrix_ffs
// (src/generate_system_matrix_ffs.cpp:60-290).
//
// The reference is ray-driven: for every SSM view i, channel j and row k it
// steps along the ray's major axis and pushes (q, value) into the column of
// each voxel (ix, iy) the ray crosses inside the central slice.  Here the
// work is voxel-driven: one wavefront per (column, view).  The wavefront
// bounds the (channel, row) window of rays that can reach the column from
// the source position of that view (conservative, in FP32), and every lane
// re-evaluates the reference's FP64 ray formula for one candidate ray at
// this column's ix (x-major rays) or iy (y-major rays).  The entry test and
// the value are the reference's own expressions, evaluated in the same
// order without contraction (-ffp-contract=off), with trigonometric inputs
// from host tables computed by the same libm as the reference, so the entry
// set, its order (ascending q) and the FP32 values are identical.
//
// Passes: count (entries per (column, view)) -> per-column exclusive scan
// over views -> exclusive scan over columns -> fill.  Output is compressed
// sparse column: col_ptr[ncol+1] (int64), idx[nnz] (int32 ray index),
// val[nnz] (FP32); view_off[ncol*n_views_sm] (int32 offset of each view's
// segment inside its column) is kept for the footprint analysis.
#include "common.hpp"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace {

struct ViewRec {          // one SSM view (src/generate_system_matrix.cpp:85-100)
    double ew0, ew1, eu0, eu1;
    double s0, s1, s2;    // ray origin (FFS-deflected source)
    double off0, off1, off2;  // focal-spot offset (zero without FFS)
};

struct BuildArgs {
    const ViewRec *views;
    const double *chA;    // -D * cos(gamma_j)
    const double *chB;    //  D * sin(gamma_j)
    const double *rowAp;  // axial position of row k
    const double *rowNrm; // sqrt(ap^2 + D^2) (no FFS)
    int32_t *cnt;         // [ncol*nsm] count / view offsets
    const int64_t *col_ptr;
    int32_t *idx;
    float *val;
    int64_t nx, ny, nsm, nch, nr;
    int ffs;
    double cx, cy, cz, vx, vy, vz;
    double D, c_ch, c_row, dgamma, drow;
};

constexpr int kWave = 64;
constexpr int kBuildBlock = 256;

// Exact reference entry test for ray (j, k) of view V at column (ix0, iy0).
// Returns true and the FP32 value when the reference pushes (q, value) into
// this column.  Mirrors src/generate_system_matrix.cpp:111-184 (no FFS) and
// src/generate_system_matrix_ffs.cpp:147-220 (flying focal spots: in-plane,
// z, or both).
__device__ __forceinline__ bool ray_entry(const BuildArgs &a, const ViewRec &V,
                                          int j, int k, int ix0, int iy0,
                                          float &out)
{
    const double A = a.chA[j], B = a.chB[j];
    const double ap = a.rowAp[k];
    double t0 = (A * V.ew0 - B * V.eu0) - ap * 0.0;
    double t1 = (A * V.ew1 - B * V.eu1) - ap * 0.0;
    double t2 = (A * 0.0 - B * 0.0) - ap * 1.0;
    double nrm;
    if (a.ffs) {
        t0 = t0 - V.off0;
        t1 = t1 - V.off1;
        t2 = t2 - V.off2;
        double acc = 0.0;
        double u = fabs(t0);
        acc += u * u;
        u = fabs(t1);
        acc += u * u;
        u = fabs(t2);
        acc += u * u;
        nrm = sqrt(acc);
    } else {
        nrm = a.rowNrm[k];
    }
    const double d0 = t0 / nrm, d1 = t1 / nrm, d2 = t2 / nrm;
    const double aax = fabs(d0), aay = fabs(d1);
    if (aax > aay) {
        const double x = (ix0 - a.cx) * a.vx;
        const double y_hat = V.s1 + d1 / d0 * (x - V.s0);
        const double j_hat = y_hat / a.vy + a.cy;
        const int jy = (int)floor(j_hat);
        if (jy < 0 || jy > (int)(a.ny - 2))
            return false;
        if (jy != iy0 && jy + 1 != iy0)
            return false;
        const double z_hat = V.s2 + d2 / d0 * (x - V.s0);
        const double k_hat = z_hat / a.vz + a.cz;
        const int kz = (int)floor(k_hat);
        const double scale = a.vx / aax;
        double tri;
        if (kz == -1)
            tri = scale * (k_hat - kz);
        else if (kz == 0)
            tri = scale * (1 - (k_hat - kz));
        else
            return false;
        out = (jy == iy0) ? (float)(tri * (1 - (j_hat - jy)))
                          : (float)(tri * (j_hat - jy));
        return true;
    } else {
        const double y = (iy0 - a.cy) * a.vy;
        const double x_hat = V.s0 + d0 / d1 * (y - V.s1);
        const double i_hat = x_hat / a.vx + a.cx;
        const int ix = (int)floor(i_hat);
        if (ix < 0 || ix > (int)(a.nx - 2))
            return false;
        if (ix != ix0 && ix + 1 != ix0)
            return false;
        const double z_hat = V.s2 + d2 / d1 * (y - V.s1);
        const double k_hat = z_hat / a.vz + a.cz;
        const int kz = (int)floor(k_hat);
        const double scale = a.vy / aay;
        double tri;
        if (kz == -1)
            tri = scale * (k_hat - kz);
        else if (kz == 0)
            tri = scale * (1 - (k_hat - kz));
        else
            return false;
        out = (ix == ix0) ? (float)(tri * (1 - (i_hat - ix)))
                          : (float)(tri * (i_hat - ix));
        return true;
    }
}

// Conservative (channel, row) window of rays that can produce an entry in
// column (ix0, iy0) at view V: every such ray crosses the box
// [x(ix0-1), x(ix0+1)] x [y(iy0-1), y(iy0+1)] x [-dz, dz); the window spans
// the fan angles and detector heights of the box corners seen from the
// source, widened by two channels and two rows.
__device__ __forceinline__ void ray_window(const BuildArgs &a,
                                           const ViewRec &V, int ix0, int iy0,
                                           int &j0, int &nj, int &k0, int &nk)
{
    const float ew0 = (float)V.ew0, ew1 = (float)V.ew1;
    const float eu0 = (float)V.eu0, eu1 = (float)V.eu1;
    const float s0 = (float)V.s0, s1 = (float)V.s1, s2 = (float)V.s2;
    float gmin = 1e30f, gmax = -1e30f, rmin = 1e30f, rmax = -1e30f;
#pragma unroll
    for (int cc = 0; cc < 4; cc++) {
        const float X = (float)((ix0 + ((cc & 1) ? 1 : -1) - a.cx) * a.vx);
        const float Y = (float)((iy0 + ((cc & 2) ? 1 : -1) - a.cy) * a.vy);
        const float px = X - s0, py = Y - s1;
        const float vw = px * ew0 + py * ew1;
        const float vu = px * eu0 + py * eu1;
        const float gam = atan2f(-vu, -vw);
        const float rho = sqrtf(vw * vw + vu * vu);
        gmin = fminf(gmin, gam);
        gmax = fmaxf(gmax, gam);
        rmin = fminf(rmin, rho);
        rmax = fmaxf(rmax, rho);
    }
    const float D = (float)a.D, dz = (float)a.vz;
    // detector rows sit at heights measured from the undeflected source: a
    // z-deflected focal spot (off2) shifts every hit row by -off2
    const float off2 = (float)V.off2;
    float hmin = 1e30f, hmax = -1e30f;
#pragma unroll
    for (int cc = 0; cc < 4; cc++) {
        const float z = ((cc & 1) ? dz : -dz) - s2;
        const float rho = (cc & 2) ? rmax : rmin;
        const float ap = -z * D / rho - off2;
        hmin = fminf(hmin, ap);
        hmax = fmaxf(hmax, ap);
    }
    const float dg = (float)a.dgamma;
    int jlo = (int)floorf(gmin / dg + (float)a.c_ch) - 2;
    int jhi = (int)ceilf(gmax / dg + (float)a.c_ch) + 2;
    int klo = (int)floorf(hmin / (float)a.drow + (float)a.c_row) - 2;
    int khi = (int)ceilf(hmax / (float)a.drow + (float)a.c_row) + 2;
    jlo = max(jlo, 0);
    jhi = min(jhi, (int)a.nch - 1);
    klo = max(klo, 0);
    khi = min(khi, (int)a.nr - 1);
    j0 = jlo;
    k0 = klo;
    nj = max(0, jhi - jlo + 1);
    nk = max(0, khi - klo + 1);
}

template <bool FILL>
__global__ __launch_bounds__(kBuildBlock) void ssm_build_kernel(BuildArgs a)
{
    const int lane = threadIdx.x % kWave;
    const int64_t wave_global =
        (int64_t)blockIdx.x * (kBuildBlock / kWave) + threadIdx.x / kWave;
    const int64_t nwaves = (int64_t)gridDim.x * (kBuildBlock / kWave);
    const int64_t npairs = a.nx * a.ny * a.nsm;
    for (int64_t pair = wave_global; pair < npairs; pair += nwaves) {
        const int64_t col = pair / a.nsm;
        const int view = (int)(pair - col * a.nsm);
        const int ix0 = (int)(col % a.nx), iy0 = (int)(col / a.nx);
        const ViewRec V = a.views[view];
        int j0, nj, k0, nk;
        ray_window(a, V, ix0, iy0, j0, nj, k0, nk);
        const int ncand = nj * nk;
        int64_t base = 0;
        int running = 0;
        if (FILL)
            base = a.col_ptr[col] + a.cnt[pair];
        const int qview = (int)(a.nr * a.nch * view);
        for (int c0 = 0; c0 < ncand; c0 += kWave) {
            const int cand = c0 + lane;
            bool hit = false;
            float v = 0.0f;
            int j = 0, k = 0;
            if (cand < ncand) {
                j = j0 + cand / nk;
                k = k0 + cand % nk;
                hit = ray_entry(a, V, j, k, ix0, iy0, v);
            }
            const uint64_t mask = __ballot(hit);
            if (FILL && hit) {
                const uint64_t below = mask & ((uint64_t(1) << lane) - 1);
                const int64_t pos = base + running + __popcll(below);
                a.idx[pos] = k + (int)a.nr * j + qview;
                a.val[pos] = v;
            }
            running += __popcll(mask);
        }
        if (!FILL && lane == 0)
            a.cnt[pair] = running;
    }
}

// Exclusive scan of the per-view counts of each column, in place; writes the
// column total.
constexpr int kScanBlock = 256;

__global__ __launch_bounds__(kScanBlock) void view_scan_kernel(
    int32_t *cnt, int64_t *col_nnz, int64_t ncol, int64_t nsm)
{
    __shared__ int64_t part[kScanBlock];
    const int t = threadIdx.x;
    const int64_t per = (nsm + kScanBlock - 1) / kScanBlock;
    for (int64_t col = blockIdx.x; col < ncol; col += gridDim.x) {
        int32_t *c = cnt + col * nsm;
        const int64_t b = t * per, e = min(nsm, b + per);
        int64_t s = 0;
        for (int64_t i = b; i < e; i++)
            s += c[i];
        part[t] = s;
        __syncthreads();
        if (t == 0) {
            int64_t run = 0;
            for (int i = 0; i < kScanBlock; i++) {
                int64_t v = part[i];
                part[i] = run;
                run += v;
            }
            col_nnz[col] = run;
        }
        __syncthreads();
        int64_t run = part[t];
        for (int64_t i = b; i < e; i++) {
            int32_t v = c[i];
            c[i] = (int32_t)run;
            run += v;
        }
        __syncthreads();
    }
}

// Exclusive scan over columns (single block): col_ptr[0..ncol].
constexpr int kColScanBlock = 1024;

__global__ __launch_bounds__(kColScanBlock) void col_scan_kernel(
    const int64_t *col_nnz, int64_t *col_ptr, int64_t ncol)
{
    __shared__ int64_t part[kColScanBlock];
    const int t = threadIdx.x;
    const int64_t per = (ncol + kColScanBlock - 1) / kColScanBlock;
    const int64_t b = t * per, e = min(ncol, b + per);
    int64_t s = 0;
    for (int64_t i = b; i < e; i++)
        s += col_nnz[i];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        int64_t run = 0;
        for (int i = 0; i < kColScanBlock; i++) {
            int64_t v = part[i];
            part[i] = run;
            run += v;
        }
        col_ptr[ncol] = run;
    }
    __syncthreads();
    int64_t run = part[t];
    for (int64_t i = b; i < e; i++) {
        col_ptr[i] = run;
        run += col_nnz[i];
    }
}

// FP16 conversion: per-column scale = max|a| / 2^15, h = a / scale.
__global__ __launch_bounds__(256) void to_fp16_kernel(
    const int64_t *col_ptr, const float *val, __half *val16, float *scale,
    int64_t ncol)
{
    __shared__ float red[256];
    for (int64_t col = blockIdx.x; col < ncol; col += gridDim.x) {
        const int64_t b = col_ptr[col], e = col_ptr[col + 1];
        float m = 0.0f;
        for (int64_t p = b + threadIdx.x; p < e; p += 256)
            m = fmaxf(m, fabsf(val[p]));
        red[threadIdx.x] = m;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s)
                red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
            __syncthreads();
        }
        const float mx = red[0];
        const float sc = mx > 0.0f ? mx / 32768.0f : 1.0f;
        if (threadIdx.x == 0)
            scale[col] = sc;
        for (int64_t p = b + threadIdx.x; p < e; p += 256)
            val16[p] = __float2half_rn(val[p] / sc);
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void from_fp16_kernel(
    const int64_t *col_ptr, const __half *val16, const float *scale,
    float *out, int64_t ncol)
{
    for (int64_t col = blockIdx.x; col < ncol; col += gridDim.x) {
        const int64_t b = col_ptr[col], e = col_ptr[col + 1];
        const float sc = scale[col];
        for (int64_t p = b + threadIdx.x; p < e; p += 256)
            out[p] = __half2float(val16[p]) * sc;
    }
}

// Footprint analysis: for every candidate view shift dv (the central-view
// differences that occur between slices), the number of FOV columns whose
// footprint S meets S shifted by dv views.  Per view the
// footprint's channel and row ranges (from the column's view segment) give a
// box; where the boxes of view v and view v - dv meet, the two sorted view
// segments are merged to test for a common ray exactly.  The result is exact.
constexpr int kFpBlock = 256;
constexpr int kFpMaxViews = 6144;

__global__ __launch_bounds__(kFpBlock) void footprint_conflict_kernel(
    const int64_t *col_ptr, const int32_t *idx, const int32_t *view_off,
    int64_t nx, int64_t ny, int64_t nsm, int64_t nch, int64_t nr, double cx,
    double cy, double vx, double vy, double fov_limit, const int64_t *shifts,
    int64_t nshift, uint32_t *conflict)
{
    __shared__ uint32_t chr[kFpMaxViews];   // ch_lo | ch_hi << 16
    __shared__ uint16_t rwr[kFpMaxViews];   // row_lo | row_hi << 8
    const int64_t ncol = nx * ny;
    const int64_t per_view = nch * nr;
    constexpr int64_t kChunk = 64;
    const int64_t nchunk = (nsm + kChunk - 1) / kChunk;
    for (int64_t col = blockIdx.x; col < ncol; col += gridDim.x) {
        const int64_t ix = col % nx, iy = col / nx;
        const double x = (ix - cx) * vx, y = (iy - cy) * vy;
        if (!((x * x + y * y) < fov_limit))
            continue;
        const int64_t b = col_ptr[col], len = col_ptr[col + 1] - b;
        const int32_t *vo = view_off + col * nsm;
        for (int64_t v = threadIdx.x; v < nsm; v += kFpBlock) {
            const int64_t s = vo[v];
            const int64_t e = (v + 1 < nsm) ? vo[v + 1] : len;
            uint32_t clo = 0xffff, chi = 0, rlo = 0xff, rhi = 0;
            for (int64_t p = s; p < e; p++) {
                const int64_t rem = idx[b + p] - v * per_view;
                const uint32_t ch = (uint32_t)(rem / nr);
                const uint32_t rw = (uint32_t)(rem % nr);
                clo = min(clo, ch);
                chi = max(chi, ch);
                rlo = min(rlo, rw);
                rhi = max(rhi, rw);
            }
            chr[v] = clo | (chi << 16);
            rwr[v] = (uint16_t)(rlo | (rhi << 8));
        }
        __syncthreads();
        for (int64_t w = threadIdx.x; w < nshift * nchunk; w += kFpBlock) {
            const int64_t si = w / nchunk, ck = w - si * nchunk;
            const int64_t dv = shifts[si];
            bool hit = false;
            const int64_t v0 = max(dv, ck * kChunk), v1 = min(nsm, (ck + 1) * kChunk);
            for (int64_t v = v0; v < v1 && !hit; v++) {
                const uint32_t c1 = chr[v], c2 = chr[v - dv];
                const uint32_t lo1 = c1 & 0xffff, hi1 = c1 >> 16;
                const uint32_t lo2 = c2 & 0xffff, hi2 = c2 >> 16;
                if (lo1 > hi1 || lo2 > hi2 || lo1 > hi2 || lo2 > hi1)
                    continue;
                const uint32_t r1 = rwr[v], r2 = rwr[v - dv];
                if ((r1 & 0xff) > (r2 >> 8) || (r2 & 0xff) > (r1 >> 8))
                    continue;
                // exact: a ray q of view v with q - dv*per_view in view v-dv
                int64_t pa = vo[v], ea = (v + 1 < nsm) ? vo[v + 1] : len;
                int64_t pb = vo[v - dv], eb = vo[v - dv + 1];
                const int64_t sh = dv * per_view;
                while (pa < ea && pb < eb) {
                    const int64_t qa = idx[b + pa], qb = (int64_t)idx[b + pb] + sh;
                    if (qa == qb) {
                        hit = true;
                        break;
                    }
                    if (qa < qb) pa++; else pb++;
                }
            }
            if (hit)
                atomicAdd(&conflict[si], 1u);
        }
        __syncthreads();
    }
}

int grid_for(int64_t work, int block_waves_or_threads, int64_t cap)
{
    (void)block_waves_or_threads;
    int64_t g = std::min<int64_t>(std::max<int64_t>(work, 1), cap);
    return (int)g;
}

}  // namespace

namespace ctbb {
int check_geom(const ctbb_geom *g)
{
    if (!g)
        return fail(CTBB_E_ARG, "geometry is null");
    if (g->nx < 2 || g->ny < 2 || g->nz < 1 || g->n_channels < 1 ||
        g->n_rows < 1 || g->readings < 1 || g->n_views_sm < 2)
        return fail(CTBB_E_ARG, "geometry sizes must be positive "
                                "(nx, ny >= 2; n_views_sm >= 2)");
    if (g->n_rows > 255 || g->n_channels > 65535)
        return fail(CTBB_E_ARG, "n_rows must be <= 255 and n_channels <= 65535");
    if (g->n_views_sm * g->n_channels * g->n_rows >= (int64_t(1) << 31))
        return fail(CTBB_E_ARG, "SSM ray index exceeds int32");
    if (g->voxel_size_x <= 0 || g->voxel_size_y <= 0 || g->voxel_size_z <= 0)
        return fail(CTBB_E_ARG, "voxel sizes must be positive");
    if (g->source_detector_distance <= g->focal_spot_radius ||
        g->focal_spot_radius <= 0)
        return fail(CTBB_E_ARG, "need 0 < focal_spot_radius < "
                                "source_detector_distance");
    return CTBB_OK;
}
}  // namespace ctbb

using namespace ctbb;

extern "C" int ctbb_ssm_build(const ctbb_geom *g, ctbb_ssm **out)
{
    CTBB_TRY(check_geom(g));
    if (!out)
        return fail(CTBB_E_ARG, "out is null");
    *out = nullptr;
    const int ffs = (g->zffs != 0 || g->phiffs != 0) ? 1 : 0;

    const int64_t nsm = g->n_views_sm, nch = g->n_channels, nr = g->n_rows;
    const int64_t ncol = g->nx * g->ny;
    const double D = g->source_detector_distance, R = g->focal_spot_radius;
    const double da = D * R * sin(g->fan_angle_increment) / (4.0 * (D - R));
    // src/generate_system_matrix_ffs.cpp:66 (the constant is the file's PI)
    const double dr = D * g->coll_slicewidth /
                      (4.0 * (D - R) * tan(g->anode_angle * 3.14159265359 / 180.0));
    const double tz = tan(g->anode_angle);

    // Host tables with the reference's libm (src/generate_system_matrix.cpp
    // :85-114, src/generate_system_matrix_ffs.cpp:112-143).
    std::vector<ViewRec> hv(nsm);
    for (int64_t i = 0; i < nsm; i++) {
        const double n_sm = (double)nsm;
        const double tube_angle = ((int)i - 0.5 * n_sm) * g->tube_angle_increment;
        const double cw = cos(tube_angle), sw = sin(tube_angle);
        const double ew[3] = {cw, sw, 0.0}, eu[3] = {-sw, cw, 0.0},
                     ez[3] = {0.0, 0.0, 1.0};
        const double zs = (double)g->table_direction * ((double)i - 0.5 * n_sm) *
                          g->tube_z_increment;
        double base[3], off[3], s[3];
        for (int e = 0; e < 3; e++)
            base[e] = R * ew[e] + zs * ez[e];
        // focal-spot deflection, src/generate_system_matrix_ffs.cpp:242-283:
        // +-da along e_u (phi), +-dr along e_w with +-dr tan(anode) along
        // e_z (z; tan() of the prm angle as given, in degrees)
        double cr = 0.0, ca = 0.0;
        if (g->phiffs && !g->zffs) {
            ca = (i % 2 == 0) ? 1.0 : -1.0;
        } else if (!g->phiffs && g->zffs) {
            cr = (i % 2 == 0) ? -1.0 : 1.0;
        } else if (ffs) {
            cr = (i % 4) < 2 ? -1.0 : 1.0;
            ca = (i % 2 == 0) ? -1.0 : 1.0;
        }
        for (int e = 0; e < 3; e++) {
            double o = 0.0;
            if (g->zffs)
                o = (cr * dr) * ew[e];
            if (g->phiffs)
                o = g->zffs ? o + (ca * da) * eu[e] : (ca * da) * eu[e];
            if (g->zffs)
                o = o + ((cr * dr) * tz) * ez[e];
            off[e] = ffs ? o : 0.0;
        }
        for (int e = 0; e < 3; e++)
            s[e] = ffs ? base[e] + off[e] : base[e];
        hv[i] = ViewRec{ew[0], ew[1], eu[0], eu[1], s[0], s[1],
                        s[2],  off[0], off[1], off[2]};
    }
    std::vector<double> hA(nch), hB(nch), hAp(nr), hN(nr);
    for (int64_t j = 0; j < nch; j++) {
        double gamma;
        if (ffs)
            gamma = ((int)j - g->center_channel_non_ffs) * g->fan_angle_increment;
        else
            gamma = (((int)j - g->center_channel_non_ffs) *
                     g->transaxial_detector_spacing) / D;
        hA[j] = -D * cos(gamma);
        hB[j] = D * sin(gamma);
    }
    for (int64_t k = 0; k < nr; k++) {
        hAp[k] = ((int)k - g->center_row) * g->axial_detector_spacing;
        hN[k] = sqrt(hAp[k] * hAp[k] + D * D);
    }

    ctbb_ssm *m = new ctbb_ssm();
    m->g = *g;
    m->ncol = ncol;
    ViewRec *d_views = nullptr;
    double *d_tab = nullptr;
    int64_t *d_colnnz = nullptr;
    auto cleanup = [&]() {
        if (d_views) (void)hipFree(d_views);
        if (d_tab) (void)hipFree(d_tab);
        if (d_colnnz) (void)hipFree(d_colnnz);
    };
    auto bail = [&](int rc) {
        cleanup();
        ctbb_ssm_free(m);
        return rc;
    };
    if (dev_alloc_n(&d_views, nsm, "SSM view table") ||
        dev_alloc_n(&d_tab, 2 * nch + 2 * nr, "SSM detector tables") ||
        dev_alloc_n(&d_colnnz, ncol, "SSM column counts") ||
        dev_alloc_n(&m->view_off, ncol * nsm, "SSM view offsets") ||
        dev_alloc_n(&m->col_ptr, ncol + 1, "SSM column pointers"))
        return bail(CTBB_E_NOMEM);
#define CTBB_HIP_B(call)                                                      \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess)                                                 \
            return bail(fail(CTBB_E_HIP, std::string(#call) + ": " +          \
                                             hipGetErrorString(e_)));         \
    } while (0)
    CTBB_HIP_B(hipMemcpy(d_views, hv.data(), nsm * sizeof(ViewRec),
                         hipMemcpyHostToDevice));
    CTBB_HIP_B(hipMemcpy(d_tab, hA.data(), nch * 8, hipMemcpyHostToDevice));
    CTBB_HIP_B(hipMemcpy(d_tab + nch, hB.data(), nch * 8, hipMemcpyHostToDevice));
    CTBB_HIP_B(hipMemcpy(d_tab + 2 * nch, hAp.data(), nr * 8,
                         hipMemcpyHostToDevice));
    CTBB_HIP_B(hipMemcpy(d_tab + 2 * nch + nr, hN.data(), nr * 8,
                         hipMemcpyHostToDevice));

    BuildArgs a{};
    a.views = d_views;
    a.chA = d_tab;
    a.chB = d_tab + nch;
    a.rowAp = d_tab + 2 * nch;
    a.rowNrm = d_tab + 2 * nch + nr;
    a.cnt = m->view_off;
    a.col_ptr = m->col_ptr;
    a.nx = g->nx; a.ny = g->ny; a.nsm = nsm; a.nch = nch; a.nr = nr;
    a.ffs = ffs;
    a.cx = g->center_voxel_x; a.cy = g->center_voxel_y; a.cz = g->center_voxel_z;
    a.vx = g->voxel_size_x; a.vy = g->voxel_size_y; a.vz = g->voxel_size_z;
    a.D = D;
    a.c_ch = g->center_channel_non_ffs;
    a.c_row = g->center_row;
    a.dgamma = ffs ? g->fan_angle_increment : g->transaxial_detector_spacing / D;
    a.drow = g->axial_detector_spacing;

    hipEvent_t ev0, ev1;
    CTBB_HIP_B(hipEventCreate(&ev0));
    CTBB_HIP_B(hipEventCreate(&ev1));
    const int64_t npairs = ncol * nsm;
    const int waves_per_block = kBuildBlock / kWave;
    const int gbuild = grid_for((npairs + waves_per_block - 1) / waves_per_block,
                                kBuildBlock, 1 << 20);
    CTBB_HIP_B(hipEventRecord(ev0, 0));
    hipLaunchKernelGGL(ssm_build_kernel<false>, dim3(gbuild), dim3(kBuildBlock),
                       0, 0, a);
    CTBB_HIP_B(hipGetLastError());
    hipLaunchKernelGGL(view_scan_kernel, dim3(grid_for(ncol, 0, 65536)),
                       dim3(kScanBlock), 0, 0, m->view_off, d_colnnz, ncol, nsm);
    CTBB_HIP_B(hipGetLastError());
    hipLaunchKernelGGL(col_scan_kernel, dim3(1), dim3(kColScanBlock), 0, 0,
                       d_colnnz, m->col_ptr, ncol);
    CTBB_HIP_B(hipGetLastError());
    int64_t nnz = 0;
    CTBB_HIP_B(hipMemcpy(&nnz, m->col_ptr + ncol, sizeof(int64_t),
                         hipMemcpyDeviceToHost));
    m->nnz = nnz;
    if (dev_alloc_n(&m->idx, std::max<int64_t>(nnz, 1), "SSM indices") ||
        dev_alloc_n(&m->val, std::max<int64_t>(nnz, 1), "SSM values"))
        return bail(CTBB_E_NOMEM);
    a.idx = m->idx;
    a.val = m->val;
    hipLaunchKernelGGL(ssm_build_kernel<true>, dim3(gbuild), dim3(kBuildBlock),
                       0, 0, a);
    CTBB_HIP_B(hipGetLastError());
    CTBB_HIP_B(hipEventRecord(ev1, 0));
    CTBB_HIP_B(hipEventSynchronize(ev1));
    float ms = 0.0f;
    CTBB_HIP_B(hipEventElapsedTime(&ms, ev0, ev1));
    m->build_ms = ms;
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
#undef CTBB_HIP_B
    cleanup();
    *out = m;
    return CTBB_OK;
}

extern "C" int ctbb_ssm_free(ctbb_ssm *m)
{
    if (!m)
        return CTBB_OK;
    if (m->col_ptr) (void)hipFree(m->col_ptr);
    if (m->idx) (void)hipFree(m->idx);
    if (m->val) (void)hipFree(m->val);
    if (m->val16) (void)hipFree(m->val16);
    if (m->col_scale) (void)hipFree(m->col_scale);
    if (m->view_off) (void)hipFree(m->view_off);
    if (m->rm_vc) (void)hipFree(m->rm_vc);
    if (m->rm_ent) (void)hipFree(m->rm_ent);
    if (m->rm_val) (void)hipFree(m->rm_val);
    delete m;
    return CTBB_OK;
}

extern "C" int ctbb_ssm_to_fp16(ctbb_ssm *m)
{
    if (!m)
        return fail(CTBB_E_ARG, "matrix is null");
    if (m->precision == CTBB_SSM_FP16)
        return CTBB_OK;
    CTBB_TRY(dev_alloc_n(&m->val16, std::max<int64_t>(m->nnz, 1), "FP16 values"));
    CTBB_TRY(dev_alloc_n(&m->col_scale, m->ncol, "FP16 column scales"));
    hipLaunchKernelGGL(to_fp16_kernel, dim3(grid_for(m->ncol, 0, 65536)),
                       dim3(256), 0, 0, m->col_ptr, m->val, m->val16,
                       m->col_scale, m->ncol);
    CTBB_HIP(hipGetLastError());
    CTBB_HIP(hipDeviceSynchronize());
    CTBB_HIP(hipFree(m->val));
    m->val = nullptr;
    m->precision = CTBB_SSM_FP16;
    // the ray-major copy holds the FP32 values: rebuild it from FP16 on use
    if (m->rm_vc) (void)hipFree(m->rm_vc);
    if (m->rm_ent) (void)hipFree(m->rm_ent);
    if (m->rm_val) (void)hipFree(m->rm_val);
    m->rm_vc = nullptr;
    m->rm_ent = nullptr;
    m->rm_val = nullptr;
    return CTBB_OK;
}

extern "C" int ctbb_ssm_info(const ctbb_ssm *m, int64_t *nnz, int *precision,
                             double *build_ms)
{
    if (!m)
        return fail(CTBB_E_ARG, "matrix is null");
    if (nnz) *nnz = m->nnz;
    if (precision) *precision = m->precision;
    if (build_ms) *build_ms = m->build_ms;
    return CTBB_OK;
}

extern "C" int ctbb_ssm_download(const ctbb_ssm *m, int64_t *col_ptr,
                                 int32_t *idx, float *val)
{
    if (!m)
        return fail(CTBB_E_ARG, "matrix is null");
    if (col_ptr)
        CTBB_HIP(hipMemcpy(col_ptr, m->col_ptr, (m->ncol + 1) * sizeof(int64_t),
                           hipMemcpyDeviceToHost));
    if (idx && m->nnz)
        CTBB_HIP(hipMemcpy(idx, m->idx, m->nnz * sizeof(int32_t),
                           hipMemcpyDeviceToHost));
    if (val && m->nnz) {
        if (m->precision == CTBB_SSM_FP32) {
            CTBB_HIP(hipMemcpy(val, m->val, m->nnz * sizeof(float),
                               hipMemcpyDeviceToHost));
        } else {
            float *tmp = nullptr;
            CTBB_TRY(dev_alloc_n(&tmp, m->nnz, "FP16 expansion"));
            hipLaunchKernelGGL(from_fp16_kernel, dim3(grid_for(m->ncol, 0, 65536)),
                               dim3(256), 0, 0, m->col_ptr, m->val16,
                               m->col_scale, tmp, m->ncol);
            hipError_t e1 = hipGetLastError();
            hipError_t e2 = hipMemcpy(val, tmp, m->nnz * sizeof(float),
                                      hipMemcpyDeviceToHost);
            (void)hipFree(tmp);
            CTBB_HIP(e1);
            CTBB_HIP(e2);
        }
    }
    return CTBB_OK;
}

extern "C" int ctbb_ssm_download_cols(const ctbb_ssm *m, int64_t c0, int64_t c1,
                                      int64_t *col_ptr, int32_t *idx, float *val)
{
    if (!m || !col_ptr)
        return fail(CTBB_E_ARG, "null argument");
    if (c0 < 0 || c1 > m->ncol || c0 > c1)
        return fail(CTBB_E_ARG, "column range out of bounds");
    CTBB_HIP(hipMemcpy(col_ptr, m->col_ptr + c0, (c1 - c0 + 1) * sizeof(int64_t),
                       hipMemcpyDeviceToHost));
    const int64_t base = col_ptr[0];
    for (int64_t i = 0; i <= c1 - c0; i++)
        col_ptr[i] -= base;
    const int64_t n = col_ptr[c1 - c0];
    if (idx && n)
        CTBB_HIP(hipMemcpy(idx, m->idx + base, n * sizeof(int32_t),
                           hipMemcpyDeviceToHost));
    if (val && n) {
        if (m->precision != CTBB_SSM_FP32)
            return fail(CTBB_E_STATE, "column download needs an FP32 matrix");
        CTBB_HIP(hipMemcpy(val, m->val + base, n * sizeof(float),
                           hipMemcpyDeviceToHost));
    }
    return CTBB_OK;
}

extern "C" int ctbb_ssm_upload(const ctbb_geom *g, const int64_t *col_ptr,
                               const int32_t *idx, const float *val,
                               ctbb_ssm **out)
{
    CTBB_TRY(check_geom(g));
    if (!col_ptr || !out)
        return fail(CTBB_E_ARG, "null argument");
    const int64_t ncol = g->nx * g->ny;
    const int64_t nnz = col_ptr[ncol];
    if (col_ptr[0] != 0 || nnz < 0)
        return fail(CTBB_E_ARG, "col_ptr must start at 0");
    if (nnz > 0 && (!idx || !val))
        return fail(CTBB_E_ARG, "null argument");
    const int64_t qmax = g->n_views_sm * g->n_channels * g->n_rows;
    for (int64_t c = 0; c < ncol; c++) {
        if (col_ptr[c + 1] < col_ptr[c])
            return fail(CTBB_E_ARG, "col_ptr must be non-decreasing");
        for (int64_t p = col_ptr[c]; p < col_ptr[c + 1]; p++) {
            if (idx[p] < 0 || idx[p] >= qmax)
                return fail(CTBB_E_ARG, "ray index out of range in column " +
                                            std::to_string(c));
            if (p > col_ptr[c] && idx[p] <= idx[p - 1])
                return fail(CTBB_E_ARG, "ray indices must ascend in column " +
                                            std::to_string(c));
        }
    }
    ctbb_ssm *m = new ctbb_ssm();
    m->g = *g;
    m->ncol = ncol;
    m->nnz = nnz;
    int rc = CTBB_OK;
    if ((rc = dev_alloc_n(&m->col_ptr, ncol + 1, "SSM column pointers")) ||
        (rc = dev_alloc_n(&m->idx, std::max<int64_t>(nnz, 1), "SSM indices")) ||
        (rc = dev_alloc_n(&m->val, std::max<int64_t>(nnz, 1), "SSM values"))) {
        ctbb_ssm_free(m);
        return rc;
    }
    hipError_t e = hipMemcpy(m->col_ptr, col_ptr, (ncol + 1) * 8,
                             hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz)
        e = hipMemcpy(m->idx, idx, nnz * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz)
        e = hipMemcpy(m->val, val, nnz * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        ctbb_ssm_free(m);
        return fail(CTBB_E_HIP, std::string("upload: ") + hipGetErrorString(e));
    }
    // view offsets for the footprint analysis
    std::vector<int32_t> vo(ncol * g->n_views_sm, 0);
    const int64_t per_view = g->n_channels * g->n_rows;
    for (int64_t c = 0; c < ncol; c++) {
        int64_t p = col_ptr[c];
        for (int64_t v = 0; v < g->n_views_sm; v++) {
            vo[c * g->n_views_sm + v] = (int32_t)(p - col_ptr[c]);
            while (p < col_ptr[c + 1] && idx[p] < (v + 1) * per_view)
                p++;
        }
    }
    if ((rc = dev_alloc_n(&m->view_off, vo.size(), "SSM view offsets"))) {
        ctbb_ssm_free(m);
        return rc;
    }
    e = hipMemcpy(m->view_off, vo.data(), vo.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        ctbb_ssm_free(m);
        return fail(CTBB_E_HIP, std::string("upload: ") + hipGetErrorString(e));
    }
    *out = m;
    return CTBB_OK;
}

// matrix.bin: per column uint64 nnz + nnz x {int32 index, float32 value}
// (src/generate_system_matrix.cpp:212-246, read at src/icd_iteration.cpp:
// 164-179).
extern "C" int ctbb_ssm_write_file(const ctbb_ssm *m, const char *path)
{
    if (!m || !path)
        return fail(CTBB_E_ARG, "null argument");
    std::vector<int64_t> cp(m->ncol + 1);
    std::vector<int32_t> idx(std::max<int64_t>(m->nnz, 1));
    std::vector<float> val(std::max<int64_t>(m->nnz, 1));
    CTBB_TRY(ctbb_ssm_download(m, cp.data(), idx.data(), val.data()));
    FILE *f = fopen(path, "wb");
    if (!f)
        return fail(CTBB_E_IO, std::string("cannot open ") + path);
    std::vector<char> buf;
    for (int64_t c = 0; c < m->ncol; c++) {
        const uint64_t n = (uint64_t)(cp[c + 1] - cp[c]);
        buf.resize(8 + n * 8);
        memcpy(buf.data(), &n, 8);
        for (uint64_t i = 0; i < n; i++) {
            memcpy(buf.data() + 8 + i * 8, &idx[cp[c] + i], 4);
            memcpy(buf.data() + 12 + i * 8, &val[cp[c] + i], 4);
        }
        if (fwrite(buf.data(), 1, buf.size(), f) != buf.size()) {
            fclose(f);
            return fail(CTBB_E_IO, std::string("short write to ") + path);
        }
    }
    if (fclose(f) != 0)
        return fail(CTBB_E_IO, std::string("close failed: ") + path);
    return CTBB_OK;
}

extern "C" int ctbb_ssm_read_file(const ctbb_geom *g, const char *path,
                                  ctbb_ssm **out)
{
    CTBB_TRY(check_geom(g));
    if (!path || !out)
        return fail(CTBB_E_ARG, "null argument");
    FILE *f = fopen(path, "rb");
    if (!f)
        return fail(CTBB_E_IO, std::string("cannot open ") + path);
    const int64_t ncol = g->nx * g->ny;
    std::vector<int64_t> cp(ncol + 1, 0);
    std::vector<int32_t> idx;
    std::vector<float> val;
    std::vector<char> buf;
    for (int64_t c = 0; c < ncol; c++) {
        uint64_t n = 0;
        if (fread(&n, 8, 1, f) != 1) {
            fclose(f);
            return fail(CTBB_E_IO, "matrix file truncated at column " +
                                       std::to_string(c));
        }
        if (n > (uint64_t)(g->n_views_sm * g->n_channels * g->n_rows)) {
            fclose(f);
            return fail(CTBB_E_IO, "implausible column length at column " +
                                       std::to_string(c));
        }
        buf.resize(n * 8);
        if (n && fread(buf.data(), 8, n, f) != n) {
            fclose(f);
            return fail(CTBB_E_IO, "matrix file truncated in column " +
                                       std::to_string(c));
        }
        for (uint64_t i = 0; i < n; i++) {
            int32_t q;
            float v;
            memcpy(&q, buf.data() + i * 8, 4);
            memcpy(&v, buf.data() + i * 8 + 4, 4);
            idx.push_back(q);
            val.push_back(v);
        }
        cp[c + 1] = cp[c] + (int64_t)n;
    }
    fclose(f);
    return ctbb_ssm_upload(g, cp.data(), idx.data(), val.data(), out);
}

extern "C" int ctbb_ssm_group_spacing(const ctbb_ssm *m,
                                      const int64_t *slice_index,
                                      int64_t max_spacing,
                                      double max_conflict_fraction,
                                      int64_t *spacing)
{
    if (!m || !slice_index || !spacing)
        return fail(CTBB_E_ARG, "null argument");
    const ctbb_geom *g = &m->g;
    const int64_t nsm = g->n_views_sm;
    if (!m->view_off)
        return fail(CTBB_E_STATE, "matrix has no view offsets");
    if (nsm > kFpMaxViews)
        return fail(CTBB_E_UNSUPPORTED, "footprint analysis supports at most " +
                                            std::to_string(kFpMaxViews) +
                                            " SSM views");
    if (max_spacing < 2)
        max_spacing = std::max<int64_t>(2, g->nz);
    // candidate shifts: central-view differences between slices
    std::vector<int64_t> shifts;
    for (int64_t k = 0; k < g->nz; k++)
        for (int64_t k2 = k + 1; k2 < g->nz; k2++) {
            int64_t dv = slice_index[k2] - slice_index[k];
            if (dv < 0) dv = -dv;
            if (dv > 0 && dv < nsm)
                shifts.push_back(dv);
        }
    std::sort(shifts.begin(), shifts.end());
    shifts.erase(std::unique(shifts.begin(), shifts.end()), shifts.end());
    std::vector<uint32_t> flags(shifts.size(), 0);
    if (!shifts.empty()) {
        int64_t *d_sh = nullptr;
        uint32_t *d_fl = nullptr;
        CTBB_TRY(dev_alloc_n(&d_sh, shifts.size(), "shifts"));
        int rc = dev_alloc_n(&d_fl, shifts.size(), "shift flags");
        if (rc) {
            (void)hipFree(d_sh);
            return rc;
        }
        hipError_t e = hipMemcpy(d_sh, shifts.data(), shifts.size() * 8,
                                 hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemset(d_fl, 0, shifts.size() * 4);
        const double fov_limit =
            (g->acquisition_fov / 2.0) * (g->acquisition_fov / 2.0);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(footprint_conflict_kernel,
                               dim3(grid_for(m->ncol, 0, 8192)), dim3(kFpBlock), 0,
                               0, m->col_ptr, m->idx, m->view_off, g->nx, g->ny, nsm,
                               g->n_channels, g->n_rows, g->center_voxel_x,
                               g->center_voxel_y, g->voxel_size_x,
                               g->voxel_size_y, fov_limit, d_sh,
                               (int64_t)shifts.size(), d_fl);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpy(flags.data(), d_fl, flags.size() * 4, hipMemcpyDeviceToHost);
        (void)hipFree(d_sh);
        (void)hipFree(d_fl);
        CTBB_HIP(e);
    }
    int64_t nfov = 0;
    {
        const double lim = (g->acquisition_fov / 2.0) * (g->acquisition_fov / 2.0);
        for (int64_t j = 0; j < g->ny; j++)
            for (int64_t i = 0; i < g->nx; i++) {
                const double x = (i - g->center_voxel_x) * g->voxel_size_x;
                const double y = (j - g->center_voxel_y) * g->voxel_size_y;
                if ((x * x + y * y) < lim)
                    nfov++;
            }
    }
    const double allowed = std::max(0.0, max_conflict_fraction) * (double)nfov;
    auto conflict = [&](int64_t dv) {
        if (dv < 0) dv = -dv;
        if (dv == 0) return true;
        if (dv >= nsm) return false;
        auto it = std::lower_bound(shifts.begin(), shifts.end(), dv);
        if (it == shifts.end() || *it != dv)
            return false;
        const uint32_t n = flags[it - shifts.begin()];
        return n > 0 && (double)n > allowed;
    };
    for (int64_t s = 2; s <= max_spacing; s++) {
        bool ok = true;
        for (int64_t k = 0; k < g->nz && ok; k++)
            for (int64_t k2 = k + s; k2 < g->nz && ok; k2 += s)
                if (conflict(slice_index[k2] - slice_index[k]))
                    ok = false;
        if (ok) {
            *spacing = s;
            return CTBB_OK;
        }
    }
    *spacing = 0;
    return fail(CTBB_E_ARG, "no conflict-free slice spacing <= " +
                                std::to_string(max_spacing));
}
