// icd.hip -- iterative coordinate descent (ICD) on gfx950.
//
// Replaces icd_iteration (src/icd_iteration.cpp:42-286) with the penalty
// steps of src/penalties.cpp:39-154 fused into the sweep kernel.
//
// State (device resident, reference sinogram layout
// r = row + nr*(channel + nch*view)):
//   e  FP32 error sinogram y - A x (the reference keeps yhat = A x in FP64
//      and forms raw - yhat on the fly, src/icd_iteration.cpp:209)
//   x  FP32 volume in the rotated frame, index c + nx*ny*k
//   w  optional FP32 statistical weights (w == 1 reproduces the reference)
//
// Sweep order: columns c = ix + nx*iy in raster order inside the acquisition
// FOV (src/icd_iteration.cpp:157-181); in each column the slices are visited
// as groups {j, j+s, j+2s, ...}, j = 0..s-1, where the spacing s is chosen so
// that the sinogram footprints of every group are pairwise disjoint (and no
// two members are penalty neighbours).  One group is one step of a
// persistent cooperative kernel: one workgroup per voxel of the group
// gathers theta1 = sum w a e and theta2 = sum w a^2 over the column's
// in-range entries (wave shuffles + LDS, fixed order), takes the penalty
// step in FP64, clamps to x >= 0, writes x and scatters e -= du a.  The
// scatter is conflict-free because the group's footprints are disjoint; a
// grid-wide barrier separates the steps.  The result equals the sequential
// sweep in the group order (an oracle replays it), up to FP32 rounding.
//
// FP64 is used for the per-voxel scalar penalty solve, the forward
// projection of the periodic resync e = y - A x, and the cost.
#include "common.hpp"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <cstring>


namespace {

constexpr double kSqrt2 = 1.41421356237309504880168872420969807;

// ---------------------------------------------------------------------------
// Penalty steps (src/penalties.cpp).  X is float (sweep) or double (checks).

__device__ __forceinline__ double nbr_weight(int i)
{
    // src/penalties.cpp:45-52: diagonal neighbours 1/sqrt(2), others 1
    return (i == 0 || i == 2 || i == 5 || i == 7) ? 1.0 / kSqrt2 : 1.0;
}

__device__ __forceinline__ double weights_scale8()
{
    double s = 0.0;
    for (int i = 0; i < 8; i++)
        s += nbr_weight(i);
    return s;
}

template <typename X>
__device__ double quadratic_step(int64_t q, double alpha, double beta,
                                 double lambda, const X *x, int64_t nx,
                                 int64_t n)
{
    const int64_t nb[8] = {q - nx - 1, q - nx,     q - nx + 1, q - 1,
                           q + 1,      q + nx - 1, q + nx,     q + nx + 1};
    const double xq = (double)x[q];
    double sum = 0.0;
    for (int i = 0; i < 8; i++)
        if (nb[i] >= 0 && nb[i] < n)
            sum += nbr_weight(i) * ((double)x[nb[i]] - xq);
    return (beta + lambda * sum) / (alpha + lambda * weights_scale8());
}

__device__ __forceinline__ double ep_func(double du, double alpha, double beta,
                                          double delta, double lambda,
                                          const double *diff, const bool *used)
{
    double sum = 0;
    for (int i = 0; i < 8; i++) {
        if (used[i]) {
            double t = (du - diff[i]);
            if (t >= 0)
                sum = sum + nbr_weight(i) * (1.0 - 1.0 / (1.0 + fabs(t / delta)));
            else
                sum = sum - nbr_weight(i) * (1.0 - 1.0 / (1.0 + fabs(t / delta)));
        }
    }
    return 2 * alpha * du - 2 * beta + lambda * sum;
}

// Bisection of src/penalties.cpp:88-139 in FP64 (tolerance 1e-8 on the
// half-bracket, stop on an exact zero), one thread.
template <typename X>
__device__ double edge_preserving_step(int64_t q, double alpha, double beta,
                                       double lambda, double delta, const X *x,
                                       int64_t nx, int64_t n)
{
    if (alpha == 0.0)
        return 0.0;
    const int64_t nb[8] = {q - nx - 1, q - nx,     q - nx + 1, q - 1,
                           q + 1,      q + nx - 1, q + nx,     q + nx + 1};
    const double xq = (double)x[q];
    double diff[8];
    bool used[8];
    double a = beta / alpha, b = beta / alpha;
    for (int i = 0; i < 8; i++) {
        used[i] = nb[i] >= 0 && nb[i] < n;
        diff[i] = used[i] ? ((double)x[nb[i]] - xq) : 0.0;
        if (used[i]) {
            if (diff[i] < a) a = diff[i];
            if (diff[i] > b) b = diff[i];
        }
    }
    const double tol = 0.00000001;
    double c = (a + b) / 2.0;
    for (int it = 0; it < 4096; it++) {
        c = (a + b) / 2.0;
        const double f_c = ep_func(c, alpha, beta, delta, lambda, diff, used);
        const double f_a = ep_func(a, alpha, beta, delta, lambda, diff, used);
        if ((f_c == 0) || (b - a) / 2 < tol)
            return c;
        const int sc = f_c < 0 ? -1 : 1, sa = f_a < 0 ? -1 : 1;
        if (sc == sa)
            a = c;
        else
            b = c;
    }
    return c;
}

template <typename X>
__device__ __forceinline__ double penalty_step(int penalty, int64_t q,
                                               double alpha, double beta,
                                               double lambda, double delta,
                                               const X *x, int64_t nx,
                                               int64_t n)
{
    if (penalty == CTBB_PENALTY_QUADRATIC)
        return quadratic_step(q, alpha, beta, lambda, x, nx, n);
    return edge_preserving_step(q, alpha, beta, lambda, delta, x, nx, n);
}

__global__ void penalty_steps_kernel(int penalty, double lambda, double delta,
                                     const double *x, int64_t nx, int64_t nvox,
                                     int64_t n, const int64_t *q,
                                     const double *alpha, const double *beta,
                                     double *du)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        du[i] = penalty_step(penalty, q[i], alpha[i], beta[i], lambda, delta,
                             x, nx, nvox);
}

// ---------------------------------------------------------------------------
// Sweep kernel.
//
// One step = (batch b, slice group j): the voxels (c, k) for every column c
// of batch b and every slice k = j, j+s, j+2s, ... .  Phase A: one workgroup
// per voxel gathers theta1 = sum w a e and theta2 = sum w a^2 over the
// column's in-range entries from the error sinogram as it stood at the start
// of the step, solves the penalty step, clamps and writes x and du.  Phase B:
// every voxel scatters e -= du a (atomic adds: columns of one batch may
// share rays at the crossings of their sinogram traces; slices of one column
// never do).  Grid-wide barriers separate the phases and the steps.

constexpr int kSweepBlock = 256;
constexpr int kWaves = kSweepBlock / 64;

struct SweepArgs {
    const int64_t *col_ptr;
    const int32_t *idx;
    const float *val;
    const __half *val16;
    const float *col_scale;
    float *e;
    const float *w;
    float *x;
    float *du;                // [max tasks per step]
    const int64_t *slice_off;
    const int32_t *cols;      // batch columns, concatenated
    const int64_t *batch_ptr; // [nbatch+1]
    int64_t step_begin, step_end;  // step = batch * spacing + j
    int64_t spacing;
    int64_t nz, nx, nxy, nvox, data_size;
    int penalty;
    double lambda, delta;
};

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

template <bool FP16>
__device__ __forceinline__ float ssm_value(const SweepArgs &a, int64_t p,
                                           float scale)
{
    return FP16 ? __half2float(a.val16[p]) * scale : a.val[p];
}

// Edge-preserving step by 65-section on one wavefront: the bracket of
// src/penalties.cpp:106-118 is cut at 64 interior points per round, one per
// lane; f is non-decreasing in c (alpha >= 0, psi increasing), so the first
// lane with f > 0 bounds the root from above.  Stops when the half-bracket
// is below 1e-8 (the reference's tolerance) or on an exact zero, and returns
// the bracket midpoint: within 1e-8 of the root, like the reference's
// bisection, in 5 rounds instead of ~27.  Called by all 64 lanes.
__device__ double edge_preserving_wave(int64_t q, double alpha, double beta,
                                       double lambda, double delta,
                                       const float *x, int64_t nx, int64_t n,
                                       int lane)
{
    if (alpha == 0.0)
        return 0.0;
    const int64_t nb[8] = {q - nx - 1, q - nx,     q - nx + 1, q - 1,
                           q + 1,      q + nx - 1, q + nx,     q + nx + 1};
    const double xq = (double)x[q];
    double diff[8];
    bool used[8];
    double a = beta / alpha, b = beta / alpha;
    for (int i = 0; i < 8; i++) {
        used[i] = nb[i] >= 0 && nb[i] < n;
        diff[i] = used[i] ? ((double)x[nb[i]] - xq) : 0.0;
        if (used[i]) {
            if (diff[i] < a) a = diff[i];
            if (diff[i] > b) b = diff[i];
        }
    }
    const double tol = 0.00000001;
    for (int round = 0; round < 64 && !((b - a) / 2 < tol); round++) {
        const double c = a + (b - a) * (double)(lane + 1) / 65.0;
        const double f = ep_func(c, alpha, beta, delta, lambda, diff, used);
        const uint64_t zero = __ballot(f == 0.0);
        if (zero)
            return __shfl(c, __ffsll((unsigned long long)zero) - 1, 64);
        const uint64_t pos = __ballot(f > 0.0);
        if (pos == 0) {
            a = __shfl(c, 63, 64);
        } else {
            const int l = __ffsll((unsigned long long)pos) - 1;
            const double cl = __shfl(c, l, 64);
            const double cp = __shfl(c, l > 0 ? l - 1 : 0, 64);
            if (l > 0)
                a = cp;
            b = cl;
        }
    }
    return (a + b) / 2;
}

// One wavefront per voxel: the sweep's solver, for checking.
__global__ void penalty_steps_wave_kernel(int penalty, double lambda,
                                          double delta, const float *x,
                                          int64_t nx, int64_t nvox, int64_t n,
                                          const int64_t *q, const double *alpha,
                                          const double *beta, double *du)
{
    const int64_t i = (int64_t)blockIdx.x;
    const int lane = threadIdx.x & 63;
    if (i >= n)
        return;
    double d;
    if (penalty == CTBB_PENALTY_QUADRATIC)
        d = quadratic_step(q[i], alpha[i], beta[i], lambda, x, nx, nvox);
    else
        d = edge_preserving_wave(q[i], alpha[i], beta[i], lambda, delta, x, nx,
                                 nvox, lane);
    if (lane == 0)
        du[i] = d;
}

constexpr int kUnroll = 8;
constexpr int kXcd = 8;   // workgroups are dealt round-robin to the 8 XCDs

// Task t of a step -> (column slot ci, group member v).  With at least 8
// columns in the batch, the ng tasks of one column go to workgroups of the
// same XCD (t = 8 s + xcd), so the column's SSM entries are fetched into
// one L2; smaller batches use t = ci * ng + v.
__device__ __forceinline__ bool task_of(int64_t t, int64_t ng, int64_t nc,
                                        int64_t &ci, int64_t &v)
{
    if (nc < kXcd) {
        ci = t / ng;
        v = t - ci * ng;
        return ci < nc;
    }
    const int64_t xcd = t % kXcd, sl = t / kXcd;
    ci = (sl / ng) * kXcd + xcd;
    v = sl - (sl / ng) * ng;
    return ci < nc;
}

struct StepArgs {
    int64_t c0, nc, j, ng;
};

// Phase A of a step: one workgroup per voxel task gathers the theta sums
// from the error sinogram (unchanged during the phase), solves the penalty
// step in FP64 (wavefront 0), clamps and writes x and du.
template <bool FP16, bool WEIGHTED>
__global__ __launch_bounds__(kSweepBlock) void icd_gather_kernel(SweepArgs a,
                                                                 StepArgs sa)
{
    __shared__ float red1[kWaves], red2[kWaves];
    const int tid = threadIdx.x, lane = tid & 63;
    const int64_t t = blockIdx.x;
    int64_t ci, v;
    if (!task_of(t, sa.ng, sa.nc, ci, v))
        return;
    const int64_t col = a.cols[sa.c0 + ci];
    const int64_t k = sa.j + a.spacing * v;
    const int64_t off = a.slice_off[k];
    const int64_t beg = a.col_ptr[col], len = a.col_ptr[col + 1] - beg;
    const float scale = FP16 ? a.col_scale[col] : 1.0f;
    const int32_t *__restrict__ idx = a.idx + beg;
    const float *__restrict__ e = a.e;
    const float *__restrict__ w = a.w;
    const int64_t ds = a.data_size;
    float t1 = 0.0f, t2 = 0.0f;
    // branch-free: every lane issues kUnroll index/value loads, then kUnroll
    // error (and weight) loads; out-of-range entries read e[0] with a zero
    // coefficient
    for (int64_t base = tid; base < len; base += kSweepBlock * kUnroll) {
        int32_t iv[kUnroll];
        float av[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const int64_t o = base + u * kSweepBlock;
            const int64_t oc = o < len ? o : len - 1;
            iv[u] = idx[oc];
            av[u] = ssm_value<FP16>(a, beg + oc, scale);
        }
        __builtin_amdgcn_sched_barrier(0);
        int64_t r[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const int64_t o = base + u * kSweepBlock;
            r[u] = (int64_t)iv[u] + off;
            const bool ok = o < len && r[u] >= 0 && r[u] < ds;
            av[u] = ok ? av[u] : 0.0f;
            r[u] = ok ? r[u] : 0;
        }
        float ev[kUnroll], wv[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            ev[u] = e[r[u]];
            wv[u] = WEIGHTED ? w[r[u]] : 1.0f;
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const float wa = WEIGHTED ? wv[u] * av[u] : av[u];
            t2 += wa * av[u];
            t1 += wa * ev[u];
        }
    }
    t1 = wave_sum(t1);
    t2 = wave_sum(t2);
    if (lane == 0) {
        red1[tid >> 6] = t1;
        red2[tid >> 6] = t2;
    }
    __syncthreads();
    if (tid < 64) {
        double th1 = 0.0, th2 = 0.0;
        for (int i = 0; i < kWaves; i++) {
            th1 += (double)red1[i];
            th2 += (double)red2[i];
        }
        const int64_t q = col + a.nxy * k;
        double du;
        if (a.penalty == CTBB_PENALTY_QUADRATIC)
            du = quadratic_step(q, th2, th1, a.lambda, a.x, a.nx, a.nvox);
        else
            du = edge_preserving_wave(q, th2, th1, a.lambda, a.delta, a.x, a.nx,
                                      a.nvox, lane);
        if (tid == 0) {
            const double xq = (double)a.x[q];
            if (du + xq < 0)
                du = -xq;
            a.x[q] = (float)(xq + du);
            a.du[t] = (float)du;
        }
    }
}

// Phase B of a step: every voxel task scatters e -= du a (atomic adds:
// columns of one batch may share rays where their sinogram traces cross;
// slices of one group in one column do not).
template <bool FP16>
__global__ __launch_bounds__(kSweepBlock) void icd_scatter_kernel(SweepArgs a,
                                                                  StepArgs sa)
{
    const int tid = threadIdx.x;
    const int64_t t = blockIdx.x;
    int64_t ci, v;
    if (!task_of(t, sa.ng, sa.nc, ci, v))
        return;
    const float du = a.du[t];
    if (du == 0.0f)
        return;
    const int64_t col = a.cols[sa.c0 + ci];
    const int64_t off = a.slice_off[sa.j + a.spacing * v];
    const int64_t beg = a.col_ptr[col], len = a.col_ptr[col + 1] - beg;
    const float scale = FP16 ? a.col_scale[col] : 1.0f;
    const int32_t *__restrict__ idx = a.idx + beg;
    float *__restrict__ e = a.e;
    const int64_t ds = a.data_size;
    for (int64_t base = tid; base < len; base += kSweepBlock * kUnroll) {
        int64_t r[kUnroll];
        float dv[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const int64_t o = base + u * kSweepBlock;
            const int64_t oc = o < len ? o : len - 1;
            r[u] = (int64_t)idx[oc] + off;
            dv[u] = -du * ssm_value<FP16>(a, beg + oc, scale);
            if (!(o < len && r[u] >= 0 && r[u] < ds))
                r[u] = -1;
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++)
            if (r[u] >= 0)
                atomicAdd(&e[r[u]], dv[u]);
    }
}

// ---------------------------------------------------------------------------
// Forward projection yhat = A x in FP64 (src/icd_iteration.cpp:63-116).

constexpr int kFwdBlock = 256;

template <typename X, bool FP16>
__global__ __launch_bounds__(kFwdBlock) void fwd_scatter_kernel(
    const int64_t *col_ptr, const int32_t *idx, const float *val,
    const __half *val16, const float *col_scale, const X *x,
    const int64_t *slice_off, double *yhat, int64_t ncol, int64_t nz,
    int64_t data_size)
{
    for (int64_t col = blockIdx.x; col < ncol; col += gridDim.x) {
        const int64_t beg = col_ptr[col], end = col_ptr[col + 1];
        const float scale = FP16 ? col_scale[col] : 1.0f;
        for (int64_t k = 0; k < nz; k++) {
            const double xv = (double)x[col + ncol * k];
            if (xv == 0.0)
                continue;
            const int64_t off = slice_off[k];
            for (int64_t p = beg + threadIdx.x; p < end; p += kFwdBlock) {
                const int64_t r = (int64_t)idx[p] + off;
                if (r > -1 && r < data_size) {
                    const double av =
                        FP16 ? (double)(__half2float(val16[p]) * scale)
                             : (double)val[p];
                    unsafeAtomicAdd(&yhat[r], xv * av);
                }
            }
        }
    }
}

__global__ void fill_f64_kernel(double *p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

constexpr int kRedBlock = 256;

__device__ double block_sum_f64(double v, double *sh)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0)
        sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kRedBlock / 64; i++)
            s += sh[i];
    __syncthreads();
    return s;
}

// e = y - yhat (optional) and per-block partial sums of w (y - yhat)^2.
__global__ __launch_bounds__(kRedBlock) void resync_finish_kernel(
    const float *y, const float *w, const double *yhat, float *e,
    int64_t data_size, double *partials)
{
    __shared__ double sh[kRedBlock / 64];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x;
         i < data_size; i += (int64_t)gridDim.x * kRedBlock) {
        const double d = (double)y[i] - yhat[i];
        if (e)
            e[i] = (float)d;
        const double ww = w ? (double)w[i] : 1.0;
        acc += ww * d * d;
    }
    const double s = block_sum_f64(acc, sh);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = s;
}

// Penalty part of the cost: sum over pairs (q, q+d), d in {1, nx-1, nx, nx+1}.
__global__ __launch_bounds__(kRedBlock) void penalty_cost_kernel(
    const float *x, int64_t nx, int64_t n, int penalty, double delta,
    double *partials)
{
    __shared__ double sh[kRedBlock / 64];
    const int64_t dd[4] = {1, nx - 1, nx, nx + 1};
    const double wd[4] = {1.0, 1.0 / kSqrt2, 1.0, 1.0 / kSqrt2};
    double acc = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * kRedBlock) {
        const double xq = (double)x[q];
        for (int t = 0; t < 4; t++) {
            if (q + dd[t] >= n)
                continue;
            const double diff = xq - (double)x[q + dd[t]];
            if (penalty == CTBB_PENALTY_QUADRATIC) {
                acc += wd[t] * diff * diff;
            } else {
                const double ad = fabs(diff);
                acc += wd[t] * (ad - delta * log1p(ad / delta));
            }
        }
    }
    const double s = block_sum_f64(acc, sh);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = s;
}

// In-range entry count of every (FOV column, slice): the exact number of
// entries a sweep gathers and scatters.
__global__ void inrange_count_kernel(const int64_t *col_ptr,
                                     const int32_t *idx, const int32_t *cols,
                                     int64_t ncols, const int64_t *slice_off,
                                     int64_t nz, int64_t data_size,
                                     unsigned long long *total)
{
    for (int64_t i = blockIdx.x; i < ncols; i += gridDim.x) {
        const int64_t col = cols[i];
        const int64_t beg = col_ptr[col], end = col_ptr[col + 1];
        unsigned long long cnt = 0;
        for (int64_t k = threadIdx.x; k < nz; k += blockDim.x) {
            const int64_t lo_v = -slice_off[k], hi_v = data_size - slice_off[k];
            int64_t lo = beg, hi = end;  // first idx >= lo_v
            while (lo < hi) {
                int64_t mid = (lo + hi) / 2;
                if (idx[mid] < lo_v) lo = mid + 1; else hi = mid;
            }
            const int64_t first = lo;
            lo = beg;
            hi = end;  // first idx >= hi_v
            while (lo < hi) {
                int64_t mid = (lo + hi) / 2;
                if (idx[mid] < hi_v) lo = mid + 1; else hi = mid;
            }
            cnt += (unsigned long long)(lo - first);
        }
        atomicAdd(total, cnt);
    }
}

int g_num_cu = 0;

int num_cu()
{
    if (g_num_cu == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&g_num_cu, hipDeviceAttributeMultiprocessorCount,
                                  dev) != hipSuccess)
            g_num_cu = 64;
    }
    return g_num_cu;
}

int red_grid(int64_t n)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + kRedBlock - 1) / kRedBlock,
                                                        4 * 1024));
}

// yhat = A x (FP64) on the recon stream.
template <typename X>
int launch_forward(const ctbb_ssm *m, const X *x, const int64_t *slice_off,
                   double *yhat, int64_t data_size, hipStream_t st)
{
    hipLaunchKernelGGL(fill_f64_kernel, dim3(4096), dim3(256), 0, st, yhat,
                       data_size);
    CTBB_HIP(hipGetLastError());
    const int grid = (int)std::min<int64_t>(m->ncol, 65536);
    if (m->precision == CTBB_SSM_FP16)
        hipLaunchKernelGGL((fwd_scatter_kernel<X, true>), dim3(grid),
                           dim3(kFwdBlock), 0, st, m->col_ptr, m->idx, m->val,
                           m->val16, m->col_scale, x, slice_off, yhat, m->ncol,
                           m->g.nz, data_size);
    else
        hipLaunchKernelGGL((fwd_scatter_kernel<X, false>), dim3(grid),
                           dim3(kFwdBlock), 0, st, m->col_ptr, m->idx, m->val,
                           m->val16, m->col_scale, x, slice_off, yhat, m->ncol,
                           m->g.nz, data_size);
    CTBB_HIP(hipGetLastError());
    return CTBB_OK;
}

// Sum of per-block partials, in block order (deterministic).
int sum_partials(const double *d_part, int n, hipStream_t st, double *out)
{
    std::vector<double> h(n);
    CTBB_HIP(hipMemcpyAsync(h.data(), d_part, n * sizeof(double),
                            hipMemcpyDeviceToHost, st));
    CTBB_HIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (int i = 0; i < n; i++)
        s += h[i];
    *out = s;
    return CTBB_OK;
}

int recon_data_term(ctbb_recon *r, bool write_e, double *data)
{
    CTBB_TRY(launch_forward(r->m, r->x, r->slice_off, r->yhat, r->data_size,
                            r->stream));
    const int g = red_grid(r->data_size);
    hipLaunchKernelGGL(resync_finish_kernel, dim3(g), dim3(kRedBlock), 0,
                       r->stream, r->y, r->w, r->yhat, write_e ? r->e : nullptr,
                       r->data_size, r->partials);
    CTBB_HIP(hipGetLastError());
    return sum_partials(r->partials, g, r->stream, data);
}

int recon_penalty_term(ctbb_recon *r, double *pen)
{
    const int g = red_grid(r->nvox);
    hipLaunchKernelGGL(penalty_cost_kernel, dim3(g), dim3(kRedBlock), 0,
                       r->stream, r->x, r->g.nx, r->nvox, r->opts.penalty,
                       r->opts.delta, r->partials);
    CTBB_HIP(hipGetLastError());
    return sum_partials(r->partials, g, r->stream, pen);
}

}  // namespace

using namespace ctbb;

extern "C" int ctbb_forward_project(const ctbb_ssm *m,
                                    const int64_t *slice_index, const double *x,
                                    double *yhat)
{
    if (!m || !slice_index || !x || !yhat)
        return fail(CTBB_E_ARG, "null argument");
    const ctbb_geom *g = &m->g;
    const int64_t nvox = g->nx * g->ny * g->nz, ds = data_size_of(g);
    std::vector<int64_t> off(g->nz);
    for (int64_t k = 0; k < g->nz; k++)
        off[k] = slice_offset(g, slice_index[k]);
    double *d_x = nullptr, *d_y = nullptr;
    int64_t *d_off = nullptr;
    int rc = CTBB_OK;
    if ((rc = dev_alloc_n(&d_x, nvox, "volume")) ||
        (rc = dev_alloc_n(&d_y, ds, "projection")) ||
        (rc = dev_alloc_n(&d_off, g->nz, "slice offsets"))) {
        if (d_x) (void)hipFree(d_x);
        if (d_y) (void)hipFree(d_y);
        return rc;
    }
    hipError_t e = hipMemcpy(d_x, x, nvox * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(d_off, off.data(), g->nz * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        rc = launch_forward(m, (const double *)d_x, d_off, d_y, ds, 0);
        if (rc == CTBB_OK)
            e = hipMemcpy(yhat, d_y, ds * 8, hipMemcpyDeviceToHost);
    }
    (void)hipFree(d_x);
    (void)hipFree(d_y);
    (void)hipFree(d_off);
    if (rc != CTBB_OK)
        return rc;
    CTBB_HIP(e);
    return CTBB_OK;
}

extern "C" int ctbb_penalty_steps(const ctbb_geom *g, const ctbb_icd_opts *o,
                                  int solver, const double *x, int64_t n,
                                  const int64_t *q, const double *alpha,
                                  const double *beta, double *du)
{
    CTBB_TRY(check_geom(g));
    if (!o || !x || !q || !alpha || !beta || !du || n < 0)
        return fail(CTBB_E_ARG, "null argument");
    if (o->penalty != CTBB_PENALTY_QUADRATIC &&
        o->penalty != CTBB_PENALTY_EDGE_PRESERVING)
        return fail(CTBB_E_ARG, "unknown penalty");
    if (o->neighbours != 8)
        return fail(CTBB_E_UNSUPPORTED, "only the 8-neighbour stencil is "
                                        "available in ctbb_penalty_steps");
    const int64_t nvox = g->nx * g->ny * g->nz;
    for (int64_t i = 0; i < n; i++)
        if (q[i] < 0 || q[i] >= nvox)
            return fail(CTBB_E_ARG, "voxel index out of range");
    if (n == 0)
        return CTBB_OK;
    double *d_x = nullptr, *d_a = nullptr, *d_b = nullptr, *d_du = nullptr;
    int64_t *d_q = nullptr;
    int rc = CTBB_OK;
    if ((rc = dev_alloc_n(&d_x, nvox, "volume")) ||
        (rc = dev_alloc_n(&d_a, n, "alpha")) || (rc = dev_alloc_n(&d_b, n, "beta")) ||
        (rc = dev_alloc_n(&d_du, n, "du")) || (rc = dev_alloc_n(&d_q, n, "q"))) {
        (void)hipFree(d_x); (void)hipFree(d_a); (void)hipFree(d_b);
        (void)hipFree(d_du); (void)hipFree(d_q);
        return rc;
    }
    hipError_t e = hipMemcpy(d_x, x, nvox * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_a, alpha, n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_b, beta, n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_q, q, n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess && solver == 0) {
        hipLaunchKernelGGL(penalty_steps_kernel, dim3((n + 255) / 256), dim3(256),
                           0, 0, o->penalty, o->lambda, o->delta, d_x, g->nx,
                           nvox, n, d_q, d_a, d_b, d_du);
        e = hipGetLastError();
    } else if (e == hipSuccess) {
        std::vector<float> xf(nvox);
        for (int64_t i = 0; i < nvox; i++)
            xf[i] = (float)x[i];
        float *d_xf = nullptr;
        if ((rc = dev_alloc_n(&d_xf, nvox, "volume (FP32)")) == CTBB_OK) {
            e = hipMemcpy(d_xf, xf.data(), nvox * 4, hipMemcpyHostToDevice);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(penalty_steps_wave_kernel, dim3((unsigned)n),
                                   dim3(64), 0, 0, o->penalty, o->lambda, o->delta,
                                   d_xf, g->nx, nvox, n, d_q, d_a, d_b, d_du);
                e = hipGetLastError();
            }
            (void)hipFree(d_xf);
        } else {
            e = hipErrorOutOfMemory;
        }
    }
    if (e == hipSuccess)
        e = hipMemcpy(du, d_du, n * 8, hipMemcpyDeviceToHost);
    (void)hipFree(d_x); (void)hipFree(d_a); (void)hipFree(d_b);
    (void)hipFree(d_du); (void)hipFree(d_q);
    CTBB_HIP(e);
    return CTBB_OK;
}

extern "C" int ctbb_recon_free(ctbb_recon *r)
{
    if (!r)
        return CTBB_OK;
    float *fl[] = {r->y, r->w, r->e, r->x};
    for (float *p : fl)
        if (p) (void)hipFree(p);
    if (r->yhat) (void)hipFree(r->yhat);
    if (r->slice_off) (void)hipFree(r->slice_off);
    if (r->cols) (void)hipFree(r->cols);
    if (r->batch_ptr) (void)hipFree(r->batch_ptr);
    if (r->du) (void)hipFree(r->du);
    if (r->partials) (void)hipFree(r->partials);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
    return CTBB_OK;
}

extern "C" int ctbb_recon_create(const ctbb_ssm *m, const float *raw,
                                 const float *weights,
                                 const int64_t *slice_index,
                                 const ctbb_icd_opts *opts, ctbb_recon **out)
{
    if (!m || !raw || !slice_index || !opts || !out)
        return fail(CTBB_E_ARG, "null argument");
    *out = nullptr;
    if (opts->penalty != CTBB_PENALTY_QUADRATIC &&
        opts->penalty != CTBB_PENALTY_EDGE_PRESERVING)
        return fail(CTBB_E_ARG, "unknown penalty (valid: quadratic, "
                                "edge-preserving)");
    if (opts->neighbours != 8)
        return fail(CTBB_E_UNSUPPORTED, "neighbours must be 8");
    if (opts->lambda < 0 || (opts->penalty == CTBB_PENALTY_EDGE_PRESERVING &&
                             !(opts->delta > 0)))
        return fail(CTBB_E_ARG, "need lambda >= 0 and delta > 0");
    const ctbb_geom *g = &m->g;
    for (int64_t k = 0; k < g->nz; k++)
        if (slice_index[k] < 0 || slice_index[k] >= g->readings)
            return fail(CTBB_E_ARG, "slice_index out of range");

    ctbb_recon *r = new ctbb_recon();
    r->m = m;
    r->g = *g;
    r->opts = *opts;
    r->data_size = data_size_of(g);
    r->nvox = g->nx * g->ny * g->nz;
    r->h_slice_index.assign(slice_index, slice_index + g->nz);
    auto bail = [&](int rc) {
        ctbb_recon_free(r);
        return rc;
    };
    hipError_t he = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
    if (he != hipSuccess)
        return bail(fail(CTBB_E_HIP, std::string("stream: ") + hipGetErrorString(he)));

    // Columns inside the acquisition FOV (src/icd_iteration.cpp:154-181),
    // grouped into batches: the image is cut into tile x tile blocks of
    // columns and batch (ly, lx) holds the column at local position (lx, ly)
    // of every tile, so the columns of a batch are at least `tile` columns
    // apart (never penalty neighbours).  Batches run in raster order of
    // (ly, lx); with tile >= max(nx, ny) every batch is a single column and
    // the sweep is the reference's raster order.
    const int64_t tile = opts->tile > 0 ? opts->tile : 64;
    if (tile < 2)
        return bail(fail(CTBB_E_ARG, "tile must be >= 2"));
    const double fov_limit = (g->acquisition_fov / 2.0) * (g->acquisition_fov / 2.0);
    auto in_fov = [&](int64_t i, int64_t j) {
        const double y = (j - g->center_voxel_y) * g->voxel_size_y;
        const double x = (i - g->center_voxel_x) * g->voxel_size_x;
        return (x * x + y * y) < fov_limit;
    };
    std::vector<int32_t> cols;
    std::vector<int64_t> bptr(1, 0);
    const int64_t ntx = (g->nx + tile - 1) / tile, nty = (g->ny + tile - 1) / tile;
    for (int64_t ly = 0; ly < tile; ly++)
        for (int64_t lx = 0; lx < tile; lx++) {
            const size_t before = cols.size();
            for (int64_t ty = 0; ty < nty; ty++)
                for (int64_t tx = 0; tx < ntx; tx++) {
                    const int64_t i = tx * tile + lx, j = ty * tile + ly;
                    if (i < g->nx && j < g->ny && in_fov(i, j))
                        cols.push_back((int32_t)(i + g->nx * j));
                }
            if (cols.size() > before)
                bptr.push_back((int64_t)cols.size());
        }
    r->ncols_fov = (int64_t)cols.size();
    r->nbatch = (int64_t)bptr.size() - 1;
    r->tile = tile;
    int64_t max_batch = 0;
    for (int64_t b = 0; b < r->nbatch; b++)
        max_batch = std::max(max_batch, bptr[b + 1] - bptr[b]);
    r->max_batch = max_batch;
    r->h_batch_ptr = bptr;
    std::vector<int64_t> off(g->nz);
    for (int64_t k = 0; k < g->nz; k++)
        off[k] = slice_offset(g, slice_index[k]);

    int rc = CTBB_OK;
    if ((rc = dev_alloc_n(&r->y, r->data_size, "sinogram")) ||
        (rc = dev_alloc_n(&r->e, r->data_size, "error sinogram")) ||
        (rc = dev_alloc_n(&r->yhat, r->data_size, "FP64 projection")) ||
        (rc = dev_alloc_n(&r->x, r->nvox, "volume")) ||
        (rc = dev_alloc_n(&r->slice_off, g->nz, "slice offsets")) ||
        (rc = dev_alloc_n(&r->cols, std::max<int64_t>(1, r->ncols_fov), "FOV columns")) ||
        (rc = dev_alloc_n(&r->partials, 4096, "reduction partials")) ||
        (rc = dev_alloc_n(&r->batch_ptr, bptr.size(), "batch pointers")) ||
        (rc = dev_alloc_n(&r->du, (max_batch + 8) * g->nz, "step updates")))
        return bail(rc);
    if (weights && (rc = dev_alloc_n(&r->w, r->data_size, "weights")))
        return bail(rc);
#define CTBB_HIP_R(call)                                                      \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess)                                                 \
            return bail(fail(CTBB_E_HIP, std::string(#call) + ": " +          \
                                             hipGetErrorString(e_)));         \
    } while (0)
    CTBB_HIP_R(hipMemcpy(r->y, raw, r->data_size * 4, hipMemcpyHostToDevice));
    CTBB_HIP_R(hipMemcpy(r->e, raw, r->data_size * 4, hipMemcpyHostToDevice));
    if (weights)
        CTBB_HIP_R(hipMemcpy(r->w, weights, r->data_size * 4, hipMemcpyHostToDevice));
    CTBB_HIP_R(hipMemset(r->x, 0, r->nvox * 4));
    CTBB_HIP_R(hipMemcpy(r->slice_off, off.data(), g->nz * 8, hipMemcpyHostToDevice));
    if (r->ncols_fov)
        CTBB_HIP_R(hipMemcpy(r->cols, cols.data(), r->ncols_fov * 4,
                             hipMemcpyHostToDevice));
    CTBB_HIP_R(hipMemcpy(r->batch_ptr, bptr.data(), bptr.size() * 8,
                         hipMemcpyHostToDevice));
    {
        // exact in-range entry count of one sweep, and SSM entries of the
        // FOV columns (algorithmic traffic figures)
        unsigned long long *d_cnt = nullptr;
        if ((rc = dev_alloc_n(&d_cnt, 1, "counter")))
            return bail(rc);
        CTBB_HIP_R(hipMemset(d_cnt, 0, 8));
        hipLaunchKernelGGL(inrange_count_kernel, dim3(1024), dim3(256), 0, 0,
                           m->col_ptr, m->idx, r->cols, r->ncols_fov, r->slice_off,
                           g->nz, r->data_size, d_cnt);
        unsigned long long visits = 0;
        hipError_t e1 = hipGetLastError();
        if (e1 == hipSuccess)
            e1 = hipMemcpy(&visits, d_cnt, 8, hipMemcpyDeviceToHost);
        (void)hipFree(d_cnt);
        CTBB_HIP_R(e1);
        r->visits_per_sweep = (int64_t)visits;
        std::vector<int64_t> cp(m->ncol + 1);
        CTBB_HIP_R(hipMemcpy(cp.data(), m->col_ptr, cp.size() * 8, hipMemcpyDeviceToHost));
        r->ssm_fov_entries = 0;
        for (int32_t c : cols)
            r->ssm_fov_entries += cp[c + 1] - cp[c];
    }
#undef CTBB_HIP_R
    r->spacing = opts->group_spacing;
    if (r->spacing <= 0) {
        const double frac = (tile >= std::max(g->nx, g->ny)) ? 0.0 : 0.01;
        rc = ctbb_ssm_group_spacing(m, slice_index, g->nz, frac, &r->spacing);
        if (rc != CTBB_OK)
            return bail(rc);
    } else if (r->spacing < 2 && g->nz > 1) {
        return bail(fail(CTBB_E_ARG, "group_spacing must be >= 2"));
    }
    if (r->spacing < 1)
        r->spacing = 1;
    *out = r;
    return CTBB_OK;
}

extern "C" int ctbb_recon_set_volume(ctbb_recon *r, const float *volume)
{
    if (!r || !volume)
        return fail(CTBB_E_ARG, "null argument");
    CTBB_HIP(hipMemcpyAsync(r->x, volume, r->nvox * 4, hipMemcpyHostToDevice,
                            r->stream));
    double d;
    return recon_data_term(r, true, &d);
}

extern "C" int ctbb_recon_get_volume(const ctbb_recon *r, float *volume)
{
    if (!r || !volume)
        return fail(CTBB_E_ARG, "null argument");
    CTBB_HIP(hipMemcpyAsync(volume, r->x, r->nvox * 4, hipMemcpyDeviceToHost,
                            r->stream));
    CTBB_HIP(hipStreamSynchronize(r->stream));
    return CTBB_OK;
}

extern "C" int ctbb_recon_get_error(const ctbb_recon *r, float *e)
{
    if (!r || !e)
        return fail(CTBB_E_ARG, "null argument");
    CTBB_HIP(hipMemcpyAsync(e, r->e, r->data_size * 4, hipMemcpyDeviceToHost,
                            r->stream));
    CTBB_HIP(hipStreamSynchronize(r->stream));
    return CTBB_OK;
}

extern "C" int ctbb_recon_resync(ctbb_recon *r)
{
    if (!r)
        return fail(CTBB_E_ARG, "null argument");
    double d;
    return recon_data_term(r, true, &d);
}

extern "C" int ctbb_recon_cost(ctbb_recon *r, double *cost)
{
    if (!r || !cost)
        return fail(CTBB_E_ARG, "null argument");
    double data = 0.0, pen = 0.0;
    CTBB_TRY(recon_data_term(r, false, &data));
    CTBB_TRY(recon_penalty_term(r, &pen));
    if (r->opts.penalty == CTBB_PENALTY_QUADRATIC)
        *cost = 0.5 * data + 0.5 * r->opts.lambda * pen;
    else
        *cost = data + r->opts.lambda * pen;
    return CTBB_OK;
}

extern "C" int ctbb_recon_iterate(ctbb_recon *r, int n, ctbb_icd_stats *stats)
{
    if (!r || n < 0)
        return fail(CTBB_E_ARG, "bad argument");
    const ctbb_ssm *m = r->m;
    SweepArgs a{};
    a.col_ptr = m->col_ptr;
    a.idx = m->idx;
    a.val = m->val;
    a.val16 = m->val16;
    a.col_scale = m->col_scale;
    a.e = r->e;
    a.w = r->w;
    a.x = r->x;
    a.slice_off = r->slice_off;
    a.cols = r->cols;
    a.spacing = r->spacing;
    a.nz = r->g.nz;
    a.nx = r->g.nx;
    a.nxy = r->g.nx * r->g.ny;
    a.nvox = r->nvox;
    a.data_size = r->data_size;
    a.penalty = r->opts.penalty;
    a.lambda = r->opts.lambda;
    a.delta = r->opts.delta;

    a.du = r->du;
    a.batch_ptr = r->batch_ptr;
    const bool fp16 = m->precision == CTBB_SSM_FP16;
    const bool weighted = r->w != nullptr;
    auto gather = fp16 ? (weighted ? icd_gather_kernel<true, true>
                                   : icd_gather_kernel<true, false>)
                       : (weighted ? icd_gather_kernel<false, true>
                                   : icd_gather_kernel<false, false>);
    auto scatter = fp16 ? icd_scatter_kernel<true> : icd_scatter_kernel<false>;

    hipEvent_t e0, e1;
    CTBB_HIP(hipEventCreate(&e0));
    CTBB_HIP(hipEventCreate(&e1));
    double sweep_ms = 0.0, resync_ms = 0.0;
    int64_t launches = 0;
    // Per-kernel durations: events around the gather and the scatter launch
    // of every `every`-th step (at most kSamples steps per iteration).
    constexpr int kSamples = 128;
    const int64_t steps_per_it = r->nbatch * std::min<int64_t>(r->spacing, r->g.nz);
    const int64_t every = std::max<int64_t>(1, steps_per_it / kSamples);
    std::vector<hipEvent_t> ev(3 * kSamples);
    for (auto &x : ev)
        CTBB_HIP(hipEventCreate(&x));
    double g_ms = 0.0, s_ms = 0.0;
    int64_t n_samp = 0;
    for (int it = 0; it < n; it++) {
        CTBB_HIP(hipEventRecord(e0, r->stream));
        int64_t step = 0, used = 0;
        for (int64_t b = 0; b < r->nbatch; b++) {
            const int64_t c0 = r->h_batch_ptr[b], nc = r->h_batch_ptr[b + 1] - c0;
            for (int64_t j = 0; j < r->spacing && j < r->g.nz; j++) {
                StepArgs sa;
                sa.c0 = c0;
                sa.nc = nc;
                sa.j = j;
                sa.ng = (r->g.nz - j + r->spacing - 1) / r->spacing;
                const int64_t ntask = nc < kXcd
                                          ? nc * sa.ng
                                          : ((nc + kXcd - 1) / kXcd) * kXcd * sa.ng;
                const bool samp = (step % every) == 0 && used < kSamples;
                if (samp)
                    CTBB_HIP(hipEventRecord(ev[3 * used], r->stream));
                hipLaunchKernelGGL(gather, dim3((unsigned)ntask), dim3(kSweepBlock),
                                   0, r->stream, a, sa);
                if (samp)
                    CTBB_HIP(hipEventRecord(ev[3 * used + 1], r->stream));
                hipLaunchKernelGGL(scatter, dim3((unsigned)ntask), dim3(kSweepBlock),
                                   0, r->stream, a, sa);
                if (samp) {
                    CTBB_HIP(hipEventRecord(ev[3 * used + 2], r->stream));
                    used++;
                }
                launches += 2;
                step++;
            }
        }
        CTBB_HIP(hipGetLastError());
        CTBB_HIP(hipEventRecord(e1, r->stream));
        CTBB_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        CTBB_HIP(hipEventElapsedTime(&ms, e0, e1));
        sweep_ms += ms;
        for (int64_t i = 0; i < used; i++) {
            float ga = 0.0f, sc = 0.0f;
            CTBB_HIP(hipEventElapsedTime(&ga, ev[3 * i], ev[3 * i + 1]));
            CTBB_HIP(hipEventElapsedTime(&sc, ev[3 * i + 1], ev[3 * i + 2]));
            g_ms += ga;
            s_ms += sc;
        }
        n_samp += used;
        r->iter_done++;
        if (r->opts.resync_every > 0 && (r->iter_done % r->opts.resync_every) == 0) {
            CTBB_HIP(hipEventRecord(e0, r->stream));
            CTBB_TRY(ctbb_recon_resync(r));
            CTBB_HIP(hipEventRecord(e1, r->stream));
            CTBB_HIP(hipEventSynchronize(e1));
            CTBB_HIP(hipEventElapsedTime(&ms, e0, e1));
            resync_ms += ms;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (auto &x : ev)
        (void)hipEventDestroy(x);
    if (stats) {
        stats->gather_ms_avg = n_samp ? g_ms / n_samp : 0.0;
        stats->scatter_ms_avg = n_samp ? s_ms / n_samp : 0.0;
        stats->iterations = n;
        stats->voxel_updates = (int64_t)n * r->ncols_fov * r->g.nz;
        stats->entry_visits = (int64_t)n * r->visits_per_sweep;
        stats->sweep_ms = sweep_ms;
        stats->resync_ms = resync_ms;
        stats->cost = NAN;
        stats->sweep_launches = launches;
        stats->ssm_entries = r->ssm_fov_entries;
    }
    return CTBB_OK;
}
