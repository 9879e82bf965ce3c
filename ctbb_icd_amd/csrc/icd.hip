// icd.hip -- iterative coordinate descent (ICD) on gfx950.
//
// Replaces icd_iteration (src/icd_iteration.cpp:42-286), with the penalty
// steps of src/penalties.cpp:39-154 evaluated per voxel in FP64.
//
// State (device resident, reference sinogram layout
// r = row + nr*(channel + nch*view)):
//   e   FP32 error sinogram y - A x (the reference keeps yhat = A x in FP64
//       and forms raw - yhat on the fly, src/icd_iteration.cpp:209); in
//       weighted sweeps kept current only where w == 0 and formed as w e / w
//       elsewhere when it is read (running cost, download)
//   ew  FP32 weighted error w e (weighted reconstructions; e itself when
//       unweighted), the only sinogram the gather reads
//   x   FP32 volume in the rotated frame, index c + nx*ny*k
//   w   optional FP32 statistical weights (w == 1 reproduces the reference)
//
// Sweep order: the FOV columns (src/icd_iteration.cpp:157-181) are cut into
// batches -- the column at one local position of every tile x tile block of
// the image -- run in raster order of that position; in each batch the
// slices are visited as groups {j, j+s, j+2s, ...}, j = 0..s-1, where the
// spacing s keeps the sinogram footprints of one column's group members
// apart (and no two members are penalty neighbours).  A step is one (batch,
// group): phase A1 gathers theta1 for every voxel of the step from the error
// sinogram as it stood at the start of the step, phase A2 solves the
// penalty step per voxel and writes x and du, phase B scatters
// e -= du a, sample-driven and conflict-free.  Across the batch's columns the
// step is Jacobi, along a column Gauss-Seidel in group order; with tile >=
// max(nx, ny) batches are single columns and the order is the reference's.
// By default the curvature of a voxel's update is the batch surrogate
// sum_r w_r a_r S_r (S_r: the summed matrix values, on data sample r, of
// every voxel of the step -- all the batch's columns, every slice of the
// group), which majorises the step's data term, so every step lowers the
// cost; single-column batches keep the reference's sum_i w_i a_i^2.  The oracle replays exactly
// this order and curvature.
//
// FP64 is used for the per-voxel sums of partials and the penalty solve, the
// forward projection of the periodic resync e = y - A x, and the cost.
#include "common.hpp"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>


using namespace ctbb;

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
    // f(a) is carried instead of re-evaluated: a only ever takes the value
    // of c, whose f was just computed with the same arguments
    double f_a = ep_func(a, alpha, beta, delta, lambda, diff, used);
    for (int it = 0; it < 4096; it++) {
        c = (a + b) / 2.0;
        const double f_c = ep_func(c, alpha, beta, delta, lambda, diff, used);
        if ((f_c == 0) || (b - a) / 2 < tol)
            return c;
        const int sc = f_c < 0 ? -1 : 1, sa = f_a < 0 ? -1 : 1;
        if (sc == sa) {
            a = c;
            f_a = f_c;
        } else {
            b = c;
        }
    }
    return c;
}

// 26-neighbour penalty (no reference counterpart: the reference stencil is
// 8 in-plane neighbours, src/penalties.cpp:39-59, "limited to 2D",
// src/icd_iteration.cpp:140).  Voxel p of rotating slice k sits at
// R(theta_k) p, so the point of slice k+1 above it is
// R(theta_k - theta_{k+1}) p; that rotation is applied as three shears with
// integer-rounded shifts (an exact bijection of the grid with an exact
// inverse).  Neighbours: the 8 in-plane voxels (no wrap), the 3x3 block of
// slice k+1 around pi_k(p) and the voxels of slice k-1 whose blocks hold p
// (pi_{k-1}^-1(p - d)); weights 1/|offset| in voxel units.  26 fixed slots,
// in-plane (dy, dx) order, then above, then below, unused slots masked.
struct Nb26 {
    const double *shear;      // [2*(nz-1)]: a_k = -tan(phi_k/2), b_k = sin(phi_k);
                              // null: the reference's 8-neighbour stencil
    int64_t nx, ny, nz;
    double cx, cy;
};

__device__ __forceinline__ double rnd_half(double v) { return floor(v + 0.5); }

__device__ __forceinline__ void rot26_fwd(const Nb26 &n, int64_t k, int64_t i, int64_t j,
                                          int64_t &io, int64_t &jo)
{
    const double a = n.shear[2 * k], b = n.shear[2 * k + 1];
    const double x = (double)i - n.cx, y = (double)j - n.cy;
    const double x1 = x + rnd_half(a * y);
    const double y2 = y + rnd_half(b * x1);
    const double x3 = x1 + rnd_half(a * y2);
    io = (int64_t)(x3 + n.cx);
    jo = (int64_t)(y2 + n.cy);
}

__device__ __forceinline__ void rot26_inv(const Nb26 &n, int64_t k, int64_t i, int64_t j,
                                          int64_t &io, int64_t &jo)
{
    const double a = n.shear[2 * k], b = n.shear[2 * k + 1];
    const double x3 = (double)i - n.cx, y2 = (double)j - n.cy;
    const double x1 = x3 - rnd_half(a * y2);
    const double y1 = y2 - rnd_half(b * x1);
    const double x = x1 - rnd_half(a * y1);
    io = (int64_t)(x + n.cx);
    jo = (int64_t)(y1 + n.cy);
}

__device__ __forceinline__ double w26(int s)
{
    // slots 0..7 in-plane ((dy, dx) order without the centre), 8..16 above,
    // 17..25 below ((dy, dx) order)
    if (s < 8)
        return (s == 0 || s == 2 || s == 5 || s == 7) ? 1.0 / kSqrt2 : 1.0;
    const int t = (s - 8) % 9, dx = t % 3 - 1, dy = t / 3 - 1;
    return 1.0 / sqrt((double)(dx * dx + dy * dy + 1));
}

// Differences x_n - x_q of the 26 slots and the mask of the slots present.
template <typename X>
__device__ __forceinline__ uint32_t diffs26(const Nb26 &n, int64_t q, const X *x,
                                            double (&diff)[26])
{
    const int64_t nxy = n.nx * n.ny, k = q / nxy, c = q - k * nxy;
    const int64_t j = c / n.nx, i = c - j * n.nx;
    const double xq = (double)x[q];
    uint32_t used = 0;
    int s = 0;
#pragma unroll
    for (int dy = -1; dy <= 1; dy++)
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
            if (!dx && !dy)
                continue;
            const int64_t ii = i + dx, jj = j + dy;
            const bool in = ii >= 0 && ii < n.nx && jj >= 0 && jj < n.ny;
            diff[s] = in ? (double)x[ii + n.nx * jj + nxy * k] - xq : 0.0;
            used |= in ? (1u << s) : 0u;
            s++;
        }
    int64_t iu = 0, ju = 0;
    const bool up = k + 1 < n.nz;
    if (up)
        rot26_fwd(n, k, i, j, iu, ju);
#pragma unroll
    for (int dy = -1; dy <= 1; dy++)
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
            const int64_t ii = iu + dx, jj = ju + dy;
            const bool in = up && ii >= 0 && ii < n.nx && jj >= 0 && jj < n.ny;
            diff[s] = in ? (double)x[ii + n.nx * jj + nxy * (k + 1)] - xq : 0.0;
            used |= in ? (1u << s) : 0u;
            s++;
        }
#pragma unroll
    for (int dy = -1; dy <= 1; dy++)
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
            int64_t ii = -1, jj = -1;
            if (k >= 1)
                rot26_inv(n, k - 1, i - dx, j - dy, ii, jj);
            const bool in = k >= 1 && ii >= 0 && ii < n.nx && jj >= 0 && jj < n.ny;
            diff[s] = in ? (double)x[ii + n.nx * jj + nxy * (k - 1)] - xq : 0.0;
            used |= in ? (1u << s) : 0u;
            s++;
        }
    return used;
}

// Quadratic: the coordinate minimiser over the neighbours present.
template <typename X>
__device__ double quadratic26_step(const Nb26 &n, int64_t q, double alpha, double beta,
                                   double lambda, const X *x)
{
    double diff[26];
    const uint32_t used = diffs26(n, q, x, diff);
    double sum = 0.0, W = 0.0;
#pragma unroll
    for (int s = 0; s < 26; s++)
        if (used & (1u << s)) {
            sum += w26(s) * diff[s];
            W += w26(s);
        }
    return (beta + lambda * sum) / (alpha + lambda * W);
}

__device__ __forceinline__ double ep26_func(double c, double alpha, double beta,
                                            double delta, double lambda,
                                            const double (&diff)[26], uint32_t used)
{
    double sum = 0.0;
#pragma unroll
    for (int s = 0; s < 26; s++)
        if (used & (1u << s)) {
            const double u = c - diff[s];
            sum += (u >= 0 ? 1.0 : -1.0) * w26(s) * (1.0 - 1.0 / (1.0 + fabs(u / delta)));
        }
    return 2 * alpha * c - 2 * beta + lambda * sum;
}

// Edge-preserving: the bisection of src/penalties.cpp:88-139 over the
// 26-neighbourhood (f(a) carried, as in edge_preserving_step).
template <typename X>
__device__ double ep26_step(const Nb26 &n, int64_t q, double alpha, double beta,
                            double lambda, double delta, const X *x)
{
    if (alpha == 0.0)
        return 0.0;
    double diff[26];
    const uint32_t used = diffs26(n, q, x, diff);
    double a = beta / alpha, b = beta / alpha;
#pragma unroll
    for (int s = 0; s < 26; s++)
        if (used & (1u << s)) {
            if (diff[s] < a) a = diff[s];
            if (diff[s] > b) b = diff[s];
        }
    const double tol = 0.00000001;
    double c = (a + b) / 2.0;
    double f_a = ep26_func(a, alpha, beta, delta, lambda, diff, used);
    for (int it = 0; it < 4096; it++) {
        c = (a + b) / 2.0;
        const double f_c = ep26_func(c, alpha, beta, delta, lambda, diff, used);
        if ((f_c == 0) || (b - a) / 2 < tol)
            return c;
        const int sc = f_c < 0 ? -1 : 1, sa = f_a < 0 ? -1 : 1;
        if (sc == sa) {
            a = c;
            f_a = f_c;
        } else {
            b = c;
        }
    }
    return c;
}

template <typename X>
__device__ __forceinline__ double penalty_step(int penalty, int64_t q,
                                               double alpha, double beta,
                                               double lambda, double delta,
                                               const X *x, int64_t nx,
                                               int64_t n, const Nb26 &nb)
{
    if (nb.shear)
        return penalty == CTBB_PENALTY_QUADRATIC
                   ? quadratic26_step(nb, q, alpha, beta, lambda, x)
                   : ep26_step(nb, q, alpha, beta, lambda, delta, x);
    if (penalty == CTBB_PENALTY_QUADRATIC)
        return quadratic_step(q, alpha, beta, lambda, x, nx, n);
    return edge_preserving_step(q, alpha, beta, lambda, delta, x, nx, n);
}

// One thread per voxel; X = double (the FP64 volume as given) or float (the
// volume as the sweep kernel holds it: the sweep calls penalty_step<float>).
template <typename X>
__global__ void penalty_steps_kernel(int penalty, double lambda, double delta,
                                     const X *x, int64_t nx, int64_t nvox,
                                     int64_t n, const int64_t *q,
                                     const double *alpha, const double *beta,
                                     double *du, Nb26 nb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        du[i] = penalty_step(penalty, q[i], alpha[i], beta[i], lambda, delta,
                             x, nx, nvox, nb);
}

// ---------------------------------------------------------------------------
// Sweep kernels (gather, solve, scatter): see the phase comments below.

constexpr int kSweepBlock = 256;
constexpr int kWaves = kSweepBlock / 64;

// The sweep state (error sinograms, volume, updates, partials, surrogate
// sums) is FP32 or FP64; the kernels take its type T as a template argument
// and the argument blocks carry the state untyped.
struct SweepArgs {
    const int64_t *col_ptr;
    const int32_t *idx;
    const float *val;
    const __half *val16;
    const float *col_scale;
    void *e;                  // T: error sinogram y - A x
    void *ew;                 // T: weighted error w (y - A x) (== e unweighted)
    const float *w;           // weights or null
    void *x;                  // T: volume
    void *du;                 // T [ng*nc]: updates of the step
    double *th2;              // [nvox] theta2 per voxel (cached)
    const int64_t *slice_off;
    Nb26 nb;                  // penalty neighbourhood (nb.shear null: 8)
    const int32_t *cols;      // batch columns, concatenated
    int64_t spacing;
    int64_t nz, nx, nxy, nvox, data_size;
    int penalty;
    double lambda, delta;
};

// Inclusive sum over runs of equal keys across the lanes of a wavefront:
// bit l of `head` is set where lane l starts a run (lane 0 always does), and
// runs are contiguous because keys arrive sorted.  After the call, the last
// lane of each run holds the sum of its run; the order of additions is fixed
// (Hillis-Steele), so the result is deterministic.
constexpr int kXcd = 8;   // workgroups are dealt round-robin to the 8 XCDs
// Steps per iteration whose three launches are timed with HIP events.
constexpr int kTimedSteps = 128;

struct StepArgs {
    int64_t c0, nc, j, ng;
};

// ---------------------------------------------------------------------------
// Phase A1 of a step: view-blocked gather.
//
// The SSM views are cut into blocks of kViewBlock views.  For every batch
// and block, the matrix entries of the batch's columns are copied with the
// columns ranked by where their footprint starts on the detector in that
// block (channel, then row), in chunks of 64 ranks; inside a chunk the
// entries are ordered (view in block, rank).  A workgroup takes one view
// block and 4 chunks (one per wavefront) and loops over the members of the
// slice group, kGatherMembers of them per pass over the chunk's entries
// (each entry is loaded once per pass and gathered at each member's sample);
// neighbouring lanes read neighbouring samples of the error sinogram, so a
// 64-lane gather touches few cache lines.  Lanes form a * (w e) at sample
// idx + slice offset (and, on the first sweep, w a S, the surrogate
// curvature); runs of equal entry keys (view in block, rank) are summed
// with DPP lane moves and the run's last
// lane adds into the wavefront's LDS accumulator of that key (one wavefront
// per chunk, in entry order: deterministic).  Per member, each lane adds its
// rank's kViewBlock accumulators in view order and writes the partial of
// (member, block, slot); phase A2 adds the partials over the blocks in a
// fixed order.  Workgroups are dispatched by view block, the blocks in
// residue classes of P = the view distance of consecutive group members in
// blocks (ib, ib + P, ib + 2P, ...): the chunks of one block share the
// error-sinogram lines of each member's band, and member m + 1 of block ib
// reads the same data views as member m of block ib + P, so those lines are
// fetched into L2 once per class instead of once per member (logical
// workgroups go to the XCDs in contiguous ranges).  FP16 matrices: values are summed unscaled and the per-column
// scale is applied in phase A2 (it factors out of both sums).
#ifndef CTBB_VIEW_BLOCK
#define CTBB_VIEW_BLOCK 4
#endif
#ifndef CTBB_GATHER_UNROLL
#define CTBB_GATHER_UNROLL 4
#endif
#ifndef CTBB_GATHER_MEMBERS
#define CTBB_GATHER_MEMBERS 4
#endif
constexpr int kViewBlock = CTBB_VIEW_BLOCK;
constexpr int kGatherUnroll = CTBB_GATHER_UNROLL;
#ifndef CTBB_GATHER_MEMBERS_FIRST
#define CTBB_GATHER_MEMBERS_FIRST 4
#endif
// Members of the slice group a wavefront takes per pass over its entries in
// the later sweeps and in the first sweep (which also sums theta2)
constexpr int kGatherMembers = CTBB_GATHER_MEMBERS;
constexpr int kGatherMembersFirst = CTBB_GATHER_MEMBERS_FIRST;
static_assert(kGatherMembers >= 1 && kGatherMembersFirst >= 1, "members per pass");

// Packed view-blocked entry: the sample in its view in the low kSampleBits
// bits, the key (view in block << 6 | rank % 64) above them.
constexpr uint32_t kSampleBits = 17, kSampleMask = (1u << kSampleBits) - 1;

struct GatherArgs {
    const uint32_t *ent;      // view-blocked copy (see above): sample in view
                              //   | (view in block << 6 | rank % 64) << kSampleBits
    const float *val;
    const __half *val16;
    int64_t plane;            // samples per view (n_rows * n_channels)
    const uint16_t *perm;     // column slot of each (b, ib, rank)
    const int64_t *ptr;       // starts of the (b, ib, chunk) entry ranges
    int64_t base;             // nib * (first FOV index of the batch)
    int64_t wbase;            // first chunk of the batch
    int64_t nw;               // 64-rank chunks of the batch
    void *part1, *part2;      // T [ng][nib][nc]
    int64_t nib;
    const int64_t *si_half;   // [nz] central view - n_sm/2
    const void *ssum;         // T: batch surrogate, sum of the step's values
                              //   per data sample (null: sum w a^2)
    int64_t nsc;              // workgroups per view block
    int64_t pchain, rows;     // dispatch: view blocks in residue classes of
                              // pchain (rows blocks per class)
    int64_t nsm, readings;
};

// View block ib of slice half-offset h holds data views [kViewBlock ib + h,
// min(kViewBlock (ib + 1), n_sm) + h); true if any of them is inside
// [0, readings).
__host__ __device__ inline bool block_valid(int64_t ib, int64_t h, int64_t nsm,
                                            int64_t readings)
{
    const int64_t v0 = ib * kViewBlock;
    const int64_t v1 = v0 + kViewBlock < nsm ? v0 + kViewBlock : nsm;
    return v0 + h < readings && v1 + h > 0;
}

template <bool FP16>
__device__ __forceinline__ float gb_value(const GatherArgs &g, int64_t p)
{
    return FP16 ? __half2float(g.val16[p]) : g.val[p];
}

template <typename T, bool FP16, bool THETA2>
__global__ __launch_bounds__(kSweepBlock) void icd_gather_kernel(SweepArgs a,
                                                                 StepArgs sa,
                                                                 GatherArgs g)
{
    constexpr int kAcc = 64 * kViewBlock;
    constexpr int M = THETA2 ? kGatherMembersFirst : kGatherMembers;
    __shared__ T accm[THETA2 ? 2 * M : M][kWaves * kAcc];
    // wave index pinned to a scalar register: the per-wave chunk bounds and
    // bases below are then scalar too
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t nwg = g.pchain * g.rows * g.nsc;
    const int64_t per_xcd = (nwg + kXcd - 1) / kXcd;
    const int64_t L = (int64_t)(blockIdx.x % kXcd) * per_xcd + blockIdx.x / kXcd;
    if (L >= nwg)
        return;
    const int64_t lb = L / g.nsc;
    const int64_t cls = lb / g.rows;
    const int64_t ib = cls + (lb - cls * g.rows) * g.pchain;
    if (ib >= g.nib)
        return;
    const int64_t w = (L - lb * g.nsc) * kWaves + wave;
    if (w >= g.nw)
        return;
    const int64_t *ptr = g.ptr + g.wbase + ib * g.nw;
    const int64_t lo = ptr[w], hi = ptr[w + 1];
    const int64_t rk = w * 64 + lane;
    const bool own = rk < sa.nc;
    const int64_t slot = own ? (int64_t)g.perm[g.base + ib * sa.nc + rk] : 0;
    const T *__restrict__ ew = static_cast<const T *>(a.ew);
    const T *__restrict__ ssum = static_cast<const T *>(g.ssum);
    const float *__restrict__ wt = a.w;
    const uint32_t ds = (uint32_t)a.data_size;
    // M members of the slice group per pass over the chunk's entries, each
    // with its own LDS accumulators (two on the first sweep, which also sums
    // theta2), so every entry is loaded once for all of them and M times as
    // many gathers are in flight
    int64_t v = 0;
    while (true) {
        int64_t vm[M];
        bool has[M];
        int32_t off[M];
#pragma unroll
        for (int m = 0; m < M; m++) {
            while (v < sa.ng &&
                   !block_valid(ib, g.si_half[sa.j + a.spacing * v], g.nsm, g.readings))
                v++;
            has[m] = v < sa.ng;
            vm[m] = v;
            off[m] = has[m] ? (int32_t)a.slice_off[sa.j + a.spacing * v] : 0;
            if (has[m])
                v++;
        }
        if (!has[0])
            break;
#pragma unroll
        for (int m = 0; m < (THETA2 ? 2 * M : M); m++)
#pragma unroll
            for (int q = 0; q < kViewBlock; q++)
                accm[m][wave * kAcc + q * 64 + lane] = T(0);
        // the chunk's entries from scalar bases with 32-bit lane offsets,
        // clamped into the chunk (a chunk is non-empty inside the loop)
        const uint32_t *ent_c = g.ent + lo;
        const float *val_c = FP16 ? nullptr : g.val + lo;
        const __half *val16_c = FP16 ? g.val16 + lo : nullptr;
        const uint32_t n = (uint32_t)(hi - lo);
        // ray index = sample in view + (first view of the block + view in
        // block) * plane; below 2^31 (checked at set-up)
        const int32_t ib_base = (int32_t)(ib * kViewBlock * g.plane);
        const int32_t plane = (int32_t)g.plane;
        for (uint32_t p0 = 0; p0 < n; p0 += 64 * kGatherUnroll) {
            int32_t iv[kGatherUnroll], kv[kGatherUnroll];
            float av[kGatherUnroll];
#pragma unroll
            for (int u = 0; u < kGatherUnroll; u++) {
                const uint32_t p = p0 + u * 64 + lane;
                const bool ok = p < n;
                const uint32_t pc = ok ? p : 0u;
                const uint32_t e0 = ent_c[pc];
                const float a0 = FP16 ? __half2float(val16_c[pc]) : val_c[pc];
                const int32_t k0 = (int32_t)(e0 >> kSampleBits);
                const int32_t i0 = (int32_t)(e0 & kSampleMask) + ib_base + (k0 >> 6) * plane;
                iv[u] = ok ? i0 : 0;
                av[u] = ok ? a0 : 0.0f;
                kv[u] = ok ? k0 : -1;
            }
            T ev[M][kGatherUnroll], sv[M][kGatherUnroll];
            float mv[M][kGatherUnroll], wv[M][kGatherUnroll];
#pragma unroll
            for (int m = 0; m < M; m++)
#pragma unroll
                for (int u = 0; u < kGatherUnroll; u++) {
                    // |ray index| + |slice offset| < 2^31 (checked at set-up)
                    const int32_t r = iv[u] + off[m];
                    const bool in = has[m] && kv[u] >= 0 && (uint32_t)r < ds;
                    const uint32_t rc = in ? (uint32_t)r : 0u;
                    mv[m][u] = in ? av[u] : 0.0f;
                    ev[m][u] = ew[rc];
                    // the surrogate sums come weighted (w S); a^2 needs w
                    wv[m][u] = (THETA2 && wt && !ssum) ? wt[rc] : 1.0f;
                    // curvature factor: the step's summed values on this
                    // sample (surrogate), or a itself (a^2 squared in FP32,
                    // as src/icd_iteration.cpp:208)
                    if (THETA2)
                        sv[m][u] = ssum ? (T)mv[m][u] * ssum[rc]
                                        : (T)(mv[m][u] * av[u]);
                }
#pragma unroll
            for (int u = 0; u < kGatherUnroll; u++) {
                if (p0 + u * 64 >= n)
                    break;
                const uint64_t head = run_heads(kv[u], lane);
                T t1[M], t2[M];
#pragma unroll
                for (int m = 0; m < M; m++) {
                    t1[m] = run_sum((T)mv[m][u] * ev[m][u], head, lane);
                    if (THETA2)
                        t2[m] = run_sum((T)wv[m][u] * sv[m][u], head, lane);
                }
                if (kv[u] >= 0 && run_end(head, lane)) {
#pragma unroll
                    for (int m = 0; m < M; m++) {
                        accm[m][wave * kAcc + kv[u]] += t1[m];
                        if (THETA2)
                            accm[M + m][wave * kAcc + kv[u]] += t2[m];
                    }
                }
            }
        }
#pragma unroll
        for (int m = 0; m < M; m++) {
            T s1 = T(0), s2 = T(0);
#pragma unroll
            for (int q = 0; q < kViewBlock; q++) {
                s1 += accm[m][wave * kAcc + q * 64 + lane];
                if (THETA2)
                    s2 += accm[M + m][wave * kAcc + q * 64 + lane];
            }
            if (own && has[m]) {
                const int64_t o = (vm[m] * g.nib + ib) * sa.nc + slot;
                static_cast<T *>(g.part1)[o] = s1;
                if (THETA2)
                    static_cast<T *>(g.part2)[o] = s2;
            }
        }
    }
}

// Phase A2 of a step: per voxel, theta1 (and, on the first sweep, theta2)
// is the FP64 sum of its partials over the view blocks (wavefront w takes
// blocks w, w+4, ...; the four sums are added in wavefront order).  Then one
// thread per voxel takes the penalty step with the reference's solver in
// FP64, clamps to x >= 0 and writes x and the update du.  A workgroup covers
// 64 slots of one group member.
template <typename T, bool FP16, bool THETA2>
__global__ __launch_bounds__(kSweepBlock) void icd_solve_kernel(SweepArgs a,
                                                                StepArgs sa,
                                                                GatherArgs g)
{
    const T *part1 = static_cast<const T *>(g.part1);
    const T *part2 = static_cast<const T *>(g.part2);
    T *x = static_cast<T *>(a.x);
    __shared__ double s1[kWaves][64], s2[kWaves][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t nch64 = (sa.nc + 63) / 64;
    const int64_t v = blockIdx.x / nch64;
    const int64_t s = (blockIdx.x - v * nch64) * 64 + lane;
    double t1 = 0.0, t2 = 0.0;
    const int64_t k = sa.j + a.spacing * v;
    // valid view blocks of this slice: a contiguous range [ib0, ib1)
    const int64_t h = g.si_half[k];
    int64_t ib0 = 0, ib1 = g.nib;
    while (ib0 < ib1 && !block_valid(ib0, h, g.nsm, g.readings))
        ib0++;
    while (ib1 > ib0 && !block_valid(ib1 - 1, h, g.nsm, g.readings))
        ib1--;
    if (s < sa.nc) {
        const int64_t o = v * g.nib * sa.nc + s;
        int64_t ib = ib0 + wave;
        for (; ib + 3 * kWaves < ib1; ib += 4 * kWaves) {
            T q1[4], q2[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                q1[u] = part1[o + (ib + u * kWaves) * sa.nc];
                q2[u] = THETA2 ? part2[o + (ib + u * kWaves) * sa.nc] : T(0);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                t1 += (double)q1[u];
                t2 += (double)q2[u];
            }
        }
        for (; ib < ib1; ib += kWaves) {
            t1 += (double)part1[o + ib * sa.nc];
            if (THETA2)
                t2 += (double)part2[o + ib * sa.nc];
        }
    }
    s1[wave][lane] = t1;
    s2[wave][lane] = t2;
    __syncthreads();
    if (wave != 0 || s >= sa.nc)
        return;
    double th1 = 0.0, th2 = 0.0;
    for (int i = 0; i < kWaves; i++) {
        th1 += s1[i][lane];
        th2 += s2[i][lane];
    }
    const int64_t col = a.cols[sa.c0 + s];
    const int64_t q = col + a.nxy * k;
    if (FP16) {
        const double sc = (double)a.col_scale[col];
        th1 *= sc;
        // sum w a^2 carries the scale twice; the surrogate's sample sums
        // are formed from scaled values, so sum w a S carries it once
        th2 *= g.ssum ? sc : sc * sc;
    }
    if (THETA2)
        a.th2[q] = th2;
    else
        th2 = a.th2[q];
    double du = penalty_step(a.penalty, q, th2, th1, a.lambda, a.delta, x, a.nx,
                             a.nvox, a.nb);
    const double xq = (double)x[q];
    if (du + xq < 0)
        du = -xq;
    x[q] = (T)(xq + du);
    static_cast<T *>(a.du)[v * sa.nc + s] = (T)du;
}

// Phase B of a step: sample-driven scatter.  Each workgroup owns the
// contiguous samples of kScatterViews consecutive data views and kScatterCh
// channels (all rows) and adds up, in LDS, the updates of every voxel of the
// step that reaches them: for a data view and a slice k of the group, the
// SSM view i = view - (c_k - n_sm/2) and the batch's transposed entries of
// that view (sorted by channel, row, then column slot; every run of one
// sample padded to a multiple of four entries) give the rays and their
// column slots.  Wavefront w owns the channels [c0 + 8w, c0 + 8w + 8) of
// every view of the workgroup: one (view, slice) item's entries for them are
// contiguous and are taken in chunks of 64 quads (each lane four entries of
// one sample, two 16-byte loads).  Lanes form sum du[slot] * a over their
// quad, runs of equal (channel, row) across lanes are summed with DPP lane
// moves and the last lane of a run adds it into LDS.  The loads of
// kItemUnroll chunks are issued before any of them is reduced, and chunks
// are applied in a fixed order by the one wavefront that owns the channels:
// no LDS atomics, deterministic sums.  The workgroup then writes each
// changed sample once; no two workgroups touch the same sample, so the
// scatter is conflict-free.  The du reads are gathers over the batch's
// slots; slots follow the Z-order curve of the tile grid, so the columns
// along a ray, whose entries make up a run, mostly share cache lines.
// Tile and unroll sizes (overridable at build time for tuning sweeps).
#ifndef CTBB_SCATTER_CH
#define CTBB_SCATTER_CH 32
#endif
#ifndef CTBB_SCATTER_VIEWS
#define CTBB_SCATTER_VIEWS 2
#endif
#ifndef CTBB_ITEM_UNROLL
#define CTBB_ITEM_UNROLL 4
#endif
constexpr int kScatterCh = CTBB_SCATTER_CH;
constexpr int kScatterViews = CTBB_SCATTER_VIEWS;
constexpr int kWaveCh = kScatterCh / kWaves;
constexpr int kItemUnroll = CTBB_ITEM_UNROLL;
// Transposed entry: slot | lkey << 14 | (ch / kScatterCh) << 26, where
// lkey = (ch % kScatterCh) * nr + row is the sample's place in the scatter
// tile of its channel group (needs kScatterCh * nr <= 4096).
constexpr uint32_t kSlotBits = 14, kLkeyBits = 12;
// Runs of one sample in the transposed copy are padded to multiples of kQuad
// entries (copies of the run's last entry with value 0), so each lane of the
// scatter takes four entries of one sample with two 16-byte loads.
constexpr int kQuad = 4;

__device__ __forceinline__ uint32_t bt_slot(uint32_t p) { return p & ((1u << kSlotBits) - 1); }
__device__ __forceinline__ uint32_t bt_lkey(uint32_t p) { return (p >> kSlotBits) & ((1u << kLkeyBits) - 1); }

struct BtArgs {
    const uint32_t *ent;     // packed slot | lkey << 14 | channel group << 26
    const float *val;        // entry values (FP32)
    const int32_t *vc_idx;   // [nbatch][nsm*nch+1] segment starts per (view, ch)
    const int64_t *ent_off;  // [nbatch] first entry of each batch
    const int64_t *si_half;  // [nz] central view - n_sm/2 of each slice
    int64_t nsm, nch, nr, readings;
};

// One 64-entry chunk of an item: reduce runs of equal samples and add them
// into the view's tile (acc: rows of nr samples per channel from c0).
template <typename T>
__device__ __forceinline__ void scatter_chunk(uint32_t e0, T d, bool ok, int lane, T *acc)
{
    const int key = ok ? (int)bt_lkey(e0) : -1;
    const uint64_t head = run_heads(key, lane);
    d = run_sum(d, head, lane);
    if (ok && run_end(head, lane))
        acc[key] += d;
}

// Dispatch order: the groups of kScatterViews data views from vlo are
// visited in residue classes of P (group g, g + P, g + 2P, ...), P = view
// distance of consecutive members of the slice group / kScatterViews: those
// groups read the same SSM views' transposed entries (for successive
// members), so the entries are fetched into L2 once per class; logical
// workgroups go to the XCDs in contiguous ranges.
struct ScatterOrder {
    int64_t vlo, vhi, nq, ncg, P, rows;
};

// LDS: kScatterViews * kScatterCh * nr values of T (dynamic).  WS: the
// first sweep's surrogate sums, written weighted, e = w (e - sum) with w =
// a.w (the gather then reads w S as one array).
template <typename T, bool WS = false>
__global__ __launch_bounds__(kSweepBlock) void icd_scatter_kernel(
    SweepArgs a, StepArgs sa, BtArgs bt, int64_t b, ScatterOrder so)
{
    extern __shared__ __align__(8) unsigned char acc_raw[];
    T *acc = reinterpret_cast<T *>(acc_raw);
    const T *du = static_cast<const T *>(a.du);
    __shared__ int any_work;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t nwg = so.P * so.rows * so.ncg;
    const int64_t per_xcd = (nwg + kXcd - 1) / kXcd;
    const int64_t L = (int64_t)(blockIdx.x % kXcd) * per_xcd + blockIdx.x / kXcd;
    if (L >= nwg)
        return;
    const int64_t ql = L / so.ncg;
    const int64_t cls = ql / so.rows, m = ql - cls * so.rows;
    const int64_t q = cls + m * so.P;
    if (q >= so.nq)
        return;
    const int64_t view0 = so.vlo + q * kScatterViews;
    const int64_t nview = min((int64_t)kScatterViews, so.vhi - view0);
    const int64_t c0 = (L - ql * so.ncg) * kScatterCh;
    const int64_t c1 = min(c0 + kScatterCh, bt.nch);
    const int64_t wc0 = min(c0 + kWaveCh * wave, c1), wc1 = min(wc0 + kWaveCh, c1);
    const int nr = (int)bt.nr;
    const int tile = (int)((c1 - c0) * nr);
    const int64_t nvc = bt.nsm * bt.nch + 1;
    const int32_t *vci = bt.vc_idx + b * nvc;
    const uint32_t *ent = bt.ent + bt.ent_off[b];
    const float *val = bt.val + bt.ent_off[b];
    if (tid == 0)
        any_work = 0;
    for (int i = tid; i < kScatterViews * tile; i += kSweepBlock)
        acc[i] = T(0);
    __syncthreads();
    // items: lane l of a round covers member vb + l / kScatterViews and view
    // u = l % kScatterViews, so one member's items come one after another
    // (the columns of neighbouring views are nearly the same: their du reads
    // hit the lines the previous item brought in)
    constexpr int kItemLanes = 64 / kScatterViews;
    for (int64_t vb = 0; vb < sa.ng; vb += kItemLanes) {
        const int u_l = lane % kScatterViews;
        const int64_t v_l = vb + lane / kScatterViews;
        int32_t lo = 0, hi = 0;
        if (u_l < nview && v_l < sa.ng && wc1 > wc0) {
            const int64_t i = view0 + u_l - bt.si_half[sa.j + a.spacing * v_l];
            if (i >= 0 && i < bt.nsm) {
                // entry offsets are multiples of kQuad: work in quads
                lo = vci[i * bt.nch + wc0] / kQuad;
                hi = vci[i * bt.nch + wc1] / kQuad;
            }
        }
        uint64_t todo = __ballot(hi > lo);
        if (todo && lane == 0)
            any_work = 1;
        // (item, 64-entry chunk) pairs in item order, kItemUnroll at a time:
        // the loads of several chunks -- of one long item or of several
        // short ones -- are in flight before any is reduced; chunks are
        // applied in stream order (deterministic).  The schedule is
        // wave-uniform and kept in scalar registers (readlane of the lane
        // that holds the item); loads use uniform base pointers with 32-bit
        // lane offsets, clamped into the chunk instead of branched around.
        int cur = -1;
        int32_t clo = 0, chi = 0;
        while (true) {
            int it[kItemUnroll];
            int32_t lo_u[kItemUnroll], n_u[kItemUnroll];
#pragma unroll
            for (int u = 0; u < kItemUnroll; u++) {
                if (clo >= chi && todo) {
                    cur = __ffsll((unsigned long long)todo) - 1;
                    todo &= todo - 1;
                    clo = __builtin_amdgcn_readlane(lo, cur);
                    chi = __builtin_amdgcn_readlane(hi, cur);
                }
                if (clo < chi) {
                    it[u] = cur;
                    lo_u[u] = clo;
                    n_u[u] = min(64, chi - clo);
                    clo += 64;
                } else {
                    it[u] = -1;
                    lo_u[u] = 0;
                    n_u[u] = 0;
                }
            }
            if (it[0] < 0)
                break;
            uint4 ev[kItemUnroll];
            float4 av[kItemUnroll];
            bool ok[kItemUnroll];
#pragma unroll
            for (int u = 0; u < kItemUnroll; u++) {
                ok[u] = lane < n_u[u];
                const uint4 *ent_u = reinterpret_cast<const uint4 *>(ent) + lo_u[u];
                const float4 *val_u = reinterpret_cast<const float4 *>(val) + lo_u[u];
                const uint32_t pc = ok[u] ? (uint32_t)lane : 0u;
                ev[u] = it[u] >= 0 ? ent_u[pc] : make_uint4(0, 0, 0, 0);
                av[u] = it[u] >= 0 ? val_u[pc] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
            T dv[kItemUnroll];
#pragma unroll
            for (int u = 0; u < kItemUnroll; u++) {
                const int32_t vm = (int32_t)vb + (it[u] < 0 ? 0 : (it[u] / kScatterViews));
                const T *du_row = du + (uint32_t)(vm * (int32_t)sa.nc);
                // the quad's four entries share one sample (padding: value 0)
                T d = du_row[bt_slot(ev[u].x)] * (T)av[u].x;
                d += du_row[bt_slot(ev[u].y)] * (T)av[u].y;
                d += du_row[bt_slot(ev[u].z)] * (T)av[u].z;
                d += du_row[bt_slot(ev[u].w)] * (T)av[u].w;
                dv[u] = ok[u] ? d : T(0);
            }
#pragma unroll
            for (int u = 0; u < kItemUnroll; u++) {
                if (it[u] < 0)
                    continue;
                scatter_chunk(ev[u].x, dv[u], ok[u], lane,
                              acc + (it[u] % kScatterViews) * tile);
            }
        }
    }
    __syncthreads();
    if (!any_work)
        return;
    for (int64_t u = 0; u < nview; u++) {
        const int64_t r0 = ((view0 + u) * bt.nch + c0) * bt.nr;
        const T *tacc = acc + u * tile;
        T *e = static_cast<T *>(a.e) + r0;
        if (WS) {
            const float *w = a.w + r0;
            for (int i = tid; i < tile; i += kSweepBlock) {
                const T d = tacc[i];
                if (d != T(0))
                    e[i] = (T)w[i] * (e[i] - d);
            }
        } else if (a.w) {
            // only w e is kept current (the gather reads nothing else); e
            // follows where w == 0, where w e cannot give it back, and is
            // formed as w e / w elsewhere when it is read (refresh_error)
            T *ew = static_cast<T *>(a.ew) + r0;
            const float *w = a.w + r0;
            for (int i = tid; i < tile; i += kSweepBlock) {
                const T d = tacc[i];
                if (d != T(0)) {
                    const float wi = w[i];
                    ew[i] = ew[i] - (T)wi * d;
                    if (wi == 0.0f)
                        e[i] = e[i] - d;
                }
            }
        } else {
            for (int i = tid; i < tile; i += kSweepBlock) {
                const T d = tacc[i];
                if (d != T(0))
                    e[i] = e[i] - d;
            }
        }
    }
}

// Transposed-batch construction: per FOV column its batch and slot.
__global__ void bt_count_kernel(const int64_t *col_ptr, const int32_t *idx,
                                const int32_t *cols, int64_t ncols,
                                const int32_t *col_batch, int64_t nr,
                                int64_t nvc, uint32_t *counts)
{
    for (int64_t i = blockIdx.x; i < ncols; i += gridDim.x) {
        const int64_t col = cols[i];
        const int64_t b = col_batch[col];
        for (int64_t p = col_ptr[col] + threadIdx.x; p < col_ptr[col + 1];
             p += blockDim.x)
            atomicAdd(&counts[b * nvc + idx[p] / nr], 1u);
    }
}

// Exclusive scan of each batch's (view, channel) counts; one block per batch.
__global__ __launch_bounds__(1024) void bt_scan_kernel(const uint32_t *counts,
                                                       int32_t *vc_idx,
                                                       int64_t *batch_tot,
                                                       int64_t nvc)
{
    __shared__ int64_t part[1024];
    const int64_t b = blockIdx.x;
    const uint32_t *c = counts + b * nvc;
    int32_t *o = vc_idx + b * nvc;
    const int t = threadIdx.x;
    const int64_t per = (nvc + 1023) / 1024;
    const int64_t lo = t * per, hi = min(nvc, lo + per);
    int64_t s = 0;
    for (int64_t i = lo; i < hi; i++)
        s += c[i];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        int64_t run = 0;
        for (int i = 0; i < 1024; i++) {
            const int64_t v = part[i];
            part[i] = run;
            run += v;
        }
        batch_tot[b] = run;
    }
    __syncthreads();
    int64_t run = part[t];
    for (int64_t i = lo; i < hi; i++) {
        o[i] = (int32_t)run;
        run += c[i];
    }
}

__global__ void bt_fill_kernel(const int64_t *col_ptr, const int32_t *idx,
                               const float *val, const __half *val16,
                               const float *col_scale, const int32_t *cols,
                               int64_t ncols, const int32_t *col_batch,
                               const int32_t *col_slot, int64_t nr, int64_t nch,
                               int64_t nvc, const int32_t *vc_idx,
                               const int64_t *ent_off, uint32_t *cursor,
                               uint32_t *ent, float *bval)
{
    for (int64_t i = blockIdx.x; i < ncols; i += gridDim.x) {
        const int64_t col = cols[i];
        const int64_t b = col_batch[col];
        const uint32_t slot = (uint32_t)col_slot[col];
        for (int64_t p = col_ptr[col] + threadIdx.x; p < col_ptr[col + 1];
             p += blockDim.x) {
            const int64_t q = idx[p];
            const int64_t key = q / nr;
            const uint32_t row = (uint32_t)(q - key * nr);
            const uint32_t ch = (uint32_t)(key % nch);
            const int64_t pos = ent_off[b] + vc_idx[b * nvc + key] +
                                atomicAdd(&cursor[b * nvc + key], 1u);
            const uint32_t lkey = (ch % kScatterCh) * (uint32_t)nr + row;
            ent[pos] = slot | (lkey << kSlotBits) | ((ch / kScatterCh) << (kSlotBits + kLkeyBits));
            bval[pos] = val16 ? __half2float(val16[p]) * col_scale[col] : val[p];
        }
    }
}

// Rank placement of a segment of up to 64 * NC entries held in registers
// (NC chunks of 64 lanes): each entry goes to the number of smaller keys,
// counted against every key of the segment by lane broadcasts.
template <int NC>
__device__ void wave_rank_place(int64_t lo, int64_t n, uint32_t *key, float *val, int lane)
{
    uint32_t k[NC];
    float v[NC];
    int rank[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int64_t i = c * 64 + lane;
        k[c] = i < n ? key[lo + i] : 0xffffffffu;
        v[c] = i < n ? val[lo + i] : 0.0f;
        rank[c] = 0;
    }
    const int nch = (int)((n + 63) / 64);
    for (int c2 = 0; c2 < nch; c2++) {
        uint32_t kc = k[0];
#pragma unroll
        for (int c = 1; c < NC; c++)
            kc = c == c2 ? k[c] : kc;
        for (int j = 0; j < 64; j++) {
            const uint32_t kj = (uint32_t)__builtin_amdgcn_readlane((int)kc, j);
#pragma unroll
            for (int c = 0; c < NC; c++)
                rank[c] += kj < k[c] ? 1 : 0;
        }
    }
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int64_t i = c * 64 + lane;
        if (i < n) {
            key[lo + rank[c]] = k[c];
            val[lo + rank[c]] = v[c];
        }
    }
}

// Sort one segment [lo, hi) of (key, value) pairs by key -- keys unique in
// the segment, below 2^32 - 1 -- with one wavefront, by rank placement for
// up to 1024 entries and by an insertion sort of lane 0 beyond that.  The
// result does not depend on the input order.
__device__ void wave_rank_sort(int64_t lo, int64_t hi, uint32_t *key, float *val, int lane)
{
    const int64_t n = hi - lo;
    if (n <= 1)
        return;
    if (n <= 128) {
        wave_rank_place<2>(lo, n, key, val, lane);
        return;
    }
    if (n <= 1024) {
        wave_rank_place<16>(lo, n, key, val, lane);
        return;
    }
    if (lane == 0)
        for (int64_t i = lo + 1; i < hi; i++) {
            const uint32_t ke = key[i];
            const float kv = val[i];
            int64_t jj = i - 1;
            while (jj >= lo && key[jj] > ke) {
                key[jj + 1] = key[jj];
                val[jj + 1] = val[jj];
                jj--;
            }
            key[jj + 1] = ke;
            val[jj + 1] = kv;
        }
}

// Sort each (view, channel) segment by (row, slot): a deterministic order.
// One wavefront per segment.
__global__ __launch_bounds__(256) void bt_sort_kernel(const int32_t *vc_idx,
                                                      const int64_t *ent_off,
                                                      int64_t nbatch, int64_t nvc,
                                                      uint32_t *ent, float *bval)
{
    const int lane = threadIdx.x & 63;
    const int64_t nseg = nbatch * (nvc - 1);
    for (int64_t sg = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64; sg < nseg;
         sg += (int64_t)gridDim.x * blockDim.x / 64) {
        const int64_t b = sg / (nvc - 1), key = sg - b * (nvc - 1);
        const int64_t lo = ent_off[b] + vc_idx[b * nvc + key];
        const int64_t hi = ent_off[b] + vc_idx[b * nvc + key + 1];
        wave_rank_sort(lo, hi, ent, bval, lane);
    }
}

// Padded length of every sorted (view, channel) segment: each run of one
// sample rounded up to a multiple of kQuad.  One thread per segment.
__global__ void bt_pad_count_kernel(const int32_t *vc_idx, const int64_t *ent_off,
                                    int64_t nbatch, int64_t nvc, const uint32_t *ent,
                                    uint32_t *pcnt)
{
    const int64_t nseg = nbatch * (nvc - 1);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nseg;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = t / (nvc - 1), key = t - b * (nvc - 1);
        const int64_t lo = ent_off[b] + vc_idx[b * nvc + key];
        const int64_t hi = ent_off[b] + vc_idx[b * nvc + key + 1];
        uint32_t n = 0, prev = 0xffffffffu, run = 0;
        for (int64_t p = lo; p < hi; p++) {
            const uint32_t k = bt_lkey(ent[p]);
            if (k != prev) {
                n += (run + kQuad - 1) / kQuad * kQuad;
                run = 0;
                prev = k;
            }
            run++;
        }
        n += (run + kQuad - 1) / kQuad * kQuad;
        pcnt[b * nvc + key] = n;
    }
}

// Copy of the sorted segments with every run padded to a multiple of kQuad
// (padding: the run's last entry with value 0).  One thread per segment.
__global__ void bt_pad_fill_kernel(const int32_t *vc_idx, const int64_t *ent_off,
                                   int64_t nbatch, int64_t nvc, const uint32_t *ent,
                                   const float *val, const int32_t *pvc,
                                   const int64_t *pent_off, uint32_t *pent, float *pval)
{
    const int64_t nseg = nbatch * (nvc - 1);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nseg;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = t / (nvc - 1), key = t - b * (nvc - 1);
        const int64_t lo = ent_off[b] + vc_idx[b * nvc + key];
        const int64_t hi = ent_off[b] + vc_idx[b * nvc + key + 1];
        int64_t d = pent_off[b] + pvc[b * nvc + key];
        uint32_t prev = 0xffffffffu, run = 0, last = 0;
        for (int64_t p = lo; p < hi; p++) {
            const uint32_t e = ent[p];
            const uint32_t k = bt_lkey(e);
            if (k != prev) {
                for (; run % kQuad; run++, d++) {
                    pent[d] = last;
                    pval[d] = 0.0f;
                }
                run = 0;
                prev = k;
            }
            pent[d] = e;
            pval[d] = val[p];
            d++;
            run++;
            last = e;
        }
        for (; run % kQuad; run++, d++) {
            pent[d] = last;
            pval[d] = 0.0f;
        }
    }
}

// p[0, n) = 0 (a plain store kernel, so that a caller may capture the sweep)
template <typename T>
__global__ __launch_bounds__(256) void zero_kernel(T *p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
        p[i] = T(0);
}

// ---------------------------------------------------------------------------
// Reductions, resync and cost.

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

// ew = w e (the FP64 product, rounded to the state type T).
template <typename T>
__global__ __launch_bounds__(kRedBlock) void weighted_error_kernel(
    const T *e, const float *w, int64_t n, T *ew)
{
    for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kRedBlock)
        ew[i] = (T)((double)w[i] * (double)e[i]);
}

// e = w e / w wherever w > 0 (the weighted sweep keeps only w e current
// there); samples with w == 0 keep their e.
template <typename T>
__global__ __launch_bounds__(kRedBlock) void error_from_weighted_kernel(
    T *e, const T *ew, const float *w, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kRedBlock) {
        const float wi = w[i];
        if (wi > 0.0f)
            e[i] = (T)((double)ew[i] / (double)wi);
    }
}

// out = (T) in: the measured sinogram as the starting error sinogram.
template <typename T>
__global__ __launch_bounds__(kRedBlock) void widen_kernel(const float *in, int64_t n, T *out)
{
    for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kRedBlock)
        out[i] = (T)in[i];
}

// Data part of the cost from the kept error sinogram: sum w e^2, FP64.
template <typename T>
__global__ __launch_bounds__(kRedBlock) void error_sq_kernel(const T *e, const float *w,
                                                             int64_t n, double *partials)
{
    __shared__ double sh[kRedBlock / 64];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kRedBlock) {
        const double d = (double)e[i];
        acc += (w ? (double)w[i] : 1.0) * d * d;
    }
    const double s = block_sum_f64(acc, sh);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = s;
}

// Penalty part of the cost.  8 neighbours (the reference's flat stencil):
// pairs (q, q+d), d in {1, nx-1, nx, nx+1}, q+d < n.  26 neighbours: per
// voxel the in-plane forward pairs (i+1, j), (i-1, j+1), (i, j+1),
// (i+1, j+1) inside the slice, then the 3x3 block above through pi_k.
template <typename T>
__global__ __launch_bounds__(kRedBlock) void penalty_cost_kernel(
    const T *x, int64_t nx, int64_t n, int penalty, double delta,
    double *partials, Nb26 nb)
{
    __shared__ double sh[kRedBlock / 64];
    const int64_t dd[4] = {1, nx - 1, nx, nx + 1};
    const double wd[4] = {1.0, 1.0 / kSqrt2, 1.0, 1.0 / kSqrt2};
    auto term = [&](double diff, double w) {
        if (penalty == CTBB_PENALTY_QUADRATIC)
            return w * diff * diff;
        const double ad = fabs(diff);
        return w * (ad - delta * log1p(ad / delta));
    };
    double acc = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * kRedBlock) {
        const double xq = (double)x[q];
        if (!nb.shear) {
            for (int t = 0; t < 4; t++) {
                if (q + dd[t] >= n)
                    continue;
                acc += term(xq - (double)x[q + dd[t]], wd[t]);
            }
            continue;
        }
        const int64_t nxy = nb.nx * nb.ny, k = q / nxy, c = q - k * nxy;
        const int64_t j = c / nb.nx, i = c - j * nb.nx;
        const int ox[4] = {1, -1, 0, 1}, oy[4] = {0, 1, 1, 1};
        for (int t = 0; t < 4; t++) {
            const int64_t ii = i + ox[t], jj = j + oy[t];
            if (ii >= 0 && ii < nb.nx && jj >= 0 && jj < nb.ny)
                acc += term(xq - (double)x[ii + nb.nx * jj + nxy * k],
                            (ox[t] && oy[t]) ? 1.0 / kSqrt2 : 1.0);
        }
        if (k + 1 < nb.nz) {
            int64_t iu, ju;
            rot26_fwd(nb, k, i, j, iu, ju);
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    const int64_t ii = iu + dx, jj = ju + dy;
                    if (ii >= 0 && ii < nb.nx && jj >= 0 && jj < nb.ny)
                        acc += term(xq - (double)x[ii + nb.nx * jj + nxy * (k + 1)],
                                    1.0 / sqrt((double)(dx * dx + dy * dy + 1)));
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

// ---------------------------------------------------------------------------
// Ray-major copy of the SSM and the sample-driven FP64 forward projection
// built on it.  Entries are grouped by (SSM view, channel) and sorted by
// (row, column) inside a group; an entry packs its column and row into one
// word, col | row << 24.  Built through per-ray counts (CSR over rays
// q = row + nr*(ch + nch*view)), of which only the (view, channel) group
// starts are kept.

constexpr uint32_t kRmColBits = 24, kRmColMask = (1u << kRmColBits) - 1;

__global__ void rm_count_kernel(const int64_t *col_ptr, const int32_t *idx,
                                int64_t ncol, uint32_t *cnt)
{
    for (int64_t c = blockIdx.x; c < ncol; c += gridDim.x)
        for (int64_t p = col_ptr[c] + threadIdx.x; p < col_ptr[c + 1]; p += blockDim.x)
            atomicAdd(&cnt[idx[p]], 1u);
}

constexpr int kScanChunk = 65536;

__global__ __launch_bounds__(256) void chunk_sum_kernel(const uint32_t *cnt,
                                                        int64_t n, int64_t *sums)
{
    __shared__ int64_t sh[256];
    const int64_t b0 = (int64_t)blockIdx.x * kScanChunk;
    int64_t s = 0;
    for (int64_t i = b0 + threadIdx.x; i < min(n, b0 + kScanChunk); i += 256)
        s += cnt[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int i = 0; i < 256; i++)
            t += sh[i];
        sums[blockIdx.x] = t;
    }
}

// exclusive scan of one chunk with a base offset: 256 threads, each a
// contiguous run of kScanChunk/256 elements
__global__ __launch_bounds__(256) void chunk_scan_kernel(const uint32_t *cnt,
                                                         int64_t n,
                                                         const int64_t *base,
                                                         int64_t *out)
{
    __shared__ int64_t sh[256];
    const int64_t b0 = (int64_t)blockIdx.x * kScanChunk;
    const int64_t per = kScanChunk / 256;
    const int64_t lo = b0 + threadIdx.x * per, hi = min(n, lo + per);
    int64_t s = 0;
    for (int64_t i = lo; i < hi; i++)
        s += cnt[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t run = base[blockIdx.x];
        for (int i = 0; i < 256; i++) {
            const int64_t v = sh[i];
            sh[i] = run;
            run += v;
        }
    }
    __syncthreads();
    int64_t run = sh[threadIdx.x];
    for (int64_t i = lo; i < hi; i++) {
        out[i] = run;
        run += cnt[i];
    }
}

__global__ void rm_fill_kernel(const int64_t *col_ptr, const int32_t *idx,
                               const float *val, const __half *val16,
                               const float *col_scale, int64_t ncol, int64_t nr,
                               unsigned long long *cursor, uint32_t *rent,
                               float *rval)
{
    for (int64_t c = blockIdx.x; c < ncol; c += gridDim.x)
        for (int64_t p = col_ptr[c] + threadIdx.x; p < col_ptr[c + 1]; p += blockDim.x) {
            const int32_t q = idx[p];
            const unsigned long long pos = atomicAdd(&cursor[q], 1ull);
            rent[pos] = (uint32_t)c | ((uint32_t)(q % nr) << kRmColBits);
            rval[pos] = val16 ? __half2float(val16[p]) * col_scale[c] : val[p];
        }
}

// Sort every ray's entries by column (the fill order is arbitrary): one
// wavefront per ray.
__global__ __launch_bounds__(256) void rm_sort_kernel(const int64_t *rptr,
                                                      const int64_t *prptr, int64_t nrays,
                                                      uint32_t *rent, float *rval)
{
    const int lane = threadIdx.x & 63;
    for (int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64; q < nrays;
         q += (int64_t)gridDim.x * blockDim.x / 64)
        wave_rank_sort(prptr[q], prptr[q] + (rptr[q + 1] - rptr[q]), rent, rval, lane);
}

// Ray counts rounded up to whole quads of entries.
__global__ void rm_quad_count_kernel(uint32_t *cnt, int64_t nrays)
{
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nrays;
         q += (int64_t)gridDim.x * blockDim.x)
        cnt[q] = (cnt[q] + kQuad - 1) / kQuad * kQuad;
}

// Padding of every ray to its quads: the ray's first column (any column of
// the ray works) with value 0.
__global__ void rm_pad_kernel(const int64_t *rptr, const int64_t *prptr, int64_t nrays,
                              int64_t nr, uint32_t *rent, float *rval)
{
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nrays;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t lo = prptr[q], n = rptr[q + 1] - rptr[q], hi = prptr[q + 1];
        const uint32_t col = n > 0 ? (rent[lo] & kRmColMask) : 0u;
        const uint32_t e = col | ((uint32_t)(q % nr) << kRmColBits);
        for (int64_t p = lo + n; p < hi; p++) {
            rent[p] = e;
            rval[p] = 0.0f;
        }
    }
}

__global__ void rm_group_kernel(const int64_t *rptr, int64_t ngroups, int64_t nr,
                                int64_t *vc)
{
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j <= ngroups;
         j += (int64_t)gridDim.x * blockDim.x)
        vc[j] = rptr[j * nr];
}

// Sample-driven FP64 projection.  One workgroup per (data view, kScatterCh
// channels), all rows: for each slice k whose view window holds the view,
// SSM view i = view - (c_k - n_sm/2); wavefront w takes channels c0+w,
// c0+w+4, ..., reads a channel's entries as quads (every ray padded to a
// multiple of four entries, one ray's quad per lane), forms
// sum x[col + nxy k] * a over the quad in FP64 per lane, sums runs of equal
// rows across lanes with DPP moves and adds each run into an FP64 LDS tile.  Fixed order, no
// atomics.  MODE 0: yhat = A x.  MODE 1: e = (float)(y - A x) (if e) and
// one partial of w (y - A x)^2 per workgroup.  LDS: kScatterCh * nr doubles.
constexpr int kProjUnroll = 4;

template <typename X, typename E, int MODE>
__global__ __launch_bounds__(kSweepBlock) void project_gather_kernel(
    const int64_t *vc_ptr, const uint32_t *rent, const float *rval, const X *x,
    const int64_t *si_half, const int32_t *k_lo, const int32_t *k_hi,
    int64_t nsm, int64_t nch, int64_t nr, int64_t nxy, int64_t ncg,
    double *yhat, const float *y, const float *w, E *e, E *ew,
    double *partials, int64_t readings, int64_t pstride, int64_t rows)
{
    extern __shared__ double pacc[];
    __shared__ double sh[kWaves];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // data views in residue classes of pstride, the view distance of
    // neighbouring slices: views v and v + pstride read the same SSM views
    // for slices k and k + 1, so consecutive workgroups share the entries in
    // L2; logical workgroups go to the XCDs in contiguous ranges
    const int64_t nwg = pstride * rows * ncg;
    const int64_t per_xcd = (nwg + kXcd - 1) / kXcd;
    const int64_t L = (int64_t)(blockIdx.x % kXcd) * per_xcd + blockIdx.x / kXcd;
    if (L >= nwg)
        return;
    const int64_t ql = L / ncg, cg = L - ql * ncg;
    const int64_t cls = ql / rows;
    const int64_t view = cls + (ql - cls * rows) * pstride;
    if (view >= readings)
        return;
    const int64_t c0 = cg * kScatterCh;
    const int64_t c1 = min(c0 + kScatterCh, nch);
    const int nsamp = (int)((c1 - c0) * nr);
    for (int i = tid; i < nsamp; i += kSweepBlock)
        pacc[i] = 0.0;
    __syncthreads();
    for (int32_t k = k_lo[view]; k < k_hi[view]; k++) {
        const int64_t i = view - si_half[k];
        if (i < 0 || i >= nsm)
            continue;
        const X *xk = x + nxy * k;
        for (int64_t c = c0 + wave; c < c1; c += kWaves) {
            const int64_t lo = vc_ptr[i * nch + c], hi = vc_ptr[i * nch + c + 1];
            double *accc = pacc + (c - c0) * nr;
            // rays are padded to quads of entries (one ray's four entries per
            // lane); kProjUnroll chunks of 64 quads in flight, applied in order
            const int64_t qlo = lo / kQuad, qhi = hi / kQuad;
            const uint4 *rent4 = reinterpret_cast<const uint4 *>(rent);
            const float4 *rval4 = reinterpret_cast<const float4 *>(rval);
            for (int64_t p0 = qlo; p0 < qhi; p0 += 64 * kProjUnroll) {
                uint4 e0[kProjUnroll];
                float4 av[kProjUnroll];
                bool ok[kProjUnroll];
#pragma unroll
                for (int u = 0; u < kProjUnroll; u++) {
                    const int64_t p = p0 + u * 64 + lane;
                    ok[u] = p < qhi;
                    e0[u] = ok[u] ? rent4[p] : make_uint4(0, 0, 0, 0);
                    av[u] = ok[u] ? rval4[p] : make_float4(0.f, 0.f, 0.f, 0.f);
                }
                double d[kProjUnroll];
#pragma unroll
                for (int u = 0; u < kProjUnroll; u++) {
                    double t = (double)xk[e0[u].x & kRmColMask] * (double)av[u].x;
                    t += (double)xk[e0[u].y & kRmColMask] * (double)av[u].y;
                    t += (double)xk[e0[u].z & kRmColMask] * (double)av[u].z;
                    t += (double)xk[e0[u].w & kRmColMask] * (double)av[u].w;
                    d[u] = ok[u] ? t : 0.0;
                }
#pragma unroll
                for (int u = 0; u < kProjUnroll; u++) {
                    if (p0 + u * 64 >= qhi)
                        break;
                    const int row = ok[u] ? (int)(e0[u].x >> kRmColBits) : -1;
                    const uint64_t head = run_heads(row, lane);
                    const double t = run_sum(d[u], head, lane);
                    if (ok[u] && run_end(head, lane))
                        accc[row] += t;
                }
            }
        }
    }
    __syncthreads();
    const int64_t r0 = (view * nch + c0) * nr;
    double part = 0.0;
    for (int i = tid; i < nsamp; i += kSweepBlock) {
        const double acc = pacc[i];
        if (MODE == 0) {
            yhat[r0 + i] = acc;
        } else {
            const double d = (double)y[r0 + i] - acc;
            const double wi = w ? (double)w[r0 + i] : 1.0;
            if (e)
                e[r0 + i] = (E)d;
            if (ew)
                ew[r0 + i] = (E)(wi * d);
            part += wi * d * d;
        }
    }
    if (MODE == 1) {
        const double s = block_sum_f64(part, sh);
        if (tid == 0)
            partials[view * ncg + cg] = s;
    }
}

// Shear parameters of the slice-to-slice rotations of the 26-neighbour
// penalty: phi_k = theta_k - theta_{k+1} = -(c_{k+1} - c_k) * dtheta for
// central views c_k (the tube angle advances by dtheta per view).
std::vector<double> shear_table(const ctbb_geom *g, const int64_t *slice_index)
{
    std::vector<double> sh(2 * std::max<int64_t>(1, g->nz - 1), 0.0);
    for (int64_t k = 0; k + 1 < g->nz; k++) {
        const double phi = -(double)(slice_index[k + 1] - slice_index[k]) *
                           g->tube_angle_increment;
        sh[2 * k] = -tan(phi / 2.0);
        sh[2 * k + 1] = sin(phi);
    }
    return sh;
}

Nb26 make_nb(const ctbb_geom *g, const double *d_shear)
{
    Nb26 n;
    n.shear = d_shear;
    n.nx = g->nx;
    n.ny = g->ny;
    n.nz = g->nz;
    n.cx = g->center_voxel_x;
    n.cy = g->center_voxel_y;
    return n;
}

// neighbours must be 8 (the reference's stencil) or 26 (square slices).
int check_neighbours(const ctbb_geom *g, int32_t neighbours)
{
    if (neighbours != 8 && neighbours != 26)
        return fail(CTBB_E_ARG, "neighbours must be 8 or 26");
    if (neighbours == 26 && (g->nx != g->ny || g->voxel_size_x != g->voxel_size_y))
        return fail(CTBB_E_UNSUPPORTED, "the 26-neighbour penalty needs square slices "
                                        "of square voxels");
    return CTBB_OK;
}

// Upload the shear table of slice_index (26 neighbours) or leave null.
int upload_shear(const ctbb_geom *g, int32_t neighbours, const int64_t *slice_index,
                 double **d_shear)
{
    *d_shear = nullptr;
    if (neighbours != 26)
        return CTBB_OK;
    if (!slice_index)
        return fail(CTBB_E_ARG, "the 26-neighbour penalty needs the slice table");
    const std::vector<double> sh = shear_table(g, slice_index);
    CTBB_TRY(dev_alloc_n(d_shear, sh.size(), "slice rotation shears"));
    hipError_t e = hipMemcpy(*d_shear, sh.data(), sh.size() * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(*d_shear);
        *d_shear = nullptr;
        return fail(CTBB_E_HIP, std::string("shear table: ") + hipGetErrorString(e));
    }
    return CTBB_OK;
}

int red_grid(int64_t n)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + kRedBlock - 1) / kRedBlock,
                                                        4 * 1024));
}

// Build (once) the ray-major copy of the matrix.  Reconstructions created
// from several host threads share the matrix: the first builds the copy
// while the others wait.
static std::mutex g_ray_major_mu;

int ensure_ray_major(const ctbb_ssm *mc)
{
    ctbb_ssm *m = const_cast<ctbb_ssm *>(mc);
    std::lock_guard<std::mutex> lock(g_ray_major_mu);
    if (m->rm_vc)
        return CTBB_OK;
    const ctbb_geom *g = &m->g;
    if (g->n_rows > 256 || g->nx * g->ny > (int64_t)kRmColMask + 1)
        return fail(CTBB_E_UNSUPPORTED, "ray-major matrix needs n_rows <= 256 and "
                                        "nx*ny <= 2^24");
    const int64_t nrays = g->n_views_sm * g->n_channels * g->n_rows;
    const int64_t ngroups = g->n_views_sm * g->n_channels;
    uint32_t *cnt = nullptr;
    int64_t *sums = nullptr, *rptr = nullptr, *prptr = nullptr;
    unsigned long long *cursor = nullptr;
    const int64_t nchunk = (nrays + kScanChunk - 1) / kScanChunk;
    int rc = CTBB_OK;
    auto cleanup = [&]() {
        if (cnt) (void)hipFree(cnt);
        if (sums) (void)hipFree(sums);
        if (rptr) (void)hipFree(rptr);
        if (prptr) (void)hipFree(prptr);
        if (cursor) (void)hipFree(cursor);
    };
    auto drop = [&]() {
        if (m->rm_vc) (void)hipFree(m->rm_vc);
        if (m->rm_ent) (void)hipFree(m->rm_ent);
        if (m->rm_val) (void)hipFree(m->rm_val);
        m->rm_vc = nullptr;
        m->rm_ent = nullptr;
        m->rm_val = nullptr;
    };
    if ((rc = dev_alloc_n(&cnt, nrays, "ray counts")) ||
        (rc = dev_alloc_n(&sums, nchunk, "chunk sums")) ||
        (rc = dev_alloc_n(&rptr, nrays + 1, "ray pointers")) ||
        (rc = dev_alloc_n(&prptr, nrays + 1, "padded ray pointers")) ||
        (rc = dev_alloc_n(&m->rm_vc, ngroups + 1, "ray-major group starts"))) {
        cleanup();
        drop();
        return rc;
    }
    const int gcol = (int)std::min<int64_t>(m->ncol, 65536);
    hipError_t e = hipMemset(cnt, 0, nrays * 4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(rm_count_kernel, dim3(gcol), dim3(256), 0, 0, m->col_ptr,
                           m->idx, m->ncol, cnt);
        hipLaunchKernelGGL(chunk_sum_kernel, dim3((unsigned)nchunk), dim3(256), 0, 0,
                           cnt, nrays, sums);
        e = hipGetLastError();
    }
    std::vector<int64_t> hs(nchunk);
    if (e == hipSuccess)
        e = hipMemcpy(hs.data(), sums, nchunk * 8, hipMemcpyDeviceToHost);
    int64_t run = 0;
    for (int64_t i = 0; i < nchunk; i++) {
        const int64_t v = hs[i];
        hs[i] = run;
        run += v;
    }
    if (e == hipSuccess)
        e = hipMemcpy(sums, hs.data(), nchunk * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(chunk_scan_kernel, dim3((unsigned)nchunk), dim3(256), 0, 0,
                           cnt, nrays, sums, rptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpy(rptr + nrays, &run, 8, hipMemcpyHostToDevice);
    if (e == hipSuccess && run != m->nnz)
        e = hipErrorUnknown;
    // every ray padded to whole quads: the same scan over the rounded counts
    if (e == hipSuccess) {
        hipLaunchKernelGGL(rm_quad_count_kernel, dim3(4096), dim3(256), 0, 0, cnt, nrays);
        hipLaunchKernelGGL(chunk_sum_kernel, dim3((unsigned)nchunk), dim3(256), 0, 0,
                           cnt, nrays, sums);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpy(hs.data(), sums, nchunk * 8, hipMemcpyDeviceToHost);
    int64_t prun = 0;
    for (int64_t i = 0; i < nchunk; i++) {
        const int64_t v = hs[i];
        hs[i] = prun;
        prun += v;
    }
    if (e == hipSuccess)
        e = hipMemcpy(sums, hs.data(), nchunk * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(chunk_scan_kernel, dim3((unsigned)nchunk), dim3(256), 0, 0,
                           cnt, nrays, sums, prptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpy(prptr + nrays, &prun, 8, hipMemcpyHostToDevice);
    if (e == hipSuccess &&
        ((rc = dev_alloc_n(&m->rm_ent, std::max<int64_t>(1, prun), "ray-major entries")) ||
         (rc = dev_alloc_n(&m->rm_val, std::max<int64_t>(1, prun), "ray-major values")))) {
        cleanup();
        drop();
        return rc;
    }
    if (e == hipSuccess) {
        (void)hipFree(cnt);
        cnt = nullptr;
        rc = dev_alloc_n(&cursor, nrays, "ray cursors");
        if (rc) {
            cleanup();
            drop();
            return rc;
        }
        e = hipMemcpy(cursor, prptr, nrays * 8, hipMemcpyDeviceToDevice);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(rm_fill_kernel, dim3(gcol), dim3(256), 0, 0, m->col_ptr,
                           m->idx, m->val, m->val16, m->col_scale, m->ncol, g->n_rows,
                           cursor, m->rm_ent, m->rm_val);
        hipLaunchKernelGGL(rm_sort_kernel, dim3(16384), dim3(256), 0, 0, rptr, prptr, nrays,
                           m->rm_ent, m->rm_val);
        hipLaunchKernelGGL(rm_pad_kernel, dim3(8192), dim3(256), 0, 0, rptr, prptr, nrays,
                           g->n_rows, m->rm_ent, m->rm_val);
        hipLaunchKernelGGL(rm_group_kernel, dim3(4096), dim3(256), 0, 0, prptr, ngroups,
                           g->n_rows, m->rm_vc);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    cleanup();
    if (e != hipSuccess) {
        drop();
        return fail(CTBB_E_HIP, std::string("ray-major matrix: ") + hipGetErrorString(e));
    }
    return CTBB_OK;
}

// Workgroups of the projection: one per (data view, kScatterCh channels).
int64_t project_grid(const ctbb_geom *g)
{
    return g->readings * ((g->n_channels + kScatterCh - 1) / kScatterCh);
}

// Per data view, the contiguous range of slices whose SSM view window holds
// the view (the central views are monotone in the slice index).
void view_slice_ranges(const ctbb_geom *g, const std::vector<int64_t> &si_half,
                       std::vector<int32_t> *klo, std::vector<int32_t> *khi)
{
    klo->assign(g->readings, 0);
    khi->assign(g->readings, 0);
    for (int64_t v = 0; v < g->readings; v++) {
        int32_t lo = -1, hi = -1;
        for (int64_t k = 0; k < g->nz; k++) {
            const int64_t i = v - si_half[k];
            if (i >= 0 && i < g->n_views_sm) {
                if (lo < 0)
                    lo = (int32_t)k;
                hi = (int32_t)k + 1;
            }
        }
        (*klo)[v] = lo < 0 ? 0 : lo;
        (*khi)[v] = lo < 0 ? 0 : hi;
    }
}

template <typename X, int MODE, typename E = float>
int launch_project(const ctbb_ssm *m, const X *x, const int64_t *d_si_half,
                   const int32_t *d_klo, const int32_t *d_khi, double *yhat,
                   const float *y, const float *w, E *e, E *ew,
                   double *partials,
                   hipStream_t st)
{
    CTBB_TRY(ensure_ray_major(m));
    const ctbb_geom *g = &m->g;
    const int64_t ncg = (g->n_channels + kScatterCh - 1) / kScatterCh;
    const size_t lds = sizeof(double) * kScatterCh * g->n_rows;
    // view distance of neighbouring slices (a dispatch order only)
    int64_t pstride = 1;
    if (g->tube_z_increment != 0.0)
        pstride = (int64_t)llround(fabs(g->voxel_size_z / g->tube_z_increment));
    pstride = std::min<int64_t>(std::max<int64_t>(pstride, 1), g->readings);
    const int64_t rows = (g->readings + pstride - 1) / pstride;
    const int64_t nwg = pstride * rows * ncg;
    const int64_t grid = kXcd * ((nwg + kXcd - 1) / kXcd);
    hipLaunchKernelGGL((project_gather_kernel<X, E, MODE>), dim3((unsigned)grid),
                       dim3(kSweepBlock), lds, st, m->rm_vc, m->rm_ent, m->rm_val, x,
                       d_si_half, d_klo, d_khi, g->n_views_sm, g->n_channels, g->n_rows,
                       g->nx * g->ny, ncg, yhat, y, w, e, ew, partials, g->readings,
                       pstride, rows);
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

template <typename T>
int recon_data_term_t(ctbb_recon *r, bool write_e, double *data)
{
    const int g = (int)project_grid(&r->g);
    T *e = static_cast<T *>(r->e), *ew = static_cast<T *>(r->ew);
    CTBB_TRY((launch_project<T, 1, T>(r->m, static_cast<const T *>(r->x), r->si_half,
                                      r->k_lo, r->k_hi, nullptr, r->y, r->w,
                                      write_e ? e : nullptr, write_e ? ew : nullptr,
                                      r->partials, r->stream)));
    if (write_e)
        r->e_stale = false;
    return sum_partials(r->partials, g, r->stream, data);
}

// Data term sum w (y - A x)^2 of the current volume (FP64 projection);
// write_e: also store e = y - A x (and w e) as the running error sinogram.
int recon_data_term(ctbb_recon *r, bool write_e, double *data)
{
    return r->fp64 ? recon_data_term_t<double>(r, write_e, data)
                   : recon_data_term_t<float>(r, write_e, data);
}

int recon_penalty_term(ctbb_recon *r, double *pen)
{
    const int g = red_grid(r->nvox);
    if (r->fp64)
        hipLaunchKernelGGL(penalty_cost_kernel<double>, dim3(g), dim3(kRedBlock), 0,
                           r->stream, static_cast<const double *>(r->x), r->g.nx, r->nvox,
                           r->opts.penalty, r->opts.delta, r->partials,
                           make_nb(&r->g, r->shear));
    else
        hipLaunchKernelGGL(penalty_cost_kernel<float>, dim3(g), dim3(kRedBlock), 0,
                           r->stream, static_cast<const float *>(r->x), r->g.nx, r->nvox,
                           r->opts.penalty, r->opts.delta, r->partials,
                           make_nb(&r->g, r->shear));
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
    std::vector<int64_t> sih(g->nz);
    for (int64_t k = 0; k < g->nz; k++)
        sih[k] = slice_offset(g, slice_index[k]) / (g->n_channels * g->n_rows);
    std::vector<int32_t> klo, khi;
    view_slice_ranges(g, sih, &klo, &khi);
    double *d_x = nullptr, *d_y = nullptr;
    int64_t *d_sih = nullptr;
    int32_t *d_klo = nullptr, *d_khi = nullptr;
    int rc = CTBB_OK;
    auto cleanup = [&]() {
        if (d_x) (void)hipFree(d_x);
        if (d_y) (void)hipFree(d_y);
        if (d_sih) (void)hipFree(d_sih);
        if (d_klo) (void)hipFree(d_klo);
        if (d_khi) (void)hipFree(d_khi);
    };
    if ((rc = dev_alloc_n(&d_x, nvox, "volume")) ||
        (rc = dev_alloc_n(&d_y, ds, "projection")) ||
        (rc = dev_alloc_n(&d_sih, g->nz, "slice view offsets")) ||
        (rc = dev_alloc_n(&d_klo, g->readings, "view slice ranges")) ||
        (rc = dev_alloc_n(&d_khi, g->readings, "view slice ranges"))) {
        cleanup();
        return rc;
    }
    hipError_t e = hipMemcpy(d_x, x, nvox * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_sih, sih.data(), g->nz * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_klo, klo.data(), g->readings * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_khi, khi.data(), g->readings * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        rc = launch_project<double, 0, float>(m, d_x, d_sih, d_klo, d_khi, d_y, nullptr, nullptr,
                                       nullptr, nullptr, nullptr, 0);
        if (rc == CTBB_OK)
            e = hipMemcpy(yhat, d_y, ds * 8, hipMemcpyDeviceToHost);
    }
    cleanup();
    if (rc != CTBB_OK)
        return rc;
    CTBB_HIP(e);
    return CTBB_OK;
}

extern "C" int ctbb_penalty_steps(const ctbb_geom *g, const ctbb_icd_opts *o,
                                  const int64_t *slice_index, int volume_fp32,
                                  const double *x, int64_t n, const int64_t *q,
                                  const double *alpha, const double *beta, double *du)
{
    CTBB_TRY(check_geom(g));
    if (!o || !x || !q || !alpha || !beta || !du || n < 0)
        return fail(CTBB_E_ARG, "null argument");
    if (o->penalty != CTBB_PENALTY_QUADRATIC &&
        o->penalty != CTBB_PENALTY_EDGE_PRESERVING)
        return fail(CTBB_E_ARG, "unknown penalty");
    CTBB_TRY(check_neighbours(g, o->neighbours));
    const int64_t nvox = g->nx * g->ny * g->nz;
    for (int64_t i = 0; i < n; i++)
        if (q[i] < 0 || q[i] >= nvox)
            return fail(CTBB_E_ARG, "voxel index out of range");
    if (n == 0)
        return CTBB_OK;
    double *d_x = nullptr, *d_a = nullptr, *d_b = nullptr, *d_du = nullptr;
    double *d_sh = nullptr;
    int64_t *d_q = nullptr;
    int rc = CTBB_OK;
    if ((rc = upload_shear(g, o->neighbours, slice_index, &d_sh)))
        return rc;
    if ((rc = dev_alloc_n(&d_x, nvox, "volume")) ||
        (rc = dev_alloc_n(&d_a, n, "alpha")) || (rc = dev_alloc_n(&d_b, n, "beta")) ||
        (rc = dev_alloc_n(&d_du, n, "du")) || (rc = dev_alloc_n(&d_q, n, "q"))) {
        (void)hipFree(d_x); (void)hipFree(d_a); (void)hipFree(d_b);
        (void)hipFree(d_du); (void)hipFree(d_q); (void)hipFree(d_sh);
        return rc;
    }
    const Nb26 nb = make_nb(g, d_sh);
    hipError_t e = hipMemcpy(d_x, x, nvox * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_a, alpha, n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_b, beta, n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_q, q, n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess && volume_fp32 == 0) {
        hipLaunchKernelGGL(penalty_steps_kernel<double>, dim3((n + 255) / 256), dim3(256),
                           0, 0, o->penalty, o->lambda, o->delta, d_x, g->nx,
                           nvox, n, d_q, d_a, d_b, d_du, nb);
        e = hipGetLastError();
    } else if (e == hipSuccess) {
        std::vector<float> xf(nvox);
        for (int64_t i = 0; i < nvox; i++)
            xf[i] = (float)x[i];
        float *d_xf = nullptr;
        if ((rc = dev_alloc_n(&d_xf, nvox, "volume (FP32)")) == CTBB_OK) {
            e = hipMemcpy(d_xf, xf.data(), nvox * 4, hipMemcpyHostToDevice);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(penalty_steps_kernel<float>, dim3((n + 255) / 256),
                                   dim3(256), 0, 0, o->penalty, o->lambda, o->delta,
                                   d_xf, g->nx, nvox, n, d_q, d_a, d_b, d_du, nb);
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
    (void)hipFree(d_du); (void)hipFree(d_q); (void)hipFree(d_sh);
    CTBB_HIP(e);
    return CTBB_OK;
}

extern "C" int ctbb_recon_free(ctbb_recon *r)
{
    if (!r)
        return CTBB_OK;
    for (void *p : {(void *)r->y, (void *)r->w, r->e, r->x})
        if (p) (void)hipFree(p);
    if (r->slice_off) (void)hipFree(r->slice_off);
    if (r->cols) (void)hipFree(r->cols);
    if (r->batch_ptr) (void)hipFree(r->batch_ptr);
    if (r->bt_ent) (void)hipFree(r->bt_ent);
    if (r->bt_val) (void)hipFree(r->bt_val);
    if (r->bt_vc) (void)hipFree(r->bt_vc);
    if (r->bt_off) (void)hipFree(r->bt_off);
    if (r->si_half) (void)hipFree(r->si_half);
    if (r->k_lo) (void)hipFree(r->k_lo);
    if (r->k_hi) (void)hipFree(r->k_hi);
    if (r->du) (void)hipFree(r->du);
    if (r->partials) (void)hipFree(r->partials);
    if (r->ew) (void)hipFree(r->ew);
    if (r->th2) (void)hipFree(r->th2);
    if (r->part1) (void)hipFree(r->part1);
    if (r->ssum) (void)hipFree(r->ssum);
    if (r->du_m1) (void)hipFree(r->du_m1);
    if (r->shear) (void)hipFree(r->shear);
    for (hipEvent_t ev : r->ev_sweep)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : r->ev_step)
        if (ev) (void)hipEventDestroy(ev);
    if (r->part2) (void)hipFree(r->part2);
    if (r->gb_ent) (void)hipFree(r->gb_ent);
    if (r->gb_val) (void)hipFree(r->gb_val);
    if (r->gb_val16) (void)hipFree(r->gb_val16);
    if (r->gb_perm) (void)hipFree(r->gb_perm);
    if (r->gb_ptr) (void)hipFree(r->gb_ptr);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
    return CTBB_OK;
}

// View-blocked batch copy (gather).  Per batch b and view block ib the
// batch's column slots are put in the order of where their footprint starts
// on the detector in the block's first view (channel, then row), so that
// consecutive ranks read neighbouring samples of the error sinogram.  Ranks
// are cut into wavefront chunks of 64; a chunk's entries are stored as
// (view in block q, rank, SSM order), each with the byte q << 6 | rank % 64.
//
// gb_key_kernel: sort key of every (FOV column, view block) -- the first
// entry's ray index within its view (row + nr * ch; 2^17 when the column has
// no entry in the block) << 14 | slot (unique keys: the order is fixed).
__global__ void gb_key_kernel(const int64_t *col_ptr, const int32_t *idx,
                              const int32_t *view_off, const int32_t *cols,
                              int64_t ncols, const int64_t *fov_base,
                              const int32_t *fov_nc, const int32_t *fov_slot,
                              int64_t nsm, int64_t nib, int64_t plane,
                              uint32_t *keys)
{
    const int64_t n = ncols * nib;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t / nib, ib = t - i * nib;
        const int64_t col = cols[i];
        const int32_t *vo = view_off + col * nsm;
        const int64_t v0 = ib * kViewBlock, v1 = min(v0 + kViewBlock, nsm);
        const int64_t e0 = vo[v0];
        const int64_t e1 = v1 < nsm ? (int64_t)vo[v1] : col_ptr[col + 1] - col_ptr[col];
        const uint32_t pos =
            e1 > e0 ? (uint32_t)(idx[col_ptr[col] + e0] % plane) : (1u << 17);
        keys[fov_base[i] + ib * fov_nc[i] + fov_slot[i]] = pos << 14 | (uint32_t)fov_slot[i];
    }
}

// One workgroup per (batch, view block): bitonic sort of the batch's keys in
// LDS (dynamic, next power of two >= the largest batch), then rank -> slot.
__global__ __launch_bounds__(1024) void gb_sort_kernel(const uint32_t *keys,
                                                       const int64_t *bptr,
                                                       int64_t nib, uint16_t *perm)
{
    extern __shared__ uint32_t sk[];
    const int64_t b = blockIdx.x / nib, ib = blockIdx.x - b * nib;
    const int64_t c0 = bptr[b], nc = bptr[b + 1] - c0;
    const int64_t o = nib * c0 + ib * nc;
    int P = 1;
    while (P < nc)
        P <<= 1;
    for (int i = threadIdx.x; i < P; i += blockDim.x)
        sk[i] = i < nc ? keys[o + i] : 0xffffffffu;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1)
        for (int jj = k >> 1; jj > 0; jj >>= 1) {
            for (int i = threadIdx.x; i < P; i += blockDim.x) {
                const int l = i ^ jj;
                if (l > i) {
                    const uint32_t x = sk[i], y = sk[l];
                    if ((x > y) == ((i & k) == 0)) {
                        sk[i] = y;
                        sk[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < nc; i += blockDim.x)
        perm[o + i] = (uint16_t)(sk[i] & 0x3fffu);
}

// Entries of column col in SSM view v (0 past the last view).
__device__ __forceinline__ int64_t gb_view_count(const int64_t *col_ptr,
                                                 const int32_t *view_off, int64_t col,
                                                 int64_t nsm, int64_t v)
{
    if (v >= nsm)
        return 0;
    const int32_t *vo = view_off + col * nsm;
    const int64_t e1 = v + 1 < nsm ? (int64_t)vo[v + 1] : col_ptr[col + 1] - col_ptr[col];
    return e1 - vo[v];
}

// Inclusive sum over the lanes of a wavefront.
__device__ __forceinline__ int64_t wave_incl_sum(int64_t x, int lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int64_t u = __shfl_up(x, off, 64);
        if (lane >= off)
            x += u;
    }
    return x;
}

// One workgroup per (batch, view block), wavefront w takes chunks w, w+4, ..:
// FILL = false writes the entry count of every chunk, FILL = true copies the
// chunk's entries to the starts the scan of the counts gave.
template <bool FILL>
__global__ __launch_bounds__(256) void gb_chunk_kernel(
    const int64_t *col_ptr, const int32_t *idx, const float *val,
    const __half *val16, const int32_t *view_off, const int32_t *cols,
    const int64_t *bptr, const int64_t *wbase, const uint16_t *perm, int64_t nsm,
    int64_t nib, uint32_t *wcnt, const int64_t *gptr, uint32_t *gent, float *gval,
    __half *gval16, int64_t plane)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t b = blockIdx.x / nib, ib = blockIdx.x - b * nib;
    const int64_t c0 = bptr[b], nc = bptr[b + 1] - c0;
    const int64_t nw = (nc + 63) / 64;
    const int64_t o = nib * c0 + ib * nc;
    for (int64_t w = wave; w < nw; w += 4) {
        const int64_t rk = w * 64 + lane;
        const bool ok = rk < nc;
        const int64_t col = ok ? cols[c0 + perm[o + rk]] : 0;
        const int64_t ci = wbase[b] + ib * nw + w;
        int64_t seg = FILL ? gptr[ci] : 0;
        for (int q = 0; q < kViewBlock; q++) {
            const int64_t v = ib * kViewBlock + q;
            const int64_t n = ok ? gb_view_count(col_ptr, view_off, col, nsm, v) : 0;
            const int64_t inc = wave_incl_sum(n, lane);
            const int64_t tot = __shfl(inc, 63, 64);
            if (FILL && n > 0) {
                const int64_t src = col_ptr[col] + view_off[col * nsm + v];
                const int64_t dst = seg + inc - n;
                const uint32_t kb = (uint32_t)(q << 6 | lane) << kSampleBits;
                for (int64_t e = 0; e < n; e++) {
                    gent[dst + e] = (uint32_t)(idx[src + e] - v * plane) | kb;
                    if (gval16)
                        gval16[dst + e] = val16[src + e];
                    else
                        gval[dst + e] = val[src + e];
                }
            }
            seg += tot;
        }
        if (!FILL && lane == 0)
            wcnt[ci] = (uint32_t)seg;
    }
}

static int build_gather_blocks(ctbb_recon *r, const std::vector<int64_t> &bptr)
{
    const ctbb_ssm *m = r->m;
    const ctbb_geom *g = &r->g;
    const int64_t nsm = g->n_views_sm, nib = (nsm + kViewBlock - 1) / kViewBlock;
    const int64_t ncols = r->ncols_fov, n = nib * ncols;
    const bool fp16 = m->precision == CTBB_SSM_FP16;
    static_assert(kViewBlock <= 4, "the entry key holds the view in block in 2 bits");
    if (g->n_rows * g->n_channels > (int64_t)kSampleMask + 1)
        return fail(CTBB_E_UNSUPPORTED, "view-blocked batches need n_rows * n_channels "
                                        "<= 2^17");
    r->nib = nib;
    std::vector<int64_t> base(ncols);
    std::vector<int32_t> nc(ncols), slot(ncols);
    r->h_wbase.assign(r->nbatch + 1, 0);
    int64_t max_nc = 1;
    for (int64_t b = 0; b < r->nbatch; b++) {
        const int64_t nb = bptr[b + 1] - bptr[b];
        max_nc = std::max(max_nc, nb);
        r->h_wbase[b + 1] = r->h_wbase[b] + nib * ((nb + 63) / 64);
        for (int64_t i = bptr[b]; i < bptr[b + 1]; i++) {
            base[i] = nib * bptr[b];
            nc[i] = (int32_t)nb;
            slot[i] = (int32_t)(i - bptr[b]);
        }
    }
    int P = 1;
    while (P < max_nc)
        P <<= 1;
    if (P > (1 << 14))
        return fail(CTBB_E_ARG, "batches of more than 16384 columns");
    const int64_t nchunks = r->h_wbase[r->nbatch];
    const int64_t nblk = r->nbatch * nib;
    int64_t *d_base = nullptr, *sums = nullptr, *d_wbase = nullptr;
    int32_t *d_nc = nullptr, *d_slot = nullptr;
    uint32_t *keys = nullptr, *wcnt = nullptr;
    const int64_t nscan = (nchunks + kScanChunk - 1) / kScanChunk;
    int rc = CTBB_OK;
    auto cleanup = [&]() {
        if (d_base) (void)hipFree(d_base);
        if (sums) (void)hipFree(sums);
        if (d_wbase) (void)hipFree(d_wbase);
        if (d_nc) (void)hipFree(d_nc);
        if (d_slot) (void)hipFree(d_slot);
        if (keys) (void)hipFree(keys);
        if (wcnt) (void)hipFree(wcnt);
    };
    const int64_t ne = std::max<int64_t>(1, r->ssm_fov_entries);
    if ((rc = dev_alloc_n(&d_base, std::max<int64_t>(1, ncols), "FOV batch bases")) ||
        (rc = dev_alloc_n(&d_nc, std::max<int64_t>(1, ncols), "FOV batch sizes")) ||
        (rc = dev_alloc_n(&d_slot, std::max<int64_t>(1, ncols), "FOV slots")) ||
        (rc = dev_alloc_n(&d_wbase, r->nbatch + 1, "chunk bases")) ||
        (rc = dev_alloc_n(&keys, std::max<int64_t>(1, n), "rank keys")) ||
        (rc = dev_alloc_n(&wcnt, std::max<int64_t>(1, nchunks), "chunk counts")) ||
        (rc = dev_alloc_n(&sums, std::max<int64_t>(1, nscan), "chunk sums")) ||
        (rc = dev_alloc_n(&r->gb_perm, std::max<int64_t>(1, n), "rank slots")) ||
        (rc = dev_alloc_n(&r->gb_ptr, nchunks + 1, "chunk pointers")) ||
        (rc = dev_alloc_n(&r->gb_ent, ne, "blocked entries")) ||
        (rc = fp16 ? dev_alloc_n(&r->gb_val16, ne, "blocked values")
                   : dev_alloc_n(&r->gb_val, ne, "blocked values"))) {
        cleanup();
        return rc;
    }
    hipError_t e = hipSuccess;
    if (ncols) {
        e = hipMemcpy(d_base, base.data(), ncols * 8, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_nc, nc.data(), ncols * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_slot, slot.data(), ncols * 4, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess)
        e = hipMemcpy(d_wbase, r->h_wbase.data(), (r->nbatch + 1) * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess && n) {
        hipLaunchKernelGGL(gb_key_kernel, dim3(4096), dim3(256), 0, 0, m->col_ptr, m->idx,
                           m->view_off, r->cols, ncols, d_base, d_nc, d_slot, nsm, nib,
                           g->n_rows * g->n_channels, keys);
        hipLaunchKernelGGL(gb_sort_kernel, dim3((unsigned)nblk), dim3(1024),
                           (size_t)P * 4, 0, keys, r->batch_ptr, nib, r->gb_perm);
        hipLaunchKernelGGL(gb_chunk_kernel<false>, dim3((unsigned)nblk), dim3(256), 0, 0,
                           m->col_ptr, m->idx, m->val, m->val16, m->view_off, r->cols,
                           r->batch_ptr, d_wbase, r->gb_perm, nsm, nib, wcnt, nullptr,
                           nullptr, nullptr, nullptr, g->n_rows * g->n_channels);
        hipLaunchKernelGGL(chunk_sum_kernel, dim3((unsigned)nscan), dim3(256), 0, 0, wcnt,
                           nchunks, sums);
        e = hipGetLastError();
    }
    std::vector<int64_t> hs(nscan);
    if (e == hipSuccess && nscan)
        e = hipMemcpy(hs.data(), sums, nscan * 8, hipMemcpyDeviceToHost);
    int64_t run = 0;
    for (int64_t i = 0; i < nscan; i++) {
        const int64_t v = hs[i];
        hs[i] = run;
        run += v;
    }
    if (e == hipSuccess && nscan)
        e = hipMemcpy(sums, hs.data(), nscan * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess && nscan) {
        hipLaunchKernelGGL(chunk_scan_kernel, dim3((unsigned)nscan), dim3(256), 0, 0, wcnt,
                           nchunks, sums, r->gb_ptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpy(r->gb_ptr + nchunks, &run, 8, hipMemcpyHostToDevice);
    if (e == hipSuccess && run != r->ssm_fov_entries)
        e = hipErrorUnknown;
    if (e == hipSuccess && n) {
        hipLaunchKernelGGL(gb_chunk_kernel<true>, dim3((unsigned)nblk), dim3(256), 0, 0,
                           m->col_ptr, m->idx, m->val, m->val16, m->view_off, r->cols,
                           r->batch_ptr, d_wbase, r->gb_perm, nsm, nib, nullptr, r->gb_ptr,
                           r->gb_ent, r->gb_val, r->gb_val16, g->n_rows * g->n_channels);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    cleanup();
    if (e != hipSuccess)
        return fail(CTBB_E_HIP, std::string("view-blocked batches: ") + hipGetErrorString(e));
    return CTBB_OK;
}

// Transposed copy of every batch's SSM entries for the sample-driven
// scatter: entries grouped by batch, then by (SSM view, channel) segment,
// sorted by (row, column slot) inside a segment.
static int build_transposed_batches(ctbb_recon *r, const std::vector<int32_t> &cols,
                                    const std::vector<int64_t> &bptr)
{
    const ctbb_ssm *m = r->m;
    const ctbb_geom *g = &r->g;
    const int64_t nbatch = r->nbatch, ncol = g->nx * g->ny;
    const int64_t nvc = g->n_views_sm * g->n_channels + 1;
    static_assert(kScatterCh * 128 <= (1 << kLkeyBits), "tile-local sample key width");
    if (g->n_rows > 128 || g->n_channels > 1024)
        return fail(CTBB_E_UNSUPPORTED, "transposed batches need n_rows <= 128 "
                                        "and n_channels <= 1024");
    if (r->max_batch > (1 << 14))
        return fail(CTBB_E_ARG, "batches of more than 16384 columns (tile too "
                                "small for this volume)");
    std::vector<int32_t> cb(ncol, -1), cs(ncol, -1);
    for (int64_t b = 0; b < nbatch; b++)
        for (int64_t i = bptr[b]; i < bptr[b + 1]; i++) {
            cb[cols[i]] = (int32_t)b;
            cs[cols[i]] = (int32_t)(i - bptr[b]);
        }
    int32_t *d_cb = nullptr, *d_cs = nullptr;
    uint32_t *d_cnt = nullptr;
    int64_t *d_tot = nullptr;
    int rc = CTBB_OK;
    auto cleanup = [&]() {
        if (d_cb) (void)hipFree(d_cb);
        if (d_cs) (void)hipFree(d_cs);
        if (d_cnt) (void)hipFree(d_cnt);
        if (d_tot) (void)hipFree(d_tot);
    };
    if ((rc = dev_alloc_n(&d_cb, ncol, "column batches")) ||
        (rc = dev_alloc_n(&d_cs, ncol, "column slots")) ||
        (rc = dev_alloc_n(&d_cnt, nbatch * nvc, "segment counts")) ||
        (rc = dev_alloc_n(&d_tot, nbatch, "batch totals")) ||
        (rc = dev_alloc_n(&r->bt_vc, nbatch * nvc, "segment index")) ||
        (rc = dev_alloc_n(&r->bt_off, nbatch, "batch offsets")) ||
        (rc = dev_alloc_n(&r->bt_ent, std::max<int64_t>(1, r->ssm_fov_entries),
                          "transposed entries")) ||
        (rc = dev_alloc_n(&r->bt_val, std::max<int64_t>(1, r->ssm_fov_entries),
                          "transposed values")) ||
        (rc = dev_alloc_n(&r->si_half, g->nz, "slice view offsets"))) {
        cleanup();
        return rc;
    }
    std::vector<int64_t> sih(g->nz);
    for (int64_t k = 0; k < g->nz; k++)
        sih[k] = r->slice_off_h[k] / (g->n_channels * g->n_rows);
    r->h_si_half = sih;
    std::vector<int32_t> klo, khi;
    view_slice_ranges(g, sih, &klo, &khi);
    if ((rc = dev_alloc_n(&r->k_lo, g->readings, "view slice ranges")) ||
        (rc = dev_alloc_n(&r->k_hi, g->readings, "view slice ranges"))) {
        cleanup();
        return rc;
    }
    hipError_t e = hipMemcpy(d_cb, cb.data(), ncol * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(r->k_lo, klo.data(), g->readings * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(r->k_hi, khi.data(), g->readings * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_cs, cs.data(), ncol * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(r->si_half, sih.data(), g->nz * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(d_cnt, 0, nbatch * nvc * 4);
    if (e == hipSuccess) {
        const int grid = (int)std::min<int64_t>(std::max<int64_t>(1, r->ncols_fov), 65536);
        hipLaunchKernelGGL(bt_count_kernel, dim3(grid), dim3(256), 0, 0, m->col_ptr,
                           m->idx, r->cols, r->ncols_fov, d_cb, g->n_rows, nvc, d_cnt);
        hipLaunchKernelGGL(bt_scan_kernel, dim3((unsigned)nbatch), dim3(1024), 0, 0,
                           d_cnt, r->bt_vc, d_tot, nvc);
        e = hipGetLastError();
    }
    std::vector<int64_t> tot(nbatch), off(nbatch);
    if (e == hipSuccess)
        e = hipMemcpy(tot.data(), d_tot, nbatch * 8, hipMemcpyDeviceToHost);
    int64_t run = 0;
    for (int64_t b = 0; b < nbatch; b++) {
        off[b] = run;
        run += tot[b];
    }
    if (e == hipSuccess && run != r->ssm_fov_entries)
        e = hipErrorUnknown;
    if (e == hipSuccess) e = hipMemcpy(r->bt_off, off.data(), nbatch * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(d_cnt, 0, nbatch * nvc * 4);
    if (e == hipSuccess) {
        const int grid = (int)std::min<int64_t>(std::max<int64_t>(1, r->ncols_fov), 65536);
        hipLaunchKernelGGL(bt_fill_kernel, dim3(grid), dim3(256), 0, 0, m->col_ptr,
                           m->idx, m->val, m->val16, m->col_scale, r->cols,
                           r->ncols_fov, d_cb, d_cs, g->n_rows, g->n_channels, nvc,
                           r->bt_vc, r->bt_off, d_cnt, r->bt_ent, r->bt_val);
        hipLaunchKernelGGL(bt_sort_kernel, dim3(16384), dim3(256), 0, 0, r->bt_vc,
                           r->bt_off, nbatch, nvc, r->bt_ent, r->bt_val);
        e = hipGetLastError();
    }
    // runs of one sample padded to multiples of kQuad entries
    if (e == hipSuccess) e = hipMemset(d_cnt, 0, nbatch * nvc * 4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(bt_pad_count_kernel, dim3(16384), dim3(256), 0, 0, r->bt_vc,
                           r->bt_off, nbatch, nvc, r->bt_ent, d_cnt);
        e = hipGetLastError();
    }
    int32_t *pvc = nullptr;
    int64_t *poff = nullptr;
    uint32_t *pent = nullptr;
    float *pval = nullptr;
    auto free_pad = [&]() {
        if (pvc) (void)hipFree(pvc);
        if (poff) (void)hipFree(poff);
        if (pent) (void)hipFree(pent);
        if (pval) (void)hipFree(pval);
    };
    if (e == hipSuccess && ((rc = dev_alloc_n(&pvc, nbatch * nvc, "segment index")) ||
                            (rc = dev_alloc_n(&poff, nbatch, "batch offsets")))) {
        free_pad();
        cleanup();
        return rc;
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(bt_scan_kernel, dim3((unsigned)nbatch), dim3(1024), 0, 0,
                           d_cnt, pvc, d_tot, nvc);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpy(tot.data(), d_tot, nbatch * 8, hipMemcpyDeviceToHost);
    run = 0;
    for (int64_t b = 0; b < nbatch; b++) {
        off[b] = run;
        run += tot[b];
    }
    if (e == hipSuccess && run >= ((int64_t)1 << 31) * kQuad)
        e = hipErrorInvalidValue;
    if (e == hipSuccess &&
        ((rc = dev_alloc_n(&pent, std::max<int64_t>(1, run), "transposed entries")) ||
         (rc = dev_alloc_n(&pval, std::max<int64_t>(1, run), "transposed values")))) {
        free_pad();
        cleanup();
        return rc;
    }
    if (e == hipSuccess) e = hipMemcpy(poff, off.data(), nbatch * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(bt_pad_fill_kernel, dim3(16384), dim3(256), 0, 0, r->bt_vc,
                           r->bt_off, nbatch, nvc, r->bt_ent, r->bt_val, pvc, poff, pent,
                           pval);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    if (e == hipSuccess) {
        (void)hipFree(r->bt_vc);
        (void)hipFree(r->bt_off);
        (void)hipFree(r->bt_ent);
        (void)hipFree(r->bt_val);
        r->bt_vc = pvc;
        r->bt_off = poff;
        r->bt_ent = pent;
        r->bt_val = pval;
    } else {
        free_pad();
    }
    cleanup();
    if (e != hipSuccess)
        return fail(CTBB_E_HIP, std::string("transposed batches: ") + hipGetErrorString(e));
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
    CTBB_TRY(check_neighbours(&m->g, opts->neighbours));
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
    r->fp64 = opts->state_fp64 != 0;
    r->sb = r->fp64 ? 8 : 4;
    auto bail = [&](int rc) {
        ctbb_recon_free(r);
        return rc;
    };
    hipError_t he = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
    if (he != hipSuccess)
        return bail(fail(CTBB_E_HIP, std::string("stream: ") + hipGetErrorString(he)));
    r->ev_step.assign(4 * kTimedSteps, nullptr);
    for (hipEvent_t *ev : {&r->ev_sweep[0], &r->ev_sweep[1]})
        if ((he = hipEventCreate(ev)) != hipSuccess)
            return bail(fail(CTBB_E_HIP, std::string("events: ") + hipGetErrorString(he)));
    for (auto &ev : r->ev_step)
        if ((he = hipEventCreate(&ev)) != hipSuccess)
            return bail(fail(CTBB_E_HIP, std::string("events: ") + hipGetErrorString(he)));

    // Columns inside the acquisition FOV (src/icd_iteration.cpp:154-181),
    // grouped into batches: the image is cut into tile x tile blocks of
    // columns and batch (ly, lx) holds the column at local position (lx, ly)
    // of every tile, so the columns of a batch are at least `tile` columns
    // apart (never penalty neighbours).  Batches run in raster order of
    // (ly, lx); with tile >= max(nx, ny) every batch is a single column and
    // the sweep is the reference's raster order.
    // default: tile 4, or the smallest tile whose batches hold at most 2^14
    // columns (the slot width of the transposed copy)
    int64_t tile = opts->tile;
    if (tile <= 0) {
        tile = 4;
        while (((g->nx + tile - 1) / tile) * ((g->ny + tile - 1) / tile) > (1 << 14))
            tile++;
    }
    if (tile < 2)
        return bail(fail(CTBB_E_ARG, "tile must be >= 2"));
    // every tile of at least max(nx, ny) gives the same single-column
    // batches in raster order; the clamp bounds the tile x tile loop below
    tile = std::min<int64_t>(tile, std::max<int64_t>({2, g->nx, g->ny}));
    const double fov_limit = (g->acquisition_fov / 2.0) * (g->acquisition_fov / 2.0);
    auto in_fov = [&](int64_t i, int64_t j) {
        const double y = (j - g->center_voxel_y) * g->voxel_size_y;
        const double x = (i - g->center_voxel_x) * g->voxel_size_x;
        return (x * x + y * y) < fov_limit;
    };
    std::vector<int32_t> cols;
    std::vector<int64_t> bptr(1, 0);
    const int64_t ntx = (g->nx + tile - 1) / tile, nty = (g->ny + tile - 1) / tile;
    // inside a batch the columns (its slots) follow the Z-order curve over
    // the tile grid: the columns along any ray through the image then have
    // nearby slots, so the scatter's reads of their updates share cache
    // lines (a batch is one Jacobi step: the order inside it does not change
    // the result)
    std::vector<std::pair<uint64_t, int64_t>> zt;
    zt.reserve(ntx * nty);
    for (int64_t ty = 0; ty < nty; ty++)
        for (int64_t tx = 0; tx < ntx; tx++) {
            uint64_t z = 0;
            for (int bit = 0; bit < 32; bit++)
                z |= (((uint64_t)tx >> bit) & 1u) << (2 * bit) |
                     (((uint64_t)ty >> bit) & 1u) << (2 * bit + 1);
            zt.emplace_back(z, ty * ntx + tx);
        }
    std::sort(zt.begin(), zt.end());
    for (int64_t ly = 0; ly < tile; ly++)
        for (int64_t lx = 0; lx < tile; lx++) {
            const size_t before = cols.size();
            for (const auto &t : zt) {
                const int64_t ty = t.second / ntx, tx = t.second - ty * ntx;
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
    if (r->data_size + g->n_views_sm * g->n_channels * g->n_rows >= ((int64_t)1 << 31))
        return bail(fail(CTBB_E_UNSUPPORTED, "sinogram of 2^31 samples or more"));
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
        (rc = dev_alloc(&r->e, r->data_size * r->sb, "error sinogram")) ||
        (rc = dev_alloc(&r->x, r->nvox * r->sb, "volume")) ||
        (rc = dev_alloc_n(&r->slice_off, g->nz, "slice offsets")) ||
        (rc = dev_alloc_n(&r->cols, std::max<int64_t>(1, r->ncols_fov), "FOV columns")) ||
        (rc = dev_alloc_n(&r->partials, std::max<int64_t>(4096, project_grid(&r->g)),
                           "reduction partials")) ||
        (rc = dev_alloc_n(&r->batch_ptr, bptr.size(), "batch pointers")) ||
        (rc = dev_alloc(&r->du, (max_batch + 8) * g->nz * r->sb, "step updates")))
        return bail(rc);
    if (weights && ((rc = dev_alloc_n(&r->w, r->data_size, "weights")) ||
                    (rc = dev_alloc(&r->ew, r->data_size * r->sb, "weighted error"))))
        return bail(rc);
#define CTBB_HIP_R(call)                                                      \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess)                                                 \
            return bail(fail(CTBB_E_HIP, std::string(#call) + ": " +          \
                                             hipGetErrorString(e_)));         \
    } while (0)
    CTBB_HIP_R(hipMemcpy(r->y, raw, r->data_size * 4, hipMemcpyHostToDevice));
    // zero start: e = y (and w e)
    if (r->fp64)
        hipLaunchKernelGGL(widen_kernel<double>, dim3(red_grid(r->data_size)), dim3(kRedBlock),
                           0, 0, r->y, r->data_size, static_cast<double *>(r->e));
    else
        hipLaunchKernelGGL(widen_kernel<float>, dim3(red_grid(r->data_size)), dim3(kRedBlock),
                           0, 0, r->y, r->data_size, static_cast<float *>(r->e));
    CTBB_HIP_R(hipGetLastError());
    if (weights) {
        CTBB_HIP_R(hipMemcpy(r->w, weights, r->data_size * 4, hipMemcpyHostToDevice));
        if (r->fp64)
            hipLaunchKernelGGL(weighted_error_kernel<double>, dim3(red_grid(r->data_size)),
                               dim3(kRedBlock), 0, 0, static_cast<const double *>(r->e),
                               r->w, r->data_size, static_cast<double *>(r->ew));
        else
            hipLaunchKernelGGL(weighted_error_kernel<float>, dim3(red_grid(r->data_size)),
                               dim3(kRedBlock), 0, 0, static_cast<const float *>(r->e),
                               r->w, r->data_size, static_cast<float *>(r->ew));
        CTBB_HIP_R(hipGetLastError());
    }
    CTBB_HIP_R(hipMemset(r->x, 0, r->nvox * r->sb));
    CTBB_HIP_R(hipMemcpy(r->slice_off, off.data(), g->nz * 8, hipMemcpyHostToDevice));
    if ((rc = upload_shear(g, opts->neighbours, slice_index, &r->shear)))
        return bail(rc);
    r->slice_off_h = off;
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
    if ((rc = build_transposed_batches(r, cols, bptr)) != CTBB_OK)
        return bail(rc);
    if ((rc = build_gather_blocks(r, bptr)) != CTBB_OK)
        return bail(rc);
    // partials and theta2 cache
    {
        const int64_t ng_max = (g->nz + r->spacing - 1) / r->spacing;
        const int64_t np = std::max<int64_t>(1, ng_max * r->nib * max_batch);
        if ((rc = dev_alloc(&r->part1, np * r->sb, "partial sums")) ||
            (rc = dev_alloc(&r->part2, np * r->sb, "partial sums")) ||
            (rc = dev_alloc_n(&r->th2, r->nvox, "theta2")))
            return bail(rc);
        r->th2_valid = false;
        // batch surrogate sample sums (one step at a time, first sweep only):
        // zeroed once here, filled by the -1 scatter of a step and set back
        // to zero over the step's data views after it;
        // single-column batches keep the plain curvature (their conflict-free
        // slice groups share no sample, where S would equal a)
        const bool single = tile >= std::max(g->nx, g->ny);
        if (!opts->plain_curvature && !single) {
            const int64_t nd = (max_batch + 8) * g->nz;
            if ((rc = dev_alloc(&r->ssum, r->data_size * r->sb, "surrogate sample sums")) ||
                (rc = dev_alloc(&r->du_m1, nd * r->sb, "surrogate scatter weights")))
                return bail(rc);
            std::vector<float> m1(nd, -1.0f);
            std::vector<double> m1d(nd, -1.0);
            const void *hm = r->fp64 ? (const void *)m1d.data() : (const void *)m1.data();
            hipError_t he2 = hipMemset(r->ssum, 0, r->data_size * r->sb);
            if (he2 == hipSuccess)
                he2 = hipMemcpy(r->du_m1, hm, nd * r->sb, hipMemcpyHostToDevice);
            if (he2 != hipSuccess)
                return bail(fail(CTBB_E_HIP, std::string("surrogate set-up: ") +
                                                 hipGetErrorString(he2)));
        }
        // scatter dispatch stride: median view distance of consecutive members
        r->view_stride.assign(std::max<int64_t>(1, r->spacing), 1);
        for (int64_t j = 0; j < r->spacing && j < g->nz; j++) {
            std::vector<int64_t> d;
            for (int64_t k = j; k + r->spacing < g->nz; k += r->spacing)
                d.push_back(std::llabs(r->h_si_half[k + r->spacing] - r->h_si_half[k]));
            if (d.empty())
                continue;
            std::nth_element(d.begin(), d.begin() + d.size() / 2, d.end());
            r->view_stride[j] = std::max<int64_t>(1, d[d.size() / 2]);
        }
    }
    // ray-major copy for the FP64 projections of set_volume / cost / resync
    if ((rc = ensure_ray_major(m)) != CTBB_OK)
        return bail(rc);
    *out = r;
    return CTBB_OK;
}

// Host <-> device copies of state buffers of n values: `bytes_host` per
// host value (4: float, 8: double), converted where the state type differs.
static int state_upload(ctbb_recon *r, void *dst, const void *src, int64_t n, int host_b)
{
    if ((size_t)host_b == r->sb) {
        CTBB_HIP(hipMemcpyAsync(dst, src, n * r->sb, hipMemcpyHostToDevice, r->stream));
        CTBB_HIP(hipStreamSynchronize(r->stream));
        return CTBB_OK;
    }
    if (host_b == 4) {   // float -> double state
        std::vector<double> t(static_cast<const float *>(src),
                              static_cast<const float *>(src) + n);
        CTBB_HIP(hipMemcpyAsync(dst, t.data(), n * 8, hipMemcpyHostToDevice, r->stream));
        CTBB_HIP(hipStreamSynchronize(r->stream));
        return CTBB_OK;
    }
    std::vector<float> t(static_cast<const double *>(src),   // double -> float state
                         static_cast<const double *>(src) + n);
    CTBB_HIP(hipMemcpyAsync(dst, t.data(), n * 4, hipMemcpyHostToDevice, r->stream));
    CTBB_HIP(hipStreamSynchronize(r->stream));
    return CTBB_OK;
}

static int state_download(const ctbb_recon *r, void *dst, const void *src, int64_t n,
                          int host_b)
{
    if ((size_t)host_b == r->sb) {
        CTBB_HIP(hipMemcpyAsync(dst, src, n * r->sb, hipMemcpyDeviceToHost, r->stream));
        CTBB_HIP(hipStreamSynchronize(r->stream));
        return CTBB_OK;
    }
    if (host_b == 4) {   // double state -> float
        std::vector<double> t(n);
        CTBB_HIP(hipMemcpyAsync(t.data(), src, n * 8, hipMemcpyDeviceToHost, r->stream));
        CTBB_HIP(hipStreamSynchronize(r->stream));
        float *d = static_cast<float *>(dst);
        for (int64_t i = 0; i < n; i++)
            d[i] = (float)t[i];
        return CTBB_OK;
    }
    std::vector<float> t(n);   // float state -> double
    CTBB_HIP(hipMemcpyAsync(t.data(), src, n * 4, hipMemcpyDeviceToHost, r->stream));
    CTBB_HIP(hipStreamSynchronize(r->stream));
    double *d = static_cast<double *>(dst);
    for (int64_t i = 0; i < n; i++)
        d[i] = t[i];
    return CTBB_OK;
}

extern "C" int ctbb_recon_set_volume(ctbb_recon *r, const float *volume)
{
    if (!r || !volume)
        return fail(CTBB_E_ARG, "null argument");
    CTBB_TRY(state_upload(r, r->x, volume, r->nvox, 4));
    double d;
    return recon_data_term(r, true, &d);
}

extern "C" int ctbb_recon_set_volume64(ctbb_recon *r, const double *volume)
{
    if (!r || !volume)
        return fail(CTBB_E_ARG, "null argument");
    CTBB_TRY(state_upload(r, r->x, volume, r->nvox, 8));
    double d;
    return recon_data_term(r, true, &d);
}

extern "C" int ctbb_recon_get_volume(const ctbb_recon *r, float *volume)
{
    if (!r || !volume)
        return fail(CTBB_E_ARG, "null argument");
    return state_download(r, volume, r->x, r->nvox, 4);
}

extern "C" int ctbb_recon_get_volume64(const ctbb_recon *r, double *volume)
{
    if (!r || !volume)
        return fail(CTBB_E_ARG, "null argument");
    return state_download(r, volume, r->x, r->nvox, 8);
}

// Bring the error sinogram e up to date from w e (weighted sweeps update
// e only where w == 0).
static int refresh_error(const ctbb_recon *r)
{
    if (!r->w || !r->e_stale)
        return CTBB_OK;
    const int g = red_grid(r->data_size);
    if (r->fp64)
        hipLaunchKernelGGL(error_from_weighted_kernel<double>, dim3(g), dim3(kRedBlock), 0,
                           r->stream, static_cast<double *>(r->e),
                           static_cast<const double *>(r->ew), r->w, r->data_size);
    else
        hipLaunchKernelGGL(error_from_weighted_kernel<float>, dim3(g), dim3(kRedBlock), 0,
                           r->stream, static_cast<float *>(r->e),
                           static_cast<const float *>(r->ew), r->w, r->data_size);
    CTBB_HIP(hipGetLastError());
    r->e_stale = false;
    return CTBB_OK;
}

extern "C" int ctbb_recon_get_error(const ctbb_recon *r, float *e)
{
    if (!r || !e)
        return fail(CTBB_E_ARG, "null argument");
    CTBB_TRY(refresh_error(r));
    return state_download(r, e, r->e, r->data_size, 4);
}

extern "C" int ctbb_recon_resync(ctbb_recon *r)
{
    if (!r)
        return fail(CTBB_E_ARG, "null argument");
    double d;
    return recon_data_term(r, true, &d);
}

extern "C" int ctbb_cost(const ctbb_ssm *m, const float *raw, const float *weights,
                         const int64_t *slice_index, const ctbb_icd_opts *opts,
                         const float *volume, double *cost)
{
    if (!m || !raw || !slice_index || !opts || !volume || !cost)
        return fail(CTBB_E_ARG, "null argument");
    if (opts->penalty != CTBB_PENALTY_QUADRATIC &&
        opts->penalty != CTBB_PENALTY_EDGE_PRESERVING)
        return fail(CTBB_E_ARG, "unknown penalty (valid: quadratic, edge-preserving)");
    const ctbb_geom *g = &m->g;
    CTBB_TRY(check_neighbours(g, opts->neighbours));
    const int64_t nvox = g->nx * g->ny * g->nz, ds = data_size_of(g);
    for (int64_t k = 0; k < g->nz; k++)
        if (slice_index[k] < 0 || slice_index[k] >= g->readings)
            return fail(CTBB_E_ARG, "slice_index out of range");
    std::vector<int64_t> sih(g->nz);
    for (int64_t k = 0; k < g->nz; k++)
        sih[k] = slice_offset(g, slice_index[k]) / (g->n_channels * g->n_rows);
    std::vector<int32_t> klo, khi;
    view_slice_ranges(g, sih, &klo, &khi);
    float *d_x = nullptr, *d_y = nullptr, *d_w = nullptr;
    int64_t *d_sih = nullptr;
    int32_t *d_klo = nullptr, *d_khi = nullptr;
    double *d_part = nullptr, *d_sh = nullptr;
    const int64_t np = std::max<int64_t>(project_grid(g), red_grid(nvox));
    int rc = CTBB_OK;
    auto cleanup = [&]() {
        for (void *p : {(void *)d_x, (void *)d_y, (void *)d_w, (void *)d_sih, (void *)d_klo,
                        (void *)d_khi, (void *)d_part, (void *)d_sh})
            if (p) (void)hipFree(p);
    };
    if ((rc = dev_alloc_n(&d_x, nvox, "volume")) ||
        (rc = dev_alloc_n(&d_y, ds, "sinogram")) ||
        (weights && (rc = dev_alloc_n(&d_w, ds, "weights"))) ||
        (rc = dev_alloc_n(&d_sih, g->nz, "slice view offsets")) ||
        (rc = dev_alloc_n(&d_klo, g->readings, "view slice ranges")) ||
        (rc = dev_alloc_n(&d_khi, g->readings, "view slice ranges")) ||
        (rc = dev_alloc_n(&d_part, np, "reduction partials")) ||
        (rc = upload_shear(g, opts->neighbours, slice_index, &d_sh))) {
        cleanup();
        return rc;
    }
    hipError_t e = hipMemcpy(d_x, volume, nvox * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_y, raw, ds * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && weights) e = hipMemcpy(d_w, weights, ds * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_sih, sih.data(), g->nz * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_klo, klo.data(), g->readings * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_khi, khi.data(), g->readings * 4, hipMemcpyHostToDevice);
    double data = 0.0, pen = 0.0;
    if (e == hipSuccess) {
        rc = launch_project<float, 1, float>(m, d_x, d_sih, d_klo, d_khi, nullptr, d_y, d_w, nullptr,
                                      nullptr, d_part, 0);
        if (rc == CTBB_OK)
            rc = sum_partials(d_part, (int)project_grid(g), 0, &data);
        if (rc == CTBB_OK) {
            const int gp = red_grid(nvox);
            hipLaunchKernelGGL(penalty_cost_kernel<float>, dim3(gp), dim3(kRedBlock), 0, 0, d_x,
                               g->nx, nvox, opts->penalty, opts->delta, d_part,
                               make_nb(g, d_sh));
            e = hipGetLastError();
            if (e == hipSuccess)
                rc = sum_partials(d_part, gp, 0, &pen);
        }
    }
    cleanup();
    if (rc != CTBB_OK)
        return rc;
    CTBB_HIP(e);
    if (opts->penalty == CTBB_PENALTY_QUADRATIC)
        *cost = 0.5 * data + 0.5 * opts->lambda * pen;
    else
        *cost = data + opts->lambda * pen;
    return CTBB_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free cost (ctbb_cost_noalloc).  Its scratch -- the slice and
// view tables of the projection, the per-workgroup partials of the data and
// penalty terms and the result slot -- lives in one file-scope device array
// per GPU, addressed through the pointer hipGetSymbolAddress returns on that
// GPU (the array's name is a host-side stand-in, never a device address).
// Every call on one GPU shares that array: calls that overlap on one GPU
// race; consecutive calls on one stream are ordered.

constexpr size_t kNoAllocScratch = size_t(8) << 20;   // bytes, per GPU
static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];
constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// Device address of this GPU's scratch array, looked up once per GPU (with
// that GPU current) on the first call there.
static int noalloc_scratch(void **out)
{
    int dev = 0;
    CTBB_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(CTBB_E_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(CTBB_E_HIP, std::string("scratch symbol address: ") +
                                        hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return CTBB_OK;
}

// Slice view offsets and, per data view, the slices whose view window holds
// it (the device form of view_slice_ranges).  One thread per data view;
// thread v < nz also writes si_half[v].
__global__ void noalloc_tables_kernel(const int64_t *slice_index, int64_t nz,
                                      int64_t readings, int64_t nsm, int64_t nffs,
                                      int ffs, int64_t *si_half, int32_t *klo,
                                      int32_t *khi)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    auto half = [&](int64_t k) {
        int64_t c = slice_index[k];
        if (ffs)
            c -= c % nffs;
        return c - nsm / 2;
    };
    if (v < nz)
        si_half[v] = half(v);
    if (v >= readings)
        return;
    int32_t lo = -1, hi = -1;
    for (int64_t k = 0; k < nz; k++) {
        const int64_t i = v - half(k);
        if (i >= 0 && i < nsm) {
            if (lo < 0)
                lo = (int32_t)k;
            hi = (int32_t)k + 1;
        }
    }
    klo[v] = lo < 0 ? 0 : lo;
    khi[v] = lo < 0 ? 0 : hi;
}

// The two partial arrays summed in block order by one thread (the order of
// sum_partials on the host), combined as in ctbb_cost.
__global__ void noalloc_finish_kernel(const double *pdata, int64_t ndata,
                                      const double *ppen, int64_t npen, int penalty,
                                      double lambda, double *cost)
{
    if (blockIdx.x != 0 || threadIdx.x != 0)
        return;
    double data = 0.0, pen = 0.0;
    for (int64_t i = 0; i < ndata; i++)
        data += pdata[i];
    for (int64_t i = 0; i < npen; i++)
        pen += ppen[i];
    *cost = penalty == CTBB_PENALTY_QUADRATIC ? 0.5 * data + 0.5 * lambda * pen
                                              : data + lambda * pen;
}

extern "C" int ctbb_cost_noalloc(const ctbb_ssm *m, const float *raw,
                                 const float *weights, const int64_t *slice_index,
                                 const ctbb_icd_opts *opts, const float *volume,
                                 double *cost, void *stream)
{
    if (!m || !raw || !slice_index || !opts || !volume || !cost)
        return fail(CTBB_E_ARG, "null argument");
    if (opts->penalty != CTBB_PENALTY_QUADRATIC &&
        opts->penalty != CTBB_PENALTY_EDGE_PRESERVING)
        return fail(CTBB_E_ARG, "unknown penalty (valid: quadratic, edge-preserving)");
    if (opts->neighbours != 8)
        return fail(CTBB_E_UNSUPPORTED, "the allocation-free cost has the 8-neighbour "
                                        "penalty only (use ctbb_cost)");
    const ctbb_geom *g = &m->g;
    const int64_t nvox = g->nx * g->ny * g->nz;
    const int64_t nproj = project_grid(g), npen = red_grid(nvox);
    // scratch layout: si_half [nz] | klo [readings] | khi [readings] |
    // data partials [nproj] | penalty partials [npen]
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t o_klo = up(8 * (size_t)g->nz);
    const size_t o_khi = o_klo + up(4 * (size_t)g->readings);
    const size_t o_pd = o_khi + up(4 * (size_t)g->readings);
    const size_t o_pp = o_pd + up(8 * (size_t)nproj);
    const size_t need = o_pp + up(8 * (size_t)npen);
    if (need > kNoAllocScratch)
        return fail(CTBB_E_ARG, "geometry needs " + std::to_string(need) +
                                    " bytes of scratch; the allocation-free cost has " +
                                    std::to_string(kNoAllocScratch) +
                                    " (readings x channel groups too large)");
    if (!m->rm_vc)
        return fail(CTBB_E_STATE, "the matrix has no ray-major copy yet: call ctbb_cost, "
                                  "ctbb_forward_project or ctbb_recon_create on it once "
                                  "(they build it) before the allocation-free cost");
    void *sp = nullptr;
    CTBB_TRY(noalloc_scratch(&sp));
    unsigned char *base = static_cast<unsigned char *>(sp);
    int64_t *d_sih = reinterpret_cast<int64_t *>(base);
    int32_t *d_klo = reinterpret_cast<int32_t *>(base + o_klo);
    int32_t *d_khi = reinterpret_cast<int32_t *>(base + o_khi);
    double *d_pd = reinterpret_cast<double *>(base + o_pd);
    double *d_pp = reinterpret_cast<double *>(base + o_pp);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t nt = std::max(g->readings, g->nz);
    hipLaunchKernelGGL(noalloc_tables_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256),
                       0, st, slice_index, g->nz, g->readings, g->n_views_sm, nffs_of(g),
                       (g->zffs || g->phiffs) ? 1 : 0, d_sih, d_klo, d_khi);
    CTBB_HIP(hipGetLastError());
    // every partial slot is written by its workgroup: no clearing needed
    CTBB_TRY((launch_project<float, 1, float>(m, volume, d_sih, d_klo, d_khi, nullptr, raw, weights,
                                       nullptr, nullptr, d_pd, st)));
    hipLaunchKernelGGL(penalty_cost_kernel<float>, dim3((unsigned)npen), dim3(kRedBlock), 0, st,
                       volume, g->nx, nvox, opts->penalty, opts->delta, d_pp,
                       make_nb(g, nullptr));
    hipLaunchKernelGGL(noalloc_finish_kernel, dim3(1), dim3(64), 0, st, d_pd, nproj, d_pp,
                       npen, opts->penalty, opts->lambda, cost);
    CTBB_HIP(hipGetLastError());
    return CTBB_OK;
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

extern "C" int ctbb_recon_cost_running(ctbb_recon *r, double *cost)
{
    if (!r || !cost)
        return fail(CTBB_E_ARG, "null argument");
    double data = 0.0, pen = 0.0;
    const int g = red_grid(r->data_size);
    CTBB_TRY(refresh_error(r));
    if (r->fp64)
        hipLaunchKernelGGL(error_sq_kernel<double>, dim3(g), dim3(kRedBlock), 0, r->stream,
                           static_cast<const double *>(r->e), r->w, r->data_size,
                           r->partials);
    else
        hipLaunchKernelGGL(error_sq_kernel<float>, dim3(g), dim3(kRedBlock), 0, r->stream,
                           static_cast<const float *>(r->e), r->w, r->data_size,
                           r->partials);
    CTBB_HIP(hipGetLastError());
    CTBB_TRY(sum_partials(r->partials, g, r->stream, &data));
    CTBB_TRY(recon_penalty_term(r, &pen));
    if (r->opts.penalty == CTBB_PENALTY_QUADRATIC)
        *cost = 0.5 * data + 0.5 * r->opts.lambda * pen;
    else
        *cost = data + r->opts.lambda * pen;
    return CTBB_OK;
}

template <typename T>
static int recon_iterate_t(ctbb_recon *r, int n, ctbb_icd_stats *stats)
{
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
    a.nb = make_nb(&r->g, r->shear);
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
    a.ew = r->w ? r->ew : r->e;
    a.th2 = r->th2;
    const bool fp16 = r->gb_val16 != nullptr;
    GatherArgs ga;
    ga.ent = r->gb_ent;
    ga.plane = r->g.n_rows * r->g.n_channels;
    ga.val = r->gb_val;
    ga.val16 = r->gb_val16;
    ga.perm = r->gb_perm;
    ga.ptr = r->gb_ptr;
    ga.part1 = r->part1;
    ga.part2 = r->part2;
    ga.nib = r->nib;
    ga.si_half = r->si_half;
    ga.ssum = r->ssum;
    ga.nsm = r->g.n_views_sm;
    ga.readings = r->g.readings;
    BtArgs bt;
    bt.ent = r->bt_ent;
    bt.val = r->bt_val;
    bt.vc_idx = r->bt_vc;
    bt.ent_off = r->bt_off;
    bt.si_half = r->si_half;
    bt.nsm = r->g.n_views_sm;
    bt.nch = r->g.n_channels;
    bt.nr = r->g.n_rows;
    bt.readings = r->g.readings;
    const int64_t ncg = (r->g.n_channels + kScatterCh - 1) / kScatterCh;
    const size_t scatter_lds = sizeof(T) * kScatterViews * kScatterCh * r->g.n_rows;

    hipEvent_t e0 = r->ev_sweep[0], e1 = r->ev_sweep[1];
    double sweep_ms = 0.0, resync_ms = 0.0;
    int64_t launches = 0;
    // Per-kernel durations: events around the three launches of every
    // `every`-th step (at most kTimedSteps steps per iteration).
    constexpr int kSamples = kTimedSteps;
    const int64_t steps_per_it = r->nbatch * std::min<int64_t>(r->spacing, r->g.nz);
    const int64_t every = std::max<int64_t>(1, steps_per_it / kSamples);
    const std::vector<hipEvent_t> &ev = r->ev_step;
    double g_ms = 0.0, v_ms = 0.0, s_ms = 0.0;
    int64_t n_samp = 0;
    for (int it = 0; it < n; it++) {
        // the first sweep also computes and caches theta2
        const bool first = !r->th2_valid;
        auto gather = fp16 ? (first ? icd_gather_kernel<T, true, true>
                                    : icd_gather_kernel<T, true, false>)
                           : (first ? icd_gather_kernel<T, false, true>
                                    : icd_gather_kernel<T, false, false>);
        auto solve = fp16 ? (first ? icd_solve_kernel<T, true, true>
                                   : icd_solve_kernel<T, true, false>)
                          : (first ? icd_solve_kernel<T, false, true>
                                   : icd_solve_kernel<T, false, false>);
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
                ga.base = r->nib * c0;
                ga.wbase = r->h_wbase[b];
                ga.nw = (nc + 63) / 64;
                ga.nsc = (ga.nw + kWaves - 1) / kWaves;
                // view blocks in residue classes of the view stride between
                // consecutive group members (in blocks): blocks ib and
                // ib + pchain -- dispatched together -- read the same data
                // views of the error sinogram for members m + 1 and m
                ga.pchain = std::min<int64_t>(
                    std::max<int64_t>(1, (r->view_stride[j] + kViewBlock / 2) / kViewBlock),
                    r->nib);
                ga.rows = (r->nib + ga.pchain - 1) / ga.pchain;
                const int64_t nwg = ga.pchain * ga.rows * ga.nsc;
                const int64_t ntask = kXcd * ((nwg + kXcd - 1) / kXcd);
                int64_t vlo = INT64_MAX, vhi = INT64_MIN;
                for (int64_t k = j; k < r->g.nz; k += r->spacing) {
                    vlo = std::min(vlo, r->h_si_half[k]);
                    vhi = std::max(vhi, r->h_si_half[k] + r->g.n_views_sm);
                }
                vlo = std::max<int64_t>(vlo, 0);
                vhi = std::min<int64_t>(vhi, r->g.readings);
                ScatterOrder so{};
                int64_t scatter_grid = 0;
                if (vhi > vlo) {
                    so.vlo = vlo;
                    so.vhi = vhi;
                    so.nq = (vhi - vlo + kScatterViews - 1) / kScatterViews;
                    so.ncg = ncg;
                    so.P = std::min<int64_t>(
                        std::max<int64_t>(1, (r->view_stride[j] + kScatterViews / 2) /
                                                 kScatterViews),
                        so.nq);
                    so.rows = (so.nq + so.P - 1) / so.P;
                    const int64_t nwg = so.P * so.rows * ncg;
                    scatter_grid = kXcd * ((nwg + kXcd - 1) / kXcd);
                }
                // first sweep with the batch surrogate: S = sum of a over every
                // voxel of the step (a scatter of -1 per voxel into the zeroed
                // S), stored as w S when the data term is weighted
                SweepArgs as = a;
                as.e = r->ssum;
                as.w = r->w;
                if (first && r->ssum && scatter_grid > 0) {
                    as.du = r->du_m1;
                    if (r->w)
                        hipLaunchKernelGGL((icd_scatter_kernel<T, true>),
                                           dim3((unsigned)scatter_grid), dim3(kSweepBlock),
                                           scatter_lds, r->stream, as, sa, bt, b, so);
                    else
                        hipLaunchKernelGGL((icd_scatter_kernel<T, false>),
                                           dim3((unsigned)scatter_grid), dim3(kSweepBlock),
                                           scatter_lds, r->stream, as, sa, bt, b, so);
                }
                const bool samp = (step % every) == 0 && used < kSamples;
                if (samp)
                    CTBB_HIP(hipEventRecord(ev[4 * used], r->stream));
                if (ntask > 0)
                    hipLaunchKernelGGL(gather, dim3((unsigned)ntask), dim3(kSweepBlock), 0,
                                       r->stream, a, sa, ga);
                if (samp)
                    CTBB_HIP(hipEventRecord(ev[4 * used + 1], r->stream));
                hipLaunchKernelGGL(solve, dim3((unsigned)(sa.ng * ((nc + 63) / 64))),
                                   dim3(kSweepBlock), 0, r->stream, a, sa, ga);
                if (samp)
                    CTBB_HIP(hipEventRecord(ev[4 * used + 2], r->stream));
                if (scatter_grid > 0)
                    hipLaunchKernelGGL(icd_scatter_kernel<T>, dim3((unsigned)scatter_grid),
                                       dim3(kSweepBlock), scatter_lds, r->stream, a, sa, bt, b,
                                       so);
                if (samp) {
                    CTBB_HIP(hipEventRecord(ev[4 * used + 3], r->stream));
                    used++;
                }
                // ... and S back to zero: the step touched data views [vlo, vhi) only
                if (first && r->ssum && scatter_grid > 0) {
                    const int64_t plane = r->g.n_channels * r->g.n_rows;
                    hipLaunchKernelGGL(zero_kernel<T>, dim3(8192), dim3(256), 0, r->stream,
                                       static_cast<T *>(r->ssum) + vlo * plane,
                                       (vhi - vlo) * plane);
                }
                launches += 3;
                step++;
            }
        }
        r->th2_valid = true;
        if (r->w)
            r->e_stale = true;
        CTBB_HIP(hipGetLastError());
        CTBB_HIP(hipEventRecord(e1, r->stream));
        CTBB_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        CTBB_HIP(hipEventElapsedTime(&ms, e0, e1));
        sweep_ms += ms;
        for (int64_t i = 0; i < used; i++) {
            float ga = 0.0f, so = 0.0f, sc = 0.0f;
            CTBB_HIP(hipEventElapsedTime(&ga, ev[4 * i], ev[4 * i + 1]));
            CTBB_HIP(hipEventElapsedTime(&so, ev[4 * i + 1], ev[4 * i + 2]));
            CTBB_HIP(hipEventElapsedTime(&sc, ev[4 * i + 2], ev[4 * i + 3]));
            g_ms += ga;
            v_ms += so;
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
    if (stats) {
        stats->gather_ms_avg = n_samp ? g_ms / n_samp : 0.0;
        stats->solve_ms_avg = n_samp ? v_ms / n_samp : 0.0;
        stats->scatter_ms_avg = n_samp ? s_ms / n_samp : 0.0;
        stats->iterations = n;
        stats->voxel_updates = (int64_t)n * r->ncols_fov * r->g.nz;
        stats->entry_visits = (int64_t)n * r->visits_per_sweep;
        stats->sweep_ms = sweep_ms;
        stats->resync_ms = resync_ms;
        stats->cost = NAN;
        stats->sweep_launches = launches;
        stats->ssm_entries = r->ssm_fov_entries;
        stats->tile = r->tile;
        stats->group_spacing = r->spacing;
    }
    return CTBB_OK;
}

extern "C" int ctbb_recon_iterate(ctbb_recon *r, int n, ctbb_icd_stats *stats)
{
    if (!r || n < 0)
        return fail(CTBB_E_ARG, "bad argument");
    return r->fp64 ? recon_iterate_t<double>(r, n, stats) : recon_iterate_t<float>(r, n, stats);
}
