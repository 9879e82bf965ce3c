// common.hpp -- internal definitions shared by the HIP translation units of
// libctbb_icd.so: error slot, HIP checks, device-resident handle layouts.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <cstdio>
#include <ostream>
#include <string>
#include <vector>

#include "../../include/ctbb_icd.h"

namespace ctbb {

// Thread-local last-error message (ctbb_last_error).
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
// Report stream (stdout, or a sink when ctbb_set_verbose(0)).
std::ostream &out();

#define CTBB_HIP(call)                                                        \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess)                                                 \
            return ::ctbb::fail(CTBB_E_HIP,                                   \
                                std::string(#call) + ": " +                   \
                                    hipGetErrorString(e_));                   \
    } while (0)

#define CTBB_TRY(call)                                                        \
    do {                                                                      \
        int rc_ = (call);                                                     \
        if (rc_ != CTBB_OK)                                                   \
            return rc_;                                                       \
    } while (0)

// Device allocation that records a clear message on failure.
int dev_alloc(void **p, size_t bytes, const char *what);

template <typename T>
int dev_alloc_n(T **p, size_t n, const char *what)
{
    return dev_alloc(reinterpret_cast<void **>(p), n * sizeof(T), what);
}

// Geometry checks shared by every entry point.
int check_geom(const ctbb_geom *g);

inline int64_t nffs_of(const ctbb_geom *g)
{
    return (int64_t(1) << g->zffs) * (int64_t(1) << g->phiffs);
}

inline int64_t data_size_of(const ctbb_geom *g)
{
    return g->readings * g->n_channels * g->n_rows;
}

// Sample offset of slice k: (central view - n_views_sm/2) * nch * nr, with
// the central view aligned down to a multiple of the FFS count
// (src/icd_iteration.cpp:88-96).
inline int64_t slice_offset(const ctbb_geom *g, int64_t central)
{
    int64_t n = nffs_of(g);
    if (g->zffs || g->phiffs)
        central -= central % n;
    return (central - g->n_views_sm / 2) * g->n_channels * g->n_rows;
}

// ---------------------------------------------------------------------------
// Wavefront run sums (segmented inclusive scans over lanes whose keys are
// sorted), shared by the sweep, projection and backprojection kernels.

// DPP lane moves: row_shr:n inside rows of 16 lanes (0 from outside the
// row), row_bcast:15 / row_bcast:31 (lane 15 of each row to the next row /
// lane 31 to rows 2 and 3), wave_shr:1 (lane l - 1; lane 0 keeps `old`).
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_mov(float v)
{
    return __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xf, false));
}

template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_mov(double v)
{
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROWS, 0xf,
                                               false);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// Inclusive sums over runs of equal keys (segmented scan): lane l adds the
// values of lanes [start, l], start = the last head at or before l.  Four
// in-row shifts, then the carries across rows (row 0 -> 1 and 2 -> 3, then
// rows 0-1 -> 2 and 3); steps no lane needs (every run shorter) are
// skipped, wave-uniformly.
template <typename T>
__device__ __forceinline__ T run_sum(T v, uint64_t head, int lane)
{
    const uint64_t upto = head & (~0ull >> (63 - lane));
    const int start = 63 - __clzll((unsigned long long)upto);
    const int dist = lane - start;
    if (!__ballot(dist >= 1))
        return v;
    T t;
    t = dpp_mov<0x111, 0xf>(v);
    if (dist >= 1) v += t;
    if (__ballot(dist >= 2)) {
        t = dpp_mov<0x112, 0xf>(v);
        if (dist >= 2) v += t;
        if (__ballot(dist >= 4)) {
            t = dpp_mov<0x114, 0xf>(v);
            if (dist >= 4) v += t;
            if (__ballot(dist >= 8)) {
                t = dpp_mov<0x118, 0xf>(v);
                if (dist >= 8) v += t;
            }
        }
    }
    // carries across rows: lanes whose run began in an earlier row
    const int row0 = lane & ~15;
    if (__ballot(start < row0)) {
        t = dpp_mov<0x142, 0xa>(v);
        if ((lane & 16) && start < row0) v += t;
        t = dpp_mov<0x143, 0xc>(v);
        if ((lane & 32) && start < 32) v += t;
    }
    return v;
}

// Head mask of a wavefront whose lanes hold sorted keys (invalid lanes, at
// the tail, carry key -1) and run-end test for lane `lane`.
__device__ __forceinline__ uint64_t run_heads(int key, int lane)
{
    const int prev = __builtin_amdgcn_update_dpp(-2, key, 0x138, 0xf, 0xf, false);
    return __ballot(lane == 0 || prev != key);
}

__device__ __forceinline__ bool run_end(uint64_t head, int lane)
{
    return lane == 63 || ((head >> (lane + 1)) & 1ull);
}

}  // namespace ctbb

// Stored system matrix, device resident, compressed sparse column.
struct ctbb_ssm {
    ctbb_geom g;
    int64_t ncol = 0;
    int64_t nnz = 0;
    int precision = CTBB_SSM_FP32;
    int64_t *col_ptr = nullptr;   // [ncol+1]
    int32_t *idx = nullptr;       // [nnz]
    float *val = nullptr;         // [nnz]   (FP32 matrices)
    __half *val16 = nullptr;      // [nnz]   (FP16 matrices)
    float *col_scale = nullptr;   // [ncol]  (FP16 matrices)
    int32_t *view_off = nullptr;  // [ncol*n_views_sm] entry offset of each
                                  // view inside its column (or null)
    double build_ms = 0.0;
    // ray-major copy, built on first use by the projection: entries grouped
    // by (view, channel), sorted by (row, column) inside a group
    int64_t *rm_vc = nullptr;     // [n_views_sm*n_channels + 1] group starts
    uint32_t *rm_ent = nullptr;   // [nnz] column | row << 24
    float *rm_val = nullptr;      // [nnz]
};

// Reconstruction context.
struct ctbb_recon {
    const ctbb_ssm *m = nullptr;
    ctbb_geom g;
    ctbb_icd_opts opts;
    int64_t data_size = 0;
    int64_t nvox = 0;
    float *y = nullptr;         // measured line integrals
    float *w = nullptr;         // weights or null
    // sweep state, FP32 or FP64 (state_fp64): e, ew, x, du, the partials,
    // the surrogate sums and the +-1 scatter weights hold `sb`-byte values
    bool fp64 = false;
    size_t sb = 4;
    void *e = nullptr;          // running error sinogram y - A x
    mutable bool e_stale = false;  // weighted sweeps ran since e was formed:
                                   // e is current only where w == 0
    void *x = nullptr;          // volume, rotated frame
    int64_t *slice_off = nullptr;  // [nz] sample offsets
    double *shear = nullptr;    // [2*(nz-1)] 26-neighbour slice rotations (or null)
    int32_t *cols = nullptr;    // FOV columns, grouped by batch
    int64_t *batch_ptr = nullptr;  // [nbatch+1] offsets into cols
    void *ew = nullptr;         // weighted error w e (weighted recons)
    void *du = nullptr;         // per-step voxel updates
    double *th2 = nullptr;      // theta2 (batch surrogate curvature) per voxel
    bool th2_valid = false;     //   [nvox], iteration-invariant, first sweep
    // batch surrogate: S[r] = sum over the voxels of one step (every column
    // of the batch, every slice of the group) of their matrix values on data
    // sample r; zero between steps (first sweep only)
    void *ssum = nullptr;         // [data_size] or null (plain curvature)
    void *du_m1 = nullptr;        // [(max_batch+8)*nz] all -1: scatter S += a
    // view-blocked batch copy of the matrix (gather): per batch b and view
    // block ib (kViewBlock views) the columns ranked by detector position,
    // in chunks of 64 ranks; a chunk's entries ordered (view in block, rank)
    uint32_t *gb_ent = nullptr;   // [fov entries] sample in view | key << 17
    float *gb_val = nullptr;      // [fov entries] FP32 value (FP32 matrices)
    __half *gb_val16 = nullptr;   // [fov entries] FP16 value (FP16 matrices)
    uint16_t *gb_perm = nullptr;  // [nib*ncols_fov] column slot of (b, ib, rank)
    int64_t *gb_ptr = nullptr;    // [chunks+1] start of (b, ib, 64-rank chunk)
    std::vector<int64_t> h_wbase; // [nbatch+1] first chunk of each batch
    int64_t nib = 0;              // view blocks
    std::vector<int64_t> view_stride;   // [spacing] scatter dispatch stride
    void *part1 = nullptr;        // per-step partials [ng][nib][max_batch]
    void *part2 = nullptr;        //   (theta1, theta2)
    int64_t ncols_fov = 0;
    int64_t nbatch = 0;
    int64_t max_batch = 0;
    int64_t tile = 0;
    std::vector<int64_t> h_batch_ptr;
    std::vector<int64_t> slice_off_h;   // [nz] sample offsets (host copy)
    std::vector<int64_t> h_si_half;     // [nz] central view - n_views_sm/2
    // transposed batches (sample-driven scatter)
    uint32_t *bt_ent = nullptr;   // packed slot | lkey << 14 | channel group << 26
    float *bt_val = nullptr;
    int32_t *bt_vc = nullptr;     // [nbatch][n_views_sm*n_channels+1]
    int64_t *bt_off = nullptr;    // [nbatch]
    int64_t *si_half = nullptr;   // [nz] device copy of h_si_half
    int32_t *k_lo = nullptr;      // [readings] slices whose view window holds
    int32_t *k_hi = nullptr;      //   the view: [k_lo, k_hi)
    int64_t visits_per_sweep = 0;
    int64_t ssm_fov_entries = 0;
    int64_t spacing = 0;
    double *partials = nullptr; // reduction partials
    int iter_done = 0;
    hipStream_t stream = nullptr;
    // timing events of ctbb_recon_iterate (created once with the context):
    // sweep start/end and, per sampled step, around gather / solve / scatter
    hipEvent_t ev_sweep[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> ev_step;
    std::vector<int64_t> h_slice_index;
};
