/*
 * ctbb_icd.h -- C ABI of the MI355X (gfx950) reconstruction hot path.
 *
 * Plain C types only: pointers, sizes, status codes.  Every entry point is
 * synchronous, returns 0 on success and a negative CTBB_E* code on failure;
 * ctbb_last_error() then holds a message (thread-local).  The library never
 * frees caller memory.  Device-resident objects (stored system matrix,
 * reconstruction context) are opaque handles released explicitly.
 *
 * Each entry point names the reference function it stands in for
 * (paths relative to the CTBB_ICD source tree).
 */
#ifndef CTBB_ICD_H
#define CTBB_ICD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------- */
#define CTBB_OK            0
#define CTBB_E_ARG        -1   /* invalid argument / inconsistent sizes   */
#define CTBB_E_HIP        -2   /* HIP runtime error                       */
#define CTBB_E_NOMEM      -3   /* device allocation failed                */
#define CTBB_E_IO         -4   /* file could not be read / written        */
#define CTBB_E_UNSUPPORTED -5  /* geometry branch not implemented         */
#define CTBB_E_STATE      -6   /* call made in the wrong state            */

/* ---- geometry --------------------------------------------------------
 * POD mirror of the numeric fields of `struct recon_params`
 * (src/recon_structs.h:25-114) that the hot path reads, after the derived
 * block of configure_recon_params (src/setup.cpp:183-237) and the slice
 * table of load_raw (src/setup.cpp:349-409) have run.  Units: cm, radians.
 */
typedef struct ctbb_geom {
    int64_t nx, ny, nz;            /* num_voxels_x/y/z                      */
    int64_t n_channels;            /* n_channels                            */
    int64_t n_rows;                /* Nrows_projection                      */
    int64_t readings;              /* Readings (views in the sinogram)      */
    int64_t n_views_sm;            /* num_views_for_system_matrix           */
    int64_t zffs, phiffs;          /* Zffs, Phiffs                          */
    int64_t table_direction;       /* table_direction (-1 or +1)            */
    double tube_angle_increment;   /* radians per view                      */
    double tube_z_increment;       /* cm per view                           */
    double focal_spot_radius;      /* R                                     */
    double source_detector_distance; /* D                                   */
    double center_channel_non_ffs;
    double center_row;
    double transaxial_detector_spacing;
    double axial_detector_spacing;
    double fan_angle_increment;    /* radians                               */
    double anode_angle;            /* degrees                               */
    double coll_slicewidth;        /* CollSlicewidth                        */
    double center_voxel_x, center_voxel_y, center_voxel_z;
    double voxel_size_x, voxel_size_y, voxel_size_z;
    double acquisition_fov;        /* cm; ICD updates only x^2+y^2 < (fov/2)^2 */
} ctbb_geom;

/* ---- library ---------------------------------------------------------- */
const char *ctbb_last_error(void);
/* Select the HIP device used by every later call from this thread. */
int ctbb_init(int device);
/* Version string of this build. */
const char *ctbb_version(void);
/* Measurement helper (no counterpart in the reference): copies `bytes` bytes
 * between two fresh device buffers with a plain copy kernel
 * (ctbb_copy_kernel) and reports its duration, so that hardware byte
 * counters of other kernels can be calibrated against a known traffic. */
int ctbb_copy_probe(int64_t bytes, double *ms);

/* ---- stored system matrix (SSM) --------------------------------------
 * Replaces generate_system_matrix (src/generate_system_matrix.h:25,
 * src/generate_system_matrix.cpp:54-209) and generate_system_matrix_ffs
 * (src/generate_system_matrix_ffs.h:25, src/generate_system_matrix_ffs.cpp:
 * 60-290: in-plane, z or both flying focal spots, deflections of
 * generate_ffs_offset, :242-283).  The matrix is built on the
 * device in compressed-sparse-column form: column c = ix + nx*iy holds
 * strictly ascending ray indices q = row + n_rows*(channel + n_channels*view)
 * over SSM-relative views [0, n_views_sm), with FP32 values (explicit zeros
 * kept), exactly the entry set the reference pushes for that column.
 */
typedef struct ctbb_ssm ctbb_ssm;

#define CTBB_SSM_FP32 0
#define CTBB_SSM_FP16 1   /* FP16 values with one FP32 scale per column  */

int ctbb_ssm_build(const ctbb_geom *g, ctbb_ssm **out);
/* Convert the values of a built matrix to FP16 + per-column FP32 scale
 * (frees the FP32 values). */
int ctbb_ssm_to_fp16(ctbb_ssm *m);
int ctbb_ssm_free(ctbb_ssm *m);
/* nnz: stored entries; precision: CTBB_SSM_*; build_ms: device time of the
 * build kernels (count + scan + fill). */
int ctbb_ssm_info(const ctbb_ssm *m, int64_t *nnz, int *precision,
                  double *build_ms);
/* Copy the matrix to host arrays: col_ptr[nx*ny+1], idx[nnz], val[nnz]
 * (FP16 matrices are expanded to FP32 as scale*h). */
int ctbb_ssm_download(const ctbb_ssm *m, int64_t *col_ptr, int32_t *idx,
                      float *val);
/* Copy the columns [c0, c1) only: col_ptr[c1-c0+1] relative to the first of
 * them, idx/val sized col_ptr[c1-c0]; pass idx = val = NULL to size them. */
int ctbb_ssm_download_cols(const ctbb_ssm *m, int64_t c0, int64_t c1,
                           int64_t *col_ptr, int32_t *idx, float *val);
/* Upload a matrix produced elsewhere (e.g. read from a matrix.bin file,
 * src/generate_system_matrix.cpp:212-303 format). */
int ctbb_ssm_upload(const ctbb_geom *g, const int64_t *col_ptr,
                    const int32_t *idx, const float *val, ctbb_ssm **out);
/* Write / read the reference's matrix.bin layout: per column a uint64 nnz
 * followed by nnz x {int32 index, float32 value}
 * (src/generate_system_matrix.cpp:212-246; read at src/icd_iteration.cpp:
 * 164-179). */
int ctbb_ssm_write_file(const ctbb_ssm *m, const char *path);
int ctbb_ssm_read_file(const ctbb_geom *g, const char *path, ctbb_ssm **out);
/* Smallest slice spacing s >= 2 such that any two slices whose indices are
 * congruent mod s have disjoint sinogram footprints (exact per-view test of
 * the built matrix) in all but a fraction max_conflict_fraction of the
 * columns inside the acquisition FOV (0 = in every column).
 * slice_index[nz] = central view of each slice (load_raw slice table). */
int ctbb_ssm_group_spacing(const ctbb_ssm *m, const int64_t *slice_index,
                           int64_t max_spacing, double max_conflict_fraction,
                           int64_t *spacing);

/* ---- forward projection ------------------------------------------------
 * yhat[r] = sum over columns c, slices k, entries (i,a) with
 * r = i + (slice_index[k] - n_views_sm/2)*n_channels*n_rows in range of
 * x[c + nx*ny*k] * a, in FP64 (src/icd_iteration.cpp:63-116, including the
 * n_ffs alignment of the slice's central view at :90-94).
 * x: host FP64 volume [nx*ny*nz]; yhat: host FP64 [readings*n_channels*n_rows].
 */
int ctbb_forward_project(const ctbb_ssm *m, const int64_t *slice_index,
                         const double *x, double *yhat);

/* ---- filtered backprojection start image -------------------------------
 * Stands in for the external wFBP program of initialize_recon_volume
 * (src/initialize_recon_volume.cpp:82-178, wfbp_initialize != 0) followed
 * by rotate_slices_fixed2rotating: every rotated slice is reconstructed in
 * its own frame by an equiangular fan-beam FBP (ramp filter along channels,
 * cos(gamma) and FDK cone pre-weights) backprojected through the matrix
 * (footprint-weighted per view, 1/L^2 weighted; views one turn apart share
 * a source angle and are weighted so that every angle counts once).
 * raw: [readings*n_channels*n_rows] (reference layout); slice_index[nz];
 * volume: host FP32 [nx*ny*nz], rotating frame, zero outside the FOV. */
int ctbb_fbp(const ctbb_ssm *m, const float *raw, const int64_t *slice_index,
             float *volume);

/* ---- ICD reconstruction context ----------------------------------------
 * Replaces icd_iteration (src/icd_iteration.h:23, src/icd_iteration.cpp:
 * 42-286) with the penalties of src/penalties.cpp:39-154 fused in.
 */
typedef struct ctbb_recon ctbb_recon;

#define CTBB_PENALTY_QUADRATIC       0
#define CTBB_PENALTY_EDGE_PRESERVING 1


typedef struct ctbb_icd_opts {
    int32_t penalty;          /* CTBB_PENALTY_*                          */
    int32_t neighbours;       /* 8: the reference's in-plane stencil
                                 (src/penalties.cpp:39-59, flat-index
                                 offsets that wrap across rows and slices).
                                 26 (no reference counterpart; square
                                 slices): the 8 in-plane neighbours without
                                 wrap and 9 in each adjacent slice, found
                                 through the rotation between the slices'
                                 frames -- voxel p of slice k meets the 3x3
                                 block around pi_k(p) in slice k+1, pi_k the
                                 three-shear integer rotation by the tube
                                 angle between the slices' central views
                                 (a bijection of the grid); weights
                                 1/|offset| in voxel units.  The quadratic
                                 step divides by the weights of the
                                 neighbours present (the exact coordinate
                                 minimiser). */
    double lambda;            /* regulariser weight (prm `lambda`)       */
    double delta;             /* edge-preserving delta (prm `delta`)     */
    int64_t group_spacing;    /* s; 0 = ctbb_ssm_group_spacing with conflict
                                 fraction 0 for single-column batches and
                                 0.01 otherwise                          */
    int32_t resync_every;     /* FP64 e = y - A x resync period in
                                 iterations (0 = never)                  */
    int32_t tile;             /* batch tile: each step updates, in every
                                 column of one batch, one slice group
                                 {j, j+s, j+2s, ...}; a batch holds the
                                 column at the same local position of every
                                 tile x tile block of columns.  Inside a
                                 column the update is Gauss-Seidel (the
                                 group's footprints are disjoint); across the
                                 batch's columns it is Jacobi (all gather
                                 from the error sinogram as it stood at the
                                 start of the step).  0 = 4 (or the
                                 smallest tile with at most 16384 columns
                                 per batch); tile >=
                                 max(nx, ny) gives single-column batches,
                                 i.e. the reference's raster order. */
    int32_t plain_curvature;  /* 0 (default): batch surrogate -- the
                                 curvature of a voxel's update is
                                 sum_r w_r a_r S_r over its data samples r,
                                 S_r = sum over every voxel of the step
                                 (all the batch's columns, all the group's
                                 slices) of their values on sample r, which
                                 majorises the data term of the whole step
                                 separably (Cauchy-Schwarz), so every step
                                 lowers the cost however the step's voxels
                                 share samples -- across columns, across
                                 slices, and inside a column where the
                                 default spacing lets footprints overlap
                                 (no reference counterpart).  Single-column
                                 batches use the reference's sum_r w_r a_r^2
                                 (their default spacing is conflict-free).
                                 1: sum_r w_r a_r^2 in every batch (plain
                                 Jacobi across the step; can overshoot when
                                 the weighted data term dominates). */
    int32_t state_fp64;       /* 0 (default): FP32 sweep state (error
                                 sinogram, volume, updates, partial sums);
                                 1: FP64 state -- x and e = y - A x kept in
                                 FP64 like the reference's image and
                                 sinogram estimate (src/icd_iteration.cpp:
                                 47-48), for converged reference runs
                                 (twice the state bytes per sweep). */
} ctbb_icd_opts;

typedef struct ctbb_icd_stats {
    int64_t iterations;       /* iterations run by the last call          */
    int64_t voxel_updates;    /* voxel updates performed                  */
    int64_t entry_visits;     /* in-range SSM entries gathered+scattered  */
    double sweep_ms;          /* device time of the ICD sweep kernels     */
    double resync_ms;         /* device time of FP64 resyncs              */
    double cost;              /* cost after the last iteration (FP64)     */
    int64_t sweep_launches;   /* kernel launches of the sweeps (partial
                                 gather, solve and scatter per step)     */
    int64_t ssm_entries;      /* SSM entries of the FOV columns (each read
                                 once per column per sweep, algorithmically) */
    double gather_ms_avg;     /* average partial-gather launch duration
                                 (HIP events around sampled steps)       */
    double scatter_ms_avg;    /* average scatter-launch duration          */
    double solve_ms_avg;      /* average solve-launch duration            */
    int64_t tile;             /* batch tile in use                        */
    int64_t group_spacing;    /* slice spacing s of the groups in use     */
} ctbb_icd_stats;

/* raw: measured line integrals y [readings*n_channels*n_rows], reference
 * layout raw[row + n_rows*channel + n_rows*n_channels*view]
 * (src/setup.cpp:335-344).  weights: per-sample statistical weights of the
 * same layout, finite and >= 0 (e.g. detected counts; not checked), or NULL
 * (unweighted, as the reference).  slice_index[nz]: central view of every
 * slice (multiple of the FFS count). */
int ctbb_recon_create(const ctbb_ssm *m, const float *raw,
                      const float *weights, const int64_t *slice_index,
                      const ctbb_icd_opts *opts, ctbb_recon **out);
int ctbb_recon_free(ctbb_recon *r);
/* Load a start volume (rotated frame, float32 [nx*ny*nz]) and set the
 * error sinogram e = y - A x in FP64 (the reference forward-projects only
 * when wfbp_initialize is set, src/icd_iteration.cpp:63; this always does). */
int ctbb_recon_set_volume(ctbb_recon *r, const float *volume);
int ctbb_recon_get_volume(const ctbb_recon *r, float *volume);
/* FP64 forms of set/get volume (exact for the FP64 state; converted to and
 * from FP32 for the FP32 state). */
int ctbb_recon_set_volume64(ctbb_recon *r, const double *volume);
int ctbb_recon_get_volume64(const ctbb_recon *r, double *volume);
/* Run n ICD iterations (src/icd_iteration.cpp:149-274). */
int ctbb_recon_iterate(ctbb_recon *r, int n, ctbb_icd_stats *stats);
/* Recompute e = y - A x in FP64 and store it as FP32. */
int ctbb_recon_resync(ctbb_recon *r);
/* Global cost of the current volume in FP64 (fresh A x, not the running e):
 * quadratic:  0.5 sum w e^2 + 0.5 lambda sum_pairs w_d (x_q - x_{q+d})^2
 * edge-pres.:     sum w e^2 +     lambda sum_pairs w_d rho(x_q - x_{q+d})
 * rho(t) = |t| - delta ln(1 + |t|/delta); 8 neighbours: pairs d in
 * {+1, nx-1, nx, nx+1} with q+d < nx*ny*nz (the forward half of the flat
 * stencil, weights 1, 1/sqrt(2), 1, 1/sqrt(2)); 26 neighbours: every pair
 * of the neighbourhood once. */
int ctbb_recon_cost(ctbb_recon *r, double *cost);

/* Cost of the current volume from the kept FP32 error sinogram, without a
 * projection: sum w e^2 (FP64 sum of the FP32 samples) plus the penalty,
 * combined as in ctbb_recon_cost.  Equal to ctbb_recon_cost up to the FP32
 * rounding of e (right after ctbb_recon_set_volume or a resync, e is the
 * rounded FP64 projection residual); for cheap comparisons of candidate
 * volumes and for monitoring.  No reference counterpart. */
int ctbb_recon_cost_running(ctbb_recon *r, double *cost);
/* Copy the running error sinogram to the host (rounded to FP32).  Weighted
 * sweeps keep w e current and e itself only where w == 0; after them, this
 * call and ctbb_recon_cost_running first form e = w e / w where w > 0 (one
 * pass over the sinogram on the reconstruction's stream).  No reference
 * counterpart (the reference keeps yhat = A x, src/icd_iteration.cpp:47-48). */
int ctbb_recon_get_error(const ctbb_recon *r, float *e);

/* FP64 cost of a volume (rotated frame, float32 [nx*ny*nz]) for raw data,
 * weights (or NULL) and the penalty of opts, without a reconstruction
 * context: the same definition as ctbb_recon_cost.  The reference has no
 * global cost; this evaluates the one its updates minimise
 * (src/penalties.cpp:61-154 with src/icd_iteration.cpp:198-249). */
int ctbb_cost(const ctbb_ssm *m, const float *raw, const float *weights,
              const int64_t *slice_index, const ctbb_icd_opts *opts,
              const float *volume, double *cost);

/* Allocation-free ctbb_cost.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same cost as
 * ctbb_cost, bit for bit, but every pointer except m and opts is DEVICE
 * memory: raw and weights (or NULL) [readings*n_channels*n_rows], slice_index
 * [nz] (int64), volume [nx*ny*nz] (FP32, rotated frame), and the result
 * *cost (one FP64 value, written by the last kernel).  The work is enqueued
 * on `stream` (a hipStream_t; NULL = the null stream) and the call returns
 * without waiting.  Scratch comes from one static 8 MiB device array per
 * GPU: geometries needing more (about readings * ceil(n_channels/32) > 1e6)
 * are rejected with CTBB_E_ARG; the array is shared by every call on that
 * GPU and keeps the last call's contents, so calls that overlap on one GPU
 * (other streams, other threads) race -- order them.  The matrix needs its
 * ray-major copy, which ctbb_cost, ctbb_forward_project and
 * ctbb_recon_create build (allocating) on first use: before that this call
 * fails with CTBB_E_STATE.  It calls no hipMalloc or hipFree. */
int ctbb_cost_noalloc(const ctbb_ssm *m, const float *raw, const float *weights,
                      const int64_t *slice_index, const ctbb_icd_opts *opts,
                      const float *volume, double *cost, void *stream);

/* ---- slice rotation -----------------------------------------------------
 * Replaces rotate_slices_rotating2fixed / rotate_slices_fixed2rotating
 * (src/rotate_slices.h:25,29; src/rotate_slices.cpp:48-154 with interp2 of
 * src/interp.h:41-76).  direction = +1: rotating -> fixed (angle
 * -tube_angle[slice_index[k]] in radians), -1: fixed -> rotating.
 * tube_angles in degrees [readings].  volume: host float32 [nx*ny*nz],
 * rotated in place.  Requires nx == ny (the reference indexes the output
 * with ny as the x stride, src/rotate_slices.cpp:128). */
int ctbb_rotate_slices(const ctbb_geom *g, const double *tube_angles,
                       const int64_t *slice_index, float *volume,
                       int direction);

/* ---- penalty steps (for checking the fused device penalty) -------------
 * Evaluates the per-voxel penalty step of src/penalties.cpp:61-154 on the
 * device for n voxels q[i] of volume x (FP64 [nx*ny*nz]) with data terms
 * alpha[i], beta[i]; writes du[i] (before the positivity clamp).  The step
 * is the device function the sweep's solve kernel calls (quadratic closed
 * form; edge-preserving: the reference's bisection in FP64, tolerance 1e-8),
 * one thread per voxel, with the neighbourhood of opts->neighbours
 * (slice_index[nz] is needed for 26, may be NULL for 8).  volume_fp32 = 0:
 * on x as given (FP64); volume_fp32 = 1: on x rounded to FP32, the volume
 * the sweep holds. */
int ctbb_penalty_steps(const ctbb_geom *g, const ctbb_icd_opts *opts,
                       const int64_t *slice_index, int volume_fp32,
                       const double *x, int64_t n, const int64_t *q,
                       const double *alpha, const double *beta, double *du);

#ifdef __cplusplus
}
#endif

#endif /* CTBB_ICD_H */
