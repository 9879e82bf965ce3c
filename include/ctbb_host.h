/*
 * ctbb_host.h -- host-side C ABI: parameter file, derived geometry, slice
 * table and the seeded synthetic raw-data generator that stands in for the
 * unavailable raw-data reader library.
 */
#ifndef CTBB_HOST_H
#define CTBB_HOST_H

#include <stdint.h>
#include "ctbb_icd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Run parameters derived from the prm file besides the geometry. */
typedef struct ctbb_run_params {
    int64_t num_iterations;
    int32_t penalty;            /* CTBB_PENALTY_*                              */
    int32_t wfbp_initialize;
    double lambda, delta;
    int64_t views_per_slice;
    int64_t num_views_per_turn;
    double z_first_view, z_start, z_end;
    double beam_width, beam_width_at_roi_edge, recon_fov;
    /* extensions (optional prm keys) */
    int32_t synthetic;          /* raw data generated, not read             */
    int32_t use_weights;        /* statistical weights = detected counts     */
    int32_t ssm_fp16;
    int32_t neighbours;
    int32_t resync_every;
    int32_t state_fp64;         /* FP64 sweep state (ctbb_icd_opts)         */
    int64_t group_spacing;
} ctbb_run_params;

/* Printing of the prm parse / geometry report (the reference prints to
 * stdout, src/parse_config.cpp:27-32, src/setup.cpp:197-240).  Default on. */
void ctbb_set_verbose(int on);

/* Parse and validate a prm file (configure_recon_params,
 * src/setup.cpp:69-238) and build the FileType-0 slice table
 * (src/setup.cpp:271-284, 349-409).  Any of the output arrays may be NULL:
 * tube_angles / table_positions [readings] (degrees / cm),
 * slice_locations [nz] (cm), slice_index [nz] (central view per slice). */
int ctbb_configure(const char *prm_path, ctbb_geom *g, ctbb_run_params *p,
                   double *tube_angles, double *table_positions,
                   double *slice_locations, int64_t *slice_index);

/* Synthetic raw data for the prm's scan (stands in for load_raw's reader,
 * src/setup.cpp:245-345): a seeded ellipsoid phantom (water cylinder +
 * ellipsoids [+ small high-contrast ellipsoids] [+ textured lung region]),
 * analytic line integrals along every ray of the scan with the SSM's ray
 * model, optional Poisson noise at the prm's I0 with y = ln(I0/max(N,1)),
 * weights = N.  raw / weights: [readings*n_channels*n_rows], reference
 * layout; weights may be NULL.  use_gpu: 1 = HIP kernel, 0 = host OpenMP. */
int ctbb_synthesize(const char *prm_path, float *raw, float *weights,
                    int use_gpu);

/* Raw data of a FileType-0 (binary) scan: replaces the binary branch of
 * load_raw (src/setup.cpp:271-302, 335-344).  Reads prm `Readings` frames
 * of n_channels x Nrows float32 samples (rows fastest) from `sinogram_path`
 * after `RawOffset` header bytes, frames back to back; raw
 * [readings*n_channels*n_rows] in the reference layout (the reference's two
 * frame transposes cancel).  Other FileTypes (PTR, IMA: the ctbb_read
 * library) return CTBB_E_UNSUPPORTED; a missing or short file CTBB_E_IO. */
int ctbb_read_raw(const char *prm_path, float *raw);

/* Ground-truth phantom sampled at the voxel centres of the rotated frame
 * (slice k rotated by its central view's angle), float32 [nx*ny*nz]. */
int ctbb_phantom_volume(const char *prm_path, float *volume);

#ifdef __cplusplus
}
#endif
#endif
