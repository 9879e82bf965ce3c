// recon_structs.hpp -- host-side parameter and data structures.
//
// Same field names and meanings as `struct recon_params` / `struct ct_data`
// of the reference (src/recon_structs.h:25-125) so that the call sequence of
// the reference driver (src/main.cpp:89-128) reads the same against this
// implementation.  Extension fields (synthetic raw data, precision and
// schedule options) are appended; they are optional prm keys.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

struct recon_params {
    // Pertinent paths
    std::string initial_recon_path;
    std::string matrix_path;
    std::string output_dir;
    std::string output_file;
    std::string potential;
    std::string recon_path;
    std::string sinogram_path;

    // Scanner geometry
    double acquisition_fov;
    double anode_angle;
    double axial_detector_spacing;
    double axial_focal_spot_shift;
    double center_channel_non_ffs;
    double center_row;
    double focal_spot_radius;
    size_t n_channels;
    size_t num_views_per_turn;
    size_t num_views_per_turn_without_ffs;
    double source_detector_distance;
    double transaxial_detector_spacing;
    double transaxial_focal_spot_shift;
    double fan_angle_increment;

    // Iterative recon parameters
    std::string penalty;
    double lambda;
    double delta;
    size_t num_iterations;

    // Scan specific parameters
    size_t Nrows;
    double first_view_angle;
    size_t num_z_ffs;
    size_t num_phi_ffs;
    size_t num_views;
    double table_feed_per_rotation;
    double tube_angle_increment;
    double z_end;
    double z_first_view;
    double z_good_slice;
    double z_matrix_start;
    double z_start;
    int table_direction;

    // Recon geometry
    double fov_radius;
    double center_voxel_x;
    double center_voxel_y;
    double center_voxel_z;
    double dx;
    double dy;
    double dz;
    size_t nx;
    size_t ny;
    size_t nz;

    size_t FileType;
    size_t FileSubType;
    size_t RawOffset;
    size_t Readings;
    size_t Zffs;
    size_t Phiffs;
    double CollSlicewidth;
    size_t wfbp_initialize;

    // Derived values
    size_t Nrows_projection;
    size_t views_per_slice;
    double voxel_size_x;
    double voxel_size_y;
    double voxel_size_z;
    double tube_z_increment;
    double beam_width;
    double beam_width_at_roi_edge;
    size_t num_views_for_system_matrix;

    size_t num_voxels_x;
    size_t num_voxels_y;
    size_t num_voxels_z;

    double slice_thickness;
    double recon_fov;

    // ---- extensions (optional keys; absent keys keep these defaults) ----
    size_t synthetic;              // 1: generate raw data (no reader library)
    size_t synthetic_seed;
    double synthetic_I0;           // 0: noiseless
    size_t synthetic_weights;      // 1: statistical weights = detected counts
    double synthetic_cylinder_radius;
    double synthetic_cylinder_mu;
    size_t synthetic_n_ellipsoids;
    double synthetic_ellipsoid_radius;   // centres within this radius
    double synthetic_axis_min, synthetic_axis_max;
    double synthetic_mu_min, synthetic_mu_max;
    size_t synthetic_n_small;
    double synthetic_small_mu_min, synthetic_small_mu_max;
    size_t synthetic_lung;         // 1: add a textured lung-like region
    size_t ssm_fp16;               // 1: FP16 SSM values + FP32 column scale
    size_t neighbours;             // 8 (reference stencil)
    size_t resync_every;           // FP64 resync period in iterations
    size_t group_spacing;          // 0: derive from the SSM footprints
    size_t tile;                   // batch tile of the sweep (0: default)
    size_t plain_curvature;        // 1: no batch surrogate (ctbb_icd_opts)
    size_t state_fp64;             // 1: FP64 sweep state (ctbb_icd_opts)
    std::string prm_path;          // the parameter file itself
};

struct ct_data {
    float *system_matrix;
    float *raw;
    float *recon_volume;
    float *final_image_stack;
    double *tube_angles;
    double *table_positions;
    double *slice_locations;
    size_t *slice_indices;
    // extension: statistical weights (null when unweighted) and the
    // device-resident stored system matrix
    float *weights;
    struct ctbb_ssm *ssm;
};

// prm handling (src/parse_config.cpp, src/setup.cpp:69-238).  The *_checked
// variants return a CTBB_* status instead of exiting; the reference-named
// functions exit(1) on error exactly where the reference does.
int parse_config_checked(const std::string &config_file, recon_params *rp);
void parse_config(std::string config_file, struct recon_params *rp);
int configure_recon_params_checked(const std::string &filename,
                                   recon_params *rp);
struct recon_params configure_recon_params(std::string filename);

// Slice table of load_raw for FileType 0 (src/setup.cpp:271-284, 349-409).
int build_slice_table_checked(recon_params *rp, ct_data *data);
