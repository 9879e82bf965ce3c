// main.cpp -- command-line driver with the reference's interface and call
// sequence (src/main.cpp:41-131): usage, --empty-prm, flags -t/-v/--timing,
// configure -> load_raw -> initialize -> system matrix (reused when the
// matrix file exists) -> ICD -> de-rotation -> clean_up.
#include <fstream>
#include <iostream>
#include <string>

#include "../../include/ctbb_icd.h"
#include "pipeline.hpp"

static void usage()
{
    std::cout << "usage: ctbb_icd [-t] [-v] [--timing] parameter_file.yaml\n"
                 "       ctbb_icd --empty-prm\n"
                 "Model-based iterative CT reconstruction (ICD) on an MI355X.\n";
    exit(0);
}

static void empty_parameter_file()
{
    std::cout
        << "# Empty parameter file for an ICD reconstruction\n"
           "# All units are CM, Radians, pixel units (unitless)\n"
           "\n"
           "# Paths\n"
           "sinogram_path:      # Required, raw projection data file\n"
           "output_dir:         # Required, output directory\n"
           "output_file:        # Required, final reconstruction file name\n"
           "initial_recon_path: # Optional, float32 start volume (rotated frame)\n"
           "matrix_path:        # Optional, existing stored matrix file\n"
           "\n"
           "# Raw data file information\n"
           "FileType:        # Required, 0 binary\n"
           "FileSubType:     # Required\n"
           "RawOffset:       # Required, byte offset of the projections\n"
           "Readings:        # Required, number of projections\n"
           "\n"
           "# Scanner geometry\n"
           "acquisition_fov:\nn_channels:\nnum_views_per_turn_without_ffs:\n"
           "focal_spot_radius:\nsource_detector_distance:\nanode_angle:\n"
           "axial_detector_spacing:\naxial_focal_spot_shift:\ncenter_channel_non_ffs:\n"
           "center_row:\ntransaxial_detector_spacing:\ntransaxial_focal_spot_shift:\n"
           "fan_angle_increment:\n"
           "\n"
           "# Scan specific parameters\n"
           "table_feed_per_rotation:\nZffs:\nPhiffs:\nNrows:\nCollSlicewidth:\n"
           "\n"
           "# Recon geometry\n"
           "recon_fov:\nnx:\nny:\nslice_thickness:\nz_start:\nz_end:\n"
           "\n"
           "# Iterative recon parameters\n"
           "wfbp_initialize:\npenalty:\nlambda:\ndelta:\nnum_iterations:\n"
           "\n"
           "# Optional (this implementation)\n"
           "synthetic:          # 1: generate seeded raw data instead of reading\n"
           "ssm_fp16:           # 1: FP16 matrix values with FP32 column scales\n"
           "resync_every:       # FP64 error-sinogram resync period (iterations)\n"
           "tile:               # batch tile of the parallel sweep (default 4)\n"
           "plain_curvature:    # 1: plain Jacobi curvature in batches (no surrogate)\n"
           "state_fp64:         # 1: FP64 image and error sinogram in the sweep\n"
        << std::endl;
    exit(0);
}

static bool exists(const std::string &name)
{
    std::ifstream f(name.c_str());
    return f.good();
}

int main(int argc, char **argv)
{
    if (argc < 2)
        usage();
    if (argc == 2 && std::string(argv[1]) == "--empty-prm")
        empty_parameter_file();
    for (int i = 1; i < argc - 1; i++) {
        const std::string a = argv[i];
        if (a != "-t" && a != "-v" && a != "--timing")
            usage();
    }
    struct ct_data data = {};
    std::string parameter_file = argv[argc - 1];
    // the prm is parsed and checked on the host first, so its messages come
    // out as the reference's do on any machine
    struct recon_params rp = configure_recon_params(parameter_file);
    rp.prm_path = parameter_file;
    if (ctbb_init(0) != CTBB_OK) {
        std::cout << "No GPU: " << ctbb_last_error() << std::endl;
        return 1;
    }
    load_raw(&rp, &data);
    const struct recon_params rp_const = rp;
    initialize_recon_volume(&rp_const, &data);
    if (!exists(rp_const.matrix_path)) {
        std::cout << "No existing matrix file found for reconstruction." << std::endl;
        if (!rp_const.Zffs && !rp_const.Phiffs)
            generate_system_matrix(&rp_const, &data);
        else
            generate_system_matrix_ffs(&rp_const, &data);
    } else {
        std::cout << "Existing matrix file FOUND." << std::endl;
        std::cout << "Using matrix file: " << rp_const.matrix_path << std::endl;
        load_system_matrix(&rp_const, &data);
    }
    icd_iteration(&rp_const, &data);
    rotate_slices_rotating2fixed(&rp_const, &data);
    clean_up(&rp, &data);
    return 0;
}
