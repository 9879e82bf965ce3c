// pipeline.cpp -- the reference's pipeline functions, same names and
// argument meaning, running the hot path on the GPU through the C ABI.
//
//   load_raw                       src/setup.cpp:245-420 (FileType-0 binary
//                                  frames, or seeded synthetic data; FileType-0
//                                  angle/table conventions)
//   initialize_recon_volume        src/initialize_recon_volume.cpp:51-76,180-183
//   generate_system_matrix[_ffs]   src/generate_system_matrix.cpp:54-209,
//                                  src/generate_system_matrix_ffs.cpp:60-239
//   icd_iteration                  src/icd_iteration.cpp:42-286
//   rotate_slices_*                src/rotate_slices.cpp:48-114
//   clean_up                       src/setup.cpp:422-453
//
// Errors print a message and exit(1), as the reference does.
#include "pipeline.hpp"

#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <vector>

#include "../../include/ctbb_host.h"
#include "../../include/ctbb_icd.h"

namespace {

bool exists(const std::string &name)
{
    std::ifstream f(name.c_str());
    return f.good();
}

bool is_dir(const std::string &path)
{
    struct stat st;
    return stat(path.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

[[noreturn]] void die(const std::string &what)
{
    std::cout << what << ": " << ctbb_last_error() << std::endl;
    exit(1);
}

ctbb_geom geom_of(const recon_params *rp)
{
    ctbb_geom g;
    memset(&g, 0, sizeof(g));
    g.nx = (int64_t)rp->num_voxels_x;
    g.ny = (int64_t)rp->num_voxels_y;
    g.nz = (int64_t)rp->num_voxels_z;
    g.n_channels = (int64_t)rp->n_channels;
    g.n_rows = (int64_t)rp->Nrows_projection;
    g.readings = (int64_t)rp->Readings;
    g.n_views_sm = (int64_t)rp->num_views_for_system_matrix;
    g.zffs = (int64_t)rp->Zffs;
    g.phiffs = (int64_t)rp->Phiffs;
    g.table_direction = rp->table_direction;
    g.tube_angle_increment = rp->tube_angle_increment;
    g.tube_z_increment = rp->tube_z_increment;
    g.focal_spot_radius = rp->focal_spot_radius;
    g.source_detector_distance = rp->source_detector_distance;
    g.center_channel_non_ffs = rp->center_channel_non_ffs;
    g.center_row = rp->center_row;
    g.transaxial_detector_spacing = rp->transaxial_detector_spacing;
    g.axial_detector_spacing = rp->axial_detector_spacing;
    g.fan_angle_increment = rp->fan_angle_increment;
    g.anode_angle = rp->anode_angle;
    g.coll_slicewidth = rp->CollSlicewidth;
    g.center_voxel_x = rp->center_voxel_x;
    g.center_voxel_y = rp->center_voxel_y;
    g.center_voxel_z = rp->center_voxel_z;
    g.voxel_size_x = rp->voxel_size_x;
    g.voxel_size_y = rp->voxel_size_y;
    g.voxel_size_z = rp->voxel_size_z;
    g.acquisition_fov = rp->acquisition_fov;
    return g;
}

std::vector<int64_t> slice_index_of(const recon_params *rp, const ct_data *d)
{
    std::vector<int64_t> si(rp->num_voxels_z);
    for (size_t k = 0; k < rp->num_voxels_z; k++)
        si[k] = (int64_t)d->slice_indices[k];
    return si;
}

}  // namespace

void load_raw(struct recon_params *rp, struct ct_data *data)
{
    if (!rp->synthetic) {
        std::cout << "Checking for raw file: " << rp->sinogram_path << std::endl;
        if (rp->FileType != 0) {
            std::cout << "Raw projection readers for FileType 1/4 (the ctbb_read "
                         "library) are not part of this build. Exiting"
                      << std::endl;
            exit(1);
        }
    }
    if (build_slice_table_checked(rp, data) != CTBB_OK)
        exit(1);
    const size_t n = rp->Readings * rp->n_channels * rp->Nrows_projection;
    data->raw = new float[n];
    const size_t nv = rp->num_voxels_x * rp->num_voxels_y * rp->num_voxels_z;
    data->recon_volume = new float[nv]();
    if (!rp->synthetic) {
        // binary frames, unweighted as in the reference
        data->weights = nullptr;
        if (ctbb_read_raw(rp->prm_path.c_str(), data->raw) != CTBB_OK)
            die("Raw data read failed");
        std::cout << "Read " << rp->Readings << " binary frames" << std::endl;
        return;
    }
    data->weights = rp->synthetic_weights ? new float[n] : nullptr;
    if (ctbb_synthesize(rp->prm_path.c_str(), data->raw, data->weights, 1) != CTBB_OK)
        die("Synthetic raw data generation failed");
    std::cout << "Generated synthetic raw data (" << rp->Readings << " views)"
              << std::endl;
}

void initialize_recon_volume(const struct recon_params *rp, struct ct_data *data)
{
    const size_t nv = rp->num_voxels_x * rp->num_voxels_y * rp->num_voxels_z;
    if (rp->initial_recon_path.compare("") != 0) {
        std::cout << "Loading initial reconstruction from file" << std::endl;
        if (!exists(rp->initial_recon_path)) {
            std::cout << "Specified initial reconstruction file not found. Exiting"
                      << std::endl;
            exit(1);
        }
        std::ifstream infile(rp->initial_recon_path, std::ios::binary);
        infile.read((char *)data->recon_volume, nv * sizeof(float));
        if (!infile) {
            std::cout << "File was not read properly." << std::endl;
            std::cout << "Could not initialize reconstruction volume from given "
                         "file. Exiting"
                      << std::endl;
            exit(1);
        }
    } else if (rp->wfbp_initialize) {
        // the external wFBP program is replaced by a GPU FBP through the
        // system matrix; it needs the matrix, so icd_iteration forms it
        std::cout << "Initial volume: GPU filtered backprojection (formed with the "
                     "system matrix before the first iteration)"
                  << std::endl;
    }
    // otherwise: zeros (allocation default)
}

static void generate(const struct recon_params *rp, struct ct_data *data)
{
    ctbb_geom g = geom_of(rp);
    ctbb_ssm *m = nullptr;
    if (ctbb_ssm_build(&g, &m) != CTBB_OK)
        die("System matrix generation failed");
    int64_t nnz = 0;
    double ms = 0.0;
    ctbb_ssm_info(m, &nnz, nullptr, &ms);
    std::cout << "System matrix: " << nnz << " entries, built in " << ms << " ms"
              << std::endl;
    const std::string path = rp->output_dir + "/matrix.bin";
    if (ctbb_ssm_write_file(m, path.c_str()) != CTBB_OK)
        std::cout << "Could not write " << path << ": " << ctbb_last_error()
                  << std::endl;
    data->ssm = m;
}

void generate_system_matrix(const struct recon_params *rp, struct ct_data *data)
{
    std::cout << "Generating system matrix (no flying focal spots)..." << std::endl;
    generate(rp, data);
}

void generate_system_matrix_ffs(const struct recon_params *rp, struct ct_data *data)
{
    std::cout << "Generating system matrix (with flying focal spots)..." << std::endl;
    generate(rp, data);
}

void load_system_matrix(const struct recon_params *rp, struct ct_data *data)
{
    ctbb_geom g = geom_of(rp);
    ctbb_ssm *m = nullptr;
    if (ctbb_ssm_read_file(&g, rp->matrix_path.c_str(), &m) != CTBB_OK)
        die("Could not read the matrix file");
    data->ssm = m;
}

void icd_iteration(const struct recon_params *rp, struct ct_data *data)
{
    std::vector<int64_t> si = slice_index_of(rp, data);
    if (rp->wfbp_initialize && rp->initial_recon_path.compare("") == 0) {
        if (ctbb_fbp(data->ssm, data->raw, si.data(), data->recon_volume) != CTBB_OK)
            die("FBP initialization failed");
        std::cout << "Initialized from the GPU FBP image" << std::endl;
    }
    if (rp->ssm_fp16 && ctbb_ssm_to_fp16(data->ssm) != CTBB_OK)
        die("FP16 conversion failed");
    ctbb_icd_opts o;
    memset(&o, 0, sizeof(o));
    o.penalty = rp->penalty == "edge-preserving" ? CTBB_PENALTY_EDGE_PRESERVING
                                                 : CTBB_PENALTY_QUADRATIC;
    o.neighbours = (int32_t)rp->neighbours;
    o.lambda = rp->lambda;
    o.delta = rp->delta;
    o.group_spacing = (int64_t)rp->group_spacing;
    o.resync_every = (int32_t)rp->resync_every;
    o.tile = (int32_t)rp->tile;
    o.plain_curvature = (int32_t)rp->plain_curvature;
    o.state_fp64 = (int32_t)rp->state_fp64;
    ctbb_recon *r = nullptr;
    if (ctbb_recon_create(data->ssm, data->raw, data->weights, si.data(), &o, &r) !=
        CTBB_OK)
        die("ICD set-up failed");
    if (ctbb_recon_set_volume(r, data->recon_volume) != CTBB_OK)
        die("ICD initial projection failed");
    const size_t nv = rp->num_voxels_x * rp->num_voxels_y * rp->num_voxels_z;
    const std::string dump_dir = rp->output_dir + "/reconstructions";
    const bool dump = is_dir(dump_dir);
    std::vector<double> xd(dump ? nv : 0);
    auto write_iterate = [&](size_t n) {
        if (!dump)
            return;
        if (ctbb_recon_get_volume(r, data->recon_volume) != CTBB_OK)
            die("volume download failed");
        for (size_t i = 0; i < nv; i++)
            xd[i] = data->recon_volume[i];
        std::ostringstream p;
        p << dump_dir << "/iteration" << n << ".rcn";
        std::ofstream f(p.str(), std::ios_base::binary);
        f.write((char *)xd.data(), nv * sizeof(double));
    };
    write_iterate(0);
    if (dump) {
        // the sinogram estimate A x0 in FP64 (src/icd_iteration.cpp:127-133).
        // It is always the projection of the start volume here; the
        // reference projects only for a wFBP start (:63) and otherwise
        // writes zeros, also for a start volume read from a file.
        const size_t ds = rp->Readings * rp->n_channels * rp->Nrows_projection;
        std::vector<double> yhat(ds);
        if (ctbb_forward_project(data->ssm, si.data(), xd.data(), yhat.data()) != CTBB_OK)
            die("sinogram estimate failed");
        std::ofstream f(dump_dir + "/sino_estimation.rcn", std::ios_base::binary);
        f.write((char *)yhat.data(), ds * sizeof(double));
        std::cout << "Wrote initial sinogram to disk." << std::endl;
    }
    for (size_t n = 0; n < rp->num_iterations; n++) {
        std::cout << "Iteration #" << n + 1 << std::endl;
        ctbb_icd_stats st;
        if (ctbb_recon_iterate(r, 1, &st) != CTBB_OK)
            die("ICD iteration failed");
        double cost = 0.0;
        ctbb_recon_cost(r, &cost);
        std::cout << st.sweep_ms / 1000.0 << " s, cost " << cost << std::endl;
        write_iterate(n + 1);
    }
    if (ctbb_recon_get_volume(r, data->recon_volume) != CTBB_OK)
        die("volume download failed");
    ctbb_recon_free(r);
}

static void rotate(const struct recon_params *rp, struct ct_data *data, int dir)
{
    ctbb_geom g = geom_of(rp);
    std::vector<int64_t> si = slice_index_of(rp, data);
    if (ctbb_rotate_slices(&g, data->tube_angles, si.data(), data->recon_volume, dir) !=
        CTBB_OK)
        die("slice rotation failed");
}

void rotate_slices_fixed2rotating(const struct recon_params *rp, struct ct_data *data)
{
    std::cout << "Rotating WFBP slices to match ICD w/ rotating grid" << std::endl;
    rotate(rp, data, -1);
}

void rotate_slices_rotating2fixed(const struct recon_params *rp, struct ct_data *data)
{
    std::cout << "Rotating ICD slices into fixed cartesian grid" << std::endl;
    rotate(rp, data, 1);
}

void clean_up(struct recon_params *rp, struct ct_data *data)
{
    std::string output_filepath;
    if (rp->output_file != "")
        output_filepath = rp->output_dir + "/" + rp->output_file;
    else
        output_filepath = rp->output_dir + "/" + "reconstruction_final.img";
    std::cout << "Writing final reconstruction to: " << output_filepath << std::endl;
    std::ofstream f(output_filepath, std::ios_base::binary);
    f.write((char *)&data->recon_volume[0],
            rp->num_voxels_x * rp->num_voxels_y * rp->num_voxels_z * sizeof(float));
    f.close();
    delete[] data->tube_angles;
    delete[] data->table_positions;
    delete[] data->raw;
    delete[] data->weights;
    delete[] data->recon_volume;
    delete[] data->slice_locations;
    delete[] data->slice_indices;
    if (data->ssm)
        ctbb_ssm_free(data->ssm);
}
