// prm.cpp -- YAML parameter file, validation, derived geometry and the
// FileType-0 slice table.
//
// Restates src/parse_config.cpp:27-111 (keys, types, FOUND / NOT FOUND
// printout), src/setup.cpp:69-238 (checks, defaults, derived values) and
// src/setup.cpp:271-284, 349-409 (angles, table positions, allowed z range,
// slice locations and central views).  yaml-cpp is not available here, so the
// parser is a flat `key: value  # comment` reader with yaml-cpp's scalar
// conversion rules: a number must consume the whole scalar, unsigned values
// reject a leading '-', integers accept 0x / leading-0 prefixes (yaml-cpp
// clears std::ios::dec), `.inf` / `.nan` are floats.  An empty value (YAML
// null) is reported NOT FOUND for every type.
#include "recon_structs.hpp"

#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>

#include "../../include/ctbb_icd.h"
#include "../../include/ctbb_host.h"

namespace ctbb {
int fail(int code, const std::string &msg);
std::ostream &out();
}  // namespace ctbb

namespace {

constexpr double kPiSetup = 3.141592653589793238;  // src/setup.cpp:35

struct Flat {
    std::map<std::string, std::string> kv;   // present, non-null keys
};

std::string trim(const std::string &s)
{
    size_t b = s.find_first_not_of(" \t\r\n");
    if (b == std::string::npos)
        return "";
    size_t e = s.find_last_not_of(" \t\r\n");
    return s.substr(b, e - b + 1);
}

bool load_flat(const std::string &path, Flat *f)
{
    std::ifstream in(path);
    if (!in.good())
        return false;
    std::string line;
    while (std::getline(in, line)) {
        // comment: '#' at line start or after whitespace, outside quotes
        bool sq = false, dq = false;
        for (size_t i = 0; i < line.size(); i++) {
            char c = line[i];
            if (c == '\'' && !dq) sq = !sq;
            else if (c == '"' && !sq) dq = !dq;
            else if (c == '#' && !sq && !dq &&
                     (i == 0 || line[i - 1] == ' ' || line[i - 1] == '\t')) {
                line.resize(i);
                break;
            }
        }
        if (trim(line).empty())
            continue;
        if (line[0] == ' ' || line[0] == '\t' || line[0] == '-')
            continue;  // nested or sequence content: not part of the flat prm
        size_t colon = line.find(':');
        if (colon == std::string::npos)
            continue;
        std::string key = trim(line.substr(0, colon));
        std::string val = trim(line.substr(colon + 1));
        if (val.size() >= 2 && ((val.front() == '"' && val.back() == '"') ||
                                (val.front() == '\'' && val.back() == '\''))) {
            val = val.substr(1, val.size() - 2);
        } else if (val.empty() || val == "~" || val == "null" || val == "Null" ||
                   val == "NULL") {
            f->kv.erase(key);
            continue;
        }
        f->kv[key] = val;
    }
    return true;
}

bool conv(const std::string &s, std::string *out)
{
    *out = s;
    return true;
}

bool conv(const std::string &s, double *out)
{
    if (s == ".inf" || s == ".Inf" || s == ".INF" || s == "+.inf") {
        *out = INFINITY;
        return true;
    }
    if (s == "-.inf" || s == "-.Inf" || s == "-.INF") {
        *out = -INFINITY;
        return true;
    }
    if (s == ".nan" || s == ".NaN" || s == ".NAN") {
        *out = NAN;
        return true;
    }
    if (s.empty() || isspace((unsigned char)s[0]))
        return false;
    char *end = nullptr;
    errno = 0;
    double v = strtod(s.c_str(), &end);
    if (end == s.c_str())
        return false;
    while (*end && isspace((unsigned char)*end))
        end++;
    if (*end)
        return false;
    *out = v;
    return true;
}

bool conv(const std::string &s, size_t *out)
{
    if (s.empty() || s[0] == '-' || isspace((unsigned char)s[0]))
        return false;
    char *end = nullptr;
    errno = 0;
    unsigned long long v = strtoull(s.c_str(), &end, 0);
    if (end == s.c_str() || errno == ERANGE)
        return false;
    while (*end && isspace((unsigned char)*end))
        end++;
    if (*end)
        return false;
    *out = (size_t)v;
    return true;
}

template <typename T>
void parse_item(const Flat &f, const char *name, T *dst, bool print)
{
    if (print) {
        ctbb::out() << name << ": ";
        for (int i = 0; i < 35 - (int)(strlen(name) + 1); i++)
            ctbb::out() << " ";
    }
    auto it = f.kv.find(name);
    T v{};
    if (it != f.kv.end() && conv(it->second, &v)) {
        *dst = v;
        if (print)
            ctbb::out() << "FOUND " << std::endl;
    } else if (print) {
        ctbb::out() << "NOT FOUND" << std::endl;
    }
}

}  // namespace

int parse_config_checked(const std::string &config_file, recon_params *rp)
{
    ctbb::out() << "Config File: " << config_file << std::endl;
    Flat f;
    if (!load_flat(config_file, &f)) {
        ctbb::out() << "Could not open configuration file " << config_file
                  << std::endl;
        return CTBB_E_IO;
    }
    ctbb::out() << std::endl;
#define P(name, T) parse_item<T>(f, #name, &rp->name, true)
    // Pertinent paths block (src/parse_config.cpp:41-47)
    P(initial_recon_path, std::string);
    P(matrix_path, std::string);
    P(output_dir, std::string);
    P(output_file, std::string);
    P(potential, std::string);
    P(recon_path, std::string);
    P(sinogram_path, std::string);
    // Scanner geometry (:50-62)
    P(acquisition_fov, double);
    P(anode_angle, double);
    P(axial_detector_spacing, double);
    P(axial_focal_spot_shift, double);
    P(center_channel_non_ffs, double);
    P(center_row, double);
    P(focal_spot_radius, double);
    P(n_channels, size_t);
    P(num_views_per_turn_without_ffs, size_t);
    P(source_detector_distance, double);
    P(transaxial_detector_spacing, double);
    P(transaxial_focal_spot_shift, double);
    P(fan_angle_increment, double);
    // Iterative recon parameters (:65-69)
    P(penalty, std::string);
    P(lambda, double);
    P(delta, double);
    P(num_iterations, size_t);
    P(num_views_for_system_matrix, size_t);
    // Scan specific parameters (:72-82)
    P(first_view_angle, double);
    P(num_z_ffs, size_t);
    P(num_phi_ffs, size_t);
    P(num_views, size_t);
    P(table_feed_per_rotation, double);
    P(tube_angle_increment, double);
    P(z_end, double);
    P(z_first_view, double);
    P(z_good_slice, double);
    P(z_matrix_start, double);
    P(z_start, double);
    // Recon geometry (:85-88)
    P(recon_fov, double);
    P(nx, size_t);
    P(ny, size_t);
    P(slice_thickness, double);
    // Additions (:91-99)
    P(FileType, size_t);
    P(FileSubType, size_t);
    P(RawOffset, size_t);
    P(Readings, size_t);
    P(Zffs, size_t);
    P(Phiffs, size_t);
    P(CollSlicewidth, double);
    P(Nrows, size_t);
    P(wfbp_initialize, size_t);
#undef P
    // Extension keys: parsed quietly so that reference prm files print and
    // parse exactly as before.
#define Q(name, T) parse_item<T>(f, #name, &rp->name, false)
    Q(synthetic, size_t);
    Q(synthetic_seed, size_t);
    Q(synthetic_I0, double);
    Q(synthetic_weights, size_t);
    Q(synthetic_cylinder_radius, double);
    Q(synthetic_cylinder_mu, double);
    Q(synthetic_n_ellipsoids, size_t);
    Q(synthetic_ellipsoid_radius, double);
    Q(synthetic_axis_min, double);
    Q(synthetic_axis_max, double);
    Q(synthetic_mu_min, double);
    Q(synthetic_mu_max, double);
    Q(synthetic_n_small, size_t);
    Q(synthetic_small_mu_min, double);
    Q(synthetic_small_mu_max, double);
    Q(synthetic_lung, size_t);
    Q(ssm_fp16, size_t);
    Q(neighbours, size_t);
    Q(resync_every, size_t);
    Q(group_spacing, size_t);
    Q(tile, size_t);
    Q(plain_curvature, size_t);
    Q(state_fp64, size_t);
#undef Q
    return CTBB_OK;
}

void parse_config(std::string config_file, struct recon_params *rp)
{
    if (parse_config_checked(config_file, rp) != CTBB_OK)
        exit(1);
}

static void set_extension_defaults(recon_params *rp)
{
    rp->synthetic = 0;
    rp->synthetic_seed = 1;
    rp->synthetic_I0 = 0.0;
    rp->synthetic_weights = 0;
    rp->synthetic_cylinder_radius = 10.0;
    rp->synthetic_cylinder_mu = 0.193;
    rp->synthetic_n_ellipsoids = 10;
    rp->synthetic_ellipsoid_radius = 8.0;
    rp->synthetic_axis_min = 0.5;
    rp->synthetic_axis_max = 4.0;
    rp->synthetic_mu_min = 0.15;
    rp->synthetic_mu_max = 0.25;
    rp->synthetic_n_small = 0;
    rp->synthetic_small_mu_min = 0.3;
    rp->synthetic_small_mu_max = 0.5;
    rp->synthetic_lung = 0;
    rp->ssm_fp16 = 0;
    rp->neighbours = 8;
    rp->resync_every = 0;
    rp->group_spacing = 0;
    rp->tile = 0;
    rp->plain_curvature = 0;
    rp->state_fp64 = 0;
}

int configure_recon_params_checked(const std::string &filename, recon_params *out)
{
    recon_params rp = {};
    set_extension_defaults(&rp);
    // Zero-allowed fields start at -1 (src/setup.cpp:91-96).
    rp.FileType = (size_t)-1;
    rp.FileSubType = (size_t)-1;
    rp.RawOffset = (size_t)-1;
    rp.Zffs = (size_t)-1;
    rp.Phiffs = (size_t)-1;
    rp.wfbp_initialize = (size_t)-1;

    {
        std::ifstream probe(filename.c_str());
        if (!probe.good()) {
            ctbb::out() << "Parameter file not found. Exiting" << std::endl;
            return CTBB_E_IO;
        }
    }
    int rc = parse_config_checked(filename, &rp);
    if (rc != CTBB_OK)
        return rc;

    bool exit_flag = false;
#define STRING_CHECK(N)                                                       \
    if (rp.N.compare("") == 0) {                                              \
        ctbb::out() << #N " not properly set in configuration. Check parameter file.\n"; \
        exit_flag = true;                                                     \
    }
#define DOUBLE_CHECK(N)                                                       \
    if (rp.N == 0.0) {                                                        \
        ctbb::out() << #N " not properly set in configuration. Check parameter file.\n"; \
        exit_flag = true;                                                     \
    }
#define SIZE_T_CHECK(N)                                                       \
    if (rp.N == (size_t)0) {                                                  \
        ctbb::out() << #N " not properly set in configuration. Check parameter file.\n"; \
        exit_flag = true;                                                     \
    }
#define ZERO_ALLOWED(N)                                                       \
    if (rp.N == (size_t)-1) {                                                 \
        ctbb::out() << #N " not properly set in configuration. Check parameter file.\n"; \
        exit_flag = true;                                                     \
    }
#define WITH_DEFAULT(N, V)                                                    \
    if (rp.N == 0.0) {                                                        \
        ctbb::out() << #N " was not explicitly set in parameter file. Using default value: " #V << std::endl; \
        rp.N = V;                                                             \
    }
#define STRING_WITH_DEFAULT(N, V)                                             \
    if (strcmp(rp.N.c_str(), "") == 0) {                                      \
        ctbb::out() << #N " was not explicitly set in parameter file. Using default value: " #V << std::endl; \
        rp.N = V;                                                             \
    }
    // src/setup.cpp:111-156
    STRING_CHECK(sinogram_path);
    STRING_CHECK(output_dir);
    STRING_CHECK(output_file);
    DOUBLE_CHECK(acquisition_fov);
    SIZE_T_CHECK(num_views_per_turn_without_ffs);
    DOUBLE_CHECK(focal_spot_radius);
    DOUBLE_CHECK(source_detector_distance);
    DOUBLE_CHECK(anode_angle);
    SIZE_T_CHECK(n_channels);
    SIZE_T_CHECK(Nrows);
    DOUBLE_CHECK(axial_detector_spacing);
    DOUBLE_CHECK(axial_focal_spot_shift);
    DOUBLE_CHECK(center_channel_non_ffs);
    DOUBLE_CHECK(center_row);
    DOUBLE_CHECK(transaxial_detector_spacing);
    DOUBLE_CHECK(transaxial_focal_spot_shift);
    DOUBLE_CHECK(fan_angle_increment);
    DOUBLE_CHECK(table_feed_per_rotation);
    WITH_DEFAULT(recon_fov, 25.0);
    WITH_DEFAULT(nx, 512);
    WITH_DEFAULT(ny, 512);
    DOUBLE_CHECK(z_start);
    DOUBLE_CHECK(z_end);
    WITH_DEFAULT(slice_thickness, 1.0);
    STRING_WITH_DEFAULT(penalty, "quadratic");
    WITH_DEFAULT(lambda, 0.3);
    WITH_DEFAULT(delta, 0.005);
    SIZE_T_CHECK(num_iterations);
    ZERO_ALLOWED(FileType);
    ZERO_ALLOWED(FileSubType);
    ZERO_ALLOWED(RawOffset);
    SIZE_T_CHECK(Readings);
    ZERO_ALLOWED(Zffs);
    ZERO_ALLOWED(Phiffs);
    ZERO_ALLOWED(wfbp_initialize);
    DOUBLE_CHECK(CollSlicewidth);
#undef STRING_CHECK
#undef DOUBLE_CHECK
#undef SIZE_T_CHECK
#undef ZERO_ALLOWED
#undef WITH_DEFAULT
#undef STRING_WITH_DEFAULT
    // src/setup.cpp:169-179
    if ((strcmp(rp.penalty.c_str(), "quadratic") != 0) &&
        (strcmp(rp.penalty.c_str(), "edge-preserving") != 0)) {
        ctbb::out() << "Requesting penalty function, \"" << rp.penalty
                  << "\", is invalid." << std::endl;
        ctbb::out() << "Valid choices are \"quadratic\" or \"edge-preserving\" "
                     "(without quotes)"
                  << std::endl;
        exit_flag = true;
    }
    if (exit_flag) {
        ctbb::out() << "Missing some required parameters.  Please review the "
                     "parameter file. Exiting."
                  << std::endl;
        return CTBB_E_ARG;
    }

    // Derived values (src/setup.cpp:183-237)
    rp.center_voxel_x = (rp.nx - 1.0) / 2.0;
    rp.center_voxel_y = (rp.ny - 1.0) / 2.0;
    rp.dx = rp.recon_fov / (double)rp.nx;
    rp.dy = rp.recon_fov / (double)rp.ny;
    rp.dz = rp.slice_thickness;
    rp.fov_radius = rp.recon_fov / 2.0;
    rp.voxel_size_x = rp.dx;
    rp.voxel_size_y = rp.dy;
    rp.voxel_size_z = rp.dz;
    rp.num_voxels_x = rp.nx;
    rp.num_voxels_y = rp.ny;
    rp.num_voxels_z = rp.nz;
    ctbb::out() << "voxel_size_z (prior): " << rp.voxel_size_z << std::endl;
    int n_ffs = (size_t)pow(2.0, (double)rp.Zffs) * (size_t)pow(2.0, (double)rp.Phiffs);
    rp.num_views_per_turn = rp.num_views_per_turn_without_ffs *
                            (size_t)pow(2.0, (double)rp.Zffs) *
                            (size_t)pow(2.0, (double)rp.Phiffs);
    rp.tube_angle_increment = 2.0 * kPiSetup / rp.num_views_per_turn;
    rp.tube_z_increment = rp.table_feed_per_rotation / rp.num_views_per_turn;
    rp.voxel_size_z = round(rp.voxel_size_z / (n_ffs * rp.tube_z_increment)) *
                      (n_ffs * rp.tube_z_increment);
    rp.views_per_slice = (size_t)(rp.voxel_size_z / rp.tube_z_increment);
    ctbb::out() << "voxel_size_z (after): " << rp.voxel_size_z << std::endl;
    rp.beam_width = (double)rp.Nrows * rp.axial_detector_spacing;
    rp.beam_width_at_roi_edge = rp.beam_width *
                                (rp.focal_spot_radius + rp.fov_radius) /
                                rp.source_detector_distance;
    ctbb::out() << "Beam width: " << rp.beam_width_at_roi_edge << std::endl;
    rp.num_views_for_system_matrix =
        (size_t)(2.0 * kPiSetup / (rp.table_feed_per_rotation * rp.tube_angle_increment) *
                 rp.beam_width_at_roi_edge);
    if (rp.num_views_for_system_matrix % 2 == 1)
        rp.num_views_for_system_matrix += 1;
    while ((rp.num_views_for_system_matrix / 2) % n_ffs > 0)
        rp.num_views_for_system_matrix = rp.num_views_for_system_matrix + 1;
    rp.Nrows_projection = rp.Nrows / (size_t)pow(2.0, rp.Zffs);
    rp.center_voxel_z = 0.0;
    if (rp.matrix_path.compare("") == 0)
        rp.matrix_path = rp.output_dir + "/matrix.bin";
    rp.num_voxels_z = (size_t)(fabs(rp.z_end - rp.z_start) / rp.voxel_size_z);
    ctbb::out() << "Nvoxelsz= " << rp.num_voxels_z << std::endl;
    ctbb::out() << "z end= " << rp.z_end << std::endl;
    ctbb::out() << "z start= " << rp.z_start << std::endl;
    *out = rp;
    return CTBB_OK;
}

struct recon_params configure_recon_params(std::string filename)
{
    recon_params rp;
    if (configure_recon_params_checked(filename, &rp) != CTBB_OK)
        exit(1);
    return rp;
}

int build_slice_table_checked(recon_params *rp, ct_data *data)
{
    if (rp->FileType != 0) {
        ctbb::out() << "Only FileType 0 angle/table conventions are available "
                     "(the PTR/IMA readers are not part of this build)."
                  << std::endl;
        return CTBB_E_UNSUPPORTED;
    }
    const size_t R = rp->Readings;
    data->tube_angles = new double[R];
    data->table_positions = new double[R];
    int direction = -1;
    // src/setup.cpp:271-284 (angles and table positions of a binary file)
    for (int i = 0; i < (int)R; i++) {
        rp->first_view_angle = 0.0;
        rp->table_direction = -1;
        data->tube_angles[i] = fmod(((360.0f / rp->num_views_per_turn) * i +
                                     rp->first_view_angle),
                                    360.0f);
        if (rp->table_direction == -1)
            data->table_positions[i] =
                ((float)rp->Readings / (float)rp->num_views_per_turn) *
                    rp->table_feed_per_rotation -
                (float)i * rp->table_feed_per_rotation / (float)rp->num_views_per_turn;
        else
            data->table_positions[i] =
                0.0f + (float)i * rp->table_feed_per_rotation / (float)rp->num_views_per_turn;
    }
    direction = rp->table_direction;
    // src/setup.cpp:349-409
    rp->table_direction = direction;
    rp->z_first_view = data->table_positions[0] / 10.0;
    for (int i = 0; i < (int)R; i++)
        data->table_positions[i] =
            rp->z_first_view + (double)i * (double)direction * rp->tube_z_increment;
    bool range_is_good = false;
    double allowed_start =
        rp->z_first_view + direction *
                               ceil(0.5 * rp->beam_width_at_roi_edge / rp->voxel_size_z) *
                               rp->voxel_size_z;
    double allowed_end =
        data->table_positions[R - 1] -
        direction * ceil(0.5 * rp->beam_width_at_roi_edge / rp->voxel_size_z) *
            rp->voxel_size_z;
    if (allowed_start > allowed_end) {
        double tmp = allowed_start;
        allowed_start = allowed_end;
        allowed_end = tmp;
    }
    if ((rp->z_start > allowed_start) && (rp->z_start < allowed_end) &&
        (rp->z_end > allowed_start) && (rp->z_end < allowed_end))
        range_is_good = true;
    if (!range_is_good) {
        ctbb::out() << "Requested reconstruction range (" << rp->z_start << " : "
                  << rp->z_end << ") is invalid" << std::endl;
        ctbb::out() << "Allowed reconstruction range is: " << allowed_start << " : "
                  << allowed_end << std::endl;
        return CTBB_E_ARG;
    }
    rp->num_voxels_z = (size_t)(fabs(rp->z_end - rp->z_start) / rp->voxel_size_z);
    data->slice_locations = new double[rp->num_voxels_z];
    data->slice_indices = new size_t[rp->num_voxels_z];
    int recon_direction =
        (rp->z_end - rp->z_start) / fabs(rp->z_end - rp->z_start);
    for (int i = 0; i < (int)rp->num_voxels_z; i++) {
        data->slice_locations[i] = rp->z_start + i * rp->voxel_size_z * recon_direction;
        size_t central_idx = 0;
        double curr = data->slice_locations[i];
        int n_ffs = (size_t)pow(2.0, (double)rp->Zffs) * (size_t)pow(2.0, (double)rp->Phiffs);
        if (rp->table_direction > 0) {
            while (central_idx < R && data->table_positions[central_idx] < curr)
                central_idx += n_ffs;
        } else {
            while (central_idx < R && data->table_positions[central_idx] > curr)
                central_idx += n_ffs;
        }
        data->slice_indices[i] = central_idx;
    }
    return CTBB_OK;
}

// ---------------------------------------------------------------------------
// C ABI (include/ctbb_host.h)

static void fill_geom(const recon_params &rp, ctbb_geom *g)
{
    memset(g, 0, sizeof(*g));
    g->nx = (int64_t)rp.num_voxels_x;
    g->ny = (int64_t)rp.num_voxels_y;
    g->nz = (int64_t)rp.num_voxels_z;
    g->n_channels = (int64_t)rp.n_channels;
    g->n_rows = (int64_t)rp.Nrows_projection;
    g->readings = (int64_t)rp.Readings;
    g->n_views_sm = (int64_t)rp.num_views_for_system_matrix;
    g->zffs = (int64_t)rp.Zffs;
    g->phiffs = (int64_t)rp.Phiffs;
    g->table_direction = rp.table_direction;
    g->tube_angle_increment = rp.tube_angle_increment;
    g->tube_z_increment = rp.tube_z_increment;
    g->focal_spot_radius = rp.focal_spot_radius;
    g->source_detector_distance = rp.source_detector_distance;
    g->center_channel_non_ffs = rp.center_channel_non_ffs;
    g->center_row = rp.center_row;
    g->transaxial_detector_spacing = rp.transaxial_detector_spacing;
    g->axial_detector_spacing = rp.axial_detector_spacing;
    g->fan_angle_increment = rp.fan_angle_increment;
    g->anode_angle = rp.anode_angle;
    g->coll_slicewidth = rp.CollSlicewidth;
    g->center_voxel_x = rp.center_voxel_x;
    g->center_voxel_y = rp.center_voxel_y;
    g->center_voxel_z = rp.center_voxel_z;
    g->voxel_size_x = rp.voxel_size_x;
    g->voxel_size_y = rp.voxel_size_y;
    g->voxel_size_z = rp.voxel_size_z;
    g->acquisition_fov = rp.acquisition_fov;
}

static void fill_run(const recon_params &rp, ctbb_run_params *p)
{
    memset(p, 0, sizeof(*p));
    p->num_iterations = (int64_t)rp.num_iterations;
    p->penalty = rp.penalty == "edge-preserving" ? CTBB_PENALTY_EDGE_PRESERVING
                                                 : CTBB_PENALTY_QUADRATIC;
    p->wfbp_initialize = (int32_t)rp.wfbp_initialize;
    p->lambda = rp.lambda;
    p->delta = rp.delta;
    p->views_per_slice = (int64_t)rp.views_per_slice;
    p->num_views_per_turn = (int64_t)rp.num_views_per_turn;
    p->z_first_view = rp.z_first_view;
    p->z_start = rp.z_start;
    p->z_end = rp.z_end;
    p->beam_width = rp.beam_width;
    p->beam_width_at_roi_edge = rp.beam_width_at_roi_edge;
    p->recon_fov = rp.recon_fov;
    p->synthetic = (int32_t)rp.synthetic;
    p->use_weights = (int32_t)rp.synthetic_weights;
    p->ssm_fp16 = (int32_t)rp.ssm_fp16;
    p->neighbours = (int32_t)rp.neighbours;
    p->resync_every = (int32_t)rp.resync_every;
    p->state_fp64 = (int32_t)rp.state_fp64;
    p->group_spacing = (int64_t)rp.group_spacing;
}

extern "C" int ctbb_configure(const char *prm_path, ctbb_geom *g,
                              ctbb_run_params *p, double *tube_angles,
                              double *table_positions, double *slice_locations,
                              int64_t *slice_index)
{
    if (!prm_path || !g)
        return ctbb::fail(CTBB_E_ARG, "null argument");
    recon_params rp;
    int rc = configure_recon_params_checked(prm_path, &rp);
    if (rc != CTBB_OK)
        return ctbb::fail(rc, "parameter file rejected (see stdout)");
    ct_data d = {};
    rc = build_slice_table_checked(&rp, &d);
    if (rc != CTBB_OK)
        return ctbb::fail(rc, "slice table rejected (see stdout)");
    fill_geom(rp, g);
    if (p)
        fill_run(rp, p);
    if (tube_angles)
        memcpy(tube_angles, d.tube_angles, rp.Readings * sizeof(double));
    if (table_positions)
        memcpy(table_positions, d.table_positions, rp.Readings * sizeof(double));
    if (slice_locations)
        memcpy(slice_locations, d.slice_locations, rp.num_voxels_z * sizeof(double));
    if (slice_index)
        for (size_t k = 0; k < rp.num_voxels_z; k++)
            slice_index[k] = (int64_t)d.slice_indices[k];
    delete[] d.tube_angles;
    delete[] d.table_positions;
    delete[] d.slice_locations;
    delete[] d.slice_indices;
    return CTBB_OK;
}

// FileType 0 (binary) raw data, src/setup.cpp:271-302 and 335-344: frame i
// of n_channels x Nrows float32 samples, rows fastest inside a channel, read
// by ctbb_read's ReadBinaryFrame at RawOffset; the reference transposes the
// frame to channel-fastest and back again, so the samples keep the file's
// order: raw[row + Nrows*(channel + n_channels*view)].  ctbb_read is not
// available here; the frame placement taken (RawOffset bytes of header,
// frames back to back, native float32) is the plain reading of those calls.
extern "C" int ctbb_read_raw(const char *prm_path, float *raw)
{
    if (!prm_path || !raw)
        return ctbb::fail(CTBB_E_ARG, "null argument");
    recon_params rp;
    int rc = configure_recon_params_checked(prm_path, &rp);
    if (rc != CTBB_OK)
        return ctbb::fail(rc, "parameter file rejected (see stdout)");
    if (rp.FileType != 0)
        return ctbb::fail(CTBB_E_UNSUPPORTED,
                          "FileType " + std::to_string(rp.FileType) +
                              ": the PTR/IMA readers of ctbb_read are not part of this build");
    FILE *f = fopen(rp.sinogram_path.c_str(), "rb");
    if (!f)
        return ctbb::fail(CTBB_E_IO, "raw file not found: " + rp.sinogram_path);
    const size_t frame = rp.n_channels * rp.Nrows_projection;
    rc = CTBB_OK;
    if (fseek(f, (long)rp.RawOffset, SEEK_SET) != 0 ||
        fread(raw, sizeof(float), frame * rp.Readings, f) != frame * rp.Readings)
        rc = ctbb::fail(CTBB_E_IO, "raw file " + rp.sinogram_path + " holds fewer than " +
                                       std::to_string(rp.Readings) + " frames of " +
                                       std::to_string(frame) + " float32 samples after " +
                                       std::to_string(rp.RawOffset) + " header bytes");
    fclose(f);
    return rc;
}
