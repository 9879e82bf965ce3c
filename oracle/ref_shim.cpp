// ref_shim.cpp -- C entry points around the reference's own penalties and
// slice-rotation translation units, compiled unmodified from the CTBB_ICD
// source tree (src/penalties.cpp, src/rotate_slices.cpp, src/interp.h) into
// oracle/_ref/libctbb_ref.so.  TEST INFRASTRUCTURE ONLY: used by tests/ to
// pin the oracle's penalty and rotation restatements bit-exactly.  Never
// linked into the product and never shipped as a source of the product.
#include <cstddef>
#include <cstdint>
#include "recon_structs.h"
#include "penalties.h"
#include "rotate_slices.h"

extern "C" {

static void fill_ip(iterative_params *ip, int64_t nx, int64_t ny, int64_t nz,
                    double alpha, double beta, double lambda, double delta)
{
    initialize_2d_weights(ip);
    ip->Nx = (size_t)nx;
    ip->Ny = (size_t)ny;
    ip->Nz = (size_t)nz;
    ip->alpha = alpha;
    ip->beta = beta;
    ip->lambda = lambda;
    ip->delta = delta;
}

double ref_quadratic(int64_t nx, int64_t ny, int64_t nz, int q, double alpha,
                     double beta, double lambda, double *x)
{
    iterative_params ip;
    fill_ip(&ip, nx, ny, nz, alpha, beta, lambda, 0.0);
    double r = quadratic(q, &ip, x);
    delete[] ip.weights;
    return r;
}

double ref_edge_preserving(int64_t nx, int64_t ny, int64_t nz, int q,
                           double alpha, double beta, double lambda,
                           double delta, double *x)
{
    iterative_params ip;
    fill_ip(&ip, nx, ny, nz, alpha, beta, lambda, delta);
    double r = edge_preserving(q, &ip, x);
    delete[] ip.weights;
    return r;
}

void ref_rotate_slices(int64_t nx, int64_t ny, int64_t nz, double *tube_angles,
                       const int64_t *slice_index, float *volume,
                       int direction)
{
    recon_params rp = {};
    rp.nx = (size_t)nx;
    rp.ny = (size_t)ny;
    rp.num_voxels_x = (size_t)nx;
    rp.num_voxels_y = (size_t)ny;
    rp.num_voxels_z = (size_t)nz;
    rp.center_voxel_x = (nx - 1.0) / 2.0;
    rp.center_voxel_y = (ny - 1.0) / 2.0;
    size_t *si = new size_t[nz];
    for (int64_t k = 0; k < nz; k++)
        si[k] = (size_t)slice_index[k];
    ct_data d = {};
    d.tube_angles = tube_angles;
    d.slice_indices = si;
    d.recon_volume = volume;
    if (direction > 0)
        rotate_slices_rotating2fixed(&rp, &d);
    else
        rotate_slices_fixed2rotating(&rp, &d);
    delete[] si;
}

}
