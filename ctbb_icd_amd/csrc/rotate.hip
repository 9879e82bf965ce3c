// rotate.hip -- slice rotation between the rotating-slice frame of the ICD
// and the fixed Cartesian grid, on gfx950.
//
// Replaces rotate_slices_rotating2fixed / rotate_slices_fixed2rotating
// (src/rotate_slices.cpp:48-114) and rotate_slice (:116-154) with interp2
// (src/interp.h:41-76).  One thread per output pixel.  The coordinate
// rotation is FP64 rounded to FP32 and the bilinear interpolation is FP32,
// evaluated in the reference's order without contraction; cos/sin of each
// slice's angle come from the host's libm, like the reference's.
#include "common.hpp"

#include <cmath>

namespace {

constexpr double kPiRot = 3.141592653589793238;  // src/rotate_slices.cpp:29

__device__ __forceinline__ float interp2(const float *a, int d1, int d2,
                                         float i1, float i2)
{
    if (i1 > (float)(d1 - 2) || (i1 < 0) || i2 > (float)(d2 - 2) || (i2 < 0))
        return 0.0f;
    const float v = i1 - floorf(i1);
    const float w = i2 - floorf(i2);
    const int f1 = (int)floorf(i1), f2 = (int)floorf(i2);
    const int a1 = f2 * d1 + f1;
    const int a2 = f2 * d1 + f1 + 1;
    const int a3 = (f2 + 1) * d1 + f1;
    const int a4 = (f2 + 1) * d1 + f1 + 1;
    return a[a1] * (1.0f - v) * (1.0f - w) + a[a2] * v * (1.0f - w) +
           a[a3] * (1.0f - v) * w + a[a4] * v * w;
}

__global__ void rotate_kernel(const float *in, float *out, const double *cs,
                              int nx, int ny, int nz, double cx, double cy)
{
    const int64_t npix = (int64_t)nx * ny;
    const int64_t total = npix * nz;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(t / npix);
        const int64_t pix = t - (int64_t)k * npix;
        const int yi = (int)(pix / nx), xi = (int)(pix % nx);
        const double c = cs[2 * k], s = cs[2 * k + 1];
        const double x = xi - cx, y = yi - cy;
        const float xp = (float)(c * x - s * y);
        const float yp = (float)(s * x + c * y);
        const int64_t out_idx = xi + (int64_t)yi * ny;
        out[(int64_t)k * npix + out_idx] =
            interp2(in + (int64_t)k * npix, nx, ny, (float)(xp + cx),
                    (float)(yp + cy));
    }
}

}  // namespace

using namespace ctbb;

extern "C" int ctbb_rotate_slices(const ctbb_geom *g, const double *tube_angles,
                                  const int64_t *slice_index, float *volume,
                                  int direction)
{
    if (!g || !tube_angles || !slice_index || !volume)
        return fail(CTBB_E_ARG, "null argument");
    if (g->nx != g->ny)
        return fail(CTBB_E_ARG, "slice rotation requires nx == ny");
    if (g->nx < 2 || g->nz < 1)
        return fail(CTBB_E_ARG, "bad volume size");
    const int64_t npix = g->nx * g->ny, n = npix * g->nz;
    std::vector<double> cs(2 * g->nz);
    for (int64_t k = 0; k < g->nz; k++) {
        if (slice_index[k] < 0 || slice_index[k] >= g->readings)
            return fail(CTBB_E_ARG, "slice_index out of range");
        double angle = -(tube_angles[slice_index[k]]) * 2.0 * kPiRot / 360.0;
        if (direction < 0)
            angle = -angle;
        cs[2 * k] = cos(angle);
        cs[2 * k + 1] = sin(angle);
    }
    float *d_in = nullptr, *d_out = nullptr;
    double *d_cs = nullptr;
    int rc = CTBB_OK;
    if ((rc = dev_alloc_n(&d_in, n, "rotation input")) ||
        (rc = dev_alloc_n(&d_out, n, "rotation output")) ||
        (rc = dev_alloc_n(&d_cs, 2 * g->nz, "rotation angles"))) {
        if (d_in) (void)hipFree(d_in);
        if (d_out) (void)hipFree(d_out);
        return rc;
    }
    hipError_t e = hipMemcpy(d_in, volume, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(d_cs, cs.data(), cs.size() * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const int grid = (int)std::min<int64_t>((n + 255) / 256, 65536);
        hipLaunchKernelGGL(rotate_kernel, dim3(grid), dim3(256), 0, 0, d_in, d_out,
                           d_cs, (int)g->nx, (int)g->ny, (int)g->nz,
                           g->center_voxel_x, g->center_voxel_y);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpy(volume, d_out, n * 4, hipMemcpyDeviceToHost);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    (void)hipFree(d_cs);
    CTBB_HIP(e);
    return CTBB_OK;
}
