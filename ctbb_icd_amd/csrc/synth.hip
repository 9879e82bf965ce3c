// synth.hip -- seeded synthetic raw data (stand-in for the raw-data reader).
//
// The reference reads projections through a closed reader library
// (src/setup.cpp:245-345).  This generator produces the same `raw` layout
// (row fastest, src/setup.cpp:335-344) for the scan a prm file describes:
// data view v has tube angle v * 2pi / vpt, source z = table_positions[v]
// (FileType-0 conventions, src/setup.cpp:271-284, 349-356) and, with flying
// focal spots, the source deflected as in the matrix build (+-da along e_u,
// +-dr along e_w and +-dr tan(anode) along e_z by view index,
// src/generate_system_matrix_ffs.cpp:242-283).  Every detector ray
// uses the SSM's ray model (channel angle, row height, FFS offset), so the
// reconstruction model and the data agree up to discretisation.
//
// Phantom: infinite water cylinder along z, ellipsoids rotated about z whose
// value replaces the background (added as mu - mu_water), optional small
// high-contrast ellipsoids, optional lung-like region (an ellipsoidal
// cylinder of low attenuation plus Gaussian blobs of sigma 0.2 cm: a
// Gaussian-filtered random texture).  Line integrals are analytic chord
// lengths (and the closed-form line integral of a Gaussian), in FP64,
// stored as FP32.  Noise: N ~ Poisson(I0 exp(-p)) with a counter-based
// random stream per ray (splitmix64; Poisson by inversion for small means,
// by Hormann's PTRS transformed rejection otherwise), y = ln(I0/max(N,1)),
// weight = N.
#include "common.hpp"

#include <cmath>
#include <cstring>
#include <iostream>
#include <vector>

#include "../../include/ctbb_host.h"
#include "../host/recon_structs.hpp"

namespace {

struct Ellipsoid {
    double cx, cy, cz, a, b, c, cphi, sphi, dmu;
};

struct Blob {
    double x, y, z, amp;
};

struct Scene {
    double cyl_r, cyl_mu;
    int n_ell;
    const Ellipsoid *ell;
    int n_blob;
    const Blob *blob;
    double blob_sigma;
};

struct ScanRays {
    int64_t readings, nch, nr, vpt;
    int ffs, zffs;                  // in-plane / z flying focal spot
    double R, D, da, dr, tz, c_ch, c_row, dgamma_nonffs, fan_inc, drow;
    const double *table_pos;
    double dtheta;
    double I0;
    uint64_t seed;
    int weights;
};

__host__ __device__ inline uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ inline double uniform01(uint64_t &s)
{
    return ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

__host__ __device__ inline double poisson(double lam, uint64_t &s)
{
    if (lam <= 0.0)
        return 0.0;
    if (lam < 30.0) {  // inversion by sequential search
        double p = exp(-lam), F = p, u = uniform01(s);
        double k = 0.0;
        while (u > F && k < 1000.0) {
            k += 1.0;
            p *= lam / k;
            F += p;
        }
        return k;
    }
    // PTRS (W. Hormann, "The transformed rejection method for generating
    // Poisson random variables", Insurance: Math. Econ. 12, 1993).
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (int it = 0; it < 1000; it++) {
        const double U = uniform01(s) - 0.5;
        const double V = uniform01(s);
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr)
            return k;
        if (k < 0.0 || (us < 0.013 && V > us))
            continue;
        if (log(V) + log(invalpha) - log(a / (us * us) + b) <=
            -lam + k * loglam - lgamma(k + 1.0))
            return k;
    }
    return floor(lam + 0.5);
}

// Chord length of the ray s + t d (|d| = 1, t >= 0) through a quadric
// A t^2 + B t + C = 0 (C < 0 inside).
__host__ __device__ inline double chord(double A, double B, double C)
{
    if (A <= 0.0)
        return 0.0;
    const double disc = B * B - 4.0 * A * C;
    if (disc <= 0.0)
        return 0.0;
    const double sq = sqrt(disc);
    double t1 = (-B - sq) / (2.0 * A), t2 = (-B + sq) / (2.0 * A);
    if (t2 <= 0.0)
        return 0.0;
    if (t1 < 0.0)
        t1 = 0.0;
    return t2 - t1;
}

__host__ __device__ double line_integral(const Scene &sc, const double s[3],
                                         const double d[3])
{
    double p = 0.0;
    // water cylinder (infinite along z)
    {
        const double A = d[0] * d[0] + d[1] * d[1];
        const double B = 2.0 * (s[0] * d[0] + s[1] * d[1]);
        const double C = s[0] * s[0] + s[1] * s[1] - sc.cyl_r * sc.cyl_r;
        p += sc.cyl_mu * chord(A, B, C);
    }
    for (int i = 0; i < sc.n_ell; i++) {
        const Ellipsoid &E = sc.ell[i];
        const double px = s[0] - E.cx, py = s[1] - E.cy, pz = s[2] - E.cz;
        const double su = (E.cphi * px + E.sphi * py) / E.a;
        const double sv = (-E.sphi * px + E.cphi * py) / E.b;
        const double sw = pz / E.c;
        const double du = (E.cphi * d[0] + E.sphi * d[1]) / E.a;
        const double dv = (-E.sphi * d[0] + E.cphi * d[1]) / E.b;
        const double dw = d[2] / E.c;
        const double A = du * du + dv * dv + dw * dw;
        const double B = 2.0 * (su * du + sv * dv + sw * dw);
        const double C = su * su + sv * sv + sw * sw - 1.0;
        p += E.dmu * chord(A, B, C);
    }
    if (sc.n_blob) {
        const double k = sc.blob_sigma * 2.5066282746310002;  // sigma sqrt(2 pi)
        const double inv2s2 = 1.0 / (2.0 * sc.blob_sigma * sc.blob_sigma);
        for (int i = 0; i < sc.n_blob; i++) {
            const Blob &G = sc.blob[i];
            const double cx = G.x - s[0], cy = G.y - s[1], cz = G.z - s[2];
            const double t = cx * d[0] + cy * d[1] + cz * d[2];
            const double d2 = cx * cx + cy * cy + cz * cz - t * t;
            if (d2 * inv2s2 < 30.0)
                p += G.amp * k * exp(-d2 * inv2s2);
        }
    }
    return p;
}

// One detector sample: ray (view, channel, row) -> (y, w).
__host__ __device__ void synth_sample(const Scene &sc, const ScanRays &r,
                                      int64_t q, float *raw, float *w)
{
    const int64_t row = q % r.nr;
    const int64_t ch = (q / r.nr) % r.nch;
    const int64_t view = q / (r.nr * r.nch);
    const double th = view * r.dtheta;
    const double cw = cos(th), sw = sin(th);
    const double ew[3] = {cw, sw, 0.0}, eu[3] = {-sw, cw, 0.0};
    // focal-spot deflection of the view (src/generate_system_matrix_ffs.cpp:
    // 242-283), as in the matrix build
    double off[3] = {0.0, 0.0, 0.0};
    if (r.zffs || r.ffs) {
        double cr = 0.0, ca = 0.0;
        if (r.ffs && !r.zffs) {
            ca = (view % 2 == 0) ? 1.0 : -1.0;
        } else if (!r.ffs && r.zffs) {
            cr = (view % 2 == 0) ? -1.0 : 1.0;
        } else {
            cr = (view % 4) < 2 ? -1.0 : 1.0;
            ca = (view % 2 == 0) ? -1.0 : 1.0;
        }
        const double ez[3] = {0.0, 0.0, 1.0};
        for (int e = 0; e < 3; e++)
            off[e] = (cr * r.dr) * ew[e] + (ca * r.da) * eu[e] + ((cr * r.dr) * r.tz) * ez[e];
    }
    const double s[3] = {r.R * ew[0] + off[0], r.R * ew[1] + off[1],
                         r.table_pos[view] + off[2]};
    const double gam = (r.ffs || r.zffs) ? (ch - r.c_ch) * r.fan_inc
                                         : (ch - r.c_ch) * r.dgamma_nonffs;
    const double cg = cos(gam), sg = sin(gam);
    const double ap = (row - r.c_row) * r.drow;
    double d[3];
    for (int e = 0; e < 3; e++)
        d[e] = -r.D * cg * ew[e] - r.D * sg * eu[e] - (e == 2 ? ap : 0.0) - off[e];
    const double n = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int e = 0; e < 3; e++)
        d[e] /= n;
    const double p = line_integral(sc, s, d);
    if (r.I0 > 0.0) {
        uint64_t st = r.seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(q + 1));
        splitmix(st);
        const double N = poisson(r.I0 * exp(-p), st);
        raw[q] = (float)log(r.I0 / (N > 1.0 ? N : 1.0));
        if (w)
            w[q] = (float)N;
    } else {
        raw[q] = (float)p;
        if (w)
            w[q] = 1.0f;
    }
}

__global__ void synth_kernel(Scene sc, ScanRays r, float *raw, float *w,
                             int64_t n)
{
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x)
        synth_sample(sc, r, q, raw, w);
}

// Host-side phantom description from the prm (seeded).
struct HostScene {
    std::vector<Ellipsoid> ell;
    std::vector<Blob> blob;
    Scene sc;
};

double urand(uint64_t &s, double lo, double hi)
{
    return lo + (hi - lo) * uniform01(s);
}

void make_scene(const recon_params &rp, HostScene *hs)
{
    uint64_t s = 0x5EEDull * (rp.synthetic_seed + 1);
    splitmix(s);
    const double mu_w = rp.synthetic_cylinder_mu;
    const double zlo = std::min(rp.z_start, rp.z_end);
    const double zhi = std::max(rp.z_start, rp.z_end);
    auto add = [&](size_t n, double rad, double amin, double amax,
                   double mumin, double mumax) {
        for (size_t i = 0; i < n; i++) {
            Ellipsoid E;
            const double r = rad * sqrt(uniform01(s));
            const double ph = urand(s, 0.0, 2.0 * M_PI);
            E.cx = r * cos(ph);
            E.cy = r * sin(ph);
            E.cz = urand(s, zlo, zhi);
            E.a = urand(s, amin, amax);
            E.b = urand(s, amin, amax);
            E.c = urand(s, amin, amax);
            const double phi = urand(s, 0.0, M_PI);
            E.cphi = cos(phi);
            E.sphi = sin(phi);
            E.dmu = urand(s, mumin, mumax) - mu_w;
            hs->ell.push_back(E);
        }
    };
    add(rp.synthetic_n_ellipsoids, rp.synthetic_ellipsoid_radius,
        rp.synthetic_axis_min, rp.synthetic_axis_max, rp.synthetic_mu_min,
        rp.synthetic_mu_max);
    add(rp.synthetic_n_small, rp.synthetic_ellipsoid_radius, 0.1, 0.4,
        rp.synthetic_small_mu_min, rp.synthetic_small_mu_max);
    if (rp.synthetic_lung) {
        // lung-like region: low-attenuation elliptic cylinder segment
        const double lx = -0.45 * rp.synthetic_cylinder_radius, ly = 0.0;
        const double la = 0.28 * rp.synthetic_cylinder_radius;
        const double lb = 0.42 * rp.synthetic_cylinder_radius;
        const double lc = 0.5 * (zhi - zlo) + 2.0;
        Ellipsoid L{lx, ly, 0.5 * (zlo + zhi), la, lb, lc, 1.0, 0.0, 0.05 - mu_w};
        hs->ell.push_back(L);
        const size_t nb = rp.synthetic_lung > 1 ? rp.synthetic_lung : 4000;
        for (size_t i = 0; i < nb; i++) {
            double x, y, z;
            do {
                x = urand(s, -1.0, 1.0);
                y = urand(s, -1.0, 1.0);
                z = urand(s, -1.0, 1.0);
            } while (x * x + y * y + z * z > 1.0);
            hs->blob.push_back(Blob{lx + 0.9 * la * x, ly + 0.9 * lb * y,
                                    0.5 * (zlo + zhi) + lc * z,
                                    urand(s, -0.03, 0.03)});
        }
    }
    hs->sc.cyl_r = rp.synthetic_cylinder_radius;
    hs->sc.cyl_mu = mu_w;
    hs->sc.n_ell = (int)hs->ell.size();
    hs->sc.ell = hs->ell.data();
    hs->sc.n_blob = (int)hs->blob.size();
    hs->sc.blob = hs->blob.data();
    hs->sc.blob_sigma = 0.2;
}

double phantom_point(const Scene &sc, double x, double y, double z)
{
    double mu = 0.0;
    if (x * x + y * y < sc.cyl_r * sc.cyl_r)
        mu += sc.cyl_mu;
    for (int i = 0; i < sc.n_ell; i++) {
        const Ellipsoid &E = sc.ell[i];
        const double px = x - E.cx, py = y - E.cy, pz = z - E.cz;
        const double u = (E.cphi * px + E.sphi * py) / E.a;
        const double v = (-E.sphi * px + E.cphi * py) / E.b;
        const double w = pz / E.c;
        if (u * u + v * v + w * w < 1.0)
            mu += E.dmu;
    }
    const double inv2s2 = 1.0 / (2.0 * sc.blob_sigma * sc.blob_sigma);
    for (int i = 0; i < sc.n_blob; i++) {
        const Blob &G = sc.blob[i];
        const double d2 = (x - G.x) * (x - G.x) + (y - G.y) * (y - G.y) +
                          (z - G.z) * (z - G.z);
        if (d2 * inv2s2 < 30.0)
            mu += G.amp * exp(-d2 * inv2s2);
    }
    return mu;
}

int load_prm(const char *path, recon_params *rp, ct_data *d)
{
    int rc = configure_recon_params_checked(path, rp);
    if (rc != CTBB_OK)
        return ctbb::fail(rc, "parameter file rejected (see stdout)");
    rc = build_slice_table_checked(rp, d);
    if (rc != CTBB_OK)
        return ctbb::fail(rc, "slice table rejected (see stdout)");
    return CTBB_OK;
}

void free_ct(ct_data *d)
{
    delete[] d->tube_angles;
    delete[] d->table_positions;
    delete[] d->slice_locations;
    delete[] d->slice_indices;
}

}  // namespace

extern "C" int ctbb_synthesize(const char *prm_path, float *raw, float *weights,
                               int use_gpu)
{
    if (!prm_path || !raw)
        return ctbb::fail(CTBB_E_ARG, "null argument");
    recon_params rp;
    ct_data d = {};
    CTBB_TRY(load_prm(prm_path, &rp, &d));
    HostScene hs;
    make_scene(rp, &hs);
    ScanRays r{};
    r.readings = (int64_t)rp.Readings;
    r.nch = (int64_t)rp.n_channels;
    r.nr = (int64_t)rp.Nrows_projection;
    r.vpt = (int64_t)rp.num_views_per_turn;
    r.ffs = rp.Phiffs ? 1 : 0;
    r.zffs = rp.Zffs ? 1 : 0;
    r.R = rp.focal_spot_radius;
    r.D = rp.source_detector_distance;
    r.da = r.D * r.R * sin(rp.fan_angle_increment) / (4.0 * (r.D - r.R));
    r.dr = r.D * rp.CollSlicewidth /
           (4.0 * (r.D - r.R) * tan(rp.anode_angle * 3.14159265359 / 180.0));
    r.tz = tan(rp.anode_angle);
    r.c_ch = rp.center_channel_non_ffs;
    r.c_row = rp.center_row;
    r.dgamma_nonffs = rp.transaxial_detector_spacing / r.D;
    r.fan_inc = rp.fan_angle_increment;
    r.drow = rp.axial_detector_spacing;
    r.dtheta = 2.0 * M_PI / (double)rp.num_views_per_turn;
    r.I0 = rp.synthetic_I0;
    r.seed = 0xC0FFEEull + 7919ull * rp.synthetic_seed;
    const int64_t n = r.readings * r.nch * r.nr;
    int rc = CTBB_OK;
    if (use_gpu) {
        Ellipsoid *d_ell = nullptr;
        Blob *d_blob = nullptr;
        double *d_tp = nullptr;
        float *d_raw = nullptr, *d_w = nullptr;
        if ((rc = ctbb::dev_alloc_n(&d_ell, std::max<size_t>(1, hs.ell.size()), "ellipsoids")) ||
            (rc = ctbb::dev_alloc_n(&d_blob, std::max<size_t>(1, hs.blob.size()), "blobs")) ||
            (rc = ctbb::dev_alloc_n(&d_tp, r.readings, "table positions")) ||
            (rc = ctbb::dev_alloc_n(&d_raw, n, "raw")) ||
            (weights && (rc = ctbb::dev_alloc_n(&d_w, n, "weights")))) {
            free_ct(&d);
            return rc;
        }
        hipError_t e = hipMemcpy(d_ell, hs.ell.data(), hs.ell.size() * sizeof(Ellipsoid),
                                 hipMemcpyHostToDevice);
        if (e == hipSuccess && !hs.blob.empty())
            e = hipMemcpy(d_blob, hs.blob.data(), hs.blob.size() * sizeof(Blob),
                          hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(d_tp, d.table_positions, r.readings * 8, hipMemcpyHostToDevice);
        Scene sc = hs.sc;
        sc.ell = d_ell;
        sc.blob = d_blob;
        r.table_pos = d_tp;
        if (e == hipSuccess) {
            hipLaunchKernelGGL(synth_kernel, dim3(16384), dim3(256), 0, 0, sc, r,
                               d_raw, d_w, n);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpy(raw, d_raw, n * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess && weights)
            e = hipMemcpy(weights, d_w, n * 4, hipMemcpyDeviceToHost);
        (void)hipFree(d_ell);
        (void)hipFree(d_blob);
        (void)hipFree(d_tp);
        (void)hipFree(d_raw);
        if (d_w)
            (void)hipFree(d_w);
        if (e != hipSuccess) {
            free_ct(&d);
            return ctbb::fail(CTBB_E_HIP, std::string("synthesize: ") + hipGetErrorString(e));
        }
    } else {
        r.table_pos = d.table_positions;
        const Scene sc = hs.sc;
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < n; q++)
            synth_sample(sc, r, q, raw, weights);
    }
    free_ct(&d);
    return CTBB_OK;
}

extern "C" int ctbb_phantom_volume(const char *prm_path, float *volume)
{
    if (!prm_path || !volume)
        return ctbb::fail(CTBB_E_ARG, "null argument");
    recon_params rp;
    ct_data d = {};
    CTBB_TRY(load_prm(prm_path, &rp, &d));
    HostScene hs;
    make_scene(rp, &hs);
    const int64_t nx = rp.num_voxels_x, ny = rp.num_voxels_y, nz = rp.num_voxels_z;
    const double dth = 2.0 * M_PI / (double)rp.num_views_per_turn;
    for (int64_t k = 0; k < nz; k++) {
        const int64_t v = (int64_t)d.slice_indices[k];
        const double th = v * dth, c = cos(th), s = sin(th);
        const double z = d.table_positions[v];
#pragma omp parallel for schedule(static)
        for (int64_t p = 0; p < nx * ny; p++) {
            const double xs = (p % nx - rp.center_voxel_x) * rp.voxel_size_x;
            const double ys = (p / nx - rp.center_voxel_y) * rp.voxel_size_y;
            const double xw = c * xs - s * ys, yw = s * xs + c * ys;
            volume[k * nx * ny + p] = (float)phantom_point(hs.sc, xw, yw, z);
        }
    }
    free_ct(&d);
    return CTBB_OK;
}
