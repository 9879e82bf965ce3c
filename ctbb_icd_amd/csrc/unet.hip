// unet.hip -- learned warm start: a 2.5D U-Net trained and run on gfx950.
//
// No counterpart in the reference (it supplies the start image that
// initialize_recon_volume, src/initialize_recon_volume.cpp:51-76, reads from
// a file or sets to zero).  Network, parameter order and layouts: see
// include/ctbb_unet.h.
//
// 3x3 convolutions are implicit GEMMs on MFMA (v_mfma_f32_32x32x8_bf16):
//   forward        out[p][co] = b[co] + sum_{tap,ci} in[p + o(tap)][ci] W[tap][co][ci]
//                  one wavefront per 32 pixels x 32 output channels
//   input grad     a forward convolution of dY with Wt[tap][ci][co] =
//                  W[8 - tap][co][ci] (flipped taps, swapped channels)
//   weight grad    dW[tap][co][ci] = sum_p dY[p][co] in[p + o(tap)][ci], split
//                  over pixel chunks (four wavefronts per chunk, combined in
//                  LDS in wavefront order), chunks added in order by a
//                  second kernel: deterministic
// MFMA operand layout (32x32x8, one k-group of 4 per half-wavefront): lane l
// holds A[l % 32][4 (l / 32) + j] and B[4 (l / 32) + j][l % 32], j = 0..3; the
// 16 results of lane l are D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
// Activations and weights are BF16 (round to nearest even), every sum FP32.
#include "common.hpp"

#include "../../include/ctbb_unet.h"

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

using namespace ctbb;

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

__host__ __device__ inline uint16_t f2bf(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u)
        return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t b)
{
    return __builtin_bit_cast(float, (uint32_t)b << 16);
}

// Input of a convolution: tensor a (ca channels) followed by tensor b (cb
// channels, optionally at half resolution, read with nearest upsampling).
struct ConvIn {
    const uint16_t *a;
    int ca;
    const uint16_t *b;
    int cb;
    int up;
};

__device__ inline v4s load_in4(const ConvIn &in, int n, int y, int x, int H, int W, int c)
{
    v4s z = {0, 0, 0, 0};
    if (y < 0 || y >= H || x < 0 || x >= W)
        return z;
    if (c < in.ca)
        return *reinterpret_cast<const v4s *>(in.a + ((((int64_t)n * H + y) * W + x) * in.ca + c));
    c -= in.ca;
    if (in.up) {
        const int H2 = H >> 1, W2 = W >> 1;
        return *reinterpret_cast<const v4s *>(
            in.b + ((((int64_t)n * H2 + (y >> 1)) * W2 + (x >> 1)) * in.cb + c));
    }
    return *reinterpret_cast<const v4s *>(in.b + ((((int64_t)n * H + y) * W + x) * in.cb + c));
}

__device__ inline uint16_t load_in1(const ConvIn &in, int n, int y, int x, int H, int W, int c)
{
    if (y < 0 || y >= H || x < 0 || x >= W)
        return 0;
    if (c < in.ca)
        return in.a[(((int64_t)n * H + y) * W + x) * in.ca + c];
    c -= in.ca;
    if (in.up) {
        const int H2 = H >> 1, W2 = W >> 1;
        return in.b[(((int64_t)n * H2 + (y >> 1)) * W2 + (x >> 1)) * in.cb + c];
    }
    return in.b[(((int64_t)n * H + y) * W + x) * in.cb + c];
}

// Forward 3x3 convolution (zero padding), + bias, optional ReLU, BF16 out.
// cin must be a multiple of 8.  Grid: 4 tiles (wavefronts) per workgroup,
// tile = (32-pixel block, 32-channel block).
__global__ __launch_bounds__(256) void conv3x3_fwd_kernel(ConvIn in, const uint16_t *w,
                                                          const float *bias, uint16_t *out,
                                                          int N, int H, int W, int cin,
                                                          int cout, int relu, int nct)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t tile = (int64_t)blockIdx.x * 4 + wave;
    const int64_t P = (int64_t)N * H * W;
    const int64_t pt = tile / nct;
    const int ct = (int)(tile - pt * nct);
    if (pt * 32 >= P)
        return;
    const int r = lane & 31, kg = lane >> 5;
    const int64_t p = pt * 32 + r;
    const bool pv = p < P;
    int n = 0, y = 0, x = 0;
    if (pv) {
        n = (int)(p / ((int64_t)H * W));
        const int rem = (int)(p - (int64_t)n * H * W);
        y = rem / W;
        x = rem - y * W;
    }
    const int co = ct * 32 + r;
    const bool cv = co < cout;
    v16f acc = {0};
    const v4s z = {0, 0, 0, 0};
    for (int tap = 0; tap < 9; tap++) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        const uint16_t *wt = w + ((int64_t)tap * cout + (cv ? co : 0)) * cin + 4 * kg;
        for (int ci = 0; ci < cin; ci += 8) {
            const v4s av = pv ? load_in4(in, n, y + dy, x + dx, H, W, ci + 4 * kg) : z;
            const v4s bv = cv ? *reinterpret_cast<const v4s *>(wt + ci) : z;
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
        }
    }
    const int oc = ct * 32 + r;
    if (oc >= cout)
        return;
    const float bb = bias ? bias[oc] : 0.0f;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int i = 8 * (q >> 2) + 4 * kg + (q & 3);
        const int64_t pix = pt * 32 + i;
        if (pix < P) {
            float v = acc[q] + bb;
            if (relu && v < 0.0f)
                v = 0.0f;
            out[pix * cout + oc] = f2bf(v);
        }
    }
}

// Tiled forward 3x3 convolution (also the input gradient, with flipped and
// transposed weights).  A workgroup computes R = WPX image rows x 32 columns
// of one image (32 pixels per wavefront row) x 32 * WCO output channels;
// wavefront (wp, wc) owns row wp and channels 32 wc .. 32 wc + 31.  The K
// dimension (9 taps x cin) is walked in stages of 16 input channels: the
// (R + 2) x 34 input patch of the stage (zero outside the image) and its 9
// x (32 WCO) x 16 weights are staged in LDS, and every wavefront issues 9
// v_mfma_f32_32x32x16_bf16, one per tap, reading the patch shifted by the
// tap.  The next stage's global loads are issued before the current stage
// is computed.  Needs W % 32 == 0, H % R == 0, cin % 16 == 0, cout %
// (32 WCO) == 0 and, with two input tensors, ca % 16 == 0.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

template <int WPX, int WCO>
__global__ __launch_bounds__(64 * WPX * WCO) void conv3x3_fwd_tiled(
    ConvIn in, const uint16_t *w, const float *bias, uint16_t *out, int N, int H, int W,
    int cin, int cout, int relu)
{
    constexpr int R = WPX, PW = 34, NPIX = (R + 2) * PW, NT = 64 * WPX * WCO;
    constexpr int CO = 32 * WCO;
    constexpr int NA = NPIX * 2;              // 16-byte loads of the patch
    constexpr int NB = 9 * CO * 2;            // 16-byte loads of the weights
    constexpr int LA = (NA + NT - 1) / NT, LB = (NB + NT - 1) / NT;
    __shared__ uint4 pa[NPIX * 2];            // [patch pixel][ci 0-7 | 8-15]
    __shared__ uint4 pb[9 * CO * 2];          // [tap][co][ci 0-7 | 8-15]
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wp = wave % WPX, wc = wave / WPX;
    // workgroup -> (image, row block, column block, channel block)
    const int ncb = cout / CO, nxb = W / 32, nyb = H / R;
    int64_t b = blockIdx.x;
    const int cb = (int)(b % ncb);
    b /= ncb;
    const int xb = (int)(b % nxb);
    b /= nxb;
    const int yb = (int)(b % nyb);
    const int n = (int)(b / nyb);
    const int y0 = yb * R, x0 = xb * 32, co0 = cb * CO;
    uint4 ra[LA], rb[LB];
    const uint4 zero4 = {0u, 0u, 0u, 0u};
    auto load_stage = [&](int ci0) {
#pragma unroll
        for (int t = 0; t < LA; t++) {
            const int i = tid + t * NT;
            uint4 v = zero4;
            if (i < NA) {
                const int pix = i >> 1, half = i & 1;
                const int y = y0 - 1 + pix / PW, x = x0 - 1 + pix % PW;
                const int c = ci0 + 8 * half;
                if (y >= 0 && y < H && x >= 0 && x < W) {
                    if (c < in.ca) {
                        v = *reinterpret_cast<const uint4 *>(
                            in.a + ((((int64_t)n * H + y) * W + x) * in.ca + c));
                    } else if (in.up) {
                        const int H2 = H >> 1, W2 = W >> 1;
                        v = *reinterpret_cast<const uint4 *>(
                            in.b + ((((int64_t)n * H2 + (y >> 1)) * W2 + (x >> 1)) * in.cb +
                                    (c - in.ca)));
                    } else {
                        v = *reinterpret_cast<const uint4 *>(
                            in.b + ((((int64_t)n * H + y) * W + x) * in.cb + (c - in.ca)));
                    }
                }
            }
            ra[t] = v;
        }
#pragma unroll
        for (int t = 0; t < LB; t++) {
            const int i = tid + t * NT;
            uint4 v = zero4;
            if (i < NB) {
                const int half = i & 1, tc = i >> 1, tap = tc / CO, co = tc - tap * CO;
                v = *reinterpret_cast<const uint4 *>(
                    w + ((int64_t)tap * cout + co0 + co) * cin + ci0 + 8 * half);
            }
            rb[t] = v;
        }
    };
    v16f acc = {0};
    const int r = lane & 31, kh = lane >> 5;
    load_stage(0);
    for (int ci0 = 0; ci0 < cin; ci0 += 16) {
        __syncthreads();
#pragma unroll
        for (int t = 0; t < LA; t++)
            if (tid + t * NT < NA)
                pa[tid + t * NT] = ra[t];
#pragma unroll
        for (int t = 0; t < LB; t++)
            if (tid + t * NT < NB)
                pb[tid + t * NT] = rb[t];
        __syncthreads();
        if (ci0 + 16 < cin)
            load_stage(ci0 + 16);
#pragma unroll
        for (int tap = 0; tap < 9; tap++) {
            const int dy = tap / 3 - 1, dx = tap % 3 - 1;
            const uint4 av = pa[((wp + 1 + dy) * PW + (r + 1 + dx)) * 2 + kh];
            const uint4 bv = pb[((tap * CO) + 32 * wc + r) * 2 + kh];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, av),
                                                          __builtin_bit_cast(v8bf, bv), acc,
                                                          0, 0, 0);
        }
    }
    const int oc = co0 + 32 * wc + r;
    const float bb = bias ? bias[oc] : 0.0f;
    const int64_t rowbase = (((int64_t)n * H + y0 + wp) * W + x0);
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int px = 8 * (q >> 2) + 4 * kh + (q & 3);
        float v = acc[q] + bb;
        if (relu && v < 0.0f)
            v = 0.0f;
        out[(rowbase + px) * cout + oc] = f2bf(v);
    }
}

// Weight gradient partials: part[chunk][tap][co][ci] = sum over the chunk's
// pixels of dY[p][co] * in[p + o(tap)][ci].  Four wavefronts take quarters
// of the chunk; their 32x32 tiles are added in LDS in wavefront order.
__global__ __launch_bounds__(256) void conv3x3_wgrad_kernel(ConvIn in, const uint16_t *dy,
                                                            int N, int H, int W, int cin,
                                                            int cout, int64_t chunk,
                                                            int nco, int nci, float *part)
{
    __shared__ float red[4][16][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int64_t b = blockIdx.x;
    const int ci_t = (int)(b % nci);
    b /= nci;
    const int co_t = (int)(b % nco);
    b /= nco;
    const int tap = (int)(b % 9);
    const int64_t ck = b / 9;
    const int dyo = tap / 3 - 1, dxo = tap % 3 - 1;
    const int64_t P = (int64_t)N * H * W;
    const int64_t q4 = chunk / 4;
    const int64_t p0 = ck * chunk + wave * q4;
    const int64_t p1 = min(p0 + q4, P);
    const int r = lane & 31, kg = lane >> 5;
    const int co = co_t * 32 + r, ci = ci_t * 32 + r;
    const bool cov = co < cout, civ = ci < cin;
    v16f acc = {0};
    // (n, y, x) of this lane's first pixel p0 + 4 kg, advanced incrementally
    int64_t pq = p0 + 4 * kg;
    int n0 = 0, y0 = 0, x0 = 0;
    if (pq < P) {
        n0 = (int)(pq / ((int64_t)H * W));
        const int rem = (int)(pq - (int64_t)n0 * H * W);
        y0 = rem / W;
        x0 = rem - y0 * W;
    }
    for (int64_t p = p0; p < p1; p += 8) {
        v4s av, bv;
        int n = n0, y = y0, x = x0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int64_t pp = p + 4 * kg + j;
            uint16_t a = 0, bb = 0;
            if (pp < p1) {
                if (cov)
                    a = dy[pp * cout + co];
                if (civ)
                    bb = load_in1(in, n, y + dyo, x + dxo, H, W, ci);
            }
            av[j] = (short)a;
            bv[j] = (short)bb;
            if (++x == W) {
                x = 0;
                if (++y == H) {
                    y = 0;
                    n++;
                }
            }
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
        // next group: 8 pixels on from this lane's first pixel
        x0 += 8;
        while (x0 >= W) {
            x0 -= W;
            if (++y0 == H) {
                y0 = 0;
                n0++;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 16; q++)
        red[wave][q][lane] = acc[q];
    __syncthreads();
    if (wave != 0)
        return;
    const int cc = ci_t * 32 + r;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int i = 8 * (q >> 2) + 4 * kg + (q & 3);
        const int oc = co_t * 32 + i;
        if (oc < cout && cc < cin) {
            const float s = ((red[0][q][lane] + red[1][q][lane]) + red[2][q][lane]) +
                            red[3][q][lane];
            part[(((int64_t)ck * 9 + tap) * cout + oc) * cin + cc] = s;
        }
    }
}

// Weight gradient of a convolution with 8 input channels (the first
// layer): part[chunk][tap][co][ci] like conv3x3_wgrad_kernel, but the MFMA's
// N dimension is (tap, input channel) -- 72 columns in three 32-column
// tiles -- so one pass over a chunk's pixels serves all 9 taps and each dY
// operand feeds three MFMAs.  Four wavefronts take quarters of the chunk;
// their tiles are added in LDS in wavefront order.  Grid: chunks x nco.
__global__ __launch_bounds__(256) void conv3x3_wgrad_c8_kernel(ConvIn in, const uint16_t *dy,
                                                               int N, int H, int W, int cout,
                                                               int64_t chunk, int nco,
                                                               float *part)
{
    __shared__ float red[4][16][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int co_t = (int)(blockIdx.x % nco);
    const int64_t ck = blockIdx.x / nco;
    const int64_t P = (int64_t)N * H * W;
    const int64_t q4 = chunk / 4;
    const int64_t p0 = ck * chunk + wave * q4;
    const int64_t p1 = min(p0 + q4, P);
    const int r = lane & 31, kg = lane >> 5;
    const int co = co_t * 32 + r;
    const bool cov = co < cout;
    bool colv[3];
    int ci_c[3], dyo[3], dxo[3];
#pragma unroll
    for (int t = 0; t < 3; t++) {
        const int c = t * 32 + r;
        colv[t] = c < 72;
        const int tap = colv[t] ? c / 8 : 0;
        ci_c[t] = c % 8;
        dyo[t] = tap / 3 - 1;
        dxo[t] = tap % 3 - 1;
    }
    v16f acc[3] = {{0}, {0}, {0}};
    // (n, y, x) of this lane's first pixel p0 + 4 kg, advanced incrementally
    int64_t pq = p0 + 4 * kg;
    int n0 = 0, y0 = 0, x0 = 0;
    if (pq < P) {
        n0 = (int)(pq / ((int64_t)H * W));
        const int rem = (int)(pq - (int64_t)n0 * H * W);
        y0 = rem / W;
        x0 = rem - y0 * W;
    }
    for (int64_t p = p0; p < p1; p += 8) {
        v4s av, bv[3];
        int n = n0, y = y0, x = x0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int64_t pp = p + 4 * kg + j;
            uint16_t a = 0, bb[3] = {0, 0, 0};
            if (pp < p1) {
                if (cov)
                    a = dy[pp * cout + co];
#pragma unroll
                for (int t = 0; t < 3; t++)
                    if (colv[t])
                        bb[t] = load_in1(in, n, y + dyo[t], x + dxo[t], H, W, ci_c[t]);
            }
            av[j] = (short)a;
#pragma unroll
            for (int t = 0; t < 3; t++)
                bv[t][j] = (short)bb[t];
            if (++x == W) {
                x = 0;
                if (++y == H) {
                    y = 0;
                    n++;
                }
            }
        }
#pragma unroll
        for (int t = 0; t < 3; t++)
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv[t], acc[t], 0, 0, 0);
        // next group: 8 pixels on from this lane's first pixel
        x0 += 8;
        while (x0 >= W) {
            x0 -= W;
            if (++y0 == H) {
                y0 = 0;
                n0++;
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 3; t++) {
#pragma unroll
        for (int q = 0; q < 16; q++)
            red[wave][q][lane] = acc[t][q];
        __syncthreads();
        if (wave == 0) {
            const int c = t * 32 + r;
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int oc = co_t * 32 + 8 * (q >> 2) + 4 * kg + (q & 3);
                if (oc < cout && c < 72) {
                    const float v = ((red[0][q][lane] + red[1][q][lane]) + red[2][q][lane]) +
                                    red[3][q][lane];
                    part[((ck * 9 + c / 8) * cout + oc) * 8 + c % 8] = v;
                }
            }
        }
        __syncthreads();
    }
}

// Tiled weight gradient: part[chunk][tap][co][ci] = sum over the chunk's
// pixels of dY[p][co] in[p + o(tap)][ci], for a 32 x 32 (co, ci) block and
// all 9 taps per workgroup.  A chunk is `crows` consecutive image rows of
// the batch (stacked images, never split: H and crows are even); it is
// walked in tiles of 2 rows x 32 columns.  Per tile the dY tile (64 pixels
// x 32 channels) and the input patch (4 x 34 pixels x 32 channels) are
// staged in LDS in their NHWC order with 16-byte loads (rows padded to 80
// bytes: the two half-wavefronts' pixels fall in different banks).  The
// MFMA's K dimension is the pixel: each lane gathers 8 consecutive pixels
// of one channel from LDS.  Wavefront w accumulates taps w, w + 4 (and 8
// for w = 0), reusing its dY operand across them.  Needs W % 32 == 0, H and
// crows even, cout % 32 == 0, cin % 32 == 0 and, with two input tensors,
// ca % 32 == 0.
__global__ __launch_bounds__(256) void conv3x3_wgrad_tiled(ConvIn in, const uint16_t *dy,
                                                           int N, int H, int W, int cin,
                                                           int cout, int crows, int nco,
                                                           int nci, float *part)
{
    constexpr int PW = 34, RS = 40;           // patch width, LDS row stride (u16)
    __shared__ __align__(16) uint16_t dys[64 * RS];      // [tile pixel][co]
    __shared__ __align__(16) uint16_t pch[4 * PW * RS];  // [patch pixel][ci]
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    int64_t b = blockIdx.x;
    const int ci_t = (int)(b % nci);
    b /= nci;
    const int co_t = (int)(b % nco);
    const int64_t ck = b / nco;
    const int co0 = co_t * 32, ci0 = ci_t * 32;
    const int64_t rows_total = (int64_t)N * H;
    const int64_t r0 = ck * crows, r1 = min(r0 + crows, rows_total);
    const int ntap = wave == 0 ? 3 : 2;
    v16f acc[3] = {{0}, {0}, {0}};
    const int r = lane & 31, kh = lane >> 5;
    const uint4 zero4 = {0u, 0u, 0u, 0u};
    for (int64_t tr = r0; tr < r1; tr += 2) {
        const int n = (int)(tr / H), y0 = (int)(tr - (int64_t)n * H);
        for (int x0 = 0; x0 < W; x0 += 32) {
            __syncthreads();
            {   // dY tile: 64 pixels x 4 loads of 8 channels
                const int px = tid >> 2, q = tid & 3;
                const int y = y0 + (px >> 5), x = x0 + (px & 31);
                const uint4 v = *reinterpret_cast<const uint4 *>(
                    dy + (((int64_t)n * H + y) * W + x) * cout + co0 + 8 * q);
                *reinterpret_cast<uint4 *>(&dys[px * RS + 8 * q]) = v;
            }
            for (int i = tid; i < 4 * PW * 4; i += 256) {   // patch: 136 pixels x 4 loads
                const int pix = i >> 2, q = i & 3;
                const int y = y0 - 1 + pix / PW, x = x0 - 1 + pix % PW;
                const int c = ci0 + 8 * q;
                uint4 v = zero4;
                if (y >= 0 && y < H && x >= 0 && x < W) {
                    if (c < in.ca) {
                        v = *reinterpret_cast<const uint4 *>(
                            in.a + ((((int64_t)n * H + y) * W + x) * in.ca + c));
                    } else if (in.up) {
                        const int H2 = H >> 1, W2 = W >> 1;
                        v = *reinterpret_cast<const uint4 *>(
                            in.b + ((((int64_t)n * H2 + (y >> 1)) * W2 + (x >> 1)) * in.cb +
                                    (c - in.ca)));
                    } else {
                        v = *reinterpret_cast<const uint4 *>(
                            in.b + ((((int64_t)n * H + y) * W + x) * in.cb + (c - in.ca)));
                    }
                }
                *reinterpret_cast<uint4 *>(&pch[pix * RS + 8 * q]) = v;
            }
            __syncthreads();
#pragma unroll
            for (int s4 = 0; s4 < 4; s4++) {
                // K block of 16 tile pixels; this lane's 8: p = 16 s4 + 8 kh + j
                const int pb = 16 * s4 + 8 * kh, ty = pb >> 5, tx = pb & 31;
                v8bf av;
#pragma unroll
                for (int j = 0; j < 8; j++)
                    av[j] = __builtin_bit_cast(__bf16, dys[(pb + j) * RS + r]);
#pragma unroll
                for (int t = 0; t < 3; t++) {
                    if (t >= ntap)
                        break;
                    const int tap = wave + 4 * t;
                    const int dyo = tap / 3 - 1, dxo = tap % 3 - 1;
                    const int base = (ty + 1 + dyo) * PW + (tx + 1 + dxo);
                    v8bf bv;
#pragma unroll
                    for (int j = 0; j < 8; j++)
                        bv[j] = __builtin_bit_cast(__bf16, pch[(base + j) * RS + r]);
                    acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[t], 0, 0, 0);
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 3; t++) {
        if (t >= ntap)
            break;
        const int tap = wave + 4 * t;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int co = 8 * (q >> 2) + 4 * kh + (q & 3);
            part[(((int64_t)ck * 9 + tap) * cout + co0 + co) * cin + ci0 + r] = acc[t][q];
        }
    }
}

// g[i] = sum over chunks of part[chunk][i]: eight running sums (chunks c,
// c + 8, c + 16, ...; eight loads in flight) added in a fixed order.
__global__ void sum_chunks_kernel(const float *part, int64_t nchunk, int64_t n, float *g)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        int64_t c = 0;
        for (; c + 8 <= nchunk; c += 8) {
#pragma unroll
            for (int k = 0; k < 8; k++)
                s[k] += part[(c + k) * n + i];
        }
        for (int k = 0; c < nchunk; c++, k++)
            s[k] += part[c * n + i];
        g[i] = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
    }
}

// Bias gradient partials: part[chunk][co] = sum over the chunk's pixels of
// dY[p][co].  Thread t takes the 8 channels 8 (t % g8) .. + 7 (one 16-byte
// load of a pixel row; g8 = cout / 8) of pixels chunk_start + t / g8,
// + 256 / g8, ...; the partial sums of a channel are added in LDS in a
// fixed order.  Needs cout a multiple of 8 dividing 2048.
constexpr int64_t kBiasChunk = 1024;

__global__ __launch_bounds__(256) void bias_grad_kernel(const uint16_t *dy, int64_t P,
                                                        int cout, float *part)
{
    __shared__ float sh[256 * 8];
    const int t = threadIdx.x;
    const int g8 = cout / 8, lanes = 256 / g8;
    const int cg = t % g8, l = t / g8;
    const int64_t p0 = (int64_t)blockIdx.x * kBiasChunk;
    const int64_t p1 = min(p0 + kBiasChunk, P);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (l < lanes)
        for (int64_t p = p0 + l; p < p1; p += lanes) {
            const uint4 v = *reinterpret_cast<const uint4 *>(dy + p * cout + 8 * cg);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                s[2 * k] += __uint_as_float(w[k] << 16);
                s[2 * k + 1] += __uint_as_float(w[k] & 0xffff0000u);
            }
        }
#pragma unroll
    for (int k = 0; k < 8; k++)
        sh[t * 8 + k] = s[k];
    __syncthreads();
    for (int c = t; c < cout; c += 256) {
        float tot = 0.0f;
        for (int i = 0; i < lanes; i++)
            tot += sh[(i * g8 + c / 8) * 8 + c % 8];
        part[(int64_t)blockIdx.x * cout + c] = tot;
    }
}

__global__ void maxpool_fwd_kernel(const uint16_t *in, uint16_t *out, int N, int H, int W,
                                   int C)
{
    const int H2 = H >> 1, W2 = W >> 1;
    const int64_t n_out = (int64_t)N * H2 * W2 * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_out;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        int64_t t = i / C;
        const int x = (int)(t % W2);
        t /= W2;
        const int y = (int)(t % H2);
        const int n = (int)(t / H2);
        float m = -INFINITY;
        for (int dy = 0; dy < 2; dy++)
            for (int dx = 0; dx < 2; dx++)
                m = fmaxf(m, bf2f(in[(((int64_t)n * H + 2 * y + dy) * W + 2 * x + dx) * C + c]));
        out[i] = f2bf(m);
    }
}

// dIn = routes dOut to the first maximum of each 2x2 window (ties: row-major
// first), everything else 0.  Writes every element of dIn.
__global__ void maxpool_bwd_kernel(const uint16_t *in, const uint16_t *dout, uint16_t *din,
                                   int N, int H, int W, int C)
{
    const int H2 = H >> 1, W2 = W >> 1;
    const int64_t n_out = (int64_t)N * H2 * W2 * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_out;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        int64_t t = i / C;
        const int x = (int)(t % W2);
        t /= W2;
        const int y = (int)(t % H2);
        const int n = (int)(t / H2);
        float m = -INFINITY;
        int best = 0;
        for (int k = 0; k < 4; k++) {
            const float v =
                bf2f(in[(((int64_t)n * H + 2 * y + k / 2) * W + 2 * x + k % 2) * C + c]);
            if (v > m) {
                m = v;
                best = k;
            }
        }
        for (int k = 0; k < 4; k++)
            din[(((int64_t)n * H + 2 * y + k / 2) * W + 2 * x + k % 2) * C + c] =
                k == best ? dout[i] : (uint16_t)0;
    }
}

// dY[i] = 0 where the ReLU output Y[i] is 0.
__global__ void relu_mask_kernel(const uint16_t *y, uint16_t *dy, int64_t n)
{
    // eight values per thread (16-byte loads; n is a multiple of 8): dy kept
    // where y > 0, i.e. the BF16 bits of y in [1, 0x7f80] (NaN and <= 0 fail)
    const uint4 *y8 = reinterpret_cast<const uint4 *>(y);
    uint4 *d8 = reinterpret_cast<uint4 *>(dy);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n / 8;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint4 a = y8[i];
        uint4 d = d8[i];
        uint32_t *dw = reinterpret_cast<uint32_t *>(&d);
        const uint32_t aw[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t lo = aw[k] & 0xffffu, hi = aw[k] >> 16;
            const uint32_t keep = ((lo - 1u) < 0x7f80u ? 0x0000ffffu : 0u) |
                                  ((hi - 1u) < 0x7f80u ? 0xffff0000u : 0u);
            dw[k] &= keep;
        }
        d8[i] = d;
    }
}

// Gradient of a concatenated input: the first ca channels are added to dA
// (in FP32, rounded to BF16); the cb upsampled channels are summed over their
// 2x2 source block into dB (half resolution), in a fixed order.  Eight
// channels per thread (16-byte loads; ca and cb are multiples of 8).
__global__ void split_concat_grad_kernel(const uint16_t *dx, int N, int H, int W, int ca,
                                         int cb, uint16_t *da, uint16_t *db)
{
    const int a8 = ca / 8, b8 = cb / 8;
    const int64_t na = (int64_t)N * H * W * a8;
    const int H2 = H >> 1, W2 = W >> 1;
    const int64_t nb = (int64_t)N * H2 * W2 * b8;
    const int C = ca + cb;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < na + nb;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i < na) {
            const int c8 = (int)(i % a8);
            const int64_t p = i / a8;
            uint4 *pa = reinterpret_cast<uint4 *>(da + p * ca + 8 * c8);
            const uint4 x = *reinterpret_cast<const uint4 *>(dx + p * C + 8 * c8);
            uint4 v = *pa;
            const uint16_t *xv = reinterpret_cast<const uint16_t *>(&x);
            uint16_t *vv = reinterpret_cast<uint16_t *>(&v);
#pragma unroll
            for (int k = 0; k < 8; k++)
                vv[k] = f2bf(bf2f(vv[k]) + bf2f(xv[k]));
            *pa = v;
        } else {
            const int64_t j = i - na;
            const int c8 = (int)(j % b8);
            int64_t t = j / b8;
            const int x = (int)(t % W2);
            t /= W2;
            const int y = (int)(t % H2);
            const int n = (int)(t / H2);
            float sacc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            for (int k = 0; k < 4; k++) {
                const uint4 q = *reinterpret_cast<const uint4 *>(
                    dx + (((int64_t)n * H + 2 * y + k / 2) * W + 2 * x + k % 2) * C + ca +
                    8 * c8);
                const uint16_t *qv = reinterpret_cast<const uint16_t *>(&q);
#pragma unroll
                for (int e = 0; e < 8; e++)
                    sacc[e] += bf2f(qv[e]);
            }
            uint4 o;
            uint16_t *ov = reinterpret_cast<uint16_t *>(&o);
#pragma unroll
            for (int e = 0; e < 8; e++)
                ov[e] = f2bf(sacc[e]);
            *reinterpret_cast<uint4 *>(db + ((((int64_t)n * H2 + y) * W2 + x) * cb + 8 * c8)) = o;
        }
    }
}

// Input packing: host layout [B][3][H][W] FP32 -> NHWC BF16 with 8 channels
// (3 slices, 5 zeros) and the FP32 centre slice [B][H][W].
__global__ void pack_input_kernel(const float *src, int N, int H, int W, uint16_t *x8,
                                  float *centre)
{
    const int64_t P = (int64_t)N * H * W;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t n = p / ((int64_t)H * W), hw = p - n * H * W;
        for (int c = 0; c < 8; c++)
            x8[p * 8 + c] = c < 3 ? f2bf(src[(n * 3 + c) * H * W + hw]) : (uint16_t)0;
        centre[p] = src[(n * 3 + 1) * H * W + hw];
    }
}

// Final 1x1 convolution + residual: out[p] = centre[p] + b + sum_c w[c] f[p][c].
__global__ void final_fwd_kernel(const uint16_t *f, const float *w, const float *b, int C,
                                 const float *centre, int64_t P, float *out)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P;
         p += (int64_t)gridDim.x * blockDim.x) {
        float s = b[0];
        for (int c = 0; c < C; c++)
            s += w[c] * bf2f(f[p * C + c]);
        out[p] = centre[p] + s;
    }
}

// g[p] = 2 (out - t) / P and per-block partial sums of (out - t)^2.
__global__ __launch_bounds__(256) void loss_grad_kernel(const float *out, const float *t,
                                                        int64_t P, float *g, double *partials)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < P;
         p += (int64_t)gridDim.x * 256) {
        const float d = out[p] - t[p];
        g[p] = 2.0f * d / (float)P;
        s += (double)d * d;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partials[blockIdx.x] = sh[0];
}

// Backward of the final layer: df[p][c] = g[p] w[c] where f[p][c] > 0 (the
// ReLU of the last 3x3 convolution), and per-block partials of
// dw[c] = sum_p g[p] f[p][c], db = sum_p g[p] ([nblk][C + 1]).
__global__ __launch_bounds__(256) void final_bwd_kernel(const uint16_t *f, const float *w,
                                                        int C, const float *g, int64_t P,
                                                        uint16_t *df, float *partials)
{
    // thread t takes the 8 channels 8 cg .. 8 cg + 7 (cg = t % g8, g8 = C / 8;
    // the stride keeps cg fixed) of pixels with 16-byte loads and stores
    __shared__ float shf[256][9];
    const int t = threadIdx.x, g8 = C / 8, cg = t % g8;
    float s[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float wc[8];
#pragma unroll
    for (int k = 0; k < 8; k++)
        wc[k] = w[8 * cg + k];
    const int64_t n = P * g8;
    for (int64_t i = (int64_t)blockIdx.x * 256 + t; i < n; i += (int64_t)gridDim.x * 256) {
        const int64_t p = i / g8;
        const float gp = g[p];
        const uint4 v = *reinterpret_cast<const uint4 *>(f + p * C + 8 * cg);
        const uint32_t in[4] = {v.x, v.y, v.z, v.w};
        uint32_t out[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float f0 = __uint_as_float(in[k] << 16);
            const float f1 = __uint_as_float(in[k] & 0xffff0000u);
            const uint32_t d0 = f0 > 0.0f ? (uint32_t)f2bf(gp * wc[2 * k]) : 0u;
            const uint32_t d1 = f1 > 0.0f ? (uint32_t)f2bf(gp * wc[2 * k + 1]) : 0u;
            out[k] = d0 | (d1 << 16);
            s[2 * k] += gp * f0;
            s[2 * k + 1] += gp * f1;
        }
        *reinterpret_cast<uint4 *>(df + p * C + 8 * cg) = make_uint4(out[0], out[1], out[2],
                                                                     out[3]);
        if (cg == 0)
            s[8] += gp;
    }
#pragma unroll
    for (int k = 0; k < 9; k++)
        shf[t][k] = s[k];
    __syncthreads();
    // channel c (and the bias term, c = C) summed over its threads in order
    for (int c = t; c <= C; c += 256) {
        const int q = c < C ? c / 8 : 0, k = c < C ? c % 8 : 8;
        float tot = 0.0f;
        for (int i = q; i < 256; i += g8)
            tot += shf[i][k];
        partials[(int64_t)blockIdx.x * (C + 1) + c] = tot;
    }
}

__global__ void adam_kernel(float *w, float *m, float *v, const float *g, int64_t n,
                            float lr, float b1, float b2, float eps, float c1, float c2)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// BF16 copies of a convolution's weights: wbf = W, wt[tap][ci][co] =
// W[8 - tap][co][ci] (the input-gradient convolution).
__global__ void refresh_weights_kernel(const float *w, int cin, int cout, uint16_t *wbf,
                                       uint16_t *wt)
{
    const int64_t n = 9LL * cin * cout;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ci = (int)(i % cin);
        const int64_t t = i / cin;
        const int co = (int)(t % cout);
        const int tap = (int)(t / cout);
        const uint16_t b = f2bf(w[i]);
        wbf[i] = b;
        if (wt)
            wt[((int64_t)(8 - tap) * cin + ci) * cout + co] = b;
    }
}

int grid_for(int64_t n)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384));
}

struct Conv {
    int cin, cout, level;      // level: spatial size H >> level
    int64_t woff, boff;        // parameter offsets
    uint16_t *wbf = nullptr;   // [9][cout][cin]
    uint16_t *wt = nullptr;    // [9][cin][cout] (null for the first layer)
    ConvIn in{};
    uint16_t *out = nullptr, *dout = nullptr;
};

}  // namespace

struct ctbb_unet {
    int L = 4, base = 32, max_batch = 1, H = 0, W = 0;
    std::vector<Conv> convs;           // encoder 0..L-1 (a, b), decoder L-2..0 (a, b)
    std::vector<uint16_t *> pool, dpool;   // [L-1] pooled encoder outputs and grads
    std::vector<uint16_t *> dskip;     // [L-1] concat gradient of the encoder outputs
    int64_t nparams = 0, final_w = 0, final_b = 0;
    float *params = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *x8 = nullptr, *dxfull = nullptr, *dfinal = nullptr;
    float *centre = nullptr, *outp = nullptr, *target = nullptr, *gout = nullptr,
          *host_in = nullptr, *wpart = nullptr, *fpart = nullptr, *bpart = nullptr;
    double *lpart = nullptr;
    int64_t wpart_n = 0;
    int step = 0;
    double last_ms = 0.0, last_flops = 0.0;
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
};

namespace {

int ualloc(ctbb_unet *u, void **p, size_t bytes, const char *what)
{
    CTBB_TRY(dev_alloc(p, std::max<size_t>(bytes, 16), what));
    u->allocs.push_back(*p);
    return CTBB_OK;
}

template <typename T>
int ualloc_n(ctbb_unet *u, T **p, size_t n, const char *what)
{
    return ualloc(u, reinterpret_cast<void **>(p), n * sizeof(T), what);
}

int64_t pixels(const ctbb_unet *u, int batch, int level)
{
    return (int64_t)batch * (u->H >> level) * (u->W >> level);
}

// 3x3 convolution in -> out (cin -> cout channels, weights [9][cout][cin]):
// the tiled kernel where the shape fits it, else the per-wavefront kernel.
void launch_conv(ctbb_unet *u, const ConvIn &in, const uint16_t *w, const float *bias,
                 uint16_t *out, int batch, int Hl, int Wl, int cin, int cout, int relu)
{
    const bool fits = Wl % 32 == 0 && cin % 16 == 0 && cout % 32 == 0 &&
                      (!in.b || in.ca % 16 == 0);
    const int64_t P = (int64_t)batch * Hl * Wl;
    if (fits && cout % 64 == 0 && Hl % 2 == 0) {
        const int64_t nwg = (int64_t)batch * (Hl / 2) * (Wl / 32) * (cout / 64);
        hipLaunchKernelGGL((conv3x3_fwd_tiled<2, 2>), dim3((unsigned)nwg), dim3(256), 0,
                           u->stream, in, w, bias, out, batch, Hl, Wl, cin, cout, relu);
    } else if (fits && Hl % 4 == 0) {
        const int64_t nwg = (int64_t)batch * (Hl / 4) * (Wl / 32) * (cout / 32);
        hipLaunchKernelGGL((conv3x3_fwd_tiled<4, 1>), dim3((unsigned)nwg), dim3(256), 0,
                           u->stream, in, w, bias, out, batch, Hl, Wl, cin, cout, relu);
    } else {
        const int nct = (cout + 31) / 32;
        const int64_t tiles = ((P + 31) / 32) * nct;
        hipLaunchKernelGGL(conv3x3_fwd_kernel, dim3((unsigned)((tiles + 3) / 4)), dim3(256),
                           0, u->stream, in, w, bias, out, batch, Hl, Wl, cin, cout, relu, nct);
    }
}

int run_conv_fwd(ctbb_unet *u, const Conv &c, int batch, double *flops)
{
    const int Hl = u->H >> c.level, Wl = u->W >> c.level;
    const int64_t P = (int64_t)batch * Hl * Wl;
    launch_conv(u, c.in, c.wbf, u->params + c.boff, c.out, batch, Hl, Wl, c.cin, c.cout, 1);
    *flops += 2.0 * P * 9.0 * c.cin * c.cout;
    CTBB_HIP(hipGetLastError());
    return CTBB_OK;
}

// Input gradient of conv c into dst ([P][cin]), from c.dout (masked).
int run_conv_dgrad(ctbb_unet *u, const Conv &c, int batch, uint16_t *dst, double *flops)
{
    const int Hl = u->H >> c.level, Wl = u->W >> c.level;
    const int64_t P = (int64_t)batch * Hl * Wl;
    ConvIn gi{c.dout, c.cout, nullptr, 0, 0};
    launch_conv(u, gi, c.wt, nullptr, dst, batch, Hl, Wl, c.cout, c.cin, 0);
    *flops += 2.0 * P * 9.0 * c.cin * c.cout;
    CTBB_HIP(hipGetLastError());
    return CTBB_OK;
}

constexpr int64_t kWgradChunk = 2048;

// The tiled weight-gradient kernel applies (image rows of a multiple of 32
// pixels, 32-channel blocks) and its chunk height: whole rows, about
// kWgradChunk pixels, even.
bool wgrad_tiled(const Conv &c, int Hl, int Wl, int *crows)
{
    *crows = std::max<int>(2, (int)(kWgradChunk / std::max(1, Wl)) & ~1);
    return Wl % 32 == 0 && Hl % 2 == 0 && c.cout % 32 == 0 && c.cin % 32 == 0 &&
           (!c.in.b || c.in.ca % 32 == 0);
}

// Chunks (partial-sum slices) of one weight-gradient launch.
int64_t wgrad_chunks(const Conv &c, int batch, int Hl, int Wl)
{
    int crows = 0;
    if (wgrad_tiled(c, Hl, Wl, &crows))
        return ((int64_t)batch * Hl + crows - 1) / crows;
    return ((int64_t)batch * Hl * Wl + kWgradChunk - 1) / kWgradChunk;
}

int run_conv_wgrad(ctbb_unet *u, const Conv &c, int batch, double *flops)
{
    const int Hl = u->H >> c.level, Wl = u->W >> c.level;
    const int64_t P = (int64_t)batch * Hl * Wl;
    const int nco = (c.cout + 31) / 32, nci = (c.cin + 31) / 32;
    const int64_t n = 9LL * c.cout * c.cin;
    int crows = 0;
    const bool tiled = wgrad_tiled(c, Hl, Wl, &crows);
    const int64_t nchunk = wgrad_chunks(c, batch, Hl, Wl);
    if (nchunk * n > u->wpart_n)
        return fail(CTBB_E_STATE, "weight-gradient scratch too small");
    if (tiled)
        hipLaunchKernelGGL(conv3x3_wgrad_tiled, dim3((unsigned)(nchunk * nco * nci)), dim3(256),
                           0, u->stream, c.in, c.dout, batch, Hl, Wl, c.cin, c.cout, crows,
                           nco, nci, u->wpart);
    else if (c.cin == 8 && !c.in.b)
        hipLaunchKernelGGL(conv3x3_wgrad_c8_kernel, dim3((unsigned)(nchunk * nco)), dim3(256),
                           0, u->stream, c.in, c.dout, batch, Hl, Wl, c.cout, kWgradChunk,
                           nco, u->wpart);
    else
        hipLaunchKernelGGL(conv3x3_wgrad_kernel, dim3((unsigned)(nchunk * 9 * nco * nci)),
                           dim3(256), 0, u->stream, c.in, c.dout, batch, Hl, Wl, c.cin, c.cout,
                           kWgradChunk, nco, nci, u->wpart);
    hipLaunchKernelGGL(sum_chunks_kernel, dim3(grid_for(n)), dim3(256), 0, u->stream,
                       u->wpart, nchunk, n, u->grad + c.woff);
    const int64_t nb = (P + kBiasChunk - 1) / kBiasChunk;
    hipLaunchKernelGGL(bias_grad_kernel, dim3((unsigned)nb), dim3(256), 0, u->stream, c.dout,
                       P, c.cout, u->bpart);
    hipLaunchKernelGGL(sum_chunks_kernel, dim3(grid_for(c.cout)), dim3(256), 0, u->stream,
                       u->bpart, nb, (int64_t)c.cout, u->grad + c.boff);
    *flops += 2.0 * P * 9.0 * c.cin * c.cout;
    CTBB_HIP(hipGetLastError());
    return CTBB_OK;
}

int refresh_bf16(ctbb_unet *u)
{
    for (auto &c : u->convs)
        hipLaunchKernelGGL(refresh_weights_kernel, dim3(grid_for(9LL * c.cin * c.cout)),
                           dim3(256), 0, u->stream, u->params + c.woff, c.cin, c.cout, c.wbf,
                           c.wt);
    CTBB_HIP(hipGetLastError());
    return CTBB_OK;
}

void mask(ctbb_unet *u, const uint16_t *y, uint16_t *dy, int64_t n)
{
    hipLaunchKernelGGL(relu_mask_kernel, dim3(grid_for(n)), dim3(256), 0, u->stream, y, dy, n);
}

// Forward pass over `batch` packed inputs (u->x8, u->centre) into u->outp.
int forward(ctbb_unet *u, int batch, double *flops)
{
    const int L = u->L;
    for (int l = 0; l < L; l++) {
        Conv &a = u->convs[2 * l], &b = u->convs[2 * l + 1];
        CTBB_TRY(run_conv_fwd(u, a, batch, flops));
        CTBB_TRY(run_conv_fwd(u, b, batch, flops));
        if (l < L - 1) {
            const int Hl = u->H >> l, Wl = u->W >> l;
            hipLaunchKernelGGL(maxpool_fwd_kernel,
                               dim3(grid_for(pixels(u, batch, l + 1) * b.cout)), dim3(256), 0,
                               u->stream, b.out, u->pool[l], batch, Hl, Wl, b.cout);
        }
    }
    for (int d = 0; d < L - 1; d++) {   // decoder position d: level L-2-d
        Conv &a = u->convs[2 * L + 2 * d], &b = u->convs[2 * L + 2 * d + 1];
        CTBB_TRY(run_conv_fwd(u, a, batch, flops));
        CTBB_TRY(run_conv_fwd(u, b, batch, flops));
    }
    const Conv &last = u->convs.back();
    hipLaunchKernelGGL(final_fwd_kernel, dim3(grid_for(pixels(u, batch, 0))), dim3(256), 0,
                       u->stream, last.out, u->params + u->final_w, u->params + u->final_b,
                       u->base, u->centre, pixels(u, batch, 0), u->outp);
    CTBB_HIP(hipGetLastError());
    return CTBB_OK;
}

int backward(ctbb_unet *u, int batch, double *flops)
{
    const int L = u->L;
    const int64_t P0 = pixels(u, batch, 0);
    // final layer
    const int nblk = (int)std::min<int64_t>((P0 + 255) / 256, 1024);
    Conv &last = u->convs.back();
    hipLaunchKernelGGL(final_bwd_kernel, dim3(nblk), dim3(256), 0, u->stream, last.out,
                       u->params + u->final_w, u->base, u->gout, P0, last.dout, u->fpart);
    hipLaunchKernelGGL(sum_chunks_kernel, dim3(1), dim3(256), 0, u->stream, u->fpart,
                       (int64_t)nblk, (int64_t)(u->base + 1), u->grad + u->final_w);
    CTBB_HIP(hipGetLastError());
    // decoder positions d = L-2 .. 0 (position d works at level L-2-d): the
    // reverse of the forward order
    for (int d = L - 2; d >= 0; d--) {
        Conv &a = u->convs[2 * L + 2 * d], &b = u->convs[2 * L + 2 * d + 1];
        // b: dout already holds the masked gradient
        CTBB_TRY(run_conv_wgrad(u, b, batch, flops));
        CTBB_TRY(run_conv_dgrad(u, b, batch, a.dout, flops));
        mask(u, a.out, a.dout, pixels(u, batch, a.level) * a.cout);
        CTBB_TRY(run_conv_wgrad(u, a, batch, flops));
        CTBB_TRY(run_conv_dgrad(u, a, batch, u->dxfull, flops));
        // split: skip part -> dskip[level], upsampled part -> producer's dout
        const int lvl = a.level;
        const Conv &prod = (d == 0) ? u->convs[2 * (L - 1) + 1]
                                    : u->convs[2 * L + 2 * (d - 1) + 1];
        CTBB_HIP(hipMemsetAsync(u->dskip[lvl], 0,
                                pixels(u, batch, lvl) * a.in.ca * sizeof(uint16_t), u->stream));
        const int64_t nsplit = pixels(u, batch, lvl) * a.in.ca +
                               pixels(u, batch, lvl + 1) * a.in.cb;
        hipLaunchKernelGGL(split_concat_grad_kernel, dim3(grid_for(nsplit)), dim3(256), 0,
                           u->stream, u->dxfull, batch, u->H >> lvl, u->W >> lvl, a.in.ca,
                           a.in.cb, u->dskip[lvl], prod.dout);
        mask(u, prod.out, prod.dout, pixels(u, batch, prod.level) * prod.cout);
        CTBB_HIP(hipGetLastError());
    }
    // encoder, deepest level first
    for (int l = L - 1; l >= 0; l--) {
        Conv &a = u->convs[2 * l], &b = u->convs[2 * l + 1];
        if (l < L - 1) {
            // dout of b = pool gradient routed back + concat (skip) gradient
            const int Hl = u->H >> l, Wl = u->W >> l;
            hipLaunchKernelGGL(maxpool_bwd_kernel,
                               dim3(grid_for(pixels(u, batch, l + 1) * b.cout)), dim3(256), 0,
                               u->stream, b.out, u->dpool[l], b.dout, batch, Hl, Wl, b.cout);
            hipLaunchKernelGGL(split_concat_grad_kernel,
                               dim3(grid_for(pixels(u, batch, l) * b.cout)), dim3(256), 0,
                               u->stream, u->dskip[l], batch, Hl, Wl, b.cout, 0, b.dout,
                               nullptr);
            mask(u, b.out, b.dout, pixels(u, batch, l) * b.cout);
        }
        CTBB_TRY(run_conv_wgrad(u, b, batch, flops));
        CTBB_TRY(run_conv_dgrad(u, b, batch, a.dout, flops));
        mask(u, a.out, a.dout, pixels(u, batch, l) * a.cout);
        CTBB_TRY(run_conv_wgrad(u, a, batch, flops));
        if (l > 0)
            CTBB_TRY(run_conv_dgrad(u, a, batch, u->dpool[l - 1], flops));
        CTBB_HIP(hipGetLastError());
    }
    return CTBB_OK;
}

int upload_inputs(ctbb_unet *u, const float *input, int batch)
{
    const int64_t P = pixels(u, batch, 0);
    CTBB_HIP(hipMemcpyAsync(u->host_in, input, P * 3 * sizeof(float), hipMemcpyHostToDevice,
                            u->stream));
    hipLaunchKernelGGL(pack_input_kernel, dim3(grid_for(P)), dim3(256), 0, u->stream,
                       u->host_in, batch, u->H, u->W, u->x8, u->centre);
    CTBB_HIP(hipGetLastError());
    return CTBB_OK;
}

}  // namespace

extern "C" int ctbb_unet_create(int levels, int base_channels, int max_batch, int height,
                                int width, uint64_t seed, ctbb_unet **out)
{
    if (!out)
        return fail(CTBB_E_ARG, "null argument");
    *out = nullptr;
    if (levels < 2 || levels > 6 || base_channels < 8 ||
        (base_channels & (base_channels - 1)) || (base_channels << (levels - 1)) > 256 ||
        max_batch < 1)
        return fail(CTBB_E_ARG, "need 2 <= levels <= 6, base_channels a power of two >= 8 "
                                "with base_channels * 2^(levels-1) <= 256, max_batch >= 1");
    const int m = 1 << (levels - 1);
    if (height < m || width < m || height % m || width % m)
        return fail(CTBB_E_ARG, "height and width must be multiples of 2^(levels-1)");
    ctbb_unet *u = new ctbb_unet();
    u->L = levels;
    u->base = base_channels;
    u->max_batch = max_batch;
    u->H = height;
    u->W = width;
    auto bail = [&](int rc) {
        ctbb_unet_free(u);
        return rc;
    };
    hipError_t he = hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking);
    if (he != hipSuccess)
        return bail(fail(CTBB_E_HIP, std::string("stream: ") + hipGetErrorString(he)));
    const int L = levels;
    std::vector<int> ch(L);
    for (int l = 0; l < L; l++)
        ch[l] = base_channels << l;
    // layer table in parameter order
    int64_t off = 0;
    auto add = [&](int cin, int cout, int level) {
        Conv c;
        c.cin = cin;
        c.cout = cout;
        c.level = level;
        c.woff = off;
        off += 9LL * cin * cout;
        c.boff = off;
        off += cout;
        u->convs.push_back(c);
    };
    for (int l = 0; l < L; l++) {
        add(l == 0 ? 8 : ch[l - 1], ch[l], l);
        add(ch[l], ch[l], l);
    }
    for (int d = L - 2; d >= 0; d--) {
        add(ch[d] + ch[d + 1], ch[d], d);
        add(ch[d], ch[d], d);
    }
    u->final_w = off;
    off += base_channels;
    u->final_b = off;
    off += 1;
    u->nparams = off;
    int rc = CTBB_OK;
    if ((rc = ualloc_n(u, &u->params, off, "U-Net parameters")) ||
        (rc = ualloc_n(u, &u->grad, off, "U-Net gradients")) ||
        (rc = ualloc_n(u, &u->m, off, "Adam moments")) ||
        (rc = ualloc_n(u, &u->v, off, "Adam moments")))
        return bail(rc);
    const int64_t P0 = pixels(u, max_batch, 0);
    // activations and gradients
    for (auto &c : u->convs) {
        const int64_t n = pixels(u, max_batch, c.level) * c.cout;
        if ((rc = ualloc_n(u, &c.out, n, "U-Net activations")) ||
            (rc = ualloc_n(u, &c.dout, n, "U-Net activation gradients")) ||
            (rc = ualloc_n(u, &c.wbf, 9LL * c.cin * c.cout, "U-Net BF16 weights")))
            return bail(rc);
        if (&c != &u->convs[0] &&
            (rc = ualloc_n(u, &c.wt, 9LL * c.cin * c.cout, "U-Net BF16 weights")))
            return bail(rc);
    }
    u->pool.assign(L - 1, nullptr);
    u->dpool.assign(L - 1, nullptr);
    u->dskip.assign(L - 1, nullptr);
    for (int l = 0; l < L - 1; l++) {
        if ((rc = ualloc_n(u, &u->pool[l], pixels(u, max_batch, l + 1) * ch[l], "pooled")) ||
            (rc = ualloc_n(u, &u->dpool[l], pixels(u, max_batch, l + 1) * ch[l], "pooled")) ||
            (rc = ualloc_n(u, &u->dskip[l], pixels(u, max_batch, l) * ch[l], "skip grads")))
            return bail(rc);
    }
    int64_t dx_max = 0, wp_max = 0;
    for (auto &c : u->convs) {
        dx_max = std::max(dx_max, pixels(u, max_batch, c.level) * c.cin);
        // either kernel's chunk count (the inputs are not wired yet)
        const int Hl = u->H >> c.level, Wl = u->W >> c.level;
        int crows = 0;
        wgrad_tiled(c, Hl, Wl, &crows);
        const int64_t nch = std::max(((int64_t)max_batch * Hl + crows - 1) / crows,
                                     (pixels(u, max_batch, c.level) + kWgradChunk - 1) /
                                         kWgradChunk);
        wp_max = std::max<int64_t>(wp_max, nch * 9 * (int64_t)c.cin * c.cout);
    }
    u->wpart_n = wp_max;
    if ((rc = ualloc_n(u, &u->x8, P0 * 8, "packed input")) ||
        (rc = ualloc_n(u, &u->dxfull, dx_max, "input gradients")) ||
        (rc = ualloc_n(u, &u->centre, P0, "centre slices")) ||
        (rc = ualloc_n(u, &u->outp, P0, "output")) ||
        (rc = ualloc_n(u, &u->target, P0, "target")) ||
        (rc = ualloc_n(u, &u->gout, P0, "output gradient")) ||
        (rc = ualloc_n(u, &u->host_in, P0 * 3, "input staging")) ||
        (rc = ualloc_n(u, &u->wpart, wp_max, "weight-gradient partials")) ||
        (rc = ualloc_n(u, &u->fpart, 1024LL * (base_channels + 1), "final partials")) ||
        (rc = ualloc_n(u, &u->bpart, ((P0 + kBiasChunk - 1) / kBiasChunk) * (int64_t)ch[L - 1],
                       "bias partials")) ||
        (rc = ualloc_n(u, &u->lpart, 1024, "loss partials")))
        return bail(rc);
    // wire inputs
    for (int l = 0; l < L; l++) {
        Conv &a = u->convs[2 * l], &b = u->convs[2 * l + 1];
        a.in = l == 0 ? ConvIn{u->x8, 8, nullptr, 0, 0}
                      : ConvIn{u->pool[l - 1], ch[l - 1], nullptr, 0, 0};
        b.in = ConvIn{a.out, a.cout, nullptr, 0, 0};
    }
    for (int d = L - 2; d >= 0; d--) {
        const int i = 2 * L + 2 * (L - 2 - d);
        Conv &a = u->convs[i], &b = u->convs[i + 1];
        const Conv &below = (d == L - 2) ? u->convs[2 * (L - 1) + 1] : u->convs[i - 1];
        a.in = ConvIn{u->convs[2 * d + 1].out, ch[d], below.out, ch[d + 1], 1};
        b.in = ConvIn{a.out, a.cout, nullptr, 0, 0};
    }
    // He initialisation (normal, std sqrt(2 / fan_in)), biases 0, final 0
    std::vector<float> h(off, 0.0f);
    std::mt19937_64 rng(seed);
    for (auto &c : u->convs) {
        std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / (9.0 * c.cin)));
        for (int64_t i = 0; i < 9LL * c.cin * c.cout; i++) {
            const int ci = (int)(i % c.cin);
            // zero padding channels of the input layer carry no signal
            h[c.woff + i] = (c.woff == 0 && ci >= 3) ? 0.0f : (float)nd(rng);
        }
    }
    he = hipMemcpy(u->params, h.data(), off * sizeof(float), hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemset(u->m, 0, off * sizeof(float));
    if (he == hipSuccess) he = hipMemset(u->v, 0, off * sizeof(float));
    if (he == hipSuccess) he = hipMemset(u->grad, 0, off * sizeof(float));
    if (he != hipSuccess)
        return bail(fail(CTBB_E_HIP, std::string("U-Net init: ") + hipGetErrorString(he)));
    if ((rc = refresh_bf16(u)) != CTBB_OK)
        return bail(rc);
    he = hipStreamSynchronize(u->stream);
    if (he != hipSuccess)
        return bail(fail(CTBB_E_HIP, std::string("U-Net init: ") + hipGetErrorString(he)));
    *out = u;
    return CTBB_OK;
}

extern "C" int ctbb_unet_free(ctbb_unet *u)
{
    if (!u)
        return CTBB_OK;
    if (u->stream)
        (void)hipStreamSynchronize(u->stream);
    for (void *p : u->allocs)
        (void)hipFree(p);
    if (u->stream)
        (void)hipStreamDestroy(u->stream);
    delete u;
    return CTBB_OK;
}

extern "C" int ctbb_unet_num_params(const ctbb_unet *u, int64_t *n)
{
    if (!u || !n)
        return fail(CTBB_E_ARG, "null argument");
    *n = u->nparams;
    return CTBB_OK;
}

extern "C" int ctbb_unet_get_params(const ctbb_unet *u, float *params)
{
    if (!u || !params)
        return fail(CTBB_E_ARG, "null argument");
    CTBB_HIP(hipStreamSynchronize(u->stream));
    CTBB_HIP(hipMemcpy(params, u->params, u->nparams * sizeof(float), hipMemcpyDeviceToHost));
    return CTBB_OK;
}

extern "C" int ctbb_unet_set_params(ctbb_unet *u, const float *params)
{
    if (!u || !params)
        return fail(CTBB_E_ARG, "null argument");
    CTBB_HIP(hipStreamSynchronize(u->stream));
    CTBB_HIP(hipMemcpy(u->params, params, u->nparams * sizeof(float), hipMemcpyHostToDevice));
    CTBB_TRY(refresh_bf16(u));
    CTBB_HIP(hipStreamSynchronize(u->stream));
    return CTBB_OK;
}

extern "C" int ctbb_unet_grad(ctbb_unet *u, const float *input, const float *target,
                              int batch, double *loss, float *grads)
{
    if (!u || !input || !target)
        return fail(CTBB_E_ARG, "null argument");
    if (batch < 1 || batch > u->max_batch)
        return fail(CTBB_E_ARG, "batch must be in [1, max_batch]");
    hipEvent_t e0, e1;
    CTBB_HIP(hipEventCreate(&e0));
    CTBB_HIP(hipEventCreate(&e1));
    const int64_t P0 = pixels(u, batch, 0);
    double flops = 0.0;
    int rc = upload_inputs(u, input, batch);
    if (rc == CTBB_OK) {
        hipError_t he = hipMemcpyAsync(u->target, target, P0 * sizeof(float),
                                       hipMemcpyHostToDevice, u->stream);
        if (he != hipSuccess)
            rc = fail(CTBB_E_HIP, std::string("target upload: ") + hipGetErrorString(he));
    }
    if (rc == CTBB_OK && hipEventRecord(e0, u->stream) != hipSuccess)
        rc = fail(CTBB_E_HIP, "event record");
    if (rc == CTBB_OK)
        rc = forward(u, batch, &flops);
    const int nblk = (int)std::min<int64_t>((P0 + 255) / 256, 1024);
    if (rc == CTBB_OK) {
        hipLaunchKernelGGL(loss_grad_kernel, dim3(nblk), dim3(256), 0, u->stream, u->outp,
                           u->target, P0, u->gout, u->lpart);
        rc = backward(u, batch, &flops);
    }
    std::vector<double> lp(nblk);
    if (rc == CTBB_OK) {
        hipError_t he = hipEventRecord(e1, u->stream);
        if (he == hipSuccess)
            he = hipMemcpyAsync(lp.data(), u->lpart, nblk * sizeof(double),
                                hipMemcpyDeviceToHost, u->stream);
        if (he == hipSuccess && grads)
            he = hipMemcpyAsync(grads, u->grad, u->nparams * sizeof(float),
                                hipMemcpyDeviceToHost, u->stream);
        if (he == hipSuccess)
            he = hipStreamSynchronize(u->stream);
        float ms = 0.0f;
        if (he == hipSuccess)
            he = hipEventElapsedTime(&ms, e0, e1);
        if (he != hipSuccess)
            rc = fail(CTBB_E_HIP, std::string("U-Net step: ") + hipGetErrorString(he));
        u->last_ms = ms;
        u->last_flops = flops;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != CTBB_OK)
        return rc;
    double s = 0.0;
    for (double x : lp)
        s += x;
    if (loss)
        *loss = s / (double)P0;
    return CTBB_OK;
}

extern "C" int ctbb_unet_adam_step(ctbb_unet *u, double lr, double beta1, double beta2,
                                   double eps)
{
    if (!u)
        return fail(CTBB_E_ARG, "null argument");
    u->step++;
    const float c1 = (float)(1.0 - std::pow(beta1, u->step));
    const float c2 = (float)(1.0 - std::pow(beta2, u->step));
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(u->nparams)), dim3(256), 0, u->stream,
                       u->params, u->m, u->v, u->grad, u->nparams, (float)lr, (float)beta1,
                       (float)beta2, (float)eps, c1, c2);
    CTBB_HIP(hipGetLastError());
    CTBB_TRY(refresh_bf16(u));
    return CTBB_OK;
}

extern "C" int ctbb_unet_infer(ctbb_unet *u, const float *input, int batch, float *output)
{
    if (!u || !input || !output)
        return fail(CTBB_E_ARG, "null argument");
    if (batch < 1 || batch > u->max_batch)
        return fail(CTBB_E_ARG, "batch must be in [1, max_batch]");
    hipEvent_t e0, e1;
    CTBB_HIP(hipEventCreate(&e0));
    CTBB_HIP(hipEventCreate(&e1));
    double flops = 0.0;
    int rc = upload_inputs(u, input, batch);
    if (rc == CTBB_OK && hipEventRecord(e0, u->stream) != hipSuccess)
        rc = fail(CTBB_E_HIP, "event record");
    if (rc == CTBB_OK)
        rc = forward(u, batch, &flops);
    if (rc == CTBB_OK) {
        hipError_t he = hipEventRecord(e1, u->stream);
        if (he == hipSuccess)
            he = hipMemcpyAsync(output, u->outp, pixels(u, batch, 0) * sizeof(float),
                                hipMemcpyDeviceToHost, u->stream);
        if (he == hipSuccess)
            he = hipStreamSynchronize(u->stream);
        float ms = 0.0f;
        if (he == hipSuccess)
            he = hipEventElapsedTime(&ms, e0, e1);
        if (he != hipSuccess)
            rc = fail(CTBB_E_HIP, std::string("U-Net inference: ") + hipGetErrorString(he));
        u->last_ms = ms;
        u->last_flops = flops;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

extern "C" int ctbb_unet_last_timing(const ctbb_unet *u, double *ms, double *conv_flops)
{
    if (!u)
        return fail(CTBB_E_ARG, "null argument");
    if (ms)
        *ms = u->last_ms;
    if (conv_flops)
        *conv_flops = u->last_flops;
    return CTBB_OK;
}
