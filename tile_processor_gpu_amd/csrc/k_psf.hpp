// k_psf.hpp - CLT-domain aberration kernels and their CltExtra records
// from sampled point-spread functions.
//
// The reference reads these from files its Java host generated
// (src/dtt8x8.cu:228-238, :364-378; the Java side uses DCT-III/DST-III,
// src/dtt8x8.cuh:40-42).  The kernel of a spatial kernel p, element [i][j]
// at x = j - ox, y = i - oy from the cell centre, is the bicomplex sum
//     K(k, l) = sum p(x, y) e^{i u_k x} e^{j u_l y},  u_k = pi (2k + 1) / 16,
// i.e. the quadrants CC = sum p cos(u_k x) cos(u_l y), SC = sum p sin cos,
// CS = sum p cos sin, SS = sum p sin sin at [l = vertical][k = horizontal]:
// the factor the kernel multiply of convert_correct_tiles (convolveTiles,
// src/TileProcessor.cuh:2778-2827) applies for a convolution with p.  With
// `center` the origin (ox, oy) is the centroid of p, which goes to the
// record's data_x / data_y (the forward CLT moves its window by it); the
// slopes dxc_dx .. dyc_dy are central differences of the centroids over the
// kernel grid (cells 16 px apart), one-sided at its edges.
//
// One wavefront per (camera, cell, color), in FP64 (the sums have a few
// hundred terms of sines and cosines of fractional arguments; FP64 keeps the
// result within FP32 rounding of the exact sums), rounded once to FP32:
//   1. centroid: each lane sums a strided share of p, p x, p y; wavefront
//      reduction;
//   2. rows: lane (i0, k) forms R_c[i][k] = sum_j p[i][j] cos(u_k x_j) and
//      R_s with sin, for rows i = i0, i0 + 8, ..., into LDS;
//   3. columns: lane (l, k) forms the 4 quadrant values of bin (l, k) from
//      R_c, R_s and cos / sin(u_l y_i).
#pragma once

#include "tp_common.hpp"

namespace tp {

constexpr int PSF_MAX = 32;  // largest PSF side

struct PsfArgs {
    const float *const *psfs;  // [cam] -> [KV][KH][3][h][w]
    float *const *kernels;     // [cam] -> [KV][KH][3][4][8][8]
    float *const *koffs;       // [cam] -> tp_clt_extra [KV][KH][3]
    int KV, KH, h, w;
    int center;
};

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// 4 wavefronts per workgroup, one unit (camera, cell, color) each
__global__ __launch_bounds__(256) void k_psf_to_clt(PsfArgs a) {
    __shared__ double rows[4][2][PSF_MAX][8];  // [wave][cos, sin][row][k]
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int per_cam = a.KV * a.KH * 3;
    const int unit = blockIdx.x * 4 + wave;
    if (unit >= 4 * per_cam) return;  // whole wavefronts only: no barrier below
    const int cam = unit / per_cam, u = unit - cam * per_cam;  // u = (ky * KH + kx) * 3 + color
    const int h = a.h, w = a.w, n = h * w;
    const float *p = a.psfs[cam] + (size_t)u * n;

    // 1. origin: the central element, or the centroid
    double ox = 0.5 * (w - 1), oy = 0.5 * (h - 1);
    if (a.center) {
        double s = 0.0, sx = 0.0, sy = 0.0;
        for (int e = lane; e < n; e += 64) {
            const int i = e / w, j = e - i * w;
            const double v = p[e];
            s += v;
            sx += v * j;
            sy += v * i;
        }
        s = wave_sum_d(s);
        sx = wave_sum_d(sx);
        sy = wave_sum_d(sy);
        if (s != 0.0) {  // a zero-sum kernel keeps the central element
            ox = sx / s;
            oy = sy / s;
        }
    }

    // 2. row sums
    const int k = lane & 7, i0 = lane >> 3;
    const double fk = (2 * k + 1) * 0.0625;  // u_k / pi
    for (int i = i0; i < h; i += 8) {
        double rc = 0.0, rs = 0.0;
        const float *pr = p + i * w;
        for (int j = 0; j < w; j++) {
            double sv, cv;
            sincospi(fk * (j - ox), &sv, &cv);
            const double v = pr[j];
            rc = fma(v, cv, rc);
            rs = fma(v, sv, rs);
        }
        rows[wave][0][i][k] = rc;
        rows[wave][1][i][k] = rs;
    }
    wave_lds_sync();

    // 3. column sums of bin (l, k)
    const int l = lane >> 3;
    const double fl = (2 * l + 1) * 0.0625;
    double cc = 0.0, scq = 0.0, cs = 0.0, ss = 0.0;
    for (int i = 0; i < h; i++) {
        double sv, cv;
        sincospi(fl * (i - oy), &sv, &cv);
        const double rc = rows[wave][0][i][k], rs = rows[wave][1][i][k];
        cc = fma(cv, rc, cc);
        scq = fma(cv, rs, scq);
        cs = fma(sv, rc, cs);
        ss = fma(sv, rs, ss);
    }
    float *kout = a.kernels[cam] + (size_t)u * 256 + l * 8 + k;
    kout[0] = (float)cc;
    kout[64] = (float)scq;
    kout[128] = (float)cs;
    kout[192] = (float)ss;

    // CltExtra: centroid (relative to the central element) and cell centre;
    // the slopes follow in k_psf_slopes
    if (lane < 4) {
        const int cell = u / 3, ky = cell / a.KH, kx = cell - ky * a.KH;
        const float v = lane == 0 ? (float)(ox - 0.5 * (w - 1))
                      : lane == 1 ? (float)(oy - 0.5 * (h - 1))
                      : lane == 2 ? (float)(16 * kx - 8)
                                  : (float)(16 * ky - 8);
        a.koffs[cam][(size_t)u * 8 + lane] = v;
    }
}

// d(centroid)/d(pixel) over the kernel grid: one thread per (camera, cell,
// color); reads data_x / data_y of the neighbouring cells, writes the four
// slope fields (disjoint from what it reads).
__global__ __launch_bounds__(256) void k_psf_slopes(float *const *koffs, int KV, int KH) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int per_cam = KV * KH * 3;
    if (t >= 4 * per_cam) return;
    const int cam = t / per_cam, u = t - cam * per_cam;
    const int color = u % 3, cell = u / 3, ky = cell / KH, kx = cell - ky * KH;
    const float *o = koffs[cam];
    auto at = [&](int y, int x, int f) { return o[((size_t)(y * KH + x) * 3 + color) * 8 + f]; };
    // central differences inside the grid, one-sided at its edges, 0 for a
    // grid one cell wide
    auto grad = [&](int f, bool horiz) -> float {
        const int nn = horiz ? KH : KV, c = horiz ? kx : ky;
        if (nn < 2) return 0.f;
        const int lo = c > 0 ? c - 1 : 0, hi = c < nn - 1 ? c + 1 : nn - 1;
        const float a0 = horiz ? at(ky, lo, f) : at(lo, kx, f);
        const float a1 = horiz ? at(ky, hi, f) : at(hi, kx, f);
        return (float)(((double)a1 - (double)a0) / (16.0 * (hi - lo)));
    };
    float *d = koffs[cam] + (size_t)u * 8;
    d[4] = grad(0, true);   // dxc_dx
    d[5] = grad(0, false);  // dxc_dy
    d[6] = grad(1, true);   // dyc_dx
    d[7] = grad(1, false);  // dyc_dy
}

}  // namespace tp
