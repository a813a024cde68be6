// tp_common.hpp - device-side constants and helpers shared by the gfx950
// kernels of libtp.so.
//
// All tables are closed forms evaluated at compile time (constexpr), so the
// library needs no initialisation pass except when a caller overrides a
// low-pass table (tp_set_lpf).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tp.h"

// Tuning knobs (build-time): minimum waves per SIMD the register allocator
// must allow, and work per workgroup.
#ifndef TP_CONVERT_WAVES
#define TP_CONVERT_WAVES 4
#endif
#ifndef TP_CORR_WAVES
#define TP_CORR_WAVES 2
#endif
#ifndef TP_CONVERT_PPB
#define TP_CONVERT_PPB 4   // tile pairs per workgroup
#endif
#ifndef TP_CORR_EPB
#define TP_CORR_EPB 64     // correlation entries per workgroup (multiple of 32)
#endif

namespace tp {

// ---------------------------------------------------------------------
// compile-time math (double precision, used only to build float tables)
// ---------------------------------------------------------------------
constexpr double kPi = 3.141592653589793238462643383279502884;

constexpr double cx_reduce(double x) {  // to [-pi, pi]
    double t = x / (2.0 * kPi);
    long n = (long)(t + (t >= 0 ? 0.5 : -0.5));
    return x - 2.0 * kPi * (double)n;
}
constexpr double cx_sin(double x) {
    x = cx_reduce(x);
    double term = x, sum = x;
    for (int n = 1; n < 22; n++) {
        term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double cx_cos(double x) {
    x = cx_reduce(x);
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 22; n++) {
        term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}
constexpr double cx_exp(double x) {  // any sign, |x| < ~700
    double a = x / 1024.0, term = 1.0, sum = 1.0;
    for (int n = 1; n < 20; n++) {
        term *= a / n;
        sum += term;
    }
    for (int i = 0; i < 10; i++) sum *= sum;
    return sum;
}

// 8-point transform matrices, [k][n]:
//   0 DCT-IV  1/2 cos(pi/8 (n+1/2)(k+1/2))
//   1 DST-IV  1/2 sin(pi/8 (n+1/2)(k+1/2))
//   2 DCT-II  1/2 cos(pi (2n+1) k / 16)      (DC row not divided by sqrt 2)
//   3 DST-II  1/2 sin(pi (2n+1)(k+1) / 16)   (last row not divided by sqrt 2)
constexpr double dtt_coef(int kind, int k, int n) {
    return kind == 0 ? 0.5 * cx_cos(kPi / 8.0 * (n + 0.5) * (k + 0.5))
         : kind == 1 ? 0.5 * cx_sin(kPi / 8.0 * (n + 0.5) * (k + 0.5))
         : kind == 2 ? 0.5 * cx_cos(kPi * (2 * n + 1) * k / 16.0)
                     : 0.5 * cx_sin(kPi * (2 * n + 1) * (k + 1) / 16.0);
}

// half sine window sin(pi (2i+1)/32), i = 0..7
constexpr double hwin(int i) { return cx_sin(kPi * (2 * i + 1) / 32.0); }

// Gaussian low-pass response at the DCT-IV bin centres (normalised to 1 at
// bin 0), outer product -> 8x8
constexpr double lpf1(double sigma, int k) {
    double s = 1.0;
    for (int x = 1; x < 8; x++)
        s += 2.0 * cx_exp(-(double)(x * x) / (2.0 * sigma)) * cx_cos(kPi * (2 * k + 1) * x / 16.0);
    return s;
}

struct HostTables {
    float dtt[4][8][8];
    float hw[8];
    float lpf[6][64];
};

constexpr HostTables make_tables() {
    HostTables t{};
    for (int kd = 0; kd < 4; kd++)
        for (int k = 0; k < 8; k++)
            for (int n = 0; n < 8; n++) t.dtt[kd][k][n] = (float)dtt_coef(kd, k, n);
    for (int i = 0; i < 8; i++) t.hw[i] = (float)hwin(i);
    const double sig[6] = {0.9, 0.9, 0.6, 0.4, 0.5, 0.9};
    for (int w = 0; w < 6; w++) {
        double f[8] = {};
        for (int k = 0; k < 8; k++) f[k] = lpf1(sig[w], k);
        for (int i = 0; i < 8; i++)
            for (int j = 0; j < 8; j++) t.lpf[w][i * 8 + j] = (float)((f[i] / f[0]) * (f[j] / f[0]));
    }
    return t;
}

constexpr HostTables kTab = make_tables();

// ---------------------------------------------------------------------
// fold map of the 16-sample lapped window onto 8 transform inputs:
// source n -> folded rind(n), with signs for the cosine / sine branches.
// ---------------------------------------------------------------------
constexpr int fold_rind(int n) { return n + 4 < 8 ? n + 4 : (n + 4 < 16 ? 11 - n : n - 12); }
constexpr int fold_rcs(int n) { return n < 4 ? 1 : -1; }
constexpr int fold_rss(int n) { return n < 12 ? 1 : -1; }
// the two sources of folded index m: a (smaller) and b
constexpr int fold_src_a(int m) { return m >= 4 ? m - 4 : 11 - m; }
constexpr int fold_src_b(int m) { return m >= 4 ? 11 - m : 12 + m; }

// ---------------------------------------------------------------------
// device geometry
// ---------------------------------------------------------------------
struct Geo {
    int W, H, KH, KV, TX, TY, TYA;
};

inline Geo make_geo(const tp_geometry &g) {
    Geo r;
    r.W = g.img_width;
    r.H = g.img_height;
    r.KH = g.kernels_hor;
    r.KV = g.kernels_vert;
    r.TX = r.W / 8;
    r.TY = r.H / 8;
    r.TYA = (r.TY + 3) & ~3;
    return r;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Give each XCD a contiguous 1/8 of the work so neighbouring tiles,
// which share image rows and aberration kernels, meet in the same L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int full = nb >> 3, rem = nb & 7, x = b & 7, i = b >> 3;
    return x * full + (x < rem ? x : rem) + i;
}

// Order LDS traffic between lanes of one wavefront.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// DPP: value of lane (lane ^ 7), i.e. mirror within each group of 8 lanes
__device__ __forceinline__ float mirror8(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141 /* row_half_mirror */,
                                                   0xf, 0xf, false));
}

// y = M x for an 8-point transform of the given kind (matrix form)
template <int KIND>
__device__ __forceinline__ void dtt8(float (&x)[8]) {
    float y[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        float s = kTab.dtt[KIND][k][0] * x[0];
#pragma unroll
        for (int n = 1; n < 8; n++) s = fmaf(kTab.dtt[KIND][k][n], x[n], s);
        y[k] = s;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = y[k];
}

// complex multiply helper
__device__ __forceinline__ void cmul(float ar, float ai, float br, float bi, float &cr, float &ci) {
    cr = fmaf(ar, br, -ai * bi);
    ci = fmaf(ar, bi, ai * br);
}

}  // namespace tp
