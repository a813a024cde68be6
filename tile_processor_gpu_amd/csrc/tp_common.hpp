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

// Launch shapes and occupancy targets, chosen by measurement on MI355X (the
// alternatives tried and their timings are listed in the design notes; the
// losing variants are not kept in the source).
#define TP_CONVERT_WAVES 4     // forward CLT: minimum waves per SIMD
#define TP_CORR_WAVES 2        // correlation: minimum waves per SIMD
#define TP_CORR_WG_WAVES 3     // correlation: wavefronts per workgroup (8 entries each per pass)
#define TP_CORR_EPB (8 * TP_CORR_WG_WAVES)  // correlation entries per workgroup
#define TP_FUSED_WAVES 1       // fused kernel, matrix-core forms: minimum waves per SIMD (compiler's choice)
#define TP_FUSED_QWAVES 3      // fused kernel, FP32 form: minimum waves per SIMD
#define TP_FUSED_WG_WAVES 4    // fused kernel: wavefronts per workgroup
#define TP_STRIP_ROWS 2        // fused kernel: tile rows per strip of the visiting order
#define TP_TEX_WG_PER_CU 5     // textures: workgroups per CU in the (list-looping) grid
#define TP_TEX_TPW 2           // textures: list entries (tiles) per workgroup at a time
#define TP_TEX_WAVES 5         // textures: minimum waves per SIMD
#define TP_COMBINE_EPW 64      // pair combination: list entries per workgroup

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
    float tw_pre[4][2];   // fast DCT-IV: e^{-i pi (4m+1)/32}
    float tw_post[4][2];  // fast DCT-IV: e^{-i pi q/8} / 2
    float c2odd[4][4];    // fast DCT-II odd part: cos(pi (2n+1)(2m+1)/16) / 2
    float c2even[3];      // cos(pi 2/16)/2, cos(pi 4/16)/2, cos(pi 6/16)/2
};

constexpr HostTables make_tables() {
    HostTables t{};
    for (int kd = 0; kd < 4; kd++)
        for (int k = 0; k < 8; k++)
            for (int n = 0; n < 8; n++) t.dtt[kd][k][n] = (float)dtt_coef(kd, k, n);
    for (int i = 0; i < 8; i++) t.hw[i] = (float)hwin(i);
    for (int m = 0; m < 4; m++) {
        t.tw_pre[m][0] = (float)cx_cos(-kPi * (4 * m + 1) / 32.0);
        t.tw_pre[m][1] = (float)cx_sin(-kPi * (4 * m + 1) / 32.0);
        t.tw_post[m][0] = (float)(0.5 * cx_cos(-kPi * m / 8.0));
        t.tw_post[m][1] = (float)(0.5 * cx_sin(-kPi * m / 8.0));
        for (int n = 0; n < 4; n++) t.c2odd[m][n] = (float)(0.5 * cx_cos(kPi * (2 * n + 1) * (2 * m + 1) / 16.0));
    }
    t.c2even[0] = (float)(0.5 * cx_cos(kPi * 2 / 16.0));
    t.c2even[1] = (float)(0.5 * cx_cos(kPi * 4 / 16.0));
    t.c2even[2] = (float)(0.5 * cx_cos(kPi * 6 / 16.0));
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

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Global-memory views of generic pointers.  Pointers read from the
// per-camera pointer arrays are generic to the compiler, which then emits
// flat loads; those count against the LDS wait counter as well and
// serialise LDS waits behind outstanding HBM loads.
typedef __attribute__((address_space(1))) const float gfloat;
typedef __attribute__((address_space(1))) const f32x4 gf32x4;
__device__ __forceinline__ gfloat *gptr(const float *p) { return (gfloat *)p; }
typedef __attribute__((address_space(1))) const tp_task gtask;
__device__ __forceinline__ gtask *gtasks(const tp_task *p) { return (gtask *)p; }
// element i of a per-camera pointer array held in global memory
template <class T>
__device__ __forceinline__ T gelem(const T *arr, int i) {
    return ((__attribute__((address_space(1))) const T *)arr)[i];
}
__device__ __forceinline__ gf32x4 *gptr4(const float *p) { return (gf32x4 *)p; }

// 16-byte store of an output that is not read again by this kernel: marked
// non-temporal, so it does not displace the inputs that neighbouring
// workgroups re-read from L2.
__device__ __forceinline__ void store4(float *p, float a, float b, float c, float d) {
    typedef __attribute__((address_space(1))) f32x4 gv4;
    const f32x4 v = {a, b, c, d};
    __builtin_nontemporal_store(v, (gv4 *)p);
}
__device__ __forceinline__ void store1(float *p, float v) {
    typedef __attribute__((address_space(1))) float gv1;
    __builtin_nontemporal_store(v, (gv1 *)p);
}

// Order LDS traffic between lanes of one wavefront.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// An 8 x 8 block in LDS written one row per lane (two float4 halves) and
// read back one column per lane.  Rows 4..7 have their halves swapped: at
// the 32-byte row stride the 8 lanes' 16-byte writes would otherwise use
// half of the banks (a 2-way conflict on every write); with the swap they
// cover all of them.  Element (row, col) is at row * 8 + (col ^ (row & 4)).
__device__ __forceinline__ void put_row8(float *blk, int row, const float (&v)[8]) {
    const int s = row & 4;
    *(float4 *)(blk + row * 8 + s) = make_float4(v[0], v[1], v[2], v[3]);
    *(float4 *)(blk + row * 8 + (s ^ 4)) = make_float4(v[4], v[5], v[6], v[7]);
}
__device__ __forceinline__ float get8(const float *blk, int row, int col) {
    return blk[row * 8 + (col ^ (row & 4))];
}
// float4 half h (columns 4 h .. 4 h + 3) of row `row`
__device__ __forceinline__ const float *half8(const float *blk, int row, int h) {
    return blk + row * 8 + ((4 * h) ^ (row & 4));
}

// DPP: value of lane (lane ^ 7), i.e. mirror within each group of 8 lanes
__device__ __forceinline__ float mirror8(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141 /* row_half_mirror */,
                                                   0xf, 0xf, false));
}

// 8-point DCT-IV, y[k] = 1/2 sum_n x[n] cos(pi/8 (n+1/2)(k+1/2)), through a
// 4-point complex DFT: v[m] = x[2m] + i x[7-2m] twiddled by
// e^{-i pi (4m+1)/32}, DFT (only +-1, +-i factors), twiddled by
// e^{-i pi q/8}; then y[2q] = Re/2, y[7-2q] = -Im/2.  46 operations
// instead of 64.
__device__ __forceinline__ void dct4_8(float (&x)[8]) {
    float tr[4], ti[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const float vr = x[2 * m], vi = x[7 - 2 * m];
        const float wr = kTab.tw_pre[m][0], wi = kTab.tw_pre[m][1];
        tr[m] = fmaf(vr, wr, -vi * wi);
        ti[m] = fmaf(vr, wi, vi * wr);
    }
    // 4-point DFT T[q] = sum_m t[m] e^{-2 pi i m q / 4}
    const float s0r = tr[0] + tr[2], s0i = ti[0] + ti[2];
    const float d0r = tr[0] - tr[2], d0i = ti[0] - ti[2];
    const float s1r = tr[1] + tr[3], s1i = ti[1] + ti[3];
    const float d1r = tr[1] - tr[3], d1i = ti[1] - ti[3];
    float Tr[4], Ti[4];
    Tr[0] = s0r + s1r; Ti[0] = s0i + s1i;
    Tr[2] = s0r - s1r; Ti[2] = s0i - s1i;
    Tr[1] = d0r + d1i; Ti[1] = d0i - d1r;   // d0 - i d1
    Tr[3] = d0r - d1i; Ti[3] = d0i + d1r;   // d0 + i d1
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const float pr = kTab.tw_post[q][0], pi = kTab.tw_post[q][1];
        float ur, ui;
        if (q == 0) {
            ur = 0.5f * Tr[0];
            ui = 0.5f * Ti[0];
        } else {
            ur = fmaf(Tr[q], pr, -Ti[q] * pi);
            ui = fmaf(Tr[q], pi, Ti[q] * pr);
        }
        x[2 * q] = ur;
        x[7 - 2 * q] = -ui;
    }
}

// 8-point DST-IV, y[k] = 1/2 sum_n x[n] sin(pi/8 (n+1/2)(k+1/2)) =
// (-1)^k DCT-IV of the reversed input.
__device__ __forceinline__ void dst4_8(float (&x)[8]) {
    float t[8];
#pragma unroll
    for (int n = 0; n < 8; n++) t[n] = x[7 - n];
    dct4_8(t);
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = (k & 1) ? -t[k] : t[k];
}

// 8-point DCT-II, y[k] = 1/2 sum_n x[n] cos(pi (2n+1) k / 16) (the DC term
// not divided by sqrt 2), by the even/odd split: the even outputs are a
// 4-point DCT-II of x[n] + x[7-n], the odd ones a 4x4 product with
// x[n] - x[7-n].  36 operations instead of 64.
__device__ __forceinline__ void dct2_8(float (&x)[8]) {
    const float u0 = x[0] + x[7], u1 = x[1] + x[6], u2 = x[2] + x[5], u3 = x[3] + x[4];
    const float v0 = x[0] - x[7], v1 = x[1] - x[6], v2 = x[2] - x[5], v3 = x[3] - x[4];
    const float a0 = u0 + u3, a1 = u1 + u2, b0 = u0 - u3, b1 = u1 - u2;
    const float h2 = kTab.c2even[0], h4 = kTab.c2even[1], h6 = kTab.c2even[2];
    x[0] = 0.5f * (a0 + a1);
    x[4] = h4 * (a0 - a1);
    x[2] = fmaf(h2, b0, h6 * b1);
    x[6] = fmaf(h6, b0, -h2 * b1);
    const float v[4] = {v0, v1, v2, v3};
#pragma unroll
    for (int m = 0; m < 4; m++) {
        float s = kTab.c2odd[m][0] * v[0];
#pragma unroll
        for (int n = 1; n < 4; n++) s = fmaf(kTab.c2odd[m][n], v[n], s);
        x[2 * m + 1] = s;
    }
}

// 8-point DST-II, y[k] = 1/2 sum_n x[n] sin(pi (2n+1)(k+1) / 16): the DCT-II
// of (-1)^n x[n], read backwards.
__device__ __forceinline__ void dst2_8(float (&x)[8]) {
    float t[8];
#pragma unroll
    for (int n = 0; n < 8; n++) t[n] = (n & 1) ? -x[n] : x[n];
    dct2_8(t);
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = t[7 - k];
}

// complex multiply helper
__device__ __forceinline__ void cmul(float ar, float ai, float br, float bi, float &cr, float &ci) {
    cr = fmaf(ar, br, -ai * bi);
    ci = fmaf(ar, bi, ai * br);
}

}  // namespace tp
