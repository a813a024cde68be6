/*
 * tp_oracle.cpp - CPU oracle of the tile processor hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness reference for the
 * HIP kernels of libtp.so and the CPU baseline timed beside them.  Only the
 * tests, the smoke check and the benchmark's CPU-baseline leg may load it;
 * the product never calls it.
 *
 * It restates, per tile, the algorithm of the reference (mfkiwl/
 * tile_processor_gpu, src/TileProcessor.cuh and src/dtt8x8.cuh) in plain
 * C++, templated on the arithmetic type: T = float is the FP32 oracle,
 * T = double the FP64 one.  Every function names the reference lines it
 * follows.  Deliberate differences from the reference (all documented in the
 * public header or below) are:
 *   - border pixels: the reference's border gather reads outside the image;
 *     here a pixel outside the image is read from the nearest pixel of the
 *     same Bayer color (clamp that keeps row/column parity);
 *   - right-edge window snap uses the intended (itlx & 1) + W - 10 (the
 *     reference's "itlx & 1 + (W - 10)" is a precedence slip);
 *   - the residual shift is formed as ((itlx + 8) - cx) + offset, which is
 *     exact in float, instead of itlx - px with px rounded in float (that
 *     loses up to 1.2e-4 px at x ~ 2600);
 *   - correlation list entries with pair > 5 or a tile outside the grid,
 *     and tasks whose (tx, ty) is outside the grid, are skipped;
 *   - imclt_rbg debug writes (tile (0,0) := 100, tile (161,111) doubled) are
 *     not reproduced;
 *   - texture rms: a negative mean-square residue (equal values across
 *     cameras) is clamped to 0 before the square root (the reference's
 *     sqrtf returns NaN there).
 * The 8-point DCT-IV/DST-IV/DCT-II/DST-II butterflies of the reference are
 * replaced by their matrices (the reference's butterflies equal the
 * textbook orthonormal DCT-IV/DST-IV, and DCT-II/DST-II with the first /
 * last row not divided by sqrt(2)); tests pin those matrices against scipy.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <vector>
#include <omp.h>

#include "../include/tp.h"

namespace {

constexpr int DTT = 8;
constexpr double PI = 3.141592653589793238462643383279502884;

struct Geo {
    int W, H, KH, KV, TX, TY, TYA;
};

Geo geo_of(const tp_geometry *g) {
    Geo r;
    r.W = g->img_width;
    r.H = g->img_height;
    r.KH = g->kernels_hor;
    r.KV = g->kernels_vert;
    r.TX = r.W / DTT;
    r.TY = r.H / DTT;
    r.TYA = (r.TY + 3) & ~3;
    return r;
}

/* ------------------------------------------------------------------ */
/* Closed-form tables (TileProcessor.cuh:374-482, :492-1005;           */
/* dtt8x8.cuh:58-86)                                                   */
/* ------------------------------------------------------------------ */

/* Fold of a 16-sample lapped window onto the 8 DCT-IV/DST-IV inputs.
 * The lapped transform kernel is cos/sin(pi/8 (n + 4.5)(k + 1/2)) for
 * n = 0..15; with m = n + 4 the DCT-IV/DST-IV kernel index, m in 8..15
 * reflects to 15 - m (cos flips sign, sin keeps it) and m in 16..19 wraps
 * to m - 16 (both flip).  Same map as get_fold_rindices
 * (TileProcessor.cuh:254-296). */
struct Fold {
    int rind[16];
    int rcs[16];  /* sign, cosine branch */
    int rss[16];  /* sign, sine branch   */
};

Fold make_fold() {
    Fold f;
    for (int n = 0; n < 16; n++) {
        int m = n + 4;
        if (m < 8) {
            f.rind[n] = m;      f.rcs[n] = 1;  f.rss[n] = 1;
        } else if (m < 16) {
            f.rind[n] = 15 - m; f.rcs[n] = -1; f.rss[n] = 1;
        } else {
            f.rind[n] = m - 16; f.rcs[n] = -1; f.rss[n] = -1;
        }
    }
    return f;
}

const Fold FOLD = make_fold();

/* half window HWINDOW[i] = sin(pi (2i+1) / 32), i = 0..7 (:376) */
double hwindow(int i) { return std::sin(PI * (2 * i + 1) / 32.0); }

/* Gaussian low-pass response sampled at the DCT-IV bin centres, normalised
 * to 1 at bin 0: f[k] ~ g(0) + 2 sum_{x=1..7} g(x) cos(pi (2k+1) x / 16),
 * g(x) = exp(-x^2 / (2 sigma)); 2-D table = outer(f, f).  This reproduces
 * lpf_data (R,B sigma 0.9, G 0.6, mono 0.4), lpf_rb_corr (0.5) and
 * lpf_corr (0.9) of TileProcessor.cuh:421-482. */
void make_lpf(double sigma, double out[64]) {
    double f[8];
    for (int k = 0; k < 8; k++) {
        double s = 1.0;
        for (int x = 1; x < 8; x++)
            s += 2.0 * std::exp(-x * x / (2.0 * sigma)) * std::cos(PI * (2 * k + 1) * x / 16.0);
        f[k] = s;
    }
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++)
            out[i * 8 + j] = (f[i] / f[0]) * (f[j] / f[0]);
}

/* alphaFade[16][256] (:492-1005, generator :316-369): fade(i) =
 * (1 - cos(pi (i + 1/2) / 16)) / 2; the 4-bit mask is the set of N,E,S,W
 * neighbours that exist; missing sides fade out.  Masks not listed keep 1,
 * mask 0 is all 0. */
double alpha_fade(int m, int i, int j) {
    auto fade = [](int t) { return 0.5 * (1.0 - std::cos(PI * (t + 0.5) / 16.0)); };
    switch (m) {
    case 0:  return 0.0;
    case 1:  return fade(15 - i);
    case 2:  return fade(j);
    case 4:  return fade(i);
    case 8:  return fade(15 - j);
    case 3:  return (j > 15 - i) ? fade(j) : fade(15 - i);
    case 6:  return (j > i) ? fade(j) : fade(i);
    case 9:  return (j > i) ? fade(15 - i) : fade(15 - j);
    case 12: return (i > 15 - j) ? fade(i) : fade(15 - j);
    default: return 1.0;
    }
}

/* 8-point transforms.  kind: 0 DCT-IV, 1 DST-IV (dctiv_nodiverg /
 * dstiv_nodiverg, dtt8x8.cuh:621-745), 2 DCT-II, 3 DST-II
 * (dttii_shared_mem_nonortho, dtt8x8.cuh:382-463). */
double dtt_coef(int kind, int k, int n) {
    switch (kind) {
    case 0: return 0.5 * std::cos(PI / 8.0 * (n + 0.5) * (k + 0.5));
    case 1: return 0.5 * std::sin(PI / 8.0 * (n + 0.5) * (k + 0.5));
    case 2: return 0.5 * std::cos(PI * (2 * n + 1) * k / 16.0);
    default: return 0.5 * std::sin(PI * (2 * n + 1) * (k + 1) / 16.0);
    }
}

template <typename T>
struct Tables {
    T dtt[4][8][8];
    T hw[8];        /* HWINDOW     */
    T hw2[8];       /* HWINDOW2    */
    T hwsq[8];      /* HWINDOW_SQ  */
    T hwsqi[8];     /* HWINDOW_SQi */
    T lpf[6][64];   /* lpf_data[4], lpf_rb_corr, lpf_corr */
    T afade[16][256];
    Tables() {
        for (int t = 0; t < 4; t++)
            for (int k = 0; k < 8; k++)
                for (int n = 0; n < 8; n++) dtt[t][k][n] = (T)dtt_coef(t, k, n);
        for (int i = 0; i < 8; i++) {
            double w = hwindow(i);
            hw[i] = (T)w;
            hw2[i] = (T)(0.5 * w);
            hwsq[i] = (T)(w * w);
            hwsqi[i] = (T)(1.0 / (w * w));
        }
        const double sig[6] = {0.9, 0.9, 0.6, 0.4, 0.5, 0.9};
        for (int t = 0; t < 6; t++) {
            double l[64];
            make_lpf(sig[t], l);
            for (int i = 0; i < 64; i++) lpf[t][i] = (T)l[i];
        }
        for (int m = 0; m < 16; m++)
            for (int i = 0; i < 16; i++)
                for (int j = 0; j < 16; j++) afade[m][i * 16 + j] = (T)alpha_fade(m, i, j);
    }
};

template <typename T>
Tables<T> &tables() {
    static Tables<T> t;
    return t;
}

/* user-set low-pass tables (tp_set_lpf counterpart); NaN = use default */
float g_lpf_user[6][64];
bool g_lpf_set[6] = {false, false, false, false, false, false};

template <typename T>
const T *lpf_table(int which, T buf[64]) {
    if (g_lpf_set[which]) {
        for (int i = 0; i < 64; i++) buf[i] = (T)g_lpf_user[which][i];
        return buf;
    }
    return tables<T>().lpf[which];
}

/* ------------------------------------------------------------------ */
/* 2-D transforms on [8][8] (row = vertical frequency/position,        */
/* column = horizontal)                                                */
/* ------------------------------------------------------------------ */

template <typename T>
void dtt_rows(T a[8][8], int kind) {  /* along the horizontal index (inc = 1) */
    const auto &M = tables<T>().dtt[kind];
    for (int r = 0; r < 8; r++) {
        T y[8];
        for (int k = 0; k < 8; k++) {
            T s = 0;
            for (int n = 0; n < 8; n++) s += M[k][n] * a[r][n];
            y[k] = s;
        }
        for (int k = 0; k < 8; k++) a[r][k] = y[k];
    }
}

template <typename T>
void dtt_cols(T a[8][8], int kind) {  /* along the vertical index (inc = 9) */
    const auto &M = tables<T>().dtt[kind];
    for (int c = 0; c < 8; c++) {
        T y[8];
        for (int k = 0; k < 8; k++) {
            T s = 0;
            for (int n = 0; n < 8; n++) s += M[k][n] * a[n][c];
            y[k] = s;
        }
        for (int k = 0; k < 8; k++) a[k][c] = y[k];
    }
}

/* pixel fetch with a Bayer-preserving clamp (border policy, see header) */
inline int clamp_bayer(int v, int n) {
    if (v < 0) return v & 1;
    if (v >= n) return n - 2 + (v & 1);
    return v;
}

/* ------------------------------------------------------------------ */
/* Forward CLT of one tile, one camera, one color                     */
/* convertCorrectTile, TileProcessor.cuh:3106-3634                     */
/* ------------------------------------------------------------------ */
template <typename T>
void convert_correct_tile(const Geo &g,
                          const tp_clt_extra *koffs,  /* one camera */
                          const float *kernels,       /* one camera */
                          const float *image, size_t dstride,
                          int color, int lpf_mask, float cx, float cy,
                          T q[4][8][8]) {
    const Tables<T> &tb = tables<T>();
    /* (1) kernel cell, :3129-3134 */
    int ktx = (int)lrintf((float)((double)cx * (1.0 / 16) + 1.0));
    int kty = (int)lrintf((float)((double)cy * (1.0 / 16) + 1.0));
    ktx = std::min(g.KH - 1, std::max(0, ktx));
    kty = std::min(g.KV - 1, std::max(0, kty));
    float kdx = cx - (float)(ktx << TP_KERNELS_LSTEP) + (float)(1 << (TP_KERNELS_LSTEP - 1));
    float kdy = cy - (float)(kty << TP_KERNELS_LSTEP) + (float)(1 << (TP_KERNELS_LSTEP - 1));
    int kfi = (ktx + kty * g.KH) * TP_NUM_COLORS + color;
    const float *kern = kernels + (size_t)kfi * 256;
    const tp_clt_extra &ex = koffs[kfi];

    /* (2) integer top-left and residual shift, :3182-3247 */
    float offx = ex.data_x + ex.dxc_dx * kdx + ex.dxc_dy * kdy;
    float offy = ex.data_y + ex.dyc_dx * kdx + ex.dyc_dy * kdy;
    float px = (cx - (float)DTT) - offx;
    float py = (cy - (float)DTT) - offy;
    int itlx = (int)std::floor(px + 0.5f);
    int itly = (int)std::floor(py + 0.5f);
    if (itlx < 0) itlx &= 1;
    if (itlx >= g.W - DTT) itlx = (itlx & 1) + (g.W - DTT - 2);
    if (itly < 0) itly &= 1;
    if (itly >= g.H - DTT) itly = (itly & 1) + (g.H - DTT - 2);
    T sh, sv;
    if (sizeof(T) == sizeof(float)) {
        sh = (T)(((float)(itlx + DTT) - cx) + offx);
        sv = (T)(((float)(itly + DTT) - cy) + offy);
    } else {
        double dox = (double)ex.data_x + (double)ex.dxc_dx * kdx + (double)ex.dxc_dy * kdy;
        double doy = (double)ex.data_y + (double)ex.dyc_dx * kdx + (double)ex.dyc_dy * kdy;
        sh = (T)(((double)itlx + DTT - (double)cx) + dox);
        sv = (T)(((double)itly + DTT - (double)cy) + doy);
    }

    /* (3) shifted windows with the fold signs embedded, :3195-3262.
     * w_s(n) = sin(pi (n + 1/2 + s) / 16) by angle addition on HWINDOW. */
    T ahc = (T)std::cos(PI * (double)sh / 16.0), ahs = (T)std::sin(PI * (double)sh / 16.0);
    T avc = (T)std::cos(PI * (double)sv / 16.0), avs = (T)std::sin(PI * (double)sv / 16.0);
    T wh[16], wv[16];
    for (int i = 0; i < 8; i++) {
        wh[i] = tb.hw[i] * ahc + tb.hw[7 - i] * ahs;
        wh[i + 8] = tb.hw[7 - i] * ahc - tb.hw[i] * ahs;
        wv[i] = tb.hw[i] * avc + tb.hw[7 - i] * avs;
        wv[i + 8] = tb.hw[7 - i] * avc - tb.hw[i] * avs;
    }

    /* (4) Bayer fold into the DCT (and for G the DST) buffers, :3289-3398.
     * One pass for R and B (1/4 of the sites), two for G (two sub-lattices,
     * summed afterwards). */
    T buf[4][8][8];
    std::memset(buf, 0, sizeof(buf));
    const int color0 = color & 1, color1 = (color >> 1) & 1;
    for (int gpass = 0; gpass <= color1; gpass++) {
        int pc = (itlx & 1) ^ (1 /* BAYER_RED_COL */ ^ color0 ^ color1 ^ gpass);
        int pr = (itly & 1) ^ (0 /* BAYER_RED_ROW */ ^ color0 ^ gpass);
        for (int r = pr; r < 16; r += 2) {
            int yy = clamp_bayer(itly + r, g.H);
            for (int c = pc; c < 16; c += 2) {
                int xx = clamp_bayer(itlx + c, g.W);
                T d = (T)image[(size_t)yy * dstride + xx] * (wv[r] * (T)FOLD.rcs[r]);
                buf[2 * gpass][FOLD.rind[r]][FOLD.rind[c]] = d * (wh[c] * (T)FOLD.rcs[c]);
                if (color1)
                    buf[2 * gpass + 1][FOLD.rind[r]][FOLD.rind[c]] = d * (wh[c] * (T)FOLD.rss[c]);
            }
        }
    }
    if (color1) {
        for (int i = 0; i < 8; i++)
            for (int j = 0; j < 8; j++) {
                buf[0][i][j] += buf[2][i][j];
                buf[1][i][j] += buf[3][i][j];
            }
    }
    /* horizontal DCT-IV (G: also DST-IV), then vertical DCT-IV, :3408-3441 */
    dtt_rows(buf[0], 0);
    if (color1) dtt_rows(buf[1], 1);
    dtt_cols(buf[0], 0);
    if (color1) dtt_cols(buf[1], 0);

    /* replicate the missing quadrants from the Bayer symmetry, :3457-3497 */
    if (!color1) {
        int neg = 1 - (((itlx & 1) ^ (1 /*RED_COL*/ ^ color)) << 1);
        for (int i = 0; i < 8; i++)
            for (int j = 0; j < 8; j++) buf[1][i][j ^ 7] = (T)neg * buf[0][i][j];
    }
    {
        int neg = 1 - (((itlx & 1) ^ (itly & 1) ^ (1 ^ 0 ^ color1)) << 1);
        for (int i = 0; i < 8; i++)
            for (int j = 0; j < 8; j++) {
                buf[3][7 - i][j ^ 7] = (T)neg * buf[0][i][j];
                buf[2][7 - i][j ^ 7] = (T)neg * buf[1][i][j];
            }
    }

    /* (5) aberration kernel multiply, convolveTiles :2778-2827 */
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) {
            T k0 = (T)kern[0 * 64 + i * 8 + j], k1 = (T)kern[1 * 64 + i * 8 + j];
            T k2 = (T)kern[2 * 64 + i * 8 + j], k3 = (T)kern[3 * 64 + i * 8 + j];
            T c0 = buf[0][i][j], c1 = buf[1][i][j], c2 = buf[2][i][j], c3 = buf[3][i][j];
            T r0 = c0 * k0; r0 -= c1 * k1; r0 -= c2 * k2; r0 += c3 * k3;
            T r1 = c1 * k0; r1 += c0 * k1; r1 -= c3 * k2; r1 -= c2 * k3;
            T r2 = c2 * k0; r2 -= c3 * k1; r2 += c0 * k2; r2 -= c1 * k3;
            T r3 = c3 * k0; r3 += c2 * k1; r3 += c1 * k2; r3 += c0 * k3;
            buf[0][i][j] = r0; buf[1][i][j] = r1; buf[2][i][j] = r2; buf[3][i][j] = r3;
        }

    /* (6) fractional shifts: shiftTileHor :2715-2745, shiftTileVert :2747-2776 */
    for (int j = 0; j < 8; j++) {
        double a = PI * (double)(sh * (T)(2 * j + 1) * (T)(1.0 / 16));
        T ch = (T)std::cos(a), s = (T)std::sin(a);
        for (int i = 0; i < 8; i++) {
            T a0 = buf[0][i][j], b0 = buf[1][i][j];
            buf[0][i][j] = a0 * ch - b0 * s;
            buf[1][i][j] = a0 * s + b0 * ch;
            T a2 = buf[2][i][j], b2 = buf[3][i][j];
            buf[2][i][j] = a2 * ch - b2 * s;
            buf[3][i][j] = a2 * s + b2 * ch;
        }
    }
    for (int i = 0; i < 8; i++) {
        double a = PI * (double)(sv * (T)(2 * i + 1) * (T)(1.0 / 16));
        T ch = (T)std::cos(a), s = (T)std::sin(a);
        for (int j = 0; j < 8; j++) {
            T a0 = buf[0][i][j], b0 = buf[2][i][j];
            buf[0][i][j] = a0 * ch - b0 * s;
            buf[2][i][j] = a0 * s + b0 * ch;
            T a1 = buf[1][i][j], b1 = buf[3][i][j];
            buf[1][i][j] = a1 * ch - b1 * s;
            buf[3][i][j] = a1 * s + b1 * ch;
        }
    }
    /* (7) optional LPF, :3572-3590 */
    if ((lpf_mask >> color) & 1) {
        T lbuf[64];
        const T *l = lpf_table<T>(color, lbuf);
        for (int qq = 0; qq < 4; qq++)
            for (int i = 0; i < 8; i++)
                for (int j = 0; j < 8; j++) buf[qq][i][j] *= l[i * 8 + j];
    }
    std::memcpy(q, buf, sizeof(buf));
}

/* convert_correct_tiles, TileProcessor.cuh:1859-1937 */
template <typename T>
void convert_correct_tiles(const Geo &g, const tp_clt_extra *const *koffs,
                           const float *const *kernels, const float *const *images,
                           const tp_task *tasks, T *const *clt, size_t dstride,
                           int num_tiles, int lpf_mask) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int t = 0; t < num_tiles; t++) {
        const tp_task &tk = tasks[t];
        if (!tk.task) continue;
        int tx = tk.txy & 0xfff, ty = tk.txy >> 16;
        if (tx >= g.TX || ty < 0 || ty >= g.TY) continue;  /* outside the tile grid */
        for (int cam = 0; cam < TP_NUM_CAMS; cam++)
            for (int color = 0; color < TP_NUM_COLORS; color++) {
                T q[4][8][8];
                convert_correct_tile<T>(g, koffs[cam], kernels[cam], images[cam], dstride, color,
                                        lpf_mask, tk.xy[cam][0], tk.xy[cam][1], q);
                T *dst = clt[cam] + ((size_t)(ty * g.TX + tx) * TP_NUM_COLORS + color) * 256;
                std::memcpy(dst, q, sizeof(q));
            }
    }
}

/* ------------------------------------------------------------------ */
/* Correlation: correlate2D :1176-1443, correlateAccumulateTiles        */
/* :2829-2890, normalizeTileAmplitude :2915-2943, corrUnfoldTile        */
/* :2950-3008                                                           */
/* ------------------------------------------------------------------ */
const int PAIRS[6][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {0, 3}, {2, 1}};

template <typename T>
void corr_unfold(int r, T q[4][8][8], T *rslt) {
    /* (2r+1)^2 output, centre at (r, r); 1/4 scale; rows i and -i from
     * quadrant rows i (CC, SC) and i-1 (CS, SS), columns j and -j from
     * quadrant columns j (CC, CS) and j-1 (SC, SS). */
    const int n = 2 * r + 1, c = r;
    const T s = (T)0.25;
    for (int i = 0; i <= r; i++) {
        for (int j = 0; j <= r; j++) {
            T a0 = q[0][i][j];
            T a1 = j ? q[1][i][j - 1] : (T)0;
            T a2 = i ? q[2][i - 1][j] : (T)0;
            T a3 = (i && j) ? q[3][i - 1][j - 1] : (T)0;
            rslt[(c + i) * n + c + j] = s * (a0 + a1 + a2 + a3);
            rslt[(c + i) * n + c - j] = s * (a0 - a1 + a2 - a3);
            rslt[(c - i) * n + c + j] = s * (a0 + a1 - a2 - a3);
            rslt[(c - i) * n + c - j] = s * (a0 - a1 - a2 + a3);
        }
    }
}

template <typename T>
void correlate_one(const Geo &g, const T *const *clt, int colors, const float scales[3],
                   float fat_zero, int tile, int pair, int radius, T *out) {
    const int cam1 = PAIRS[pair][0], cam2 = PAIRS[pair][1];
    T C[4][8][8];
    std::memset(C, 0, sizeof(C));
    T lb1[64], lb2[64];
    const T *lpf_rb = lpf_table<T>(4, lb1);
    const T *lpf_c = lpf_table<T>(5, lb2);
    for (int color = 0; color < colors; color++) {
        const T *a = clt[cam1] + ((size_t)tile * TP_NUM_COLORS + color) * 256;
        const T *b = clt[cam2] + ((size_t)tile * TP_NUM_COLORS + color) * 256;
        T sc = (T)scales[color];
        for (int i = 0; i < 64; i++) {
            T a0 = a[i], a1 = a[64 + i], a2 = a[128 + i], a3 = a[192 + i];
            T b0 = b[i], b1 = b[64 + i], b2 = b[128 + i], b3 = b[192 + i];
            /* conj(A) * B in the bicomplex algebra i^2 = j^2 = -1, ij = ji */
            T r0 = a0 * b0; r0 += a1 * b1; r0 += a2 * b2; r0 += a3 * b3;
            T r1 = -a1 * b0; r1 += a0 * b1; r1 -= a3 * b2; r1 += a2 * b3;
            T r2 = -a2 * b0; r2 -= a3 * b1; r2 += a0 * b2; r2 += a1 * b3;
            T r3 = a3 * b0; r3 -= a2 * b1; r3 -= a1 * b2; r3 += a0 * b3;
            C[0][i >> 3][i & 7] += sc * r0;
            C[1][i >> 3][i & 7] += sc * r1;
            C[2][i >> 3][i & 7] += sc * r2;
            C[3][i >> 3][i & 7] += sc * r3;
        }
        if (color == 1) {  /* low-pass R+B only, :1263-1304 */
            for (int qq = 0; qq < 4; qq++)
                for (int i = 0; i < 64; i++) C[qq][i >> 3][i & 7] *= lpf_rb[i];
        }
    }
    /* normalise amplitude with fat zero, :2915-2943 */
    T fz2 = (T)fat_zero * (T)fat_zero;
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) {
            T s2 = fz2 + C[0][i][j] * C[0][i][j] + C[1][i][j] * C[1][i][j] +
                   C[2][i][j] * C[2][i][j] + C[3][i][j] * C[3][i][j];
            T sc = (T)1 / std::sqrt(s2);
            for (int qq = 0; qq < 4; qq++) C[qq][i][j] *= sc;
        }
    for (int qq = 0; qq < 4; qq++)  /* :1337-1347 */
        for (int i = 0; i < 64; i++) C[qq][i >> 3][i & 7] *= lpf_c[i];
    /* back to pixel domain: vertical then horizontal DCT-II/DST-II, :1368-1388 */
    for (int qq = 0; qq < 4; qq++) dtt_cols(C[qq], ((qq >> 1) & 1) ? 3 : 2);
    for (int qq = 0; qq < 4; qq++) dtt_rows(C[qq], (qq & 1) ? 3 : 2);
    corr_unfold<T>(radius, C, out);
}

template <typename T>
void correlate2D(const Geo &g, const T *const *clt, int colors, const float scales[3],
                 float fat_zero, size_t n, const int *idx, size_t stride, int radius,
                 T *corrs) {
#pragma omp parallel for schedule(dynamic, 64)
    for (long e = 0; e < (long)n; e++) {
        int v = idx[e];
        int tile = v >> TP_CORR_NTILE_SHIFT;
        int pair = v & ((1 << TP_CORR_NTILE_SHIFT) - 1);
        if (pair >= TP_NUM_PAIRS || tile < 0 || tile >= g.TX * g.TY) continue;
        correlate_one<T>(g, clt, colors, scales, fat_zero, tile, pair, radius,
                         corrs + (size_t)e * stride);
    }
}

/* ------------------------------------------------------------------ */
/* Inverse CLT of one tile/color to a 16x16 windowed block             */
/* imclt :3700-3838 (horizontal first), imclt8threads :3847-4017        */
/* (vertical first).                                                    */
/* ------------------------------------------------------------------ */
template <typename T>
void imclt_tile_ws(const T *src /* [4][8][8] */, T wgt[16][16], T sum[16][16],
                   int horizontal_first) {
    const Tables<T> &tb = tables<T>();
    T q[4][8][8];
    std::memcpy(q, src, sizeof(q));
    if (horizontal_first) {
        dtt_rows(q[0], 0); dtt_rows(q[2], 0); dtt_rows(q[1], 1); dtt_rows(q[3], 1);
        dtt_cols(q[0], 0); dtt_cols(q[1], 0); dtt_cols(q[2], 1); dtt_cols(q[3], 1);
    } else {
        dtt_cols(q[0], 0); dtt_cols(q[1], 0); dtt_cols(q[2], 1); dtt_cols(q[3], 1);
        dtt_rows(q[0], 0); dtt_rows(q[2], 0); dtt_rows(q[1], 1); dtt_rows(q[3], 1);
    }
    /* unfold (idct_signs :399-420, imclt_indx9 :392): output row r and column
     * c read folded row rind[r], column rind[c]; quadrant q takes the sine
     * signs in the direction(s) where it holds DST terms. */
    for (int r = 0; r < 16; r++) {
        T wr = tb.hw2[r < 8 ? r : 15 - r];
        for (int c = 0; c < 16; c++) {
            T hwc = tb.hw2[c < 8 ? c : 15 - c];
            T d[4];
            for (int qq = 0; qq < 4; qq++) {
                int sv = (qq & 2) ? FOLD.rss[r] : FOLD.rcs[r];
                int sh = (qq & 1) ? FOLD.rss[c] : FOLD.rcs[c];
                d[qq] = (T)(sv * sh) * q[qq][FOLD.rind[r]][FOLD.rind[c]];
            }
            sum[r][c] = (d[0] + d[1]) + (d[2] + d[3]);
            wgt[r][c] = wr * hwc;  /* HWINDOW2[row] * HWINDOW2[col] */
        }
    }
}

/* imclt8threads without accumulation: out = w * d */
template <typename T>
void imclt_tile(const T *src, T out[16][16], int horizontal_first) {
    T w[16][16], s[16][16];
    imclt_tile_ws<T>(src, w, s, horizontal_first);
    for (int r = 0; r < 16; r++)
        for (int c = 0; c < 16; c++) out[r][c] = w[r][c] * s[r][c];
}

/* w * d + acc rounded toward -infinity (the reference's __fmaf_rd), from
 * the exact product in double. */
inline float fma_rd(float w, float d, float acc) {
    double e = (double)w * (double)d + (double)acc;
    float f = (float)e;
    if ((double)f > e) f = std::nextafter(f, -INFINITY);
    return f;
}
inline double fma_rd(double w, double d, double acc) { return std::fma(w, d, acc); }

/* imclt_rbg :2588-2712 - one color, one of 4 parity passes */
template <typename T>
void imclt_rbg(const Geo &g, const T *clt, T *rbg, int color, int v_off, int h_off,
               size_t dstride) {
    T *plane = rbg + dstride * (size_t)(g.H + DTT) * color;
    const int pass = (v_off << 1) + h_off;
    const int thalf = (g.TX + (h_off ^ 1)) >> 1;
    const int vhalf = (g.TY + (v_off ^ 1)) >> 1;
#pragma omp parallel for schedule(dynamic, 16)
    for (int n = 0; n < thalf * vhalf; n++) {
        int ty = (n / thalf) * 2 + v_off, tx = (n % thalf) * 2 + h_off;
        if (ty >= g.TY) continue;
        const T *src = clt + ((size_t)(ty * g.TX + tx) * TP_NUM_COLORS + color) * 256;
        T w[16][16], m[16][16];
        imclt_tile_ws<T>(src, w, m, 1);
        /* pass 0 starts from zero, the others read the plane back; the
         * reference accumulates with a round-down FMA */
        for (int r = 0; r < 16; r++) {
            T *row = plane + (size_t)(ty * DTT + r) * dstride + tx * DTT;
            for (int c = 0; c < 16; c++) {
                T acc = pass ? row[c] : (T)0;
                row[c] = fma_rd(w[r][c], m[r][c], acc);
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Textures: textures_accumulate :2207-2585 with debayer_shot           */
/* :4019-4144, debayer :4146-4199 and tile_combine_rgba :4202-4668       */
/* ------------------------------------------------------------------ */
struct TexParams {
    int colors;
    float min_shot, scale_shot, diff_sigma, diff_threshold, min_agree;
    float weights[3];
    int dust_remove;
    int keep_weights;
};

inline int sym16(int i) { return i < 8 ? i : 15 - i; }

/* debayer_shot: un-window (1/w^2), 3x3 debayer (interior 14x14, border
 * copied), shot-noise companding, re-window.  The reference's debayer
 * arithmetic uses double literals, so it is evaluated in double. */
template <typename T>
void debayer_shot(int rb_mode, float min_shot, float scale_shot, const T src[16][16], T dst[16][16]) {
    const Tables<T> &tb = tables<T>();
    T tmp[16][16];
    for (int r = 0; r < 16; r++)
        for (int c = 0; c < 16; c++) tmp[r][c] = src[r][c] * tb.hwsqi[sym16(c)] * tb.hwsqi[sym16(r)];
    for (int r = 0; r < 16; r++)
        for (int c = 0; c < 16; c++) {
            if (r == 0 || r == 15 || c == 0 || c == 15) {
                dst[r][c] = tmp[r][c];
                continue;
            }
            double cross = (double)tmp[r - 1][c] + (double)tmp[r + 1][c] + (double)tmp[r][c - 1] + (double)tmp[r][c + 1];
            if (rb_mode) {
                double diag = (double)tmp[r - 1][c - 1] + (double)tmp[r - 1][c + 1] +
                              (double)tmp[r + 1][c - 1] + (double)tmp[r + 1][c + 1];
                dst[r][c] = (T)(0.0625 * diag + 0.125 * cross + 0.25 * (double)tmp[r][c]);
            } else {
                dst[r][c] = (T)(0.125 * cross + 0.5 * (double)tmp[r][c]);
            }
        }
    if (scale_shot > 0.0f) {
        T k = (T)1 / std::sqrt((T)min_shot);
        for (int r = 0; r < 16; r++)
            for (int c = 0; c < 16; c++) {
                T d = dst[r][c];
                dst[r][c] = (T)scale_shot * ((d > (T)min_shot) ? std::sqrt(d) : k * d);
            }
    }
    for (int r = 0; r < 16; r++)
        for (int c = 0; c < 16; c++) dst[r][c] *= tb.hwsq[sym16(c)] * tb.hwsq[sym16(r)];
}

/* tile_combine_rgba for one pixel.  D[cam][color]: debayered, M: not
 * debayered.  out: colors values, alpha, 4 port weights, colors+1 rms.
 *
 * The consensus makes two discrete choices: the two best ports of the
 * pairwise weights (:4359-4376) and, with dust_remove, the worst port of the
 * re-computed weights (:4463-4472).  For the checker they can be overridden
 * and their robustness reported:
 *   alt bit 0: blend the best port with the THIRD best instead of the second;
 *   alt bit 1: treat the second-lowest port as the worst;
 *   ties (if not NULL): ties[0] = (w_2nd - w_3rd) / (dw_2nd + dw_3rd) of the
 *     pairwise port weights, ties[1] = (w_2nd_lowest - w_lowest) / (dw + dw)
 *     of the re-computed ones (+inf without dust_remove), where dw bounds how
 *     far a weight moves when every D value moves by eta * max|D| of the
 *     pixel (first order in the colour differences plus their square, then
 *     exactly through the exponential) plus eta of the weight itself.  A
 *     ratio <= 1 means an implementation whose D values and arithmetic are
 *     accurate only to eta may make the other choice at this pixel. */
template <typename T>
void combine_pixel(const TexParams &p, const float port_offsets[4][2], const T D[4][3],
                   const T M[4][3], int r, int c, T *out, int alt = 0, double eta = 0.0,
                   double *ties = nullptr) {
    const Tables<T> &tb = tables<T>();
    const int colors = p.colors;
    /* pairs in the reference's (i < j) loop order */
    int pp[6][2];
    T dist2r[6];
    const T ksigma = (T)(1.0 / (2.0 * (double)p.diff_sigma * (double)p.diff_sigma));
    int n = 0;
    for (int i = 0; i < 4; i++)
        for (int j = i + 1; j < 4; j++) {
            T dx = (T)port_offsets[j][0] - (T)port_offsets[i][0];
            T dy = (T)port_offsets[j][1] - (T)port_offsets[i][1];
            pp[n][0] = i;
            pp[n][1] = j;
            dist2r[n++] = ksigma / (dx * dx + dy * dy);
        }
    T *rgba = out;
    T *alpha = out + colors;
    T *pw = out + colors + 1;
    T *crms = out + colors + 5;
    if (p.keep_weights) {
        T sw = 0;
        for (int col = 0; col < colors; col++) {
            T s0 = 0, s1 = 0, s2 = 0;
            for (int cam = 0; cam < 4; cam++) {
                s0 += 1;
                s1 += D[cam][col];
                s2 += D[cam][col] * D[cam][col];
            }
            T mse = (s0 * s2 - s1 * s1) / (s0 * s0);
            mse = std::max(mse, (T)0);  /* rounding residue of equal values */
            crms[col] = std::sqrt(mse);
            sw += (T)p.weights[col] * mse;
        }
        crms[colors] = std::sqrt(sw);
    }
    const T wnd2 = tb.hwsq[sym16(r)] * tb.hwsq[sym16(c)];
    const T wnd2_inv = (T)1 / wnd2;
    /* perturbation bound of every D value (ties only) */
    double dD = 0.0, dpw[4] = {0, 0, 0, 0};
    if (ties) {
        for (int cam = 0; cam < 4; cam++)
            for (int col = 0; col < colors; col++) dD = std::max(dD, std::fabs((double)D[cam][col]));
        dD *= eta;
    }
    const double wi = (double)wnd2_inv;
    for (int cam = 0; cam < 4; cam++) pw[cam] = 0;
    for (int ip = 0; ip < 6; ip++) {
        T d = 0;
        double dd = 0.0;
        for (int col = 0; col < colors; col++) {
            T dc = (D[pp[ip][0]][col] - D[pp[ip][1]][col]) * wnd2_inv;
            d += (T)p.weights[col] * dc * dc;
            const double x = std::fabs((double)dc), dx = 2.0 * dD * wi;
            dd += (double)p.weights[col] * (2.0 * x * dx + dx * dx);
        }
        const T e = std::exp(-dist2r[ip] * d);
        d = e + (T)TP_FAT_ZERO_WEIGHT;
        pw[pp[ip][0]] += d;
        pw[pp[ip][1]] += d;
        const double de = (double)e * std::expm1((double)dist2r[ip] * dd);
        dpw[pp[ip][0]] += de;
        dpw[pp[ip][1]] += de;
    }
    int b1 = 0;
    T best = pw[0];
    for (int cam = 1; cam < 4; cam++)
        if (pw[cam] > best) { b1 = cam; best = pw[cam]; }
    int b2 = (b1 == 0) ? 1 : 0;
    best = pw[b2];
    for (int cam = b2 + 1; cam < 4; cam++)
        if (cam != b1 && pw[cam] > best) { b2 = cam; best = pw[cam]; }
    /* the third best: the better of the two remaining ports */
    int b3 = -1;
    for (int cam = 0; cam < 4; cam++)
        if (cam != b1 && cam != b2 && (b3 < 0 || pw[cam] > pw[b3])) b3 = cam;
    if (ties) {
        for (int cam = 0; cam < 4; cam++) dpw[cam] += eta * (double)pw[cam];
        const double m = (double)pw[b2] - (double)pw[b3], s = dpw[b2] + dpw[b3];
        ties[0] = s > 0 ? m / s : (m > 0 ? INFINITY : 0.0);
        ties[1] = INFINITY;
    }
    if (alt & 1) b2 = b3;
    const T w1 = pw[b1] / (pw[b1] + pw[b2]);
    const T w2 = (T)1 - w1;
    for (int col = 0; col < colors; col++) rgba[col] = w1 * D[b1][col] + w2 * D[b2][col];
    /* perturbation of the blend: its inputs and its weight w1 */
    const double sp = (double)pw[b1] + (double)pw[b2];
    const double dw1 = ((double)pw[b2] * dpw[b1] + (double)pw[b1] * dpw[b2]) / (sp * sp);
    double dq[4] = {0, 0, 0, 0};
    for (int cam = 0; cam < 4; cam++) {
        T d2 = 0;
        double dd = 0.0;
        for (int col = 0; col < colors; col++) {
            T dc = (D[cam][col] - rgba[col]) * wnd2_inv;
            d2 += (T)p.weights[col] * dc * dc;
            const double db = dD + std::fabs((double)D[b1][col] - (double)D[b2][col]) * dw1;
            const double x = std::fabs((double)dc), dx = (dD + db) * wi;
            dd += (double)p.weights[col] * (2.0 * x * dx + dx * dx);
        }
        const T e = std::exp(-ksigma * d2);
        pw[cam] = e + (T)TP_FAT_ZERO_WEIGHT;
        dq[cam] = (double)e * std::expm1((double)ksigma * dd) + eta * (double)pw[cam];
    }
    if (p.dust_remove) {
        int worst = 0;
        T wv = pw[0];
        for (int cam = 1; cam < 4; cam++)
            if (pw[cam] < wv) { worst = cam; wv = pw[cam]; }
        int next = -1;  /* the second lowest */
        for (int cam = 0; cam < 4; cam++)
            if (cam != worst && (next < 0 || pw[cam] < pw[next])) next = cam;
        if (ties) {
            const double m = (double)pw[next] - (double)pw[worst], s = dq[next] + dq[worst];
            ties[1] = s > 0 ? m / s : (m > 0 ? INFINITY : 0.0);
        }
        if (alt & 2) {
            worst = next;
            wv = pw[worst];
        }
        T avg = -wv;
        for (int cam = 0; cam < 4; cam++) avg += pw[cam];
        avg /= (T)3;
        T scale = (T)1 + wv * (avg - wv) / (avg * avg * (T)3);
        for (int cam = 0; cam < 4; cam++)
            if (cam != worst) pw[cam] *= scale;
        pw[worst] *= wv / avg;
    }
    /* output color: port-weighted average of the non-debayered tiles */
    T k = 0;
    for (int cam = 0; cam < 4; cam++) k += pw[cam];
    k = (T)1 / k;
    for (int col = 0; col < colors; col++) {
        T acc = 0;
        for (int cam = 0; cam < 4; cam++) acc += k * pw[cam] * M[cam][col];
        rgba[col] = acc;
    }
    T a = 0;
    for (int cam = 0; cam < 4; cam++) a += pw[cam];
    *alpha = wnd2 * a / (T)4;
}

/* one texture tile: slices [colors + 1 (+ 4 + colors + 1)][16][16]; alt, eta
 * and ties[2][256] as for combine_pixel */
template <typename T>
void texture_tile(const Geo &g, const T *const *clt, int tile, const TexParams &p,
                  const float port_offsets[4][2], T *slices, int alt = 0, double eta = 0.0,
                  double *ties = nullptr) {
    const int colors = p.colors;
    T M[4][3][16][16], D[4][3][16][16];
    T lb[64];
    for (int cam = 0; cam < 4; cam++)
        for (int col = 0; col < colors; col++) {
            const T *src = clt[cam] + ((size_t)tile * TP_NUM_COLORS + col) * 256;
            const T *lpf = lpf_table<T>(colors > 1 ? col : 3, lb);
            T q[256];
            for (int i = 0; i < 256; i++) q[i] = src[i] * lpf[i & 63];
            imclt_tile<T>(q, M[cam][col], 0);
            if (colors > 1)
                debayer_shot<T>(col < 2, p.min_shot, p.scale_shot, M[cam][col], D[cam][col]);
            else
                std::memcpy(D[cam][col], M[cam][col], sizeof(D[cam][col]));
        }
    const int nsl = colors + 1 + 4 + colors + 1;
    for (int r = 0; r < 16; r++)
        for (int c = 0; c < 16; c++) {
            T dp[4][3], mp[4][3], out[12];
            for (int cam = 0; cam < 4; cam++)
                for (int col = 0; col < colors; col++) {
                    dp[cam][col] = D[cam][col][r][c];
                    mp[cam][col] = M[cam][col][r][c];
                }
            for (int i = 0; i < 12; i++) out[i] = 0;
            double tt[2];
            combine_pixel<T>(p, port_offsets, dp, mp, r, c, out, alt, eta, ties ? tt : nullptr);
            for (int s = 0; s < nsl; s++) slices[s * 256 + r * 16 + c] = out[s];
            if (ties) {
                ties[r * 16 + c] = tt[0];
                ties[256 + r * 16 + c] = tt[1];
            }
        }
}

/* textures_accumulate over a list: non-overlapping tiles and/or the
 * alpha-faded overlap-add into the woi-relative RGBA image. */
template <typename T>
void textures_accumulate(const Geo &g, const int *woi, const T *const *clt, size_t n,
                         const int *idx, const float port_offsets[4][2], const TexParams &p,
                         size_t rbga_stride, T *rbga, size_t tex_stride, T *tiles) {
    const int colors = p.colors;
    const int nsl_all = colors + 1 + 4 + colors + 1;
    const int nsl_tile = p.keep_weights ? nsl_all : colors + 1;
    /* tiles of one list never overlap in the image only within a parity
     * pass; accumulate serially to keep the order-independent sum exact */
    std::vector<T> sl(nsl_all * 256);
    for (size_t e = 0; e < n; e++) {
        const int code = idx[e];
        if (!(code & (1 << TP_LIST_TEXTURE_BIT))) continue;
        const int tile = code >> TP_CORR_NTILE_SHIFT;
        if (tile < 0 || tile >= g.TX * g.TY) continue;
        texture_tile<T>(g, clt, tile, p, port_offsets, sl.data());
        if (tiles && tex_stride) {
            for (int s = 0; s < nsl_tile; s++)
                for (int i = 0; i < 256; i++) tiles[e * tex_stride + s * 256 + i] = sl[s * 256 + i];
        }
        const int tcode = code & TP_TASK_TEXTURE_BITS;
        if (!tcode || !rbga || !rbga_stride) continue;
        if (tcode != TP_TASK_TEXTURE_BITS)
            for (int s = 0; s <= colors; s++)
                for (int i = 0; i < 256; i++) sl[s * 256 + i] *= tables<T>().afade[tcode][i];
        const size_t slice_stride = rbga_stride * (size_t)(woi[3] + 1) * 8;
        const int ty = tile / g.TX, tx = tile % g.TX;
        const int x0 = (tx - woi[0]) * 8, y0 = (ty - woi[1]) * 8;
        for (int s = 0; s <= colors; s++)
            for (int r = 0; r < 16; r++)
                for (int c = 0; c < 16; c++)
                    rbga[s * slice_stride + (size_t)(y0 + r) * rbga_stride + x0 + c] += sl[s * 256 + r * 16 + c];
    }
}

/* prepare_texture_list :1643-1713 (with the child kernels :1716-1854),
 * sequentially.  Mutates the texture bits of tasks (N, E, S, W neighbours
 * present), fills the 4 parity sub-lists (inner tiles from the start,
 * border tiles from the end), num_texture_tiles[8] and woi (x, y, width,
 * height). */
void prepare_texture_list(const Geo &g, tp_task *tasks, int num_tiles, int *indices, int *num_tex,
                          int *woi, int width, int height) {
    for (int row = 0; row < height; row++)
        for (int col = 0; col < width && col < g.TX; col++) indices[col + row * g.TX] = 0;
    std::vector<int> occ(g.TX * g.TY, 0);
    for (int t = 0; t < num_tiles; t++)
        if (tasks[t].task & TP_TASK_TEXTURE_BITS) {
            int x = tasks[t].txy & 0xffff, y = tasks[t].txy >> 16;
            occ[x + y * g.TX] = 1;
        }
    woi[0] = g.TX; woi[1] = g.TY; woi[2] = 0; woi[3] = 0;
    for (int t = 0; t < num_tiles; t++) {
        int task = tasks[t].task;
        if (!(task & TP_TASK_TEXTURE_BITS)) continue;
        int x = tasks[t].txy & 0xffff, y = tasks[t].txy >> 16;
        woi[0] = std::min(woi[0], x); woi[1] = std::min(woi[1], y);
        woi[2] = std::max(woi[2], x); woi[3] = std::max(woi[3], y);
        int d = 0;
        if (y > 0 && occ[x + (y - 1) * g.TX]) d |= 1;
        if (x < g.TX - 1 && occ[x + 1 + y * g.TX]) d |= 2;
        if (y < g.TY - 1 && occ[x + (y + 1) * g.TX]) d |= 4;
        if (x > 0 && occ[x - 1 + y * g.TX]) d |= 8;
        tasks[t].task = ((task ^ d) & TP_TASK_TEXTURE_BITS) ^ task;
    }
    for (int i = 0; i < 8; i++) num_tex[i] = 0;
    const int quarter = g.TX * (g.TYA >> 2);
    for (int t = 0; t < num_tiles; t++) {
        int task = tasks[t].task & TP_TASK_TEXTURE_BITS;
        if (!task) continue;
        int x = tasks[t].txy & 0xffff, y = tasks[t].txy >> 16;
        int border = (x == woi[0]) || (y == woi[1]) || (x == g.TX - 1) || (y == woi[3]);
        int head = 0, slot = 0;
        if (x & 1) { head += quarter; slot += 2; }
        if (y & 1) { head += 2 * quarter; slot += 4; }
        if (border) { head += quarter - 1; slot += 1; }
        int pos = num_tex[slot]++;
        indices[head + (border ? -pos : pos)] = task | ((x + y * g.TX) << TP_CORR_NTILE_SHIFT) | (1 << TP_LIST_TEXTURE_BIT);
    }
    woi[2] += 1 - woi[0];
    woi[3] += 1 - woi[1];
}

/* ------------------------------------------------------------------ */
/* Pair combination and sub-pixel disparity (not in the reference: its  */
/* Java host does this).  Each pair's tile is mapped onto the horizontal */
/* disparity axis (pairs 0,1 horizontal: as is; 2,3 vertical: transposed;*/
/* 4 (0,3) diagonal: C[x + y][x - y]; 5 (2,1) anti-diagonal:             */
/* C[y - x][x + y], offsets from the centre) and the mapped tiles are    */
/* summed.  A residual disparity dd then peaks at (0, +dd).              */
/* ------------------------------------------------------------------ */
inline int pair_src(int pair, int r, int cy, int cx, int &sy, int &sx) {
    switch (pair) {
    case 0: case 1: sy = cy; sx = cx; break;
    case 2: case 3: sy = cx; sx = cy; break;
    case 4: sy = cx + cy; sx = cx - cy; break;
    default: sy = cy - cx; sx = cx + cy; break;
    }
    return (sy >= -r && sy <= r && sx >= -r && sx <= r);
}

template <typename T>
void corr_combine(const T *const *rows, const int *pairs, int npairs, int r, T *out) {
    const int n = 2 * r + 1;
    for (int y = -r; y <= r; y++)
        for (int x = -r; x <= r; x++) {
            T s = 0;
            for (int k = 0; k < npairs; k++) {
                int sy, sx;
                if (pair_src(pairs[k], r, y, x, sy, sx)) s += rows[k][(sy + r) * n + sx + r];
            }
            out[(y + r) * n + x + r] = s;
        }
}

/* Disparity from the centre row of a combined tile: integer argmax, then
 * parabola through the 3 samples (mode 0) or centroid of the positive part
 * of the samples within +-2 of the argmax (mode 1).  out: disparity,
 * strength (peak value). */
template <typename T>
void corr_disparity(const T *tile, int r, int mode, T *out) {
    const int n = 2 * r + 1;
    const T *row = tile + r * n;
    int best = 0;
    for (int x = 1; x < n; x++)
        if (row[x] > row[best]) best = x;
    T d = (T)(best - r);
    if (mode == 0) {
        if (best > 0 && best < n - 1) {
            T a = row[best - 1], b = row[best], c = row[best + 1];
            T den = a - 2 * b + c;
            if (den < 0) d += (T)0.5 * (a - c) / den;
        }
    } else {
        T sw = 0, sx = 0;
        for (int x = std::max(0, best - 2); x <= std::min(n - 1, best + 2); x++) {
            T w = std::max(row[x], (T)0);
            sw += w;
            sx += w * (T)(x - r);
        }
        if (sw > 0) d = sx / sw;
    }
    out[0] = d;
    out[1] = row[best];
}

}  // namespace

/* ====================================================================== */
/* C ABI (host pointers) for the Python test harness                       */
/* ====================================================================== */
extern "C" {

int oracle_num_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

/* tables for tests: kind 0 dtt matrices [4][64] (double), 1 fold
 * rind/rcs/rss [48] (as double), 2 windows hw,hw2,hwsq,hwsqi [32],
 * 3 lpf [6][64], 4 alpha fade [16][256] */
int oracle_table(int kind, double *out) {
    const Tables<double> &tb = tables<double>();
    switch (kind) {
    case 0:
        for (int t = 0; t < 4; t++)
            for (int k = 0; k < 8; k++)
                for (int n = 0; n < 8; n++) out[t * 64 + k * 8 + n] = tb.dtt[t][k][n];
        return 256;
    case 1:
        for (int i = 0; i < 16; i++) {
            out[i] = FOLD.rind[i];
            out[16 + i] = FOLD.rcs[i];
            out[32 + i] = FOLD.rss[i];
        }
        return 48;
    case 2:
        for (int i = 0; i < 8; i++) {
            out[i] = tb.hw[i];
            out[8 + i] = tb.hw2[i];
            out[16 + i] = tb.hwsq[i];
            out[24 + i] = tb.hwsqi[i];
        }
        return 32;
    case 3:
        for (int t = 0; t < 6; t++)
            for (int i = 0; i < 64; i++) out[t * 64 + i] = tb.lpf[t][i];
        return 384;
    case 4:
        for (int m = 0; m < 16; m++)
            for (int i = 0; i < 256; i++) out[m * 256 + i] = tb.afade[m][i];
        return 4096;
    }
    return -1;
}

int oracle_set_lpf(int which, const float *lpf) {
    if (which < 0 || which > 5) return -1;
    if (!lpf) {
        g_lpf_set[which] = false;
        return 0;
    }
    std::memcpy(g_lpf_user[which], lpf, sizeof(float) * 64);
    g_lpf_set[which] = true;
    return 0;
}

/* single-tile forward CLT (for unit tests): q[4][8][8] */
void oracle_convert_tile_f64(const tp_geometry *geom, const tp_clt_extra *koffs,
                             const float *kernels, const float *image, size_t dstride,
                             int color, int lpf_mask, float cx, float cy, double *q) {
    Geo g = geo_of(geom);
    convert_correct_tile<double>(g, koffs, kernels, image, dstride, color, lpf_mask, cx, cy,
                                 (double(*)[8][8])q);
}
void oracle_convert_tile_f32(const tp_geometry *geom, const tp_clt_extra *koffs,
                             const float *kernels, const float *image, size_t dstride,
                             int color, int lpf_mask, float cx, float cy, float *q) {
    Geo g = geo_of(geom);
    convert_correct_tile<float>(g, koffs, kernels, image, dstride, color, lpf_mask, cx, cy,
                                (float(*)[8][8])q);
}

void oracle_convert_correct_tiles_f32(const tp_geometry *geom, const tp_clt_extra *const *koffs,
                                      const float *const *kernels, const float *const *images,
                                      const tp_task *tasks, float *const *clt, size_t dstride,
                                      int num_tiles, int lpf_mask) {
    convert_correct_tiles<float>(geo_of(geom), koffs, kernels, images, tasks, clt, dstride,
                                 num_tiles, lpf_mask);
}
void oracle_convert_correct_tiles_f64(const tp_geometry *geom, const tp_clt_extra *const *koffs,
                                      const float *const *kernels, const float *const *images,
                                      const tp_task *tasks, double *const *clt, size_t dstride,
                                      int num_tiles, int lpf_mask) {
    convert_correct_tiles<double>(geo_of(geom), koffs, kernels, images, tasks, clt, dstride,
                                  num_tiles, lpf_mask);
}

void oracle_correlate2D_f32(const tp_geometry *geom, const float *const *clt, int colors,
                            float s0, float s1, float s2, float fat_zero, size_t n,
                            const int *idx, size_t stride, int radius, float *corrs) {
    const float sc[3] = {s0, s1, s2};
    correlate2D<float>(geo_of(geom), clt, colors, sc, fat_zero, n, idx, stride, radius, corrs);
}
void oracle_correlate2D_f64(const tp_geometry *geom, const double *const *clt, int colors,
                            float s0, float s1, float s2, float fat_zero, size_t n,
                            const int *idx, size_t stride, int radius, double *corrs) {
    const float sc[3] = {s0, s1, s2};
    correlate2D<double>(geo_of(geom), clt, colors, sc, fat_zero, n, idx, stride, radius, corrs);
}

/* one-tile inverse CLT: out[16][16] */
void oracle_imclt_tile_f64(const double *clt, double *out, int horizontal_first) {
    imclt_tile<double>(clt, (double(*)[16])out, horizontal_first);
}

void oracle_imclt_rbg_f32(const tp_geometry *geom, const float *clt, float *rbg, int color,
                          int v_off, int h_off, size_t dstride) {
    imclt_rbg<float>(geo_of(geom), clt, rbg, color, v_off, h_off, dstride);
}
void oracle_imclt_rbg_f64(const tp_geometry *geom, const double *clt, double *rbg, int color,
                          int v_off, int h_off, size_t dstride) {
    imclt_rbg<double>(geo_of(geom), clt, rbg, color, v_off, h_off, dstride);
}


static TexParams tex_params(int colors, float min_shot, float scale_shot, float diff_sigma,
                            float diff_threshold, float min_agree, float w0, float w1, float w2,
                            int dust_remove, int keep_weights) {
    TexParams p;
    p.colors = colors; p.min_shot = min_shot; p.scale_shot = scale_shot;
    p.diff_sigma = diff_sigma; p.diff_threshold = diff_threshold; p.min_agree = min_agree;
    p.weights[0] = w0; p.weights[1] = w1; p.weights[2] = w2;
    p.dust_remove = dust_remove; p.keep_weights = keep_weights;
    return p;
}

/* textures_accumulate (host arrays; clt precision selects the variant) */
void oracle_textures_accumulate_f64(const tp_geometry *geom, const int *woi, const double *const *clt,
                                    size_t n, const int *idx, const float *port_offsets, int colors,
                                    float min_shot, float scale_shot, float diff_sigma,
                                    float diff_threshold, float min_agree, float w0, float w1,
                                    float w2, int dust_remove, int keep_weights, size_t rbga_stride,
                                    double *rbga, size_t tex_stride, double *tiles) {
    TexParams p = tex_params(colors, min_shot, scale_shot, diff_sigma, diff_threshold, min_agree,
                             w0, w1, w2, dust_remove, keep_weights);
    textures_accumulate<double>(geo_of(geom), woi, clt, n, idx, (const float(*)[2])port_offsets, p,
                                rbga_stride, rbga, tex_stride, tiles);
}
void oracle_textures_accumulate_f32(const tp_geometry *geom, const int *woi, const float *const *clt,
                                    size_t n, const int *idx, const float *port_offsets, int colors,
                                    float min_shot, float scale_shot, float diff_sigma,
                                    float diff_threshold, float min_agree, float w0, float w1,
                                    float w2, int dust_remove, int keep_weights, size_t rbga_stride,
                                    float *rbga, size_t tex_stride, float *tiles) {
    TexParams p = tex_params(colors, min_shot, scale_shot, diff_sigma, diff_threshold, min_agree,
                             w0, w1, w2, dust_remove, keep_weights);
    textures_accumulate<float>(geo_of(geom), woi, clt, n, idx, (const float(*)[2])port_offsets, p,
                               rbga_stride, rbga, tex_stride, tiles);
}

/* For the checker: every valid entry's slices [colors + 1 + 4 + colors + 1]
 * [256] (tile form, nothing overlap-added) with the consensus choices
 * overridden by alt, and (ties != NULL) the tie ratios [2][256] of each
 * pixel (combine_pixel).  Rows of invalid entries are not written. */
void oracle_texture_ties_f64(const tp_geometry *geom, const double *const *clt, size_t n,
                             const int *idx, const float *port_offsets, int colors, float min_shot,
                             float scale_shot, float diff_sigma, float diff_threshold,
                             float min_agree, float w0, float w1, float w2, int dust_remove,
                             int alt, double eta, double *slices, double *ties) {
    TexParams p = tex_params(colors, min_shot, scale_shot, diff_sigma, diff_threshold, min_agree,
                             w0, w1, w2, dust_remove, 1);
    const Geo g = geo_of(geom);
    const int nsl = 2 * colors + 6;
#pragma omp parallel for schedule(dynamic, 4)
    for (long e = 0; e < (long)n; e++) {
        const int code = idx[e];
        if (!(code & (1 << TP_LIST_TEXTURE_BIT))) continue;
        const int tile = code >> TP_CORR_NTILE_SHIFT;
        if (tile < 0 || tile >= g.TX * g.TY) continue;
        texture_tile<double>(g, clt, tile, p, (const float(*)[2])port_offsets,
                             slices + (size_t)e * nsl * 256, alt, eta,
                             ties ? ties + (size_t)e * 512 : nullptr);
    }
}

void oracle_prepare_texture_list(const tp_geometry *geom, tp_task *tasks, int num_tiles,
                                 int *indices, int *num_tex, int *woi, int width, int height) {
    prepare_texture_list(geo_of(geom), tasks, num_tiles, indices, num_tex, woi, width, height);
}

/* single debayer_shot / combine for unit tests */
void oracle_debayer_shot_f64(int rb_mode, float min_shot, float scale_shot, const double *src,
                             double *dst) {
    debayer_shot<double>(rb_mode, min_shot, scale_shot, (const double(*)[16])src, (double(*)[16])dst);
}
void oracle_combine_pixel_f64(int colors, float diff_sigma, float w0, float w1, float w2,
                              int dust_remove, int keep_weights, const float *port_offsets,
                              const double *D, const double *M, int r, int c, double *out,
                              int alt, double eta, double *ties) {
    TexParams p = tex_params(colors, 10.f, 3.f, diff_sigma, 10.f, 3.f, w0, w1, w2, dust_remove,
                             keep_weights);
    combine_pixel<double>(p, (const float(*)[2])port_offsets, (const double(*)[3])D,
                          (const double(*)[3])M, r, c, out, alt, eta, ties);
}

/* combine: corrs rows of the 6 pairs of each tile (row index = tile * 6 +
 * pair in a tile-major list) -> out[tile][(2r+1)^2] */
void oracle_corr_combine_f64(const double *corrs, size_t stride, int ntiles, int pair_mask,
                             int radius, double *out) {
    const int n = 2 * radius + 1;
    for (int t = 0; t < ntiles; t++) {
        const double *rows[6];
        int pairs[6], np = 0;
        for (int p = 0; p < 6; p++)
            if ((pair_mask >> p) & 1) {
                rows[np] = corrs + ((size_t)t * 6 + p) * stride;
                pairs[np++] = p;
            }
        corr_combine<double>(rows, pairs, np, radius, out + (size_t)t * n * n);
    }
}

void oracle_corr_disparity_f64(const double *tiles, size_t stride, int ntiles, int radius, int mode,
                               double *out) {
    for (int t = 0; t < ntiles; t++) corr_disparity<double>(tiles + (size_t)t * stride, radius, mode, out + 2 * t);
}

/* ------------------------------------------------------------------------
 * Disparity-refinement CNN (not in the reference).  Plain loops in double;
 * the parameters are doubles (the tests pass the BF16-rounded values the GPU
 * kernels read).  Layout: W1[32][7][9], b1[32], W2[32][32][9], b2[32],
 * W3[2][32][9], b3[2]; tiles 15 x 15 with zero padding.
 * ---------------------------------------------------------------------- */
static const int NN_P = 225;

static void nn_conv(const double *in, int cin, const double *w, const double *b, int cout, bool relu,
                    double *out) {
    for (int co = 0; co < cout; co++)
        for (int y = 0; y < 15; y++)
            for (int x = 0; x < 15; x++) {
                double acc = b[co];
                for (int ci = 0; ci < cin; ci++)
                    for (int r = 0; r < 9; r++) {
                        const int yy = y + r / 3 - 1, xx = x + r % 3 - 1;
                        if (yy < 0 || yy >= 15 || xx < 0 || xx >= 15) continue;
                        acc += (double)w[(co * cin + ci) * 9 + r] * in[ci * NN_P + yy * 15 + xx];
                    }
                out[co * NN_P + y * 15 + x] = relu ? std::max(acc, 0.0) : acc;
            }
}

/* softmax over z0, disparity = sum p (x - 7), confidence logit = sum p z1 */
static void nn_head(const double *z, double *prob, double &disp, double &cl) {
    double mx = -1e300;
    for (int i = 0; i < NN_P; i++) mx = std::max(mx, z[i]);
    double s = 0, sx = 0, sc = 0;
    for (int i = 0; i < NN_P; i++) {
        prob[i] = std::exp(z[i] - mx);
        s += prob[i];
    }
    for (int i = 0; i < NN_P; i++) {
        prob[i] /= s;
        sx += prob[i] * (i % 15 - 7);
        sc += prob[i] * z[NN_P + i];
    }
    disp = sx;
    cl = sc;
}

/* BF16 (round to nearest even) of a value first rounded to float */
static double bf16_of(double v) {
    float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return f;
}

/* out[n][2] = (disparity, confidence); acts[n][14850] = h1, h2, z (or NULL);
 * bf16_acts: the input and both hidden layers rounded to BF16 before they
 * enter the next convolution (the matrix-core inference kernel's precision) */
void oracle_nn_forward_ex(const double *prm, const float *x, int n, double *acts, double *out,
                          int bf16_acts) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int t = 0; t < n; t++) {
        std::vector<double> in(7 * NN_P), h1(32 * NN_P), h2(32 * NN_P), z(2 * NN_P), prob(NN_P);
        for (int i = 0; i < 7 * NN_P; i++) in[i] = x[(size_t)t * 7 * NN_P + i];
        if (bf16_acts)
            for (auto &v : in) v = bf16_of(v);
        nn_conv(in.data(), 7, prm + 0, prm + 2016, 32, true, h1.data());
        if (bf16_acts)
            for (auto &v : h1) v = bf16_of(v);
        nn_conv(h1.data(), 32, prm + 2048, prm + 11264, 32, true, h2.data());
        if (bf16_acts)
            for (auto &v : h2) v = bf16_of(v);
        nn_conv(h2.data(), 32, prm + 11296, prm + 11872, 2, false, z.data());
        double disp, cl;
        nn_head(z.data(), prob.data(), disp, cl);
        out[2 * t] = disp;
        out[2 * t + 1] = 1.0 / (1.0 + std::exp(-cl));
        if (acts) {
            double *a = acts + (size_t)t * 14850;
            std::copy(h1.begin(), h1.end(), a);
            std::copy(h2.begin(), h2.end(), a + 32 * NN_P);
            std::copy(z.begin(), z.end(), a + 64 * NN_P);
        }
    }
}

void oracle_nn_forward(const double *prm, const float *x, int n, double *acts, double *out) {
    oracle_nn_forward_ex(prm, x, n, acts, out, 0);
}

/* Gradient of sum over tiles with a finite target of scale * (huber_delta(d
 * - target) + conf_weight * BCE(sigmoid(cl), |d - target| < 0.5)), the BCE
 * label held constant; added to grads[11874], loss added to loss[0]. */
void oracle_nn_backward(const double *prm, const float *x, int n, const float *target,
                        double conf_weight, double delta, double scale, double *grads, double *loss) {
    const double *W2 = prm + 2048, *W3 = prm + 11296;
    for (int t = 0; t < n; t++) {
        if (!std::isfinite(target[t])) continue;
        std::vector<double> in(7 * NN_P), h1(32 * NN_P), h2(32 * NN_P), z(2 * NN_P), prob(NN_P);
        for (int i = 0; i < 7 * NN_P; i++) in[i] = x[(size_t)t * 7 * NN_P + i];
        nn_conv(in.data(), 7, prm + 0, prm + 2016, 32, true, h1.data());
        nn_conv(h1.data(), 32, W2, prm + 11264, 32, true, h2.data());
        nn_conv(h2.data(), 32, W3, prm + 11872, 2, false, z.data());
        double disp, cl;
        nn_head(z.data(), prob.data(), disp, cl);
        const double e = disp - target[t];
        const double hd = std::min(std::max(e, -delta), delta);
        const double sg = 1.0 / (1.0 + std::exp(-cl));
        const double lab = std::fabs(e) < 0.5 ? 1.0 : 0.0;
        const double gc = conf_weight * (sg - lab);
        const double ae = std::fabs(e);
        const double hub = ae <= delta ? 0.5 * e * e : delta * (ae - 0.5 * delta);
        const double bce = -(lab * std::log(std::max(sg, 1e-12)) + (1 - lab) * std::log(std::max(1 - sg, 1e-12)));
        loss[0] += scale * (hub + conf_weight * bce);
        // dz
        std::vector<double> dz(2 * NN_P), dh2(32 * NN_P, 0.0), dh1(32 * NN_P, 0.0);
        for (int i = 0; i < NN_P; i++) {
            dz[i] = scale * prob[i] * (hd * ((i % 15 - 7) - disp) + gc * (z[NN_P + i] - cl));
            dz[NN_P + i] = scale * gc * prob[i];
        }
        // layer 3
        for (int co = 0; co < 2; co++)
            for (int p = 0; p < NN_P; p++) {
                const double d = dz[co * NN_P + p];
                grads[11872 + co] += d;
                const int y = p / 15, xx0 = p % 15;
                for (int ci = 0; ci < 32; ci++)
                    for (int r = 0; r < 9; r++) {
                        const int yy = y + r / 3 - 1, xx = xx0 + r % 3 - 1;
                        if (yy < 0 || yy >= 15 || xx < 0 || xx >= 15) continue;
                        const int q = ci * NN_P + yy * 15 + xx;
                        grads[11296 + (co * 32 + ci) * 9 + r] += d * h2[q];
                        dh2[q] += d * W3[(co * 32 + ci) * 9 + r];
                    }
            }
        for (int i = 0; i < 32 * NN_P; i++)
            if (!(h2[i] > 0)) dh2[i] = 0;
        // layer 2
        for (int co = 0; co < 32; co++)
            for (int p = 0; p < NN_P; p++) {
                const double d = dh2[co * NN_P + p];
                if (d == 0) continue;
                grads[11264 + co] += d;
                const int y = p / 15, xx0 = p % 15;
                for (int ci = 0; ci < 32; ci++)
                    for (int r = 0; r < 9; r++) {
                        const int yy = y + r / 3 - 1, xx = xx0 + r % 3 - 1;
                        if (yy < 0 || yy >= 15 || xx < 0 || xx >= 15) continue;
                        const int q = ci * NN_P + yy * 15 + xx;
                        grads[2048 + (co * 32 + ci) * 9 + r] += d * h1[q];
                        dh1[q] += d * W2[(co * 32 + ci) * 9 + r];
                    }
            }
        for (int i = 0; i < 32 * NN_P; i++)
            if (!(h1[i] > 0)) dh1[i] = 0;
        // layer 1
        for (int co = 0; co < 32; co++)
            for (int p = 0; p < NN_P; p++) {
                const double d = dh1[co * NN_P + p];
                if (d == 0) continue;
                grads[2016 + co] += d;
                const int y = p / 15, xx0 = p % 15;
                for (int ci = 0; ci < 7; ci++)
                    for (int r = 0; r < 9; r++) {
                        const int yy = y + r / 3 - 1, xx = xx0 + r % 3 - 1;
                        if (yy < 0 || yy >= 15 || xx < 0 || xx >= 15) continue;
                        grads[(co * 7 + ci) * 9 + r] += d * in[ci * NN_P + yy * 15 + xx];
                    }
            }
    }
}

}  // extern "C"
