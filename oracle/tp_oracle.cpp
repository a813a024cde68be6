/*
 * tp_oracle.cpp - CPU oracle of the tile processor hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness judge for the
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
 *     not reproduced.
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

}  // extern "C"
