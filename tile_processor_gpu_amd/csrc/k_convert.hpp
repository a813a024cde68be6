// k_convert.hpp - forward CLT with aberration correction (plain FP32).
//
// Replaces convert_correct_tiles / convertCorrectTile
// (src/TileProcessor.cuh:1859-1937, :3106-3634).
//
// Work decomposition (gfx950, wave64): a "unit" is one (tile, camera, color)
// 16x16 -> 4 x 8x8 transform.  8 lanes own a unit (lane = one row or column
// of the 8x8 folded block), so a wavefront runs 8 units side by side.  The
// staged kernel k_convert_cam gives a workgroup 32 consecutive tiles of one
// camera, 8 per wavefront; a wavefront converts R, B, then G of its 8 tiles.
//
// Per unit:
//   1. lane = folded column k gathers its 8 (R, B) or 16 (G) Bayer samples,
//      applies the shifted sine windows with the fold signs;
//   2. vertical DCT-IV in registers;
//   3. 8x8 transpose through LDS (lane = folded row l);
//   4. horizontal DCT-IV (and DST-IV for G) in registers;
//   5. the missing quadrants are replicated from the Bayer symmetry
//      (register reversal + one DPP mirror across the unit's lanes);
//   6. bicomplex multiply by the aberration kernel (2 x float4 loads per
//      quadrant), horizontal then vertical phase rotation, optional LPF;
//   7. the 1 KB block goes out through LDS as whole 128-byte lines.
#pragma once

#include "tp_common.hpp"
#include "mfma_dct.hpp"

namespace tp {

struct ConvertArgs {
    Geo g;
    const float *const *koffs;
    const float *const *kernels;
    const float *const *images;
    const tp_task *tasks;
    float *const *clt;
    size_t dstride;
    int num_tiles;
    int lpf_mask;
    // this lane's camera: entries of the per-camera pointer arrays, read
    // once per kernel (or work item) by resolve_cam rather than per unit
    const float *l_koffs, *l_kern, *l_img;
    float *l_clt;
};

__device__ __forceinline__ void resolve_cam(ConvertArgs &a, int cam) {
    a.l_koffs = gelem(a.koffs, cam);
    a.l_kern = gelem(a.kernels, cam);
    a.l_img = gelem(a.images, cam);
    a.l_clt = a.clt ? gelem(a.clt, cam) : nullptr;
}

struct LpfTab {
    float v[6][64];
};
constexpr LpfTab make_lpf_tab() {
    LpfTab t{};
    for (int w = 0; w < 6; w++)
        for (int i = 0; i < 64; i++) t.v[w][i] = kTab.lpf[w][i];
    return t;
}
// low-pass tables: lpf_data[R,B,G,mono], lpf_rb_corr, lpf_corr
__constant__ LpfTab c_lpf = make_lpf_tab();

// Processing order -> task index (fused kernel).  The grid is cut into
// strips of TP_STRIP_ROWS tile rows; within a strip, tiles are visited by column pairs
// (rows of a pair top to bottom), so that tiles sharing a 16x16-px
// aberration kernel cell meet in one workgroup or in workgroups that run at
// the same time.  Identity past the last complete strip; a last odd column
// is visited top to bottom.  A bijection on [0, num_tiles) for any TX.
__device__ __forceinline__ int swizzle_strip(int s, int TX, int num_tiles) {
    constexpr int SH = TP_STRIP_ROWS;
    const int strip = SH * TX;
    if (s >= (num_tiles / strip) * strip) return s;
    const int st = s / strip, q = s - st * strip;
    const int full = (TX >> 1) * (2 * SH);
    int r, c;
    if (q < full) {
        const int cp = q / (2 * SH), sub = q - cp * (2 * SH);
        r = sub >> 1;
        c = 2 * cp + (sub & 1);
    } else {
        r = q - full;
        c = TX - 1;
    }
    return st * strip + r * TX + c;
}

// Per-lane constants of the horizontal fold: folded column li is fed by
// source columns ca (the smaller) and cb; for each local column parity p:
// the column and its window coefficients (w_s(c) = P cos(pi s/16) +
// Q sin(pi s/16) = sin(pi (c + 1/2 + s)/16)) with the cosine- and
// sine-branch fold signs applied.
struct LaneFold {
    int col0, col1;                // source column of local parity 0 / 1
    float Pc0, Qc0, Pc1, Qc1;      // cosine branch (sign rcs)
    float Ps0, Qs0, Ps1, Qs1;      // sine branch (sign rss)
};

__device__ __forceinline__ void win_coef(int c, float &P, float &Q) {
    const int i = c < 8 ? c : 15 - c;
    const int j = c < 8 ? 7 - c : c - 8;
    float hi = 0.f, hj = 0.f;
#pragma unroll
    for (int t = 0; t < 8; t++) {  // runtime index into the 8-entry table (once per kernel)
        hi = (i == t) ? kTab.hw[t] : hi;
        hj = (j == t) ? kTab.hw[t] : hj;
    }
    P = hi;
    Q = c < 8 ? hj : -hj;
}

__device__ __forceinline__ LaneFold lane_fold(int li) {
    const int ca = li >= 4 ? li - 4 : 11 - li;
    const int cb = li >= 4 ? 11 - li : 12 + li;  // opposite parity
    float Pa, Qa, Pb, Qb;
    win_coef(ca, Pa, Qa);
    win_coef(cb, Pb, Qb);
    const float sca = ca < 4 ? 1.f : -1.f, ssa = ca < 12 ? 1.f : -1.f;
    const float scb = cb < 4 ? 1.f : -1.f, ssb = cb < 12 ? 1.f : -1.f;
    const bool a_even = (ca & 1) == 0;
    LaneFold f;
    f.col0 = a_even ? ca : cb;
    f.col1 = a_even ? cb : ca;
    f.Pc0 = a_even ? sca * Pa : scb * Pb;
    f.Qc0 = a_even ? sca * Qa : scb * Qb;
    f.Pc1 = a_even ? scb * Pb : sca * Pa;
    f.Qc1 = a_even ? scb * Qb : sca * Qa;
    f.Ps0 = a_even ? ssa * Pa : ssb * Pb;
    f.Qs0 = a_even ? ssa * Qa : ssb * Qb;
    f.Ps1 = a_even ? ssb * Pb : ssa * Pa;
    f.Qs1 = a_even ? ssb * Qb : ssa * Qa;
    return f;
}

// Matrix-core constants of the split-precision forward DCT (DCT = 1: FP16,
// 2: BF16 hi + lo), per lane; DCT = 0 (FP32 butterflies) needs none.
template <int DCT>
struct CvtMfma {
    MfmaConst<DCT> c4, s4;
};
template <>
struct CvtMfma<0> {};

template <int DCT>
__device__ __forceinline__ CvtMfma<DCT> cvt_mfma(int lane) {
    CvtMfma<DCT> t;
    if constexpr (DCT != 0) {
        t.c4 = mfma_const<DCT, 0>(lane);
        t.s4 = mfma_const<DCT, 1>(lane);
    }
    return t;
}

// The loads that start a unit: the tile centre for this camera, then the
// CltExtra of its kernel cell (a dependent load).  The staged kernel issues
// them one and two pairs ahead so that their latency overlaps other work.
// Each of the unit's 8 lanes loads one word of a record (the vector memory
// path is the forward CLT's busiest unit, and it is paid per lane and
// byte), and the words are broadcast across the unit's lanes with
// ds_swizzle where they are used.
struct UnitPre {
    int task;  // the task word (0: an empty task, nothing stored)
    int txy;   // tile x | y << 16
    float cx, cy;
    int wc;    // this lane's word of the task record (lane li & 3: task, txy, x, y)
    int we;    // this lane's word of the CltExtra record (word li)
};

// Word K of the lane's 8-lane unit (lane (lane & ~7) | K): swizzle in
// bit-mask mode, and_mask 0x18, or_mask K.
template <int K>
__device__ __forceinline__ int unit_word(int v) {
    return __builtin_amdgcn_ds_swizzle(v, 0x18 | (K << 5));
}

__device__ __forceinline__ void unit_centre(const ConvertArgs &a, int cam, int task_num, UnitPre &u) {
    const int li = __lane_id() & 3;
    const int w = li < 2 ? li : 2 * cam + li;  // task, txy, xy[cam][0], xy[cam][1]
    u.wc = ((__attribute__((address_space(1))) const int *)(gtasks(a.tasks) + task_num))[w];
}

// task, txy and centre from the loaded task words
__device__ __forceinline__ void unit_centre_bcast(UnitPre &u) {
    u.task = unit_word<0>(u.wc);
    u.txy = unit_word<1>(u.wc);
    u.cx = __int_as_float(unit_word<2>(u.wc));
    u.cy = __int_as_float(unit_word<3>(u.wc));
}

// kernel cell of a centre, :3129-3134
__device__ __forceinline__ size_t kernel_cell(const Geo &g, float cx, float cy, int color, float &kdx,
                                              float &kdy) {
    int ktx = (int)__builtin_rintf(fmaf(cx, 0.0625f, 1.0f));
    int kty = (int)__builtin_rintf(fmaf(cy, 0.0625f, 1.0f));
    ktx = min(g.KH - 1, max(0, ktx));
    kty = min(g.KV - 1, max(0, kty));
    kdx = (cx - (float)(ktx << 4)) + 8.0f;
    kdy = (cy - (float)(kty << 4)) + 8.0f;
    return (size_t)(ktx + kty * g.KH) * TP_NUM_COLORS + color;
}

// Broadcasts the task words, then loads this lane's word of the CltExtra of
// the unit's kernel cell.
template <int COLOR>
__device__ __forceinline__ void unit_extra(const ConvertArgs &a, int cam, UnitPre &u) {
    unit_centre_bcast(u);
    float kdx, kdy;
    const size_t kfi = kernel_cell(a.g, u.cx, u.cy, COLOR, kdx, kdy);
    u.we = __float_as_int(gptr(a.l_koffs + kfi * 8)[__lane_id() & 7]);
}

// Phase 1 of a unit: kernel cell, integer window position and residual
// shift, and the unit's Bayer samples (R, B: 8 per lane; G: 16).  The sample
// loads are issued here and first used by unit_compute, so a caller can issue
// them early (the staged kernel issues the next unit's ahead of this unit's
// stores: loads and stores share one completion counter, so a wait for loads
// issued after stores is also a wait for the stores).
struct UnitFetch {
    float p[16];     // R, B: p[l] = sample of folded row l; G: p[l] from row ra, p[8 + l] from row rb
    float sh, sv;    // residual shifts (px)
    int itlx, itly;  // window top-left
    int kfi;         // kernel cell * 3 + color
};

template <int COLOR>
__device__ __forceinline__ void unit_fetch(const ConvertArgs &a, const LaneFold &lf, const UnitPre &up,
                                           UnitFetch &F) {
    constexpr int color0 = COLOR & 1, color1 = (COLOR >> 1) & 1;
    const Geo &g = a.g;
    const float cx = up.cx, cy = up.cy;

    // (1) kernel cell (recomputed from the centre: a few integer operations)
    float kdx, kdy;
    F.kfi = (int)kernel_cell(g, cx, cy, COLOR, kdx, kdy);
    // CltExtra data_x, data_y, dxc_dx, dxc_dy, dyc_dx, dyc_dy
    const float e0 = __int_as_float(unit_word<0>(up.we)), e1 = __int_as_float(unit_word<1>(up.we));
    const float e4 = __int_as_float(unit_word<4>(up.we)), e5 = __int_as_float(unit_word<5>(up.we));
    const float e6 = __int_as_float(unit_word<6>(up.we)), e7 = __int_as_float(unit_word<7>(up.we));
    const float offx = fmaf(e5, kdy, fmaf(e4, kdx, e0));
    const float offy = fmaf(e7, kdy, fmaf(e6, kdx, e1));

    // (2) integer top-left, residual shift, :3182-3247
    const float px = (cx - 8.0f) - offx;
    const float py = (cy - 8.0f) - offy;
    int itlx = (int)floorf(px + 0.5f);
    int itly = (int)floorf(py + 0.5f);
    if (itlx < 0) itlx &= 1;
    if (itlx >= g.W - 8) itlx = (itlx & 1) + (g.W - 10);
    if (itly < 0) itly &= 1;
    if (itly >= g.H - 8) itly = (itly & 1) + (g.H - 10);
    F.sh = ((float)(itlx + 8) - cx) + offx;
    F.sv = ((float)(itly + 8) - cy) + offy;
    F.itlx = itlx;
    F.itly = itly;

    // Pixel offsets (32-bit).  Interior windows (the usual case) are linear
    // in the row; windows reaching past the right/bottom edge read the
    // nearest pixel of the same Bayer color (itlx, itly >= 0 always).
    gfloat *img = gptr(a.l_img);
    auto pix = [&](int o) -> float { return img[o]; };
    const int ds = (int)a.dstride;
    const bool inside = (itlx + 15 < g.W) && (itly + 15 < g.H);
    auto rowoff = [&](int r) -> int {
        const int y = itly + r;
        return (inside ? y : min(y, g.H - 2 + (y & 1))) * ds;
    };
    auto coloff = [&](int c) -> int {
        const int x = itlx + c;
        return inside ? x : min(x, g.W - 2 + (x & 1));
    };
    const int x0 = coloff(lf.col0), x1 = coloff(lf.col1);
    if (!color1) {
        const int pc = (itlx & 1) ^ (1 ^ color0);
        const int pr = (itly & 1) ^ color0;
        const int xo = pc ? x1 : x0;
#pragma unroll
        for (int l = 0; l < 8; l++) {
            const int ra = l >= 4 ? l - 4 : 11 - l, rb = l >= 4 ? 11 - l : 12 + l;
            F.p[l] = pix(rowoff((ra & 1) == pr ? ra : rb) + xo);
        }
    } else {
        // G sites: (row + col + itlx + itly) even; both source rows used
        const int S = (itlx + itly) & 1;
#pragma unroll
        for (int l = 0; l < 8; l++) {
            const int ra = l >= 4 ? l - 4 : 11 - l, rb = l >= 4 ? 11 - l : 12 + l;
            const int pca = (ra + S) & 1;  // column parity for row ra; row rb takes the other
            F.p[l] = pix(rowoff(ra) + (pca ? x1 : x0));
            F.p[8 + l] = pix(rowoff(rb) + (pca ? x0 : x1));
        }
    }
}

// Phase 2 of a unit: windows, fold, DCTs, replication, kernel, phase shift.
// On return lane li holds row l = li of the 4 quadrants in q0..q3.  lds_c,
// lds_s: this unit's two 8 x 8 LDS staging blocks (the 8 units of a
// wavefront are ustride floats apart).  DCT selects the 8-point
// DCT-IV/DST-IV stages: FP32 butterflies (0) or split-precision matrix-core
// contractions across the wavefront's 8 units (1: FP16 hi + lo, 2: BF16
// hi + lo).
template <int COLOR, int DCT = 0>
__device__ __forceinline__ void unit_compute(const ConvertArgs &a, const LaneFold &lf, float *lds_c,
                                             float *lds_s, int li, const UnitFetch &F,
                                             float (&q0)[8], float (&q1)[8], float (&q2)[8],
                                             float (&q3)[8], const CvtMfma<DCT> &mt = {},
                                             int ustride = 72) {
    constexpr int color0 = COLOR & 1, color1 = (COLOR >> 1) & 1;
    const int itlx = F.itlx, itly = F.itly;
    const float sh = F.sh, sv = F.sv;
    const float *kern = a.l_kern + (size_t)F.kfi * 256;  // read through gptr4
    float ahs, ahc, avs, avc;
    sincospif(sh * 0.0625f, &ahs, &ahc);
    sincospif(sv * 0.0625f, &avs, &avc);

    // (3) windows.  Vertical: all 16 source rows (compile-time n); the fold
    // sign (cosine branch) is folded in.
    float wv[16];
#pragma unroll
    for (int n = 0; n < 16; n++) {
        const float P = n < 8 ? kTab.hw[n] : kTab.hw[15 - n];
        const float Q = n < 8 ? kTab.hw[7 - n] : -kTab.hw[n - 8];
        wv[n] = (float)fold_rcs(n) * fmaf(P, avc, Q * avs);
    }
    // Horizontal: this lane's two source columns, by local parity.
    const float whc0 = fmaf(lf.Pc0, ahc, lf.Qc0 * ahs), whc1 = fmaf(lf.Pc1, ahc, lf.Qc1 * ahs);
    const float whs0 = fmaf(lf.Ps0, ahc, lf.Qs0 * ahs), whs1 = fmaf(lf.Ps1, ahc, lf.Qs1 * ahs);

    // (4) fold: f[l] = folded (l, li)
    float f[8], fs[8];
    if (!color1) {
        const int pc = (itlx & 1) ^ (1 ^ color0);
        const int pr = (itly & 1) ^ color0;
        const float wc = pc ? whc1 : whc0;
#pragma unroll
        for (int l = 0; l < 8; l++) {
            const int ra = l >= 4 ? l - 4 : 11 - l, rb = l >= 4 ? 11 - l : 12 + l;
            const float w = (ra & 1) == pr ? wv[ra] : wv[rb];
            f[l] = F.p[l] * w * wc;
        }
    } else {
        const int S = (itlx + itly) & 1;
#pragma unroll
        for (int l = 0; l < 8; l++) {
            const int ra = l >= 4 ? l - 4 : 11 - l, rb = l >= 4 ? 11 - l : 12 + l;
            const int pca = (ra + S) & 1;
            const float da = F.p[l] * wv[ra];
            const float db = F.p[8 + l] * wv[rb];
            f[l] = fmaf(da, pca ? whc1 : whc0, db * (pca ? whc0 : whc1));
            fs[l] = fmaf(da, pca ? whs1 : whs0, db * (pca ? whs0 : whs1));
        }
    }

    // (5) vertical DCT-IV (column li), (6) transpose through LDS: column li
    // goes to row li of the staging block, then lane li reads its row li
    // (element k from lane k).
    const int lane = __lane_id();
    float *stage_c = lds_c - (lane >> 3) * ustride, *stage_s = lds_s - (lane >> 3) * ustride;
    // (rows written unswizzled: measured 3% faster here than the bank-spread
    // layout of put_row8, whose extra address arithmetic costs more than the
    // write conflicts in this kernel)
    if constexpr (DCT == 0) {
        dct4_8(f);
        if (color1) dct4_8(fs);
        *(float4 *)(lds_c + li * 8) = make_float4(f[0], f[1], f[2], f[3]);
        *(float4 *)(lds_c + li * 8 + 4) = make_float4(f[4], f[5], f[6], f[7]);
        if (color1) {
            *(float4 *)(lds_s + li * 8) = make_float4(fs[0], fs[1], fs[2], fs[3]);
            *(float4 *)(lds_s + li * 8 + 4) = make_float4(fs[4], fs[5], fs[6], fs[7]);
        }
    } else {
        mfma_dct8_to_lds<DCT>(f, mt.c4, stage_c, lane, ustride);
        if (color1) mfma_dct8_to_lds<DCT>(fs, mt.c4, stage_s, lane, ustride);
    }
    wave_lds_sync();
    float cc[8], sc[8];
#pragma unroll
    for (int k = 0; k < 8; k++) cc[k] = lds_c[k * 8 + li];
    if (color1) {
#pragma unroll
        for (int k = 0; k < 8; k++) sc[k] = lds_s[k * 8 + li];
    }
    wave_lds_sync();

    // (7) horizontal DCT-IV (and DST-IV for G)
    if constexpr (DCT == 0) {
        dct4_8(cc);
        if (color1) dst4_8(sc);
    } else {
        mfma_dct8_to_lds<DCT>(cc, mt.c4, stage_c, lane, ustride);
        if (color1) mfma_dct8_to_lds<DCT>(sc, mt.s4, stage_s, lane, ustride);
        wave_lds_sync();
#pragma unroll
        for (int k = 0; k < 8; k++) cc[k] = lds_c[li * 8 + k];
        if (color1) {
#pragma unroll
            for (int k = 0; k < 8; k++) sc[k] = lds_s[li * 8 + k];
        }
        wave_lds_sync();
    }

    // (8) replicate quadrants, :3457-3497.  Lane li is row l = li.
    const float nd = (float)(1 - ((((itlx & 1) ^ (itly & 1) ^ (1 ^ color1))) << 1));
    if (!color1) {
        const float nh = (float)(1 - (((itlx & 1) ^ (1 ^ COLOR)) << 1));
        float m[8];
#pragma unroll
        for (int k = 0; k < 8; k++) m[k] = mirror8(cc[k]);  // row 7-l
#pragma unroll
        for (int k = 0; k < 8; k++) {
            q0[k] = cc[k];
            q1[k] = nh * cc[7 - k];
            q3[k] = nd * m[7 - k];
            q2[k] = (nd * nh) * m[k];
        }
    } else {
        float m1[8], m2[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            m1[k] = mirror8(cc[k]);
            m2[k] = mirror8(sc[k]);
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            q0[k] = cc[k];
            q1[k] = sc[k];
            q3[k] = nd * m1[7 - k];
            q2[k] = nd * m2[7 - k];
        }
    }

    // (9) aberration kernel (bicomplex product), convolveTiles :2778-2827;
    // one half row (4 bins) of the 4 kernel quadrants at a time
    {
        gf32x4 *kp = gptr4(kern + li * 8);
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const f32x4 K0 = kp[h], K1 = kp[16 + h], K2 = kp[32 + h], K3 = kp[48 + h];
#pragma unroll
            for (int kk = 0; kk < 4; kk++) {
                const int k = 4 * h + kk;
                auto el = [kk](const f32x4 &v) { return v[kk]; };
                const float k0 = el(K0), k1 = el(K1), k2 = el(K2), k3 = el(K3);
                const float c0 = q0[k], c1 = q1[k], c2 = q2[k], c3 = q3[k];
                q0[k] = fmaf(c3, k3, fmaf(-c2, k2, fmaf(-c1, k1, c0 * k0)));
                q1[k] = fmaf(-c2, k3, fmaf(-c3, k2, fmaf(c0, k1, c1 * k0)));
                q2[k] = fmaf(-c1, k3, fmaf(c0, k2, fmaf(-c3, k1, c2 * k0)));
                q3[k] = fmaf(c0, k3, fmaf(c1, k2, fmaf(c2, k1, c3 * k0)));
            }
        }
    }

    // (10) phase rotations: horizontal by pi sh (2k+1)/16 (shiftTileHor
    // :2715-2745), then vertical by pi sv (2l+1)/16 (shiftTileVert :2747-2776).
    {
        float er[8], ei[8];  // e^{i (2k+1) a}, a = pi sh / 16
        const float e1r = ahc, e1i = ahs;
        float e2r, e2i, e4r, e4i, e8r, e8i;
        cmul(e1r, e1i, e1r, e1i, e2r, e2i);
        cmul(e2r, e2i, e2r, e2i, e4r, e4i);
        cmul(e4r, e4i, e4r, e4i, e8r, e8i);
        er[0] = e1r; ei[0] = e1i;
        cmul(e1r, e1i, e2r, e2i, er[1], ei[1]);        // 3
        cmul(er[1], ei[1], e2r, e2i, er[2], ei[2]);    // 5
        cmul(er[2], ei[2], e2r, e2i, er[3], ei[3]);    // 7
#pragma unroll
        for (int t = 0; t < 4; t++) cmul(er[t], ei[t], e8r, e8i, er[t + 4], ei[t + 4]);  // 9..15
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float ch = er[k], s = ei[k];
            const float a0 = q0[k], b0 = q1[k], a2 = q2[k], b2 = q3[k];
            q0[k] = fmaf(a0, ch, -b0 * s);
            q1[k] = fmaf(a0, s, b0 * ch);
            q2[k] = fmaf(a2, ch, -b2 * s);
            q3[k] = fmaf(a2, s, b2 * ch);
        }
        float vs, vc;
        sincospif(sv * (float)(2 * li + 1) * 0.0625f, &vs, &vc);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float a0 = q0[k], b0 = q2[k], a1 = q1[k], b1 = q3[k];
            q0[k] = fmaf(a0, vc, -b0 * vs);
            q2[k] = fmaf(a0, vs, b0 * vc);
            q1[k] = fmaf(a1, vc, -b1 * vs);
            q3[k] = fmaf(a1, vs, b1 * vc);
        }
    }
    // (11) optional LPF, :3572-3590
    if ((a.lpf_mask >> COLOR) & 1) {
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float lv = c_lpf.v[COLOR][li * 8 + k];
            q0[k] *= lv; q1[k] *= lv; q2[k] *= lv; q3[k] *= lv;
        }
    }
}

// One unit (tile task_num, camera cam, color COLOR) up to the optional LPF:
// unit_fetch then unit_compute.  pre: the unit's centre and CltExtra if the
// caller loaded them ahead (else they are loaded here).
template <int COLOR, int DCT = 0>
__device__ __forceinline__ void convert_unit(const ConvertArgs &a, const LaneFold &lf, float *lds_c,
                                             float *lds_s, int li, int cam, int task_num,
                                             float (&q0)[8], float (&q1)[8], float (&q2)[8],
                                             float (&q3)[8], const CvtMfma<DCT> &mt = {},
                                             int ustride = 72, const UnitPre *pre = nullptr) {
    UnitPre up;
    if (pre) {
        up = *pre;
    } else {
        unit_centre(a, cam, task_num, up);
        unit_extra<COLOR>(a, cam, up);
    }
    UnitFetch F;
    unit_fetch<COLOR>(a, lf, up, F);
    // all sample loads are issued before the first use (the scheduling
    // barrier keeps the compiler from pairing each load with its use, which
    // serialises the HBM latencies when it is short of registers)
    __builtin_amdgcn_sched_barrier(0);
    unit_compute<COLOR, DCT>(a, lf, lds_c, lds_s, li, F, q0, q1, q2, q3, mt, ustride);
}

// LDS slice of one unit in the staged kernel: its 8 x 8 transposition
// staging, then its 1 KB output block.
constexpr int kSlice = 256;

// Tile index of a task, or -1 for a task outside the tile grid (skipped).
__device__ __forceinline__ int task_tile(const Geo &g, int txy) {
    const int tx = txy & 0xfff, ty = txy >> 16;
    return (tx < g.TX && ty >= 0 && ty < g.TY) ? ty * g.TX + tx : -1;
}

// Row li of the 4 quadrants of one 8x8x4 CLT block ([q][l][k]) at dst.
__device__ __forceinline__ void store_clt_rows(float *dst, int li, const float (&q0)[8],
                                               const float (&q1)[8], const float (&q2)[8],
                                               const float (&q3)[8]) {
    dst += li * 8;
    store4(dst, q0[0], q0[1], q0[2], q0[3]);
    store4(dst + 4, q0[4], q0[5], q0[6], q0[7]);
    store4(dst + 64, q1[0], q1[1], q1[2], q1[3]);
    store4(dst + 68, q1[4], q1[5], q1[6], q1[7]);
    store4(dst + 128, q2[0], q2[1], q2[2], q2[3]);
    store4(dst + 132, q2[4], q2[5], q2[6], q2[7]);
    store4(dst + 192, q3[0], q3[1], q3[2], q3[3]);
    store4(dst + 196, q3[4], q3[5], q3[6], q3[7]);
}

// Output block of a unit, [q][l][k], :3600-3630; st: whether it is stored
// (an empty task or one outside the tile grid is skipped).
template <int COLOR>
__device__ __forceinline__ float *unit_dst(const ConvertArgs &a, int txy, bool valid, bool &st) {
    const int tile = task_tile(a.g, txy);
    st = valid && tile >= 0;
    return a.l_clt + ((size_t)max(tile, 0) * TP_NUM_COLORS + COLOR) * 256;
}

// The unit's 1 KB block goes through LDS so that store j writes float4
// j * 8 + li: every store instruction covers whole 128-byte lines of each
// unit (non-temporal stores of partial lines cost extra writes).  stage_block
// writes the block into the unit's slice (last read before unit_compute's
// final wavefront fence), store_block copies it to memory.
__device__ __forceinline__ void stage_block(float *slice, int li, const float (&q0)[8],
                                            const float (&q1)[8], const float (&q2)[8],
                                            const float (&q3)[8]) {
    const float *q[4] = {q0, q1, q2, q3};
#pragma unroll
    for (int qq = 0; qq < 4; qq++) {
        *(float4 *)(slice + qq * 64 + li * 8) = make_float4(q[qq][0], q[qq][1], q[qq][2], q[qq][3]);
        *(float4 *)(slice + qq * 64 + li * 8 + 4) = make_float4(q[qq][4], q[qq][5], q[qq][6], q[qq][7]);
    }
    wave_lds_sync();
}
// Write sink of the blocks that are not stored (empty tasks, tiles outside
// the grid, slots past the end of the list): their stores go here instead
// of being branched around, so every path issues the same memory
// operations and the compiler's wait counts for the loads issued before
// them do not have to assume the shorter path (which would make a wait for
// those loads also wait for the stores).  Never read.
__device__ __align__(256) float g_convert_sink[256];

__device__ __forceinline__ void store_block(const float *slice, float *dst, int li, bool st) {
    float *base = st ? dst : g_convert_sink;
    // store j writes unit j's whole 1 KB block: lane l copies its float4 l
    // (the block's address from unit j's first lane)
    const int lane = __lane_id();
    const float *s0 = slice - (lane >> 3) * kSlice;  // unit 0's slice
    const uint64_t b64 = (uint64_t)(uintptr_t)base;
    const uint32_t blo = (uint32_t)b64, bhi = (uint32_t)(b64 >> 32);
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const uint64_t bj = ((uint64_t)__builtin_amdgcn_readlane(bhi, 8 * j) << 32) |
                            (uint32_t)__builtin_amdgcn_readlane(blo, 8 * j);
        const f32x4 v = *(const f32x4 *)(s0 + j * kSlice + lane * 4);
        store4((float *)(uintptr_t)bj + lane * 4, v[0], v[1], v[2], v[3]);
    }
    wave_lds_sync();  // the slice is the next unit's staging
}

// R, B and G of the wavefront's 8 units (u0: their centres, task words
// loaded; v0: whether the units' slots are real list entries).  Each
// color's samples are requested before the previous color's stores: loads
// and stores share one completion counter, so a wait for samples issued
// after the stores would also wait for the stores.
__device__ __forceinline__ void rgb_units(const ConvertArgs &la, const LaneFold &lf, float *slice, int li,
                                          int cam, const UnitPre &u0, bool v0) {
    UnitPre ur = u0, ub = u0, ug = u0;
    unit_extra<0>(la, cam, ur);
    unit_extra<1>(la, cam, ub);
    unit_extra<2>(la, cam, ug);
    __builtin_amdgcn_sched_barrier(0);
    const bool valid = v0 && ur.task != 0;
    UnitFetch F;
    float q0[8], q1[8], q2[8], q3[8];
    bool st;
    float *dst;
    unit_fetch<0>(la, lf, ur, F);
    __builtin_amdgcn_sched_barrier(0);
    unit_compute<0>(la, lf, slice, slice + 72, li, F, q0, q1, q2, q3, {}, kSlice);
    dst = unit_dst<0>(la, ur.txy, valid, st);
    stage_block(slice, li, q0, q1, q2, q3);
    unit_fetch<1>(la, lf, ub, F);
    __builtin_amdgcn_sched_barrier(0);
    store_block(slice, dst, li, st);
    unit_compute<1>(la, lf, slice, slice + 72, li, F, q0, q1, q2, q3, {}, kSlice);
    dst = unit_dst<1>(la, ub.txy, valid, st);
    stage_block(slice, li, q0, q1, q2, q3);
    unit_fetch<2>(la, lf, ug, F);
    __builtin_amdgcn_sched_barrier(0);
    store_block(slice, dst, li, st);
    unit_compute<2>(la, lf, slice, slice + 72, li, F, q0, q1, q2, q3, {}, kSlice);
    dst = unit_dst<2>(la, ug.txy, valid, st);
    stage_block(slice, li, q0, q1, q2, q3);
    store_block(slice, dst, li, st);
}

// Blocks of k_convert_cam per camera: patches of 8 list columns x 4 list
// rows of TX entries (n >= 1).
__host__ __device__ inline int convert_blocks(int n, int TX) {
    return ((n - 1) / TX / 4 + 1) * ((TX + 7) / 8);
}

// Staged forward CLT.  The list is read as rows of TX entries (a raster
// list: tile rows); workgroup b converts a patch of 8 columns x 4 rows of it
// for camera b / nb (nb = convert_blocks), wavefront w row w of the patch.
// The grid runs camera after camera, patch row after patch row, so the
// workgroups resident at a time write one contiguous stretch of one
// camera's CLT (3 KB per tile) and read one camera's image and kernels,
// rather than four stretches in the four cameras' buffers, and the 4 tile
// rows of a patch share kernel cells and image lines within the workgroup.
// A wavefront's 8 units are 8 neighbouring tiles of one camera, so units
// that share an aberration kernel cell (2 x 2 tiles per cell) or image lines
// load the same addresses in the same instruction, and the camera's pointers
// are uniform.  Slots past the end of a row or of the list convert the last
// task and store nothing.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TP_CONVERT_WAVES)))
void k_convert_cam(ConvertArgs a) {
    __shared__ __align__(16) float lds[4][8][kSlice];  // [wave][unit]: staging, then output block
    const int wave = threadIdx.x >> 6;
    const int TX = a.g.TX, cpr = (TX + 7) / 8;  // 8-entry column groups per list row
    const int nb = convert_blocks(a.num_tiles, TX);
    const int cam = __builtin_amdgcn_readfirstlane((int)blockIdx.x / nb);
    const int blk = (int)blockIdx.x - cam * nb;
    const int prow = blk / cpr, pcol = blk - prow * cpr;
    const int lane = threadIdx.x & 63;
    const int unit = lane >> 3;
    const int li = lane & 7;
    const int col = pcol * 8 + unit;
    const int s = (prow * 4 + wave) * TX + col;
    const bool valid = col < TX && s < a.num_tiles;
    float *slice = &lds[wave][unit][0];
    const LaneFold lf = lane_fold(li);
    ConvertArgs la = a;
    resolve_cam(la, cam);
    UnitPre u0;
    unit_centre(la, cam, valid ? s : a.num_tiles - 1, u0);
    rgb_units(la, lf, slice, li, cam, u0, valid);
}

}  // namespace tp
