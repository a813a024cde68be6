// k_imclt.hpp - inverse CLT of color planes with overlap-add (plain FP32).
//
// Replaces imclt_rbg (src/TileProcessor.cuh:2588-2712) and imclt
// (:3700-3838).  A unit is one tile of one color plane; 8 lanes per unit:
//   1. lane = row l loads row l of the 4 quadrants;
//   2. horizontal DCT-IV (quadrants 0, 2) / DST-IV (1, 3) in registers;
//   3. transpose through LDS (lane = column k);
//   4. vertical DCT-IV (0, 1) / DST-IV (2, 3);
//   5. unfold: lane k produces the two output columns that fold onto k,
//      for all 16 rows, windowed by HWINDOW2(row) * HWINDOW2(col).
//
// k_imclt_pass reproduces one reference pass (RMW of the plane in HBM).
// k_imclt_strip does the whole frame in one launch: a workgroup owns a strip
// of 120 output columns of one plane over a band of rows, transforms each
// tile of the strip once, and writes every output pixel once, as the sum of
// the 4 tile outputs that cover it.
#pragma once

#include <type_traits>

#include "tp_common.hpp"

namespace tp {

// HWINDOW2[i] = sin(pi (2i+1)/32) / 2 for a runtime i
__device__ __forceinline__ float hw2_sel(int i) {
    float h = 0.f;
#pragma unroll
    for (int t = 0; t < 8; t++) h = (i == t) ? 0.5f * kTab.hw[t] : h;
    return h;
}

// 1 / HWINDOW2[i] for a runtime i (the un-windowing factor of one axis)
__device__ __forceinline__ float ihw2_sel(int i) {
    float h = 0.f;
#pragma unroll
    for (int t = 0; t < 8; t++) h = (i == t) ? 2.0f / kTab.hw[t] : h;
    return h;
}

// Steps 2-4 on one unit; q* in: row li, out: column li of each quadrant.
// L: the unit's staging, NQ x 72 floats; NQ quadrants are transposed at a
// time (NQ = 1 keeps the LDS footprint of a workgroup small).
template <int NQ = 4>
__device__ __forceinline__ void imclt_transform(float (&q0)[8], float (&q1)[8], float (&q2)[8],
                                                float (&q3)[8], float *L, int li) {
    static_assert(NQ == 1 || NQ == 2 || NQ == 4, "quadrants per transposition");
    dct4_8(q0);
    dst4_8(q1);
    dct4_8(q2);
    dst4_8(q3);
    float *q[4] = {q0, q1, q2, q3};
#pragma unroll
    for (int q0i = 0; q0i < 4; q0i += NQ) {
#pragma unroll
        for (int j = 0; j < NQ; j++) {
            const float *v = q[q0i + j];
            const float row[8] = {v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]};
            put_row8(L + j * 72, li, row);
        }
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < NQ; j++)
#pragma unroll
            for (int i = 0; i < 8; i++) q[q0i + j][i] = get8(L + j * 72, i, li);
        wave_lds_sync();
    }
    dct4_8(q0);
    dct4_8(q1);
    dst4_8(q2);
    dst4_8(q3);
}

// Step 5: calls emit(row, col, weight, sum) for the 32 outputs of lane li
// (with SLOT: emit(row, col, slot, weight, sum), slot a std::integral_constant
// 0 for the lane's first column, 1 for its second).
// weight = HWINDOW2(row) * HWINDOW2(col); with UNWIN it is
// 1 / (HWINDOW2(row) * HWINDOW2(col)) / 16, i.e. the windowed output
// divided by HWINDOW_SQ(row) * HWINDOW_SQ(col) (HWINDOW_SQ = HWINDOW^2 =
// 4 HWINDOW2^2): the un-windowed pixel that debayering starts from.
template <bool UNWIN = false, bool SLOT = false, typename F>
__device__ __forceinline__ void imclt_unfold(const float (&t0)[8], const float (&t1)[8],
                                             const float (&t2)[8], const float (&t3)[8], int li,
                                             F &&emit) {
    const int ca = li >= 4 ? li - 4 : 11 - li;
    const int cb = li >= 4 ? 11 - li : 12 + li;
    const int sa = ca < 8 ? ca : 15 - ca, sb = cb < 8 ? cb : 15 - cb;
    // column c's sum is cc A + sc B (signs cc, sc = +-1) = cc (A + cc sc B):
    // cc goes into the column weight, exactly
    const float cca = ca < 4 ? 1.f : -1.f, sca = ca < 12 ? 1.f : -1.f;
    const float ccb = cb < 4 ? 1.f : -1.f, scb = cb < 12 ? 1.f : -1.f;
    const float wca = cca * (UNWIN ? 0.0625f * ihw2_sel(sa) : hw2_sel(sa));
    const float wcb = ccb * (UNWIN ? 0.0625f * ihw2_sel(sb) : hw2_sel(sb));
    const float rha = cca * sca, rhb = ccb * scb;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int rho = fold_rind(r);
        const float cr = (float)fold_rcs(r), sr = (float)fold_rss(r);
        const float A = fmaf(cr, t0[rho], sr * t2[rho]);
        const float B = fmaf(cr, t1[rho], sr * t3[rho]);
        const float wr = UNWIN ? 2.0f / kTab.hw[r < 8 ? r : 15 - r] : 0.5f * kTab.hw[r < 8 ? r : 15 - r];
        if constexpr (SLOT) {
            emit(r, ca, std::integral_constant<int, 0>(), wr * wca, fmaf(rha, B, A));
            emit(r, cb, std::integral_constant<int, 1>(), wr * wcb, fmaf(rhb, B, A));
        } else {
            emit(r, ca, wr * wca, fmaf(rha, B, A));
            emit(r, cb, wr * wcb, fmaf(rhb, B, A));
        }
    }
}

__device__ __forceinline__ void load_rows(const float *src, int li, float (&q0)[8], float (&q1)[8],
                                          float (&q2)[8], float (&q3)[8]) {
    gf32x4 *p = gptr4(src + li * 8);  // global loads (src is a device-memory CLT block)
    float *q[4] = {q0, q1, q2, q3};
#pragma unroll
    for (int qq = 0; qq < 4; qq++) {
        const f32x4 u = p[qq * 16], v = p[qq * 16 + 1];
        q[qq][0] = u[0]; q[qq][1] = u[1]; q[qq][2] = u[2]; q[qq][3] = u[3];
        q[qq][4] = v[0]; q[qq][5] = v[1]; q[qq][6] = v[2]; q[qq][7] = v[3];
    }
}

// One reference pass: tiles of parity (v_off, h_off), 8 per wavefront.
__global__ __launch_bounds__(64) void k_imclt_pass(Geo g, const float *clt, float *rbg, int color,
                                                   int v_off, int h_off, size_t dstride) {
    __shared__ __align__(16) float lds[8][4][72];
    const int lane = threadIdx.x;
    const int unit = lane >> 3, li = lane & 7;
    const int th = (g.TX + (h_off ^ 1)) >> 1;
    const int tv = (g.TY + (v_off ^ 1)) >> 1;
    int n = blockIdx.x * 8 + unit;
    const bool valid = n < th * tv;
    if (!valid) n = 0;
    const int ty = (n / th) * 2 + v_off, tx = (n % th) * 2 + h_off;
    const bool first = (v_off | h_off) == 0;
    float q0[8], q1[8], q2[8], q3[8];
    load_rows(clt + ((size_t)(ty * g.TX + tx) * TP_NUM_COLORS + color) * 256, li, q0, q1, q2, q3);
    imclt_transform(q0, q1, q2, q3, &lds[unit][0][0], li);
    float *plane = rbg + dstride * (size_t)(g.H + 8) * color + (size_t)(ty * 8) * dstride + tx * 8;
    if (valid) {
        imclt_unfold(q0, q1, q2, q3, li, [&](int r, int c, float w, float d) {
            float *p = plane + (size_t)r * dstride + c;
            *p = first ? w * d : fmaf(w, d, *p);
        });
    }
}

// Whole frame by strips of tile columns: a workgroup owns IMS_OWN column
// groups of 8 output columns of one plane (the tiles starting on them plus
// the tile on the strip's left edge: IMS_TILES tiles per tile row) over
// `band` row groups (odd), and walks down them two tile rows per step.
// Every tile of the strip is transformed once (tiles on a strip's left edge
// and the tile row above a band are the only repeats), all 32 unit slots of
// the 4 wavefronts are used, and each output pixel is the sum of the 4
// windowed tile outputs that cover it, gathered from LDS and written once.
// LDS: 5 half-slots of 8 tile rows x IMS_TILES tiles x 16 columns; a step
// writes the top and bottom halves of its two tile rows into 4 of them, the
// fifth holds the bottom half of the previous step's second row.
constexpr int IMS_TILES = 16, IMS_OWN = IMS_TILES - 1, IMS_ROW = 16 * IMS_TILES, IMS_HS = 8 * IMS_ROW;
// row groups per band: 3 steps.  Smaller bands repeat more tile rows (one
// per band) but give more, shorter workgroups; 5 measured fastest on the
// full frame (0.2365 ms against 0.2383 / 0.2442 for 3 / 9)
constexpr int IMS_BAND = 5;

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
void k_imclt_strip(Geo g, const float *const *clt, float *const *rbg, size_t dstride, int band) {
    __shared__ __align__(16) float hs[5 * IMS_HS];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int unit = lane >> 3, li = lane & 7;
    const int PW = g.W + 8, PH = g.H + 8;
    const int nstrip = gridDim.x, nband = gridDim.y;
    // plane-major, strips innermost, in dispatch order (not remapped per XCD:
    // one front over the plane is 1% faster than an eighth of it per XCD)
    const int lin = blockIdx.x + nstrip * (blockIdx.y + nband * blockIdx.z);
    const int strip = lin % nstrip, rest = lin / nstrip;
    const int bnd = rest % nband, pidx = rest / nband;
    const int color = pidx % TP_NUM_COLORS, cam = pidx / TP_NUM_COLORS;
    const int G0 = strip * IMS_OWN;  // first owned column group
    const int J0 = bnd * band;       // first owned row group
    const float *cl = clt[cam];
    float *plane = rbg[cam] + dstride * (size_t)PH * color;
    // this lane's unit: tile row (wave >> 1) of the step, strip tile lt
    const int lt = (wave & 1) * 8 + unit;
    const int tx = G0 - 1 + lt;
    // gather: this thread's strip column and row parity
    const int xl = threadIdx.x & 127, hrow = threadIdx.x >> 7;
    const int cg = min(xl >> 3, IMS_OWN - 1), xc = xl & 7;  // columns past the strip read column group 14
    const int X = 8 * G0 + xl;
    const bool xok = xl < 8 * IMS_OWN && X < PW;
    const int gl = 16 * cg + xc + 8;    // left tile (cg), column xc + 8
    const int gr = 16 * (cg + 1) + xc;                  // right tile (cg + 1), column xc
    int carry = 0;
    const int steps = (band + 1) >> 1;
    for (int s = 0; s < steps; s++) {
        const int Ta = J0 - 1 + 2 * s;  // the step's tile rows Ta, Ta + 1
        // half-slots: Ta top c+1, Ta bottom c+2, Ta+1 top c+3, Ta+1 bottom c+4
        {
            const int ty = Ta + (wave >> 1);
            const int top = (carry + 1 + 2 * (wave >> 1)) % 5, bot = (carry + 2 + 2 * (wave >> 1)) % 5;
            const bool ok = ty >= 0 && ty < g.TY && tx >= 0 && tx < g.TX;
            float q0[8], q1[8], q2[8], q3[8];
            load_rows(cl + ((size_t)(ok ? ty * g.TX + tx : 0) * TP_NUM_COLORS + color) * 256, li, q0, q1, q2,
                      q3);
            // transposition staging: row `unit` of this wavefront's half of
            // the top half-slot (written by the unfold below, after it)
            float *stg = hs + top * IMS_HS + unit * IMS_ROW + (wave & 1) * (IMS_ROW / 2) + 8 * unit;
            imclt_transform<1>(q0, q1, q2, q3, stg, li);
            float *pt = hs + top * IMS_HS + 16 * lt, *pb = hs + bot * IMS_HS + 16 * lt;
            imclt_unfold(q0, q1, q2, q3, li, [&](int r, int c, float w, float d) {
                (r < 8 ? pt : pb)[(r & 7) * IMS_ROW + c] = ok ? w * d : 0.f;
            });
        }
        __syncthreads();
        // row group Ta (from the second step: the previous bottom half and
        // Ta's top half) and row group Ta + 1 (Ta's bottom, Ta+1's top)
        const float *up0 = hs + carry * IMS_HS, *lo0 = hs + ((carry + 1) % 5) * IMS_HS;
        const float *up1 = hs + ((carry + 2) % 5) * IMS_HS, *lo1 = hs + ((carry + 3) % 5) * IMS_HS;
#pragma unroll
        for (int gsel = 0; gsel < 2; gsel++) {
            const int grp = Ta + gsel;
            if (grp < J0) continue;  // the step before the band's first row group
            const float *up = gsel ? up1 : up0, *lo = gsel ? lo1 : lo0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int yr = hrow + 2 * k;
                const float v = (up[yr * IMS_ROW + gl] + up[yr * IMS_ROW + gr]) +
                                (lo[yr * IMS_ROW + gl] + lo[yr * IMS_ROW + gr]);
                const int Y = 8 * grp + yr;
                if (xok && grp < J0 + band && Y < PH) store1(plane + (size_t)Y * dstride + X, v);
            }
        }
        __syncthreads();
        carry = (carry + 4) % 5;
    }
}

}  // namespace tp
