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
// k_imclt_frame is output-stationary: a workgroup owns a 64x64 block of one
// plane, runs the 4 parity passes of the (at most 9x9) tiles that touch it
// with the sums kept in LDS, and writes the block once.
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

// Whole frame, output-stationary.  grid = (64x64 blocks, color, camera).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TP_IMCLT_WAVES)))
void k_imclt_frame(Geo g, const float *const *clt, float *const *rbg,
                                                     size_t dstride) {
    // the block's sums (pitch AP), then 8 rows of discard space: a lane's
    // outputs outside the block (whole 8-row halves of a column) are sent
    // there, which costs no per-output test
    constexpr int AP = 65, JUNK = 64 * AP;
    __shared__ __align__(16) float acc[JUNK + 8 * AP];
    __shared__ __align__(16) float lds[4][8][1][72];  // one quadrant per transposition
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int unit = lane >> 3, li = lane & 7;
    const int PW = g.W + 8, PH = g.H + 8;
    const int nbx = (PW + 63) / 64;
    // XCD-aware order: each XCD takes a contiguous range of (block, color,
    // camera), so neighbouring blocks, which share the tiles on their
    // border, are converted under one L2
    const int nblk = gridDim.x;
    const int lin = xcd_remap(blockIdx.x + nblk * (blockIdx.y + gridDim.y * blockIdx.z),
                              nblk * gridDim.y * gridDim.z);
    const int blk = lin % nblk, pidx = lin / nblk;
    const int bx = blk % nbx, by = blk / nbx;
    const int color = pidx % gridDim.y, cam = pidx / gridDim.y;
    const int X0 = bx * 64, Y0 = by * 64;
    for (int i = threadIdx.x; i < 64 * AP; i += 256) acc[i] = 0.f;
    __syncthreads();
    const float *cl = clt[cam];
    for (int pass = 0; pass < 4; pass++) {
        const int v = pass >> 1, h = pass & 1;
        // tiles of this parity touching the block: ty in [8by-1, 8by+7]
        const int ty0 = 8 * by - 1 + ((8 * by - 1 - v) & 1 ? 1 : 0);
        const int tx0 = 8 * bx - 1 + ((8 * bx - 1 - h) & 1 ? 1 : 0);
        const int nty = (8 * by + 7 - ty0) / 2 + 1, ntx = (8 * bx + 7 - tx0) / 2 + 1;
        const int slot = wave * 8 + unit;
        for (int s = slot; s < nty * ntx; s += 32) {
            const int ty = ty0 + 2 * (s / ntx), tx = tx0 + 2 * (s % ntx);
            const bool valid = ty >= 0 && ty < g.TY && tx >= 0 && tx < g.TX;
            float q0[8], q1[8], q2[8], q3[8];
            const int tyc = valid ? ty : 0, txc = valid ? tx : 0;
            load_rows(cl + ((size_t)(tyc * g.TX + txc) * TP_NUM_COLORS + color) * 256, li, q0, q1, q2, q3);
            imclt_transform<1>(q0, q1, q2, q3, &lds[wave][unit][0][0], li);
            if (valid) {
                const int oy = ty * 8 - Y0, ox = tx * 8 - X0;  // -8 .. 56
                // the lane's 2 columns (as in imclt_unfold) x 2 row halves:
                // start of each in acc, or in the discard rows
                const int ca = li >= 4 ? li - 4 : 11 - li, cb = li >= 4 ? 11 - li : 12 + li;
                int base[2][2];
#pragma unroll
                for (int cs = 0; cs < 2; cs++)
#pragma unroll
                    for (int hh = 0; hh < 2; hh++) {
                        const int c = cs ? cb : ca, x = ox + c, y0 = oy + 8 * hh;
                        base[cs][hh] = ((unsigned)x < 64u && (unsigned)y0 < 64u) ? y0 * AP + x : JUNK + c;
                    }
                imclt_unfold<false, true>(q0, q1, q2, q3, li, [&](int r, int, auto cs, float w, float d) {
                    float *p = acc + base[decltype(cs)::value][r >> 3] + (r & 7) * AP;
                    *p = fmaf(w, d, *p);
                });
            }
        }
        __syncthreads();
    }
    float *plane = rbg[cam] + dstride * (size_t)PH * color;
    for (int i = threadIdx.x; i < 64 * 64; i += 256) {
        const int y = Y0 + (i >> 6), x = X0 + (i & 63);
        if (y < PH && x < PW) store1(plane + (size_t)y * dstride + x, acc[(i >> 6) * AP + (i & 63)]);
    }
}

}  // namespace tp
