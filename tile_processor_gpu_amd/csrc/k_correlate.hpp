// k_correlate.hpp - pairwise phase correlation (plain FP32).
//
// Replaces correlate2D (src/TileProcessor.cuh:1176-1443) with
// correlateAccumulateTiles (:2829-2890), normalizeTileAmplitude (:2915-2943)
// and corrUnfoldTile (:2950-3008).
//
// A unit is one list entry (tile, pair); 8 lanes own a unit, 8 units per
// wavefront, TP_CORR_WG_WAVES (3) wavefronts per workgroup, each workgroup a
// contiguous chunk of the list (so the 6 pairs of a tile, adjacent in the
// usual tile-major list, share one L2).  Per unit:
//   1. lane = row l loads its row of the 4 quadrants of both cameras per
//      color (2 x float4 per quadrant) and accumulates scale * conj(A) B;
//      after color 1 (B) the sum is low-passed (lpf_rb_corr);
//   2. amplitude normalisation with fat zero, lpf_corr;
//   3. horizontal DCT-II / DST-II in registers, transpose through LDS,
//      vertical DCT-II / DST-II (lane = column j);
//   4. unfold to (2r+1)^2: column j combines its own column with column j-1
//      (one DPP row shift) into output columns r+j and r-j, staged in LDS;
//      the wavefront then stores its 8 output rows with contiguous stores.
// LDS per wavefront: unit u's transposition blocks at CORR_UNIT_STRIDE u
// (296 floats, 8 banks apart: the column reads of 4 units in one 32-lane
// pass fall in 32 distinct banks), then the 8 staged 15x15 tiles
// contiguous at 225 u (the 16-byte reads of the output block are aligned
// and conflict-free).
#pragma once

#include "tp_common.hpp"

namespace tp {

struct CorrArgs {
    float *const *clt;
    int colors;
    float scale[3];
    float fat_zero;
    int num_corr;
    const int *idx;
    size_t stride;
    int radius;
    float *corrs;
    int ntiles;             // tiles per camera buffer (tilesx * tilesy)
    int entries_per_block;  // multiple of 32
};

// value of lane - 1 within each row of 16 lanes (lane 0 of a row: 0)
__device__ __forceinline__ float lane_prev(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111 /* row_shr:1 */,
                                                      0xf, 0xf, false));
}

// Accumulate sum_color scale * conj(A) B for row li of one entry
// (correlateAccumulateTiles :2829-2890, R+B low-pass :1263-1304).
// load(color, which, qq, h) returns floats 4h..4h+3 of row li of quadrant qq
// of camera `which` (0: first camera of the pair, 1: second).
template <int COLORS, class Load>
__device__ __forceinline__ void corr_accumulate(Load load, const float *scale, int li, float (&c0)[8],
                                                float (&c1)[8], float (&c2)[8], float (&c3)[8]) {
#pragma unroll
    for (int k = 0; k < 8; k++) c0[k] = c1[k] = c2[k] = c3[k] = 0.f;
#pragma unroll
    for (int color = 0; color < COLORS; color++) {
        const float sc = scale[color];
        // two halves of the row (k = 0..3, 4..7) keep 32 loaded floats live
#pragma unroll
        for (int h = 0; h < 2; h++) {
            f32x4 A[4], B[4];
#pragma unroll
            for (int qq = 0; qq < 4; qq++) {
                A[qq] = load(color, 0, qq, h);
                B[qq] = load(color, 1, qq, h);
            }
#pragma unroll
            for (int kk = 0; kk < 4; kk++) {
                const int k = 4 * h + kk;
                const float a0 = A[0][kk], a1 = A[1][kk], a2 = A[2][kk], a3 = A[3][kk];
                const float b0 = B[0][kk], b1 = B[1][kk], b2 = B[2][kk], b3 = B[3][kk];
                // conj(A) * B, bicomplex with i^2 = j^2 = -1, ij = ji
                const float r0 = fmaf(a3, b3, fmaf(a2, b2, fmaf(a1, b1, a0 * b0)));
                const float r1 = fmaf(a2, b3, fmaf(-a3, b2, fmaf(a0, b1, -a1 * b0)));
                const float r2 = fmaf(a1, b3, fmaf(a0, b2, fmaf(-a3, b1, -a2 * b0)));
                const float r3 = fmaf(a0, b3, fmaf(-a1, b2, fmaf(-a2, b1, a3 * b0)));
                c0[k] = fmaf(sc, r0, c0[k]);
                c1[k] = fmaf(sc, r1, c1[k]);
                c2[k] = fmaf(sc, r2, c2[k]);
                c3[k] = fmaf(sc, r3, c3[k]);
            }
        }
        if (color == 1) {
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const float lv = c_lpf.v[4][li * 8 + k];
                c0[k] *= lv; c1[k] *= lv; c2[k] *= lv; c3[k] *= lv;
            }
        }
    }
}

// Normalise, back-transform and unfold one entry (row li of its 8 lanes):
// amplitude normalisation with fat zero and lpf_corr (:2915-2943,
// :1337-1347), horizontal DCT-II/DST-II, transpose through L, vertical
// DCT-II/DST-II, unfold (:2950-3008) into the full 15x15 tile staged at T
// (row-major, centre (7, 7)).  L: 4 * 72 floats private to the entry; T:
// 225 floats, written after every lane of the wavefront has read its L (T
// may overlap other entries' L).
__device__ __forceinline__ void corr_finish(float (&c0)[8], float (&c1)[8], float (&c2)[8],
                                            float (&c3)[8], float *L, float *T, int li, float fz2) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float s2 = fmaf(c3[k], c3[k], fmaf(c2[k], c2[k], fmaf(c1[k], c1[k], fmaf(c0[k], c0[k], fz2))));
        const float s = __frsqrt_rn(s2) * c_lpf.v[5][li * 8 + k];
        c0[k] *= s; c1[k] *= s; c2[k] *= s; c3[k] *= s;
    }
    // horizontal DCT-II (q0, q2) / DST-II (q1, q3) on row li
    dct2_8(c0);
    dst2_8(c1);
    dct2_8(c2);
    dst2_8(c3);
    // transpose: row li -> LDS, read column li
    put_row8(L, li, c0);
    put_row8(L + 72, li, c1);
    put_row8(L + 144, li, c2);
    put_row8(L + 216, li, c3);
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < 8; i++) {
        c0[i] = get8(L, i, li);
        c1[i] = get8(L + 72, i, li);
        c2[i] = get8(L + 144, i, li);
        c3[i] = get8(L + 216, i, li);
    }
    wave_lds_sync();
    // vertical DCT-II (q0, q1) / DST-II (q2, q3) on column li
    dct2_8(c0);
    dct2_8(c1);
    dst2_8(c2);
    dst2_8(c3);
    // unfold.  Lane li is column j; q1, q3 come from column j-1.
    float p1[8], p3[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        p1[i] = lane_prev(c1[i]);
        p3[i] = lane_prev(c3[i]);
    }
    const int j = li;
    const float s4 = 0.25f;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const float a0 = c0[i];
        const float a1 = j ? p1[i] : 0.f;
        const float a2 = i ? c2[i - 1] : 0.f;
        const float a3 = (i && j) ? p3[i - 1] : 0.f;
        const float t1 = a0 + a2, t2 = a0 - a2, u1 = a1 + a3, u2 = a1 - a3;
        T[(7 + i) * 15 + 7 + j] = s4 * (t1 + u1);
        T[(7 + i) * 15 + 7 - j] = s4 * (t1 - u1);
        T[(7 - i) * 15 + 7 + j] = s4 * (t2 + u2);
        T[(7 - i) * 15 + 7 - j] = s4 * (t2 - u2);
    }
}

// Offsets into the staged 15x15 tile of output elements lane + 64 t of a
// (2r+1)^2 tile (-1: past the end).
__device__ __forceinline__ void corr_out_offsets(int lane, int r, int *src_off) {
    const int n = 2 * r + 1;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int o = lane + 64 * t;
        src_off[t] = o < n * n ? ((o / n) + 7 - r) * 15 + (o % n) + 7 - r : -1;
    }
}

__device__ __forceinline__ int pair_cam1(int pair) {
    return pair == 0 ? 0 : pair == 1 ? 2 : pair == 2 ? 0 : pair == 3 ? 1 : pair == 4 ? 0 : 2;
}
__device__ __forceinline__ int pair_cam2(int pair) {
    return pair == 0 ? 1 : pair == 1 ? 3 : pair == 2 ? 2 : pair == 3 ? 3 : pair == 4 ? 3 : 1;
}

constexpr int CORR_UNIT_STRIDE = 296;

template <int COLORS>
__global__ __launch_bounds__(64 * TP_CORR_WG_WAVES) __attribute__((amdgpu_waves_per_eu(TP_CORR_WAVES)))
void k_correlate2D(CorrArgs a) {
    static_assert(8 * CORR_UNIT_STRIDE >= 8 * 225, "staging fits the wavefront's region");
    __shared__ __align__(16) float lds[TP_CORR_WG_WAVES][8 * CORR_UNIT_STRIDE];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int unit = lane >> 3;
    const int li = lane & 7;
    // blocks in list order, not remapped per XCD: the workgroups resident at
    // a time then read and write one stretch of the buffers, 4.5% faster in
    // the frame than a contiguous eighth of the list per XCD
    const int blk = blockIdx.x;
    const int e0 = blk * a.entries_per_block;
    const int e1 = min(a.num_corr, e0 + a.entries_per_block);
    const float fz2 = a.fat_zero * a.fat_zero;
    int src_off[4];
    corr_out_offsets(lane, a.radius, src_off);
    // rows of radius 7 at stride 225 are contiguous: whole-block stores
    const bool dense = a.radius == 7 && a.stride == 225 &&
                       (((uintptr_t)a.corrs) & 15) == 0;

    // the 4 cameras' CLT pointers, loaded once (uniform, scalar loads); a
    // lane picks its pair's two by selection rather than by a dependent
    // per-lane load between the list entry and the CLT rows
    const float *cp0 = gelem(a.clt, 0), *cp1 = gelem(a.clt, 1), *cp2 = gelem(a.clt, 2), *cp3 = gelem(a.clt, 3);
    auto cam_ptr = [&](int c) { return c == 0 ? cp0 : c == 1 ? cp1 : c == 2 ? cp2 : cp3; };
    for (int eb = e0 + wave * 8; eb < e1; eb += 8 * TP_CORR_WG_WAVES) {
        int e = eb + unit;
        bool valid = e < e1;
        if (!valid) e = e1 - 1;
        const int v = a.idx[e];
        int tile = v >> TP_CORR_NTILE_SHIFT;
        int pair = v & ((1 << TP_CORR_NTILE_SHIFT) - 1);
        if (pair >= TP_NUM_PAIRS || tile < 0 || tile >= a.ntiles) {  // skipped entries
            valid = false;
            pair = 0;
            tile = 0;
        }
        const size_t off = (size_t)tile * TP_NUM_COLORS * 256 + li * 8;
        const float *ca = cam_ptr(pair_cam1(pair)) + off, *cb = cam_ptr(pair_cam2(pair)) + off;
        auto load = [&](int color, int which, int qq, int h) -> f32x4 {
            return gptr4((which ? cb : ca) + color * 256)[qq * 16 + h];
        };
        float c0[8], c1[8], c2[8], c3[8];
        corr_accumulate<COLORS>(load, a.scale, li, c0, c1, c2, c3);
        float *W = &lds[wave][0];
        corr_finish(c0, c1, c2, c3, W + CORR_UNIT_STRIDE * unit, W + 225 * unit, li, fz2);
        wave_lds_sync();
        const int vflag = valid ? 1 : 0;
        if (dense && __all(vflag)) {
            // 8 consecutive full 15 x 15 rows, contiguous in memory (7200 B,
            // 16-byte aligned) and in LDS: float4 q of the block
            float *dst = a.corrs + (size_t)eb * 225;
            for (int q = lane; q < 450; q += 64) {
                const f32x4 v = *(const f32x4 *)(W + 4 * q);
                store4(dst + 4 * q, v[0], v[1], v[2], v[3]);
            }
            wave_lds_sync();
            continue;
        }
        for (int u = 0; u < 8; u++) {
            if (!__shfl(vflag, u * 8)) continue;
            const float *src = W + 225 * u;
            float *dst = a.corrs + (size_t)(eb + u) * a.stride;
#pragma unroll
            for (int t = 0; t < 4; t++)
                if (src_off[t] >= 0) store1(dst + lane + 64 * t, src[src_off[t]]);
        }
        wave_lds_sync();
    }
}

}  // namespace tp
