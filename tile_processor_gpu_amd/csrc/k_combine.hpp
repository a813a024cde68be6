// k_combine.hpp - six-pair combination and sub-pixel disparity (plain FP32).
//
// Not in the reference (its Java host did this; the reference stops at the
// per-pair tiles of correlate2D, src/TileProcessor.cuh:1421-1430).  Each
// pair's (2r+1)^2 tile is mapped onto the horizontal disparity axis and the
// mapped tiles of a tile's pairs are summed:
//   pairs 0 (0,1), 1 (2,3): horizontal baseline, as is
//   pairs 2 (0,2), 3 (1,3): vertical baseline, transposed
//   pair 4 (0,3): diagonal (+1,+1), out(y, x) = in(x + y, x - y)
//   pair 5 (2,1): diagonal (+1,-1), out(y, x) = in(y - x, x + y)
// (offsets from the tile centre), so a residual disparity dd peaks at
// (0, +dd) for every pair.
#pragma once

#include "tp_common.hpp"

namespace tp {

__device__ __forceinline__ bool pair_src(int pair, int r, int cy, int cx, int &sy, int &sx) {
    if (pair <= 1) { sy = cy; sx = cx; }
    else if (pair <= 3) { sy = cx; sx = cy; }
    else if (pair == 4) { sy = cx + cy; sx = cx - cy; }
    else { sy = cy - cx; sx = cx + cy; }
    return sy >= -r && sy <= r && sx >= -r && sx <= r;
}

struct CombineArgs {
    int num_corr;
    const int *idx;
    const float *corrs;
    size_t stride;
    int radius;
    int pair_mask;
    int ntiles;
    float *out;
    size_t out_stride;
};

// a workgroup scans 256 consecutive list entries; each group of entries of
// one tile that starts there is combined by one wavefront (lanes = outputs)
__global__ __launch_bounds__(256) void k_corr_combine(CombineArgs a) {
    __shared__ int heads[256];
    __shared__ int nheads;
    const int e0 = blockIdx.x * 256;
    const int e = e0 + threadIdx.x;
    if (threadIdx.x == 0) nheads = 0;
    __syncthreads();
    if (e < a.num_corr) {
        const int t = a.idx[e] >> TP_CORR_NTILE_SHIFT;
        if (e == 0 || (a.idx[e - 1] >> TP_CORR_NTILE_SHIFT) != t) heads[atomicAdd(&nheads, 1)] = e;
    }
    __syncthreads();
    const int nh = nheads;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = a.radius, n = 2 * r + 1;
    for (int h = wave; h < nh; h += 4) {
        const int g0 = heads[h];
        const int tile = a.idx[g0] >> TP_CORR_NTILE_SHIFT;
        if (tile < 0 || tile >= a.ntiles) continue;
        int ge = g0;
        while (ge + 1 < a.num_corr && ge + 1 < g0 + 16 && (a.idx[ge + 1] >> TP_CORR_NTILE_SHIFT) == tile) ge++;
        for (int o = lane; o < n * n; o += 64) {
            const int cy = o / n - r, cx = o % n - r;
            float s = 0.f;
            for (int k = g0; k <= ge; k++) {
                const int pair = a.idx[k] & ((1 << TP_CORR_NTILE_SHIFT) - 1);
                if (pair >= TP_NUM_PAIRS || !((a.pair_mask >> pair) & 1)) continue;
                int sy, sx;
                if (pair_src(pair, r, cy, cx, sy, sx)) s += a.corrs[(size_t)k * a.stride + (sy + r) * n + sx + r];
            }
            a.out[(size_t)tile * a.out_stride + o] = s;
        }
    }
}

// disparity from the centre row: argmax, then parabola (mode 0) or the
// centroid of the positive samples within +-2 (mode 1).  out[i] = (d, peak)
__global__ void k_corr_disparity(const float *tiles, size_t stride, int n_tiles, int radius, int mode,
                                 float *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_tiles) return;
    const int n = 2 * radius + 1;
    const float *row = tiles + (size_t)i * stride + radius * n;
    int best = 0;
    float bv = row[0];
    for (int x = 1; x < n; x++)
        if (row[x] > bv) { best = x; bv = row[x]; }
    float d = (float)(best - radius);
    if (mode == 0) {
        if (best > 0 && best < n - 1) {
            const float a0 = row[best - 1], c0 = row[best + 1];
            const float den = a0 - 2.f * bv + c0;
            if (den < 0.f) d += 0.5f * (a0 - c0) / den;
        }
    } else {
        float sw = 0.f, sx = 0.f;
        for (int x = max(0, best - 2); x <= min(n - 1, best + 2); x++) {
            const float w = fmaxf(row[x], 0.f);
            sw += w;
            sx = fmaf(w, (float)(x - radius), sx);
        }
        if (sw > 0.f) d = sx / sw;
    }
    out[2 * i] = d;
    out[2 * i + 1] = bv;
}

}  // namespace tp
