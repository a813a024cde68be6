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

// A workgroup scans TP_COMBINE_EPW consecutive list entries; each group of
// entries of one tile that starts there is combined by one wavefront.  The
// entry words (the workgroup's and the 15 after them, which a group
// starting near the end may reach) are kept in LDS, so finding a group's end and its pairs
// costs no global loads.  The group's pair tiles (consecutive rows of the
// correlation output) are copied into LDS with coalesced loads, 6 entries at
// a time (all of a chunk's loads issued before any is used), then each lane
// sums its outputs (lane + 64 t) over the mapped pair tiles.  The source
// offset of each of a lane's outputs in each of the 4 mappings is computed
// once; an output with no source reads a zero kept after the staged rows.
__global__ __launch_bounds__(256) void k_corr_combine(CombineArgs a) {
    constexpr int CHUNK = 6;   // entries staged at a time (the 6 pairs of a full task)
    constexpr int MAXG = 16;   // entries of one tile considered
    constexpr int ZS = CHUNK * 225;  // the zero after the staged rows
    constexpr int EPW = TP_COMBINE_EPW;  // list entries scanned per workgroup
    __shared__ int heads[EPW];
    __shared__ int ent[EPW + MAXG - 1];
    __shared__ int nheads;
    __shared__ float stage[4][ZS + 2];
    const int e0 = blockIdx.x * EPW;
    const int nent = min(EPW + MAXG - 1, a.num_corr - e0);  // entry words held
    if (threadIdx.x == 0) nheads = 0;
    for (int j = threadIdx.x; j < nent; j += 256) ent[j] = a.idx[e0 + j];
    const int prev = e0 > 0 ? a.idx[e0 - 1] >> TP_CORR_NTILE_SHIFT : 0;
    __syncthreads();
    for (int j = threadIdx.x; j < EPW && e0 + j < a.num_corr; j += 256) {
        const int t = ent[j] >> TP_CORR_NTILE_SHIFT;
        const int tp = j ? ent[j - 1] >> TP_CORR_NTILE_SHIFT : prev;
        if (e0 + j == 0 || tp != t) heads[atomicAdd(&nheads, 1)] = j;
    }
    __syncthreads();
    const int nh = nheads;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = a.radius, n = 2 * r + 1, nn = n * n;
    float *L = stage[wave];
    if (lane == 0) L[ZS] = 0.f;  // ordered before its first read by the chunk loop's LDS sync
    // offset in a pair tile of output lane + 64 t under mapping c (pairs
    // 0-1, 2-3, 4, 5), or -1, as 16-bit fields: mappings 0, 1 in src01[t],
    // 2, 3 in src23[t] (registers, selected by the wave-uniform mapping)
    int src01[4], src23[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int o = lane + 64 * t;
        const int cy = o / n - r, cx = o % n - r;
        int f[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            int sy, sx;
            const bool ok = o < nn && pair_src(c < 2 ? 2 * c : c + 2, r, cy, cx, sy, sx);
            f[c] = ok ? (sy + r) * n + sx + r : -1;
        }
        src01[t] = (f[0] & 0xffff) | (f[1] << 16);
        src23[t] = (f[2] & 0xffff) | (f[3] << 16);
    }
    // lane's outputs summed over k < ne staged pair tiles of L
    auto accumulate = [&](int c0, int ne, float (&s)[4]) {
        for (int k = 0; k < ne; k++) {
            const int pair = ent[c0 + k] & ((1 << TP_CORR_NTILE_SHIFT) - 1);
            if (pair >= TP_NUM_PAIRS || !((a.pair_mask >> pair) & 1)) continue;
            const int c = pair < 2 ? 0 : pair < 4 ? 1 : pair - 2;
            const unsigned sh = (c & 1) * 16;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int m = __builtin_amdgcn_sbfe(c < 2 ? src01[t] : src23[t], sh, 16);
                s[t] += L[m < 0 ? ZS : k * nn + m];
            }
        }
    };
    if (a.stride == (size_t)nn) {
        // Contiguous rows, software-pipelined: the first chunk of the
        // wavefront's next group is requested before the current group is
        // summed.  Later chunks of a group (more than CHUNK entries, only
        // in irregular lists) are copied synchronously.
        constexpr int PER_LANE = (CHUNK * 225 + 63) / 64;
        // next valid group at or after h (step 4): start j0, end je, tile
        auto next_group = [&](int &h, int &j0, int &je, int &tile) {
            for (; h < nh; h += 4) {
                j0 = heads[h];
                tile = ent[j0] >> TP_CORR_NTILE_SHIFT;
                if (tile < 0 || tile >= a.ntiles) continue;
                je = j0;
                while (je + 1 < nent && je + 1 < j0 + MAXG && (ent[je + 1] >> TP_CORR_NTILE_SHIFT) == tile) je++;
                return true;
            }
            return false;
        };
        // all loads of a chunk issued at once (addresses clamped into it)
        auto issue = [&](int c0, int cnt, float (&v)[PER_LANE]) {
            gfloat *sp = gptr(a.corrs + (size_t)(e0 + c0) * nn);
#pragma unroll
            for (int i = 0; i < PER_LANE; i++) v[i] = sp[min(lane + 64 * i, cnt - 1)];
        };
        float v[PER_LANE];
        int h = wave, j0 = 0, je = 0, tile = 0;
        bool have = next_group(h, j0, je, tile);
        if (have) issue(j0, min(CHUNK, je + 1 - j0) * nn, v);
        while (have) {
            const int ne0 = min(CHUNK, je + 1 - j0);
#pragma unroll
            for (int i = 0; i < PER_LANE; i++)
                if (lane + 64 * i < ne0 * nn) L[lane + 64 * i] = v[i];
            wave_lds_sync();
            int h2 = h + 4, j0n = 0, jen = 0, tilen = 0;
            const bool haven = next_group(h2, j0n, jen, tilen);
            if (haven) issue(j0n, min(CHUNK, jen + 1 - j0n) * nn, v);
            float s[4] = {0.f, 0.f, 0.f, 0.f};
            accumulate(j0, ne0, s);
            wave_lds_sync();  // L is rewritten next
            for (int c0 = j0 + CHUNK; c0 <= je; c0 += CHUNK) {
                const int cnt = min(CHUNK, je + 1 - c0) * nn;
                gfloat *sp = gptr(a.corrs + (size_t)(e0 + c0) * nn);
                for (int o = lane; o < cnt; o += 64) L[o] = sp[o];
                wave_lds_sync();
                accumulate(c0, min(CHUNK, je + 1 - c0), s);
                wave_lds_sync();
            }
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int o = lane + 64 * t;
                if (o < nn) a.out[(size_t)tile * a.out_stride + o] = s[t];
            }
            h = h2;
            j0 = j0n;
            je = jen;
            tile = tilen;
            have = haven;
        }
        return;
    }
    for (int h = wave; h < nh; h += 4) {
        const int j0 = heads[h];  // group start, relative to e0
        const int tile = ent[j0] >> TP_CORR_NTILE_SHIFT;
        if (tile < 0 || tile >= a.ntiles) continue;
        int je = j0;
        while (je + 1 < nent && je + 1 < j0 + MAXG && (ent[je + 1] >> TP_CORR_NTILE_SHIFT) == tile) je++;
        float s[4] = {0.f, 0.f, 0.f, 0.f};
        for (int c0 = j0; c0 <= je; c0 += CHUNK) {
            const int ne = min(CHUNK, je + 1 - c0);
            // copy rows e0 + c0 .. (nn floats each) into L[k][nn]
            if (a.stride == (size_t)nn) {  // contiguous rows: one block
                // all loads issued before the first LDS store (addresses
                // clamped into the block, so none is conditional)
                constexpr int PER_LANE = (CHUNK * 225 + 63) / 64;
                gfloat *src = gptr(a.corrs + (size_t)(e0 + c0) * nn);
                const int cnt = ne * nn;
                float v[PER_LANE];
#pragma unroll
                for (int i = 0; i < PER_LANE; i++) v[i] = src[min(lane + 64 * i, cnt - 1)];
#pragma unroll
                for (int i = 0; i < PER_LANE; i++)
                    if (lane + 64 * i < cnt) L[lane + 64 * i] = v[i];
            } else {
                for (int k = 0; k < ne; k++)
                    for (int o = lane; o < nn; o += 64) L[k * nn + o] = a.corrs[(size_t)(e0 + c0 + k) * a.stride + o];
            }
            wave_lds_sync();
            accumulate(c0, ne, s);
            wave_lds_sync();  // L is rewritten by the next chunk
        }
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int o = lane + 64 * t;
            if (o < nn) a.out[(size_t)tile * a.out_stride + o] = s[t];
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
