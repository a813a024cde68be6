// k_texture.hpp - texture generation (plain FP32).
//
// Replaces textures_accumulate (src/TileProcessor.cuh:2207-2585) with
// imclt8threads (:3847-4017), debayer_shot (:4019-4144), debayer
// (:4146-4199) and tile_combine_rgba (:4202-4668), and the texture-list
// kernels of generate_RBGA / prepare_texture_list (:1446-1854).
//
// One workgroup (4 wavefronts) per texture tile:
//   1. the 4 cameras x colors units run the inverse CLT of the low-passed
//      CLT (8 lanes per unit); the unfold writes the un-windowed pixels
//      U[cam][color][16][16] to LDS (the window folded into its weights);
//   2. thread = pixel: 3x3 debayer of U, shot-noise companding,
//      window -> D of its own pixel, in registers;
//   3. the same thread: 4-camera consensus (pairwise port weights, best-pair
//      blend, per-port weights, dust removal), colour from U, alpha;
//      non-overlapping tile slices and/or the alpha-faded overlap-add into
//      the RGBA image.
// The list kernels replace CUDA dynamic parallelism: the host enqueues them
// in order and later grids read their sizes from device memory, so the
// whole generate_RBGA sequence is stream-ordered and graph-capturable.
#pragma once

#include <climits>

#include "tp_common.hpp"
#include "k_imclt.hpp"

namespace tp {

struct TexArgs {
    Geo g;
    const int *woi;
    const float *const *clt;
    int n;              // list length (quarter == 0)
    // quarter > 0 (generate_RBGA): idx is one parity's quarter of the texture
    // list, inner entries from its start and border entries at its end, with
    // the device counts num_ptr[0] (inner) and num_ptr[1] (border); both are
    // processed (tiles of one parity never overlap)
    int quarter;
    const int *num_ptr;
    const int *idx;
    const float *ports; // [4][2]
    int colors;
    float min_shot, scale_shot, diff_sigma, diff_threshold, min_agree;
    float w[3];
    int dust_remove, keep_weights;
    // generate_RBGA's first parity pass: the RGBA image was just cleared, so
    // its overlap-add 0 + x is a plain store (no read of the image)
    int first_pass;
    size_t rbga_stride;
    float *rbga;
    size_t tex_stride;
    float *tiles;
};

__device__ __forceinline__ float hw_sq(int i) {  // HWINDOW_SQ, runtime index
    float h = 0.f;
#pragma unroll
    for (int t = 0; t < 8; t++) h = (i == t) ? kTab.hw[t] * kTab.hw[t] : h;
    return h;
}

// alphaFade[m][i][j] (:492-1005): fade(t) = (1 - cos(pi (t + 1/2)/16)) / 2
__device__ __forceinline__ float alpha_fade(int m, int i, int j) {
    auto fade = [](int t) { return 0.5f * (1.0f - cospif((t + 0.5f) * 0.0625f)); };
    switch (m) {
    case 0: return 0.f;
    case 1: return fade(15 - i);
    case 2: return fade(j);
    case 4: return fade(i);
    case 8: return fade(15 - j);
    case 3: return (j > 15 - i) ? fade(j) : fade(15 - i);
    case 6: return (j > i) ? fade(j) : fade(i);
    case 9: return (j > i) ? fade(15 - i) : fade(15 - j);
    case 12: return (i > 15 - j) ? fade(i) : fade(15 - j);
    default: return 1.f;
    }
}

// LDS of the workgroup's TPW texture tiles.  U planes are TEX_MS
// floats apart: the imclt unfold writes columns {0-3, 8-11} (or {4-7,
// 12-15}) of one row in each of 8 planes at once, and the 4-float skew
// spreads those over all 32 banks.
constexpr int TEX_MS = 260;
// Per-launch constants, computed once per workgroup: the pair factors
// -log2(e) / (2 sigma^2 |port_j - port_i|^2) of the 6 port pairs, -log2(e) /
// (2 sigma^2) and 1 / sqrt(min_shot).
enum { TC_PAIR = 0, TC_KSIG = 6, TC_KSHOT = 7, TC_N = 8 };
template <int TPW>
struct TexShared {
    static constexpr int UNIT_WAVES = (TPW * 12 + 7) / 8;  // wavefronts of phase 1 (3 colors)
    float U[TPW * 12 * TEX_MS];  // [tile][cam * colors + color][256], un-windowed
    float cst[TC_N];
    float scratch[UNIT_WAVES][8][1][72];  // phase 1: imclt transposes (one quadrant at a time)
};

// Per-pixel constants of thread t (pixel r = t / 16, c = t % 16), the same
// for every tile: the window weights and the 3 x 3 debayer neighbourhood
// (edge pixels use their own value only).
struct TexPix {
    int r, c;
    bool edge;
    float wsq, wnd2_inv;
    float sqw[3];  // sqrt(w[color]) (the colour weights are not negative)
    // plane offset of the 3 x 3 neighbourhood's corner: pixel t + 16 dr + dc
    // is nb0 + 16 (dr + 1) + dc + 1.  Edge pixels use t itself (their
    // neighbour reads stay inside TexShared and are not used).
    int nb0;
};

__device__ __forceinline__ TexPix tex_pix(int t, const float *w) {
    TexPix q;
#pragma unroll
    for (int c = 0; c < 3; c++) q.sqw[c] = sqrtf(fmaxf(w[c], 0.f));
    q.r = t >> 4;
    q.c = t & 15;
    auto sym = [](int i) { return i < 8 ? i : 15 - i; };
    q.wsq = hw_sq(sym(q.r)) * hw_sq(sym(q.c));
    q.wnd2_inv = 1.0f / q.wsq;
    q.edge = q.r == 0 || q.r == 15 || q.c == 0 || q.c == 15;
    q.nb0 = q.edge ? t : t - 17;
    return q;
}

// TexShared::cst (thread 0; the caller synchronises)
__device__ __forceinline__ void tex_constants(const TexArgs &a, float *cst) {
    constexpr float kLog2e = 1.4426950408889634f;
    const float ksigma = 1.0f / (2.0f * a.diff_sigma * a.diff_sigma);
    int k = 0;
    for (int i = 0; i < 4; i++)
        for (int j = i + 1; j < 4; j++, k++) {
            const float dx = a.ports[2 * j] - a.ports[2 * i], dy = a.ports[2 * j + 1] - a.ports[2 * i + 1];
            cst[TC_PAIR + k] = -kLog2e * (ksigma / (dx * dx + dy * dy));
        }
    cst[TC_KSIG] = -kLog2e * ksigma;
    cst[TC_KSHOT] = 1.0f / sqrtf(a.min_shot);
}

// Phases 2 and 3 of list entry e (code, tile) for thread t = pixel; U: the
// entry's 4 x colors un-windowed planes in LDS.  KEEPW: the per-port
// weights and colour spreads are output (keep_weights).  Without them the
// debayered values only enter weighted colour distances
// sum_col w[col] (x - y)^2, so they are scaled by sqrt(w[col]) once (the
// colour weights are not negative) and the distances are plain sums of
// squares.
template <int COLORS, bool KEEPW>
__device__ __forceinline__ void texture_pixel(const TexArgs &a, int e, int code, int tile,
                                              const float *U, const float *cst, const TexPix &px) {
    constexpr int colors = COLORS;
    const int t = threadIdx.x, r = px.r, c = px.c;
    const float wsq = px.wsq;
    // The overlap-add's read of the RGBA image (written by an earlier pass)
    // is requested here, so that its latency passes under the pixel work.
    const int tcode = code & TP_TASK_TEXTURE_BITS;
    const bool to_rgba = tcode && a.rbga && a.rbga_stride;
    float *dst_rgba = nullptr;
    size_t slice = 0;
    float prev[4] = {0.f, 0.f, 0.f, 0.f};
    if (to_rgba) {
        const int ty = tile / a.g.TX, tx = tile - ty * a.g.TX;
        slice = a.rbga_stride * (size_t)(a.woi[3] + 1) * 8;
        dst_rgba = a.rbga + (size_t)((ty - a.woi[1]) * 8 + r) * a.rbga_stride + (tx - a.woi[0]) * 8 + c;
        if (!a.first_pass)
#pragma unroll
            for (int s = 0; s <= colors; s++) prev[s] = dst_rgba[s * slice];
    }
    // phase 2: debayer_shot (colors > 1): 3 x 3 debayer of the un-windowed
    // pixels (edge pixels unchanged), shot-noise companding, window; else
    // D = windowed U.  Thread t keeps D and U of its own pixel.
    float D[4][3], Uv[4][3];
#pragma unroll
    for (int cam = 0; cam < 4; cam++)
#pragma unroll
        for (int col = colors; col < 3; col++) D[cam][col] = Uv[cam][col] = 0.f;
    if (colors > 1) {
        const float k = cst[TC_KSHOT];
#pragma unroll
        for (int s = 0; s < 4 * colors; s++) {
            const float *x = U + s * TEX_MS;
            const int col = s % colors;
            float v[9];
#pragma unroll
            for (int j = 0; j < 9; j++) v[j] = j == 4 ? x[t] : x[px.nb0 + 16 * (j / 3) + j % 3];
            const float cross = (v[1] + v[7]) + (v[3] + v[5]);
            float d;
            if (col < 2) {
                const float diag = (v[0] + v[2]) + (v[6] + v[8]);
                d = fmaf(0.0625f, diag, fmaf(0.125f, cross, 0.25f * v[4]));
            } else {
                d = fmaf(0.125f, cross, 0.5f * v[4]);
            }
            d = px.edge ? v[4] : d;
            if (a.scale_shot > 0.0f) d = a.scale_shot * (d > a.min_shot ? __builtin_amdgcn_sqrtf(d) : k * d);
            D[s / colors][col] = d * (KEEPW ? wsq : px.sqw[col]);
            Uv[s / colors][col] = v[4];
        }
    } else {
#pragma unroll
        for (int s = 0; s < 4; s++) {
            Uv[s][0] = U[s * TEX_MS + t];
            D[s][0] = Uv[s][0] * (KEEPW ? wsq : px.sqw[0]);
        }
    }

    // phase 3: consensus per pixel
    float out[12];
#pragma unroll
    for (int i = 0; i < 12; i++) out[i] = 0.f;
    float *rgba = out, *pw = out + colors + 1, *crms = out + colors + 5;
    if (KEEPW) {
        float sw = 0.f;
        for (int col = 0; col < colors; col++) {
            // Spread of the 4 cameras around their mean, two-pass: the same
            // value as (s0 s2 - s1^2) / s0^2 (src/TileProcessor.cuh:4286-4296)
            // without that form's cancellation when the cameras nearly agree.
            const float mean = 0.25f * ((D[0][col] + D[1][col]) + (D[2][col] + D[3][col]));
            float ss = 0.f;
#pragma unroll
            for (int cam = 0; cam < 4; cam++) {
                const float e = D[cam][col] - mean;
                ss = fmaf(e, e, ss);
            }
            const float mse = 0.25f * ss;
            crms[col] = sqrtf(mse);
            sw = fmaf(a.w[col], mse, sw);
        }
        crms[colors] = sqrtf(sw);
    }
    // The colour distances below are sum_col w[col] ((x - y) / window^2)^2;
    // with KEEPW, D is divided by window^2 once (else it holds
    // sqrt(w[col]) x already).
    if (KEEPW) {
#pragma unroll
        for (int col = 0; col < 3; col++)
#pragma unroll
            for (int cam = 0; cam < 4; cam++) D[cam][col] *= px.wnd2_inv;
    }
    auto wdist = [&](float dc, float acc, int col) {
        return KEEPW ? fmaf(a.w[col] * dc, dc, acc) : fmaf(dc, dc, acc);
    };
    float p[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0, k = 0; i < 4; i++)
#pragma unroll
        for (int j = i + 1; j < 4; j++, k++) {
            float d = 0.f;
#pragma unroll
            for (int col = 0; col < colors; col++) {
                d = wdist(D[i][col] - D[j][col], d, col);
            }
            d = __builtin_amdgcn_exp2f(cst[TC_PAIR + k] * d) + (float)TP_FAT_ZERO_WEIGHT;
            p[i] += d;
            p[j] += d;
        }
    int b1 = 0;
    float best = p[0];
#pragma unroll
    for (int cam = 1; cam < 4; cam++)
        if (p[cam] > best) { b1 = cam; best = p[cam]; }
    int b2 = b1 == 0 ? 1 : 0;
    best = p[b2];
#pragma unroll
    for (int cam = 1; cam < 4; cam++)
        if (cam > (b1 == 0 ? 1 : 0) && cam != b1 && p[cam] > best) { b2 = cam; best = p[cam]; }
    const float pb1 = b1 == 0 ? p[0] : b1 == 1 ? p[1] : b1 == 2 ? p[2] : p[3];
    const float pb2 = b2 == 0 ? p[0] : b2 == 1 ? p[1] : b2 == 2 ? p[2] : p[3];
    const float w1 = pb1 * __builtin_amdgcn_rcpf(pb1 + pb2), w2 = 1.f - w1;
    float blend[3];
#pragma unroll
    for (int col = 0; col < 3; col++) {
        float d1 = 0.f, d2 = 0.f;
#pragma unroll
        for (int cam = 0; cam < 4; cam++) {
            d1 = cam == b1 ? D[cam][col] : d1;
            d2 = cam == b2 ? D[cam][col] : d2;
        }
        blend[col] = fmaf(w1, d1, w2 * d2);
    }
#pragma unroll
    for (int cam = 0; cam < 4; cam++) {
        float d2 = 0.f;
#pragma unroll
        for (int col = 0; col < colors; col++) {
            d2 = wdist(D[cam][col] - blend[col], d2, col);
        }
        p[cam] = __builtin_amdgcn_exp2f(cst[TC_KSIG] * d2) + (float)TP_FAT_ZERO_WEIGHT;
    }
    if (a.dust_remove) {
        int worst = 0;
        float wv = p[0];
#pragma unroll
        for (int cam = 1; cam < 4; cam++)
            if (p[cam] < wv) { worst = cam; wv = p[cam]; }
        const float avg = ((p[0] + p[1]) + (p[2] + p[3]) - wv) * (1.0f / 3.0f);
        const float ia = __builtin_amdgcn_rcpf(avg);
        const float scale = 1.f + wv * (avg - wv) * (ia * ia) * (1.0f / 3.0f);
        const float fworst = wv * ia;
#pragma unroll
        for (int cam = 0; cam < 4; cam++) p[cam] *= (cam == worst) ? fworst : scale;
    }
    const float psum = (p[0] + p[1]) + (p[2] + p[3]);
    const float kn = wsq * __builtin_amdgcn_rcpf(psum);  // the window of U folded in
    for (int col = 0; col < colors; col++) {
        float acc = 0.f;
#pragma unroll
        for (int cam = 0; cam < 4; cam++) acc = fmaf(p[cam], Uv[cam][col], acc);
        rgba[col] = kn * acc;
    }
    out[colors] = wsq * psum * 0.25f;
#pragma unroll
    for (int cam = 0; cam < 4; cam++) pw[cam] = p[cam];
    const int nsl_all = 2 * colors + 6;
    if (a.tiles && a.tex_stride) {
        const int nsl = KEEPW ? nsl_all : colors + 1;
        float *dst = a.tiles + (size_t)e * a.tex_stride + t;
        for (int s = 0; s < nsl; s++) dst[s * 256] = out[s];
    }
    if (!to_rgba) return;
    const float f = tcode == TP_TASK_TEXTURE_BITS ? 1.f : alpha_fade(tcode, r, c);
#pragma unroll
    for (int s = 0; s <= colors; s++) dst_rgba[s * slice] = a.first_pass ? out[s] * f : prev[s] + out[s] * f;
}

// The list entries of a launch: n in all; in a parity quarter the first
// n_in are the inner entries from the quarter's start, the others the
// border entries at its end (device counts, capped at the quarter).
struct TexList {
    int n, n_in, n_bd;
};
__device__ __forceinline__ TexList tex_list(const TexArgs &a) {
    TexList l{a.n, a.n, 0};
    if (a.quarter) {
        l.n_in = min(a.num_ptr[0], a.quarter);
        l.n_bd = min(a.num_ptr[1], a.quarter - l.n_in);
        l.n = l.n_in + l.n_bd;
    }
    return l;
}
// Position in a.idx of entry e < l.n
__device__ __forceinline__ int entry_pos(const TexArgs &a, const TexList &l, int e) {
    if (!a.quarter) return e;
    return e < l.n_in ? e : a.quarter - l.n_bd + (e - l.n_in);
}
// List words of entries e0 .. e0 + TPW - 1 (0 past the end): lane k < TPW
// loads entry e0 + k's with a vector load, which the caller reads with
// list_word once it is needed (so a wait for it is not also a wait for LDS
// operations, as a scalar load's would be).
template <int TPW>
__device__ __forceinline__ int list_words(const TexArgs &a, const TexList &l, int e0) {
    const int e = e0 + min((int)__lane_id(), TPW - 1);
    return e < l.n ? a.idx[entry_pos(a, l, e)] : 0;
}
__device__ __forceinline__ int list_word(int words, int k) { return __builtin_amdgcn_readlane(words, k); }

// List entries e0 .. e0 + TPW - 1 (of n) by the whole workgroup:
// phase 1 runs the units of all its tiles at once (8 lanes each), then the
// threads take the tiles' pixels one tile after the other.  Entries without
// a texture bit, off the grid or past n are skipped (uniformly).
template <int COLORS, int TPW, bool KEEPW>
__device__ __forceinline__ void texture_tiles(const TexArgs &a, int e0, int words, TexShared<TPW> &sh,
                                              const TexPix &px) {
    constexpr int colors = COLORS, UPT = 4 * COLORS;  // units per tile
    int code[TPW], tile[TPW];
    bool ok[TPW], any = false;
#pragma unroll
    for (int k = 0; k < TPW; k++) {
        code[k] = list_word(words, k);
        tile[k] = code[k] >> TP_CORR_NTILE_SHIFT;
        ok[k] = (code[k] & (1 << TP_LIST_TEXTURE_BIT)) && tile[k] >= 0 && tile[k] < a.g.TX * a.g.TY;
        any = any || ok[k];
    }
    if (!any) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int slot = wave * 8 + (lane >> 3), li = lane & 7;

    // phase 1: U = imclt(CLT * lpf), un-windowed
    if (slot < TPW * UPT) {
        const int k = slot / UPT, u = slot - k * UPT;
        const int cam = u / colors, col = u % colors;
        const int lw = colors > 1 ? col : 3;
        bool okk = ok[0];
        int tk = tile[0];
#pragma unroll
        for (int j = 1; j < TPW; j++) {
            okk = k == j ? ok[j] : okk;
            tk = k == j ? tile[j] : tk;
        }
        if (okk) {
            float q0[8], q1[8], q2[8], q3[8];
            load_rows(a.clt[cam] + ((size_t)tk * TP_NUM_COLORS + col) * 256, li, q0, q1, q2, q3);
#pragma unroll
            for (int kk = 0; kk < 8; kk++) {
                const float l = c_lpf.v[lw][li * 8 + kk];
                q0[kk] *= l; q1[kk] *= l; q2[kk] *= l; q3[kk] *= l;
            }
            imclt_transform<1>(q0, q1, q2, q3, &sh.scratch[wave][lane >> 3][0][0], li);
            float *m = sh.U + slot * TEX_MS;
            imclt_unfold<true>(q0, q1, q2, q3, li, [&](int r, int c, float w, float d) { m[r * 16 + c] = w * d; });
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TPW; k++)
        if (ok[k]) texture_pixel<COLORS, KEEPW>(a, e0 + k, code[k], tile[k], sh.U + k * UPT * TEX_MS, sh.cst, px);
}

// A fixed grid of workgroups loops over the list (its length may only be
// known on the device), TPW entries at a time.
template <int COLORS, int TPW, bool KEEPW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TP_TEX_WAVES)))
void k_textures_accumulate(TexArgs a) {
    __shared__ __align__(16) TexShared<TPW> sh;
    const TexList l = tex_list(a);
    const TexPix px = tex_pix(threadIdx.x, a.w);
    if (threadIdx.x == 0) tex_constants(a, sh.cst);
    __syncthreads();
    // the next entries' list words are requested a round ahead: the CLT
    // loads that start each round then wait for one memory latency, not two
    const int step = gridDim.x * TPW;
    int words = list_words<TPW>(a, l, blockIdx.x * TPW);
    for (int e = blockIdx.x * TPW; e < l.n; e += step) {
        const int cur = words;
        words = list_words<TPW>(a, l, e + step);
        texture_tiles<COLORS, TPW, KEEPW>(a, e, cur, sh, px);
        __syncthreads();  // the LDS of these entries is rewritten by the next ones
    }
}

// ---------------- texture list (generate_RBGA / prepare_texture_list) ----------------

// zero the tile occupancy map (kept in the first tilesx*height entries of
// the index buffer, like the reference) and reset woi / counters
__global__ void k_tex_clear(int *indices, int width, int height, int TX, int TY, int *woi, int *num) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int w = min(width, TX);
    if (i < w * height) indices[(i / w) * TX + (i % w)] = 0;
    if (i == 0) {
        woi[0] = TX; woi[1] = TY; woi[2] = 0; woi[3] = 0;
    }
    if (i < 8) num[i] = 0;
}

__device__ __forceinline__ bool task_xy(const tp_task &tk, int TX, int TY, int &x, int &y) {
    x = tk.txy & 0xffff;
    y = tk.txy >> 16;
    return x < TX && y >= 0 && y < TY;
}

__global__ void k_tex_mark(const tp_task *tasks, int n, int *indices, int TX, int TY) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int x, y;
    if (!(tasks[t].task & TP_TASK_TEXTURE_BITS) || !task_xy(tasks[t], TX, TY, x, y)) return;
    indices[x + y * TX] = 1;
}

// min / max over the wavefront's active lanes (inactive lanes pass the
// identity), then one atomic per wavefront
__device__ __forceinline__ int wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

// Launched with TP_NEIGHBOR_THREADS threads per block: the window of
// interest is reduced per wavefront, then per block through LDS, so one
// block makes one set of 4 same-address atomics (these serialise in L2).
#define TP_NEIGHBOR_THREADS 1024
__global__ __launch_bounds__(TP_NEIGHBOR_THREADS) void k_tex_neighbors(tp_task *tasks, int n,
                                                                       const int *indices, int *woi,
                                                                       int TX, int TY) {
    __shared__ int red[4][TP_NEIGHBOR_THREADS / 64];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int task = t < n ? tasks[t].task : 0;
    int x = 0, y = 0;
    const bool act = t < n && (task & TP_TASK_TEXTURE_BITS) && task_xy(tasks[t], TX, TY, x, y);
    const int x0 = wave_min(act ? x : INT_MAX), y0 = wave_min(act ? y : INT_MAX);
    const int x1 = wave_max(act ? x : INT_MIN), y1 = wave_max(act ? y : INT_MIN);
    const int wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (__lane_id() == 0) {
        red[0][wv] = x0;
        red[1][wv] = y0;
        red[2][wv] = x1;
        red[3][wv] = y1;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int l = threadIdx.x;
        const int bx0 = wave_min(l < nw ? red[0][l] : INT_MAX), by0 = wave_min(l < nw ? red[1][l] : INT_MAX);
        const int bx1 = wave_max(l < nw ? red[2][l] : INT_MIN), by1 = wave_max(l < nw ? red[3][l] : INT_MIN);
        if (l == 0 && bx0 != INT_MAX) {
            atomicMin(woi + 0, bx0);
            atomicMin(woi + 1, by0);
            atomicMax(woi + 2, bx1);
            atomicMax(woi + 3, by1);
        }
    }
    if (!act) return;
    int d = 0;
    if (y > 0 && indices[x + (y - 1) * TX]) d |= 1;
    if (x < TX - 1 && indices[x + 1 + y * TX]) d |= 2;
    if (y < TY - 1 && indices[x + (y + 1) * TX]) d |= 4;
    if (x > 0 && indices[x - 1 + y * TX]) d |= 8;
    tasks[t].task = ((task ^ d) & TP_TASK_TEXTURE_BITS) ^ task;
}

// Launched with TP_NEIGHBOR_THREADS threads per block.  Positions in the
// 8 sub-lists: lane rank within the wavefront (ballot), wavefront offset
// within the block (LDS prefix), block offset from one atomic per
// (block, sub-list); tiles of a block take consecutive positions in thread
// order.
__global__ __launch_bounds__(TP_NEIGHBOR_THREADS) void k_tex_gen_list(const tp_task *tasks, int n,
                                                                      int *indices, int *num,
                                                                      const int *woi, int TX, int TY,
                                                                      int TYA) {
    __shared__ int wcount[TP_NEIGHBOR_THREADS / 64][8];
    __shared__ int bbase[8];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int task = t < n ? (tasks[t].task & TP_TASK_TEXTURE_BITS) : 0;
    int x = 0, y = 0;
    const bool act = task && task_xy(tasks[t], TX, TY, x, y);
    const int quarter = TX * (TYA >> 2);
    const bool border = act && ((x == woi[0]) || (y == woi[1]) || (x == TX - 1) || (y == woi[3]));
    int head = 0, slot = 0;
    if (x & 1) { head += quarter; slot += 2; }
    if (y & 1) { head += 2 * quarter; slot += 4; }
    if (border) { head += quarter - 1; slot += 1; }
    const int lane = __lane_id(), wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int rank = 0;
#pragma unroll
    for (int sl = 0; sl < 8; sl++) {
        const unsigned long long m = __ballot(act && slot == sl);
        if (act && slot == sl) rank = __popcll(m & ((1ull << lane) - 1));
        if (lane == 0) wcount[wv][sl] = __popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < 8) {  // exclusive prefix over wavefronts, one atomic per sub-list
        const int sl = threadIdx.x;
        int acc = 0;
        for (int w = 0; w < nw; w++) {
            const int c = wcount[w][sl];
            wcount[w][sl] = acc;
            acc += c;
        }
        bbase[sl] = acc ? atomicAdd(num + sl, acc) : 0;
    }
    __syncthreads();
    if (!act) return;
    const int pos = bbase[slot] + wcount[wv][slot] + rank;
    if (pos >= quarter) return;  // only with duplicate tiles in the task list
    indices[head + (border ? -pos : pos)] = task | ((x + y * TX) << TP_CORR_NTILE_SHIFT) | (1 << TP_LIST_TEXTURE_BIT);
}

__global__ void k_tex_finish_woi(int *woi) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        woi[2] += 1 - woi[0];
        woi[3] += 1 - woi[1];
    }
}

// zero the RGBA image rows of the woi: (woi_w + 1) * 8 columns,
// (woi_h + 1) * 8 rows per slice (slice-major rows).  One workgroup per
// TP_CLEAR_ROWS rows, grid for the largest possible woi; VEC: 16-byte
// stores (the host checks that the base and the stride allow them).
#define TP_CLEAR_ROWS 8
template <bool VEC>
__global__ __launch_bounds__(256) void k_tex_clear_rbga(float *rbga, size_t stride, const int *woi,
                                                        int slices) {
    const int wpx = (woi[2] + 1) * 8, hpx = (woi[3] + 1) * 8;
    const int row0 = blockIdx.x * TP_CLEAR_ROWS;
    const int nrow = min(TP_CLEAR_ROWS, hpx * slices - row0);
    if (nrow <= 0) return;
    const int q = VEC ? wpx / 4 : wpx;  // wpx is a multiple of 8
    for (int r = 0; r < nrow; r++) {
        float *p = rbga + (size_t)(row0 + r) * stride;
        for (int x = threadIdx.x; x < q; x += 256) {
            if (VEC)
                ((float4 *)p)[x] = make_float4(0.f, 0.f, 0.f, 0.f);
            else
                p[x] = 0.f;
        }
    }
}

}  // namespace tp
