// k_nn.hpp - disparity-refinement CNN: input assembly, forward, backward,
// Adam.  Not in the reference (its Java host took a parabolic/centroid
// argmax of the correlation); this network replaces that sub-pixel step.
//
// Network (per tile, 15 x 15 pixels, zero padding):
//   x  [7][225]   combined correlation + the 6 pair tiles mapped onto the
//                 horizontal disparity axis (as tp_corr_combine maps them)
//   h1 = relu(conv3x3(x; W1 [32][7][9], b1))
//   h2 = relu(conv3x3(h1; W2 [32][32][9], b2))
//   z  = conv3x3(h2; W3 [2][32][9], b3)           z0: logits, z1: confidence map
//   p  = softmax(z0) over the 225 pixels
//   disparity  = sum p (x - 7)                    (x = column, residual in px)
//   confidence = sigmoid(sum p z1)
// Parameters are BF16 (one flat array, layout NN_OFF_*), arithmetic FP32.
//
// Training loss per tile (mean over the tiles with a finite target):
//   huber_delta(disparity - target) + conf_weight * BCE(confidence, |err| < 0.5)
// with the BCE target held constant.
//
// One workgroup of 256 threads handles one tile at a time (thread = pixel,
// threads 225..255 only help with loads and reductions).  Activations and
// gradients of a tile live in LDS (<= 64 KB).  The backward pass keeps each
// workgroup's weight-gradient sums in registers across its tiles and adds
// them to the global gradient once at the end.
#pragma once

#include "tp_common.hpp"
#include "k_combine.hpp"

namespace tp {

constexpr int NN_C = 32, NN_IN = 7, NN_P = 225;
constexpr int NN_OFF_W1 = 0, NN_OFF_B1 = 2016, NN_OFF_W2 = 2048, NN_OFF_B2 = 11264,
              NN_OFF_W3 = 11296, NN_OFF_B3 = 11872, NN_NPARAMS = 11874;
constexpr int NN_ACTS = 2 * NN_C * NN_P + 2 * NN_P;  // h1, h2, z per tile

__device__ __forceinline__ float bf16f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ uint16_t f2bf16(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// offset of neighbour r (3x3, row-major, r = 4 is the centre)
__host__ __device__ constexpr int nbr_off(int r) { return (r / 3 - 1) * 15 + (r % 3 - 1); }

// bit r set when neighbour r of pixel (y, x) lies inside the 15 x 15 tile
__device__ __forceinline__ unsigned nbr_mask_yx(int y, int x) {
    unsigned m = 0;
#pragma unroll
    for (int r = 0; r < 9; r++) {
        const int yy = y + r / 3 - 1, xx = x + r % 3 - 1;
        if (yy >= 0 && yy < 15 && xx >= 0 && xx < 15) m |= 1u << r;
    }
    return m;
}

// value of channel plane `in` at neighbour r of pixel p (0 outside)
__device__ __forceinline__ float nbr(const float *in, int p, unsigned m, int r) {
    const int q = min(NN_P - 1, max(0, p + nbr_off(r)));
    const float v = in[q];
    return ((m >> r) & 1u) ? v : 0.f;
}

// acc[co] += sum_ci sum_r w[co][ci][r] in[ci][p + off_r]
template <int CIN, int COUT>
__device__ __forceinline__ void conv_px(const float *in, int p, unsigned m,
                                        const uint16_t *__restrict__ w, float *acc) {
    for (int ci = 0; ci < CIN; ci++) {
        float v[9];
#pragma unroll
        for (int r = 0; r < 9; r++) v[r] = nbr(in + ci * NN_P, p, m, r);
#pragma unroll
        for (int co = 0; co < COUT; co++)
#pragma unroll
            for (int r = 0; r < 9; r++) acc[co] = fmaf(bf16f(w[(co * CIN + ci) * 9 + r]), v[r], acc[co]);
    }
}

// block-wide sums of up to 3 values (256 threads)
__device__ __forceinline__ void block_sum3(float &a, float &b, float &c, float *red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
        c += __shfl_xor(c, o);
    }
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red[w] = a;
        red[4 + w] = b;
        red[8 + w] = c;
    }
    __syncthreads();
    a = red[0] + red[1] + red[2] + red[3];
    b = red[4] + red[5] + red[6] + red[7];
    c = red[8] + red[9] + red[10] + red[11];
}

__device__ __forceinline__ float block_max(float a, float *red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a = fmaxf(a, __shfl_xor(a, o));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[12 + (threadIdx.x >> 6)] = a;
    __syncthreads();
    return fmaxf(fmaxf(red[12], red[13]), fmaxf(red[14], red[15]));
}

// Softmax statistics of the head for pixel p: returns p_q (probability of
// this pixel), the disparity and the confidence logit.
__device__ __forceinline__ void head_stats(float z0, float z1, bool px, int xcol, float *red,
                                           float &prob, float &disp, float &cl) {
    const float mx = block_max(px ? z0 : -INFINITY, red);
    const float e = px ? __expf(z0 - mx) : 0.f;
    float s = e, sx = e * (float)(xcol - 7), sc = e * z1;
    block_sum3(s, sx, sc, red);
    const float inv = 1.0f / s;
    prob = e * inv;
    disp = sx * inv;
    cl = sc * inv;
}

// ---------------------------------------------------------------- input
// x[t][0] = sum of the mapped pair tiles, x[t][1 + p] = mapped tile of pair
// p (zero when the pair is not set in the task or the tile is off-grid).
// corrs: rows 6 t + p, 15 x 15 (radius 7).
__global__ void k_nn_input(const tp_task *tasks, int num_tasks, int tiles_x, int tiles_y,
                           const float *corrs, size_t corr_stride, float *x) {
    const int t = blockIdx.x;
    if (t >= num_tasks) return;
    const tp_task tk = tasks[t];
    const int tx = tk.txy & 0xfff, ty = tk.txy >> 16;
    const bool ok = tx < tiles_x && ty >= 0 && ty < tiles_y;
    for (int o = threadIdx.x; o < NN_P; o += blockDim.x) {
        const int cy = o / 15 - 7, cx = o % 15 - 7;
        float sum = 0.f;
#pragma unroll
        for (int p = 0; p < TP_NUM_PAIRS; p++) {
            float v = 0.f;
            int sy, sx;
            if (ok && ((tk.task >> (TP_TASK_CORR_BITS + p)) & 1) && pair_src(p, 7, cy, cx, sy, sx))
                v = corrs[(size_t)(6 * t + p) * corr_stride + (sy + 7) * 15 + sx + 7];
            x[((size_t)t * NN_IN + 1 + p) * NN_P + o] = v;
            sum += v;
        }
        x[(size_t)t * NN_IN * NN_P + o] = sum;
    }
}

// -------------------------------------------------------------- forward
// out[t] = (disparity, confidence); acts (optional) = h1, h2, z per tile.
__global__ __launch_bounds__(256) void k_nn_forward(const uint16_t *__restrict__ prm,
                                                    const float *__restrict__ x, int n, float *acts,
                                                    float *out) {
    __shared__ float A[NN_C * NN_P];
    __shared__ float B[NN_C * NN_P];
    __shared__ float red[16];
    const int t = threadIdx.x;
    const bool px = t < NN_P;
    const int y = px ? t / 15 : 0, xc = px ? t % 15 : 0;
    const unsigned m = px ? nbr_mask_yx(y, xc) : 0u;
    for (int tile = blockIdx.x; tile < n; tile += gridDim.x) {
        const float *xt = x + (size_t)tile * NN_IN * NN_P;
        for (int i = t; i < NN_IN * NN_P; i += 256) A[i] = xt[i];
        __syncthreads();
        float *at = acts ? acts + (size_t)tile * NN_ACTS : nullptr;
        float acc[NN_C];
        if (px) {
#pragma unroll
            for (int co = 0; co < NN_C; co++) acc[co] = bf16f(prm[NN_OFF_B1 + co]);
            conv_px<NN_IN, NN_C>(A, t, m, prm + NN_OFF_W1, acc);
#pragma unroll
            for (int co = 0; co < NN_C; co++) {
                const float h = fmaxf(acc[co], 0.f);
                B[co * NN_P + t] = h;
                if (at) at[co * NN_P + t] = h;
            }
        }
        __syncthreads();
        if (px) {
#pragma unroll
            for (int co = 0; co < NN_C; co++) acc[co] = bf16f(prm[NN_OFF_B2 + co]);
            conv_px<NN_C, NN_C>(B, t, m, prm + NN_OFF_W2, acc);
            // A (the input) was last read before the previous barrier
#pragma unroll
            for (int co = 0; co < NN_C; co++) {
                const float h = fmaxf(acc[co], 0.f);
                A[co * NN_P + t] = h;
                if (at) at[(NN_C + co) * NN_P + t] = h;
            }
        }
        __syncthreads();
        float z[2] = {bf16f(prm[NN_OFF_B3]), bf16f(prm[NN_OFF_B3 + 1])};
        if (px) conv_px<NN_C, 2>(A, t, m, prm + NN_OFF_W3, z);
        if (at && px) {
            at[2 * NN_C * NN_P + t] = z[0];
            at[2 * NN_C * NN_P + NN_P + t] = z[1];
        }
        float prob, disp, cl;
        head_stats(z[0], z[1], px, xc, red, prob, disp, cl);
        if (t == 0) {
            out[2 * tile] = disp;
            out[2 * tile + 1] = 1.0f / (1.0f + __expf(-cl));
        }
        __syncthreads();  // A and B are rewritten by the next tile
    }
}

// ------------------------------------------------------------- backward
struct NNBackArgs {
    const uint16_t *prm;
    const float *x;
    const float *acts;
    const float *target;  // residual disparity per tile; NaN: tile ignored
    int n;
    float conf_weight;
    float huber_delta;
    float scale;          // 1 / number of tiles with a finite target
    float *grads;         // NN_NPARAMS, accumulated
    float *loss;          // [1], accumulated
};

__global__ __launch_bounds__(256) void k_nn_backward(NNBackArgs a) {
    __shared__ float P[NN_C * NN_P];  // h2, then dY2 (in place)
    __shared__ float Q[NN_C * NN_P];  // h1, then dY1 (in place)
    __shared__ float S[NN_IN * NN_P];  // dz (2 planes), then the input x
    __shared__ float red[16];
    const int t = threadIdx.x;
    const bool px = t < NN_P;
    const int y = px ? t / 15 : 0, xc = px ? t % 15 : 0;
    const unsigned m = px ? nbr_mask_yx(y, xc) : 0u;
    // weight-gradient sums of this workgroup
    const int ci2 = t >> 3, cb2 = (t & 7) * 4;  // dW2: [cb2 .. cb2+3][ci2][9]
    float g2[4][9];
    float g3[3], g1[8];
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f;  // db1 (threads 192..223), db2 (224..255), db3 (0, 1)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int r = 0; r < 9; r++) g2[j][r] = 0.f;
#pragma unroll
    for (int k = 0; k < 3; k++) g3[k] = 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++) g1[k] = 0.f;
    const uint16_t *W2 = a.prm + NN_OFF_W2, *W3 = a.prm + NN_OFF_W3;

    for (int tile = blockIdx.x; tile < a.n; tile += gridDim.x) {
        const float tgt = a.target[tile];
        if (!(fabsf(tgt) < INFINITY)) continue;  // uniform: NaN/inf targets are ignored
        const float *at = a.acts + (size_t)tile * NN_ACTS;
        for (int i = t; i < NN_C * NN_P; i += 256) {
            Q[i] = at[i];
            P[i] = at[NN_C * NN_P + i];
        }
        // (0) head
        const float z0 = px ? at[2 * NN_C * NN_P + t] : 0.f;
        const float z1 = px ? at[2 * NN_C * NN_P + NN_P + t] : 0.f;
        float prob, disp, cl;
        head_stats(z0, z1, px, xc, red, prob, disp, cl);  // contains barriers
        const float e = disp - tgt;
        const float hd = fminf(fmaxf(e, -a.huber_delta), a.huber_delta);
        const float sg = 1.0f / (1.0f + __expf(-cl));
        const float lab = fabsf(e) < 0.5f ? 1.f : 0.f;
        const float gc = a.conf_weight * (sg - lab);
        if (px) {
            S[t] = a.scale * prob * (hd * (float)(xc - 7) - hd * disp + gc * (z1 - cl));
            S[NN_P + t] = a.scale * gc * prob;
        }
        if (t == 0) {
            const float ae = fabsf(e);
            const float hub = ae <= a.huber_delta ? 0.5f * e * e : a.huber_delta * (ae - 0.5f * a.huber_delta);
            const float bce = -(lab * logf(fmaxf(sg, 1e-12f)) + (1.f - lab) * logf(fmaxf(1.f - sg, 1e-12f)));
            atomicAdd(a.loss, a.scale * (hub + a.conf_weight * bce));
        }
        __syncthreads();
        // (1) dW3 [2][32][9], db3: 576 + 2 sums over pixels of dz * h2 neighbours
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int i = t + 256 * k;
            if (i < 576) {
                const int co = i / 288, ci = (i / 9) % 32, r = i % 9;
                float s = 0.f;
                for (int yy = 0; yy < 15; yy++)
                    for (int xx = 0; xx < 15; xx++) {
                        const int p = yy * 15 + xx;
                        const int sy = yy + r / 3 - 1, sx = xx + r % 3 - 1;
                        if (sy >= 0 && sy < 15 && sx >= 0 && sx < 15)
                            s = fmaf(S[co * NN_P + p], P[ci * NN_P + sy * 15 + sx], s);
                    }
                g3[k] += s;
            }
        }
        if (t < 2) {  // db3
            float s = 0.f;
            for (int p = 0; p < NN_P; p++) s += S[t * NN_P + p];
            gb3 += s;
        }
        __syncthreads();
        // (2) dY2 = (W3^T * dz) . [h2 > 0], in place over h2 (own pixel only)
        if (px) {
            float acc[NN_C];
#pragma unroll
            for (int ci = 0; ci < NN_C; ci++) acc[ci] = 0.f;
#pragma unroll
            for (int co = 0; co < 2; co++) {
                float v[9];
#pragma unroll
                for (int r = 0; r < 9; r++) v[r] = nbr(S + co * NN_P, t, m, 8 - r);  // dz[q - off_r]
#pragma unroll
                for (int ci = 0; ci < NN_C; ci++)
#pragma unroll
                    for (int r = 0; r < 9; r++) acc[ci] = fmaf(bf16f(W3[(co * NN_C + ci) * 9 + r]), v[r], acc[ci]);
            }
#pragma unroll
            for (int ci = 0; ci < NN_C; ci++)
                P[ci * NN_P + t] = P[ci * NN_P + t] > 0.f ? acc[ci] : 0.f;
        }
        __syncthreads();
        // input x into S (dz no longer needed)
        for (int i = t; i < NN_IN * NN_P; i += 256) S[i] = a.x[(size_t)tile * NN_IN * NN_P + i];
        // (3) dW2 [32][32][9]: thread (ci2, cb2..cb2+3); db2: threads 224..255
        for (int yy = 0; yy < 15; yy++)
            for (int xx = 0; xx < 15; xx++) {
                const int p = yy * 15 + xx;
                const unsigned mm = nbr_mask_yx(yy, xx);
                float h[9];
#pragma unroll
                for (int r = 0; r < 9; r++) h[r] = nbr(Q + ci2 * NN_P, p, mm, r);
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const float d = P[(cb2 + j) * NN_P + p];
#pragma unroll
                    for (int r = 0; r < 9; r++) g2[j][r] = fmaf(d, h[r], g2[j][r]);
                }
            }
        if (t >= 224) {  // db2 channel t - 224
            float s = 0.f;
            for (int p = 0; p < NN_P; p++) s += P[(t - 224) * NN_P + p];
            gb2 += s;
        }
        __syncthreads();
        // (4) dY1 = (W2^T * dY2) . [h1 > 0], in place over h1 (own pixel only)
        if (px) {
            float acc[NN_C];
#pragma unroll
            for (int ci = 0; ci < NN_C; ci++) acc[ci] = 0.f;
            for (int co = 0; co < NN_C; co++) {
                float v[9];
#pragma unroll
                for (int r = 0; r < 9; r++) v[r] = nbr(P + co * NN_P, t, m, 8 - r);
#pragma unroll
                for (int ci = 0; ci < NN_C; ci++)
#pragma unroll
                    for (int r = 0; r < 9; r++) acc[ci] = fmaf(bf16f(W2[(co * NN_C + ci) * 9 + r]), v[r], acc[ci]);
            }
#pragma unroll
            for (int ci = 0; ci < NN_C; ci++)
                Q[ci * NN_P + t] = Q[ci * NN_P + t] > 0.f ? acc[ci] : 0.f;
        }
        __syncthreads();
        // (5) dW1 [32][7][9] (8 per thread) and db1 (threads 192..223)
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int i = t + 256 * k;
            if (i < 2016) {
                const int co = i / 63, ci = (i / 9) % 7, r = i % 9;
                float s = 0.f;
                for (int yy = 0; yy < 15; yy++)
                    for (int xx = 0; xx < 15; xx++) {
                        const int sy = yy + r / 3 - 1, sx = xx + r % 3 - 1;
                        if (sy >= 0 && sy < 15 && sx >= 0 && sx < 15)
                            s = fmaf(Q[co * NN_P + yy * 15 + xx], S[ci * NN_P + sy * 15 + sx], s);
                    }
                g1[k] += s;
            }
        }
        if (t >= 192 && t < 224) {
            float s = 0.f;
            for (int p = 0; p < NN_P; p++) s += Q[(t - 192) * NN_P + p];
            gb1 += s;
        }
        __syncthreads();  // P, Q, S are rewritten by the next tile
    }
    // add this workgroup's sums to the global gradient
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int r = 0; r < 9; r++) atomicAdd(&a.grads[NN_OFF_W2 + ((cb2 + j) * NN_C + ci2) * 9 + r], g2[j][r]);
#pragma unroll
    for (int k = 0; k < 3; k++)
        if (t + 256 * k < 576) atomicAdd(&a.grads[NN_OFF_W3 + t + 256 * k], g3[k]);
#pragma unroll
    for (int k = 0; k < 8; k++)
        if (t + 256 * k < 2016) atomicAdd(&a.grads[NN_OFF_W1 + t + 256 * k], g1[k]);
    if (t >= 192 && t < 224) atomicAdd(&a.grads[NN_OFF_B1 + t - 192], gb1);
    if (t >= 224) atomicAdd(&a.grads[NN_OFF_B2 + t - 224], gb2);
    if (t < 2) atomicAdd(&a.grads[NN_OFF_B3 + t], gb3);
}

// ----------------------------------------------------- inference (MFMA)
// The same network with BF16 activations and all three convolutions on
// the matrix cores (v_mfma_f32_32x32x16_bf16, FP32 accumulation):
//   - the two 32-channel 3x3 convolutions as implicit GEMMs, D[co][pixel] =
//     sum_k W[co][k] X[k][pixel], k = (3x3 neighbour r, input channel);
//   - the 2-channel output convolution as a 1x1 product G[pixel][(co, r)] =
//     sum_ci h2[pixel][ci] W3[co][ci][r] (18 of 32 columns used), then
//     z[co](p) = b3[co] + sum_r G[p + off_r][(co, r)] (a 9-tap shift-add of
//     G, zero outside the tile: h2 is zero there).
// The head (softmax, disparity, confidence) runs on the vector units.
//
// Activations live in LDS channel-last, so the 8 consecutive channels a
// lane contributes to one K slice are one 16-byte load.  The MFMA columns
// are 256 pixel slots, slot s = pixel (y, x) = (s / 16, s % 16): slots with
// x = 15 or y = 15 are past the tile (computed, never stored).  The inputs
// of the 3x3 convolutions (x, h1) lie on a padded grid of 16 positions per
// row, slot s at position s + 17 (NN_POFF): column 15 of row y is also
// column -1 of row y + 1, and rows -1 and 15 stay zero, so the neighbour r
// of every slot is at position s + 17 + 16 (r / 3 - 1) + r % 3 - 1, in 0..289.
// h2 and G are held by slot.  Operand layouts (32x32x16): lane i holds
// A[i % 32][8 (i / 32) + e], B[8 (i / 32) + e][i % 32] (e = 0..7) and
// D[8 (v / 4) + 4 (i / 32) + v % 4][i % 32] (v = 0..15).  Each wavefront
// takes 2 of a tile's 8 blocks of 32 slots and runs their two MFMA chains
// interleaved (independent accumulators); the weights stay in registers.
//
// LDS banks (gfx950, measured with the bank-conflict counter): a 16-byte
// read serves each half-wavefront in one pass when lanes 0..15 of the half
// hit 16 distinct 16-byte bank groups (address / 16 mod 16) and lane j + 16
// hits the same group as lane j; a 16-byte write serves 8 lanes per pass
// (groups address / 16 mod 8); an 8-byte write 16 lanes (address / 8 mod
// 16); 4-byte accesses 32 lanes (address / 4 mod 32).  The layouts below
// keep every access of the loop conflict-free on these rules.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// padded grid: 16 positions per row, slot s at s + NN_POFF; positions past
// the last one read (289) hold the stores of the slots past the tile
constexpr int NN_PW = 16, NN_POFF = NN_PW + 1, NN_PJUNK = 290, NN_PROWS = NN_PJUNK + 1;
// G row pitch in floats (18 used): odd, so the 9-tap reads of 32
// consecutive slots fall in 32 distinct banks
constexpr int NN_GS = 21;
// position pitch of the 32-channel activations in BF16 elements: 80 bytes,
// 5 bank groups (odd), so 16 consecutive positions fall in 16 distinct
// groups and positions 16 apart (lanes j and j + 16) in the same one
constexpr int NN_CS = 40;

__device__ __forceinline__ bool nn_slot_in_tile(int s) { return (s & 15) < 15 && s < 15 * NN_PW; }

// A fragment: weights W[co][ci0 + e][r] of layer (CIN inputs), e = 0..7
// (channels >= cin_valid or r >= 9 give 0)
template <int CIN>
__device__ __forceinline__ u32x4 nn_wfrag(const uint16_t *__restrict__ w, int co, int ci0, int r,
                                          int cin_valid) {
    uint32_t h[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int ci = ci0 + e;
        h[e] = (r < 9 && ci < cin_valid) ? (uint32_t)w[(co * CIN + ci) * 9 + r] : 0u;
    }
    u32x4 v;
#pragma unroll
    for (int d = 0; d < 4; d++) v[d] = h[2 * d] | (h[2 * d + 1] << 16);
    return v;
}

// B fragment of the 1x1 output product, K step s: W3[co][ci][r] at
// k = ci = 16 s + 8 (lane / 32) + e, column n = lane % 32 = co * 9 + r
// (columns 18..31 zero)
__device__ __forceinline__ u32x4 nn_w3frag(const uint16_t *__restrict__ prm, int lane, int s) {
    const int n = lane & 31, co = n / 9, r = n - 9 * co;
    uint32_t h[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int ci = 16 * s + 8 * (lane >> 5) + e;
        h[e] = n < 18 ? (uint32_t)prm[NN_OFF_W3 + (co * NN_C + ci) * 9 + r] : 0u;
    }
    u32x4 v;
#pragma unroll
    for (int d = 0; d < 4; d++) v[d] = h[2 * d] | (h[2 * d + 1] << 16);
    return v;
}

// (a, b) as two BF16 in one word, round to nearest even (one
// v_cvt_pk_bf16_f32)
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short s16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t bf16pair(float a, float b) {
    const f32x2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// ReLU of two BF16 in one word: as 16-bit integers the negative values
// (and -0) are the negative ones, so max(x, 0) per half (v_pk_max_i16)
__device__ __forceinline__ uint32_t relu_bf16x2(uint32_t v) {
    const s16x2 z = {0, 0};
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2, v), z));
}

__device__ __forceinline__ f32x16 nn_mfma(const u32x4 &a, const u32x4 &b, const f32x16 &c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// Two 32-slot blocks (slots s0, s1 of this lane's column) of a 3x3
// convolution: in = padded LDS activations with C channels per position
// (pitch 8 for C = 8, NN_CS for C = 32); NK chunks of 16 K values; A =
// weight fragments.  k ordering per chunk kc: C = 8: r = 2 kc + i / 32,
// channels 0..7; C = 32: r = kc / 2, channels (kc % 2) 16 + 8 (i / 32) .. + 7.
// The B fragments of chunk kc + 2 are requested while chunks kc and kc + 1
// are multiplied (the LDS latency is not covered by other wavefronts at 2
// waves per SIMD).
template <int C, int NK>
__device__ __forceinline__ void nn_conv_2blocks(const uint16_t *in, const u32x4 (&A)[NK], int s0, int s1,
                                                int lane, f32x16 &acc0, f32x16 &acc1) {
    constexpr int CS = C == 8 ? 8 : NN_CS;
    constexpr int D = 2;  // prefetch distance in chunks
    const int half = lane >> 5;
    auto fetch = [&](int kc, u32x4 &b0, u32x4 &b1) {
        int r, c0;
        if (C == 8) {
            r = 2 * kc + half;
            c0 = 0;
        } else {
            r = kc >> 1;
            c0 = (kc & 1) * 16 + 8 * half;
        }
        b0 = u32x4{0u, 0u, 0u, 0u};
        b1 = u32x4{0u, 0u, 0u, 0u};
        if (r < 9) {
            const int d = NN_POFF + (r / 3 - 1) * NN_PW + (r % 3 - 1);
            b0 = *(const u32x4 *)(in + (s0 + d) * CS + c0);
            b1 = *(const u32x4 *)(in + (s1 + d) * CS + c0);
        }
    };
    u32x4 b0[D + 1], b1[D + 1];
#pragma unroll
    for (int kc = 0; kc < D && kc < NK; kc++) fetch(kc, b0[kc], b1[kc]);
#pragma unroll
    for (int kc = 0; kc < NK; kc++) {
        if (kc + D < NK) fetch(kc + D, b0[(kc + D) % (D + 1)], b1[(kc + D) % (D + 1)]);
        // the loads of chunk kc + D stay ahead of the products of chunk kc
        __builtin_amdgcn_sched_barrier(0);
        acc0 = nn_mfma(A[kc], b0[kc % (D + 1)], acc0);
        acc1 = nn_mfma(A[kc], b1[kc % (D + 1)], acc1);
        __builtin_amdgcn_sched_barrier(0);
    }
}

// The layer's biases in the accumulator layout (lane i: channels
// 8 j + 4 (i / 32) .. + 3 at elements 4 j .. 4 j + 3): the first product of
// a block's chain adds them, so the epilogue does not.
__device__ __forceinline__ f32x16 nn_bias_acc(const float *bias, int lane) {
    const int half = lane >> 5;
    f32x16 v;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const float4 b = *(const float4 *)(bias + 8 * j + 4 * half);
        v[4 * j] = b.x;
        v[4 * j + 1] = b.y;
        v[4 * j + 2] = b.z;
        v[4 * j + 3] = b.w;
    }
    return v;
}

// ReLU (the accumulator holds the bias already), BF16, channel-last store
// of one 32-slot block at out[idx * NN_CS] (the caller sends slots past the
// tile to a junk position).  Lane i holds channels 8 j + 4 (i / 32) .. + 3 (j = 0..3) of
// its slot; the two wavefront halves swap half their words
// (v_permlane32_swap), after which the lower half holds channels 0..7 and
// 16..23 and the upper half 8..15 and 24..31, each run stored as one
// 16-byte write.
__device__ __forceinline__ void nn_store_block(uint16_t *out, const f32x16 &acc, int idx, int lane) {
    const int half = lane >> 5;
    uint32_t w[8];  // w[2 j], w[2 j + 1]: channels 8 j + 4 half + (0, 1), (2, 3)
#pragma unroll
    for (int j = 0; j < 4; j++) {
        w[2 * j] = relu_bf16x2(bf16pair(acc[4 * j], acc[4 * j + 1]));
        w[2 * j + 1] = relu_bf16x2(bf16pair(acc[4 * j + 2], acc[4 * j + 3]));
    }
    // swap(x, y): lower half keeps x and gets the upper half's x; upper half
    // gets the lower half's y and keeps y
    u32x4 run[2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const auto r0 = __builtin_amdgcn_permlane32_swap(w[4 * c], w[4 * c + 2], false, false);
        const auto r1 = __builtin_amdgcn_permlane32_swap(w[4 * c + 1], w[4 * c + 3], false, false);
        run[c] = u32x4{r0[0], r1[0], r0[1], r1[1]};
    }
    uint16_t *o = out + idx * NN_CS + 8 * half;
    *(u32x4 *)o = run[0];
    *(u32x4 *)(o + 16) = run[1];
}

// All-lane sum / max of a wavefront: butterfly within each row of 16 lanes
// by DPP (quad swaps 1 and 2, half-row mirror, row mirror: every lane of a
// row ends with the row's result), then the 4 rows' results by readlane.
template <int CTRL>
__device__ __forceinline__ float nn_dpp(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, false));
}
__device__ __forceinline__ float nn_row_lane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ float nn_wave_sum(float v) {
    v += nn_dpp<0xB1>(v);   // quad_perm 1,0,3,2
    v += nn_dpp<0x4E>(v);   // quad_perm 2,3,0,1
    v += nn_dpp<0x141>(v);  // row_half_mirror
    v += nn_dpp<0x140>(v);  // row_mirror
    return (nn_row_lane(v, 0) + nn_row_lane(v, 16)) + (nn_row_lane(v, 32) + nn_row_lane(v, 48));
}
__device__ __forceinline__ float nn_wave_max(float v) {
    v = fmaxf(v, nn_dpp<0xB1>(v));
    v = fmaxf(v, nn_dpp<0x4E>(v));
    v = fmaxf(v, nn_dpp<0x141>(v));
    v = fmaxf(v, nn_dpp<0x140>(v));
    return fmaxf(fmaxf(nn_row_lane(v, 0), nn_row_lane(v, 16)), fmaxf(nn_row_lane(v, 32), nn_row_lane(v, 48)));
}

// 2 workgroups per CU (69 KB of LDS each): 2 waves per SIMD, at most 256
// registers per lane
// Where the inference reads its input: x [n][7][225] (tp_nn_input's
// output), or, with FROM_CORRS, the correlation rows 6 task + pair (radius 7)
// and the task words, mapped and summed as k_nn_input does.
struct NNInferIn {
    const float *x;
    const tp_task *tasks;
    const float *corrs;
    size_t corr_stride;
    int tiles_x, tiles_y;
};

template <bool FROM_CORRS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_nn_infer(const uint16_t *__restrict__ prm, NNInferIn in, int n, float *out) {
    __shared__ __align__(16) uint16_t X0[NN_PROWS * 8];   // x, padded, channel 7 zero
    __shared__ __align__(16) uint16_t H1[NN_PROWS * NN_CS];
    // h2 on the padded grid as h1: the positions of the slots past the tile
    // and the ring stay 0, so G is 0 there
    __shared__ __align__(16) uint16_t H2[NN_PROWS * NN_CS];
    // G by slot + NN_POFF, with NN_POFF zero rows before and after: the
    // 9-tap shift-add reads its neighbours unmasked (outside the tile: 0)
    __shared__ __align__(16) float Gs[(256 + 2 * NN_POFF) * NN_GS];
    __shared__ __align__(16) float bias[2][NN_C];       // b1, b2
    __shared__ float red[16];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int co = lane & 31, half = lane >> 5;
    if (t < 2 * NN_C) bias[t >> 5][t & 31] = bf16f(prm[(t < NN_C ? NN_OFF_B1 : NN_OFF_B2 - NN_C) + t]);
    for (int i = t; i < NN_PROWS * 4; i += 256) ((uint32_t *)X0)[i] = 0u;
    for (int i = t; i < NN_PROWS * NN_CS / 2; i += 256) ((uint32_t *)H2)[i] = 0u;
    for (int i = t; i < NN_PROWS * NN_CS / 2; i += 256) ((uint32_t *)H1)[i] = 0u;
    for (int i = t; i < (256 + 2 * NN_POFF) * NN_GS; i += 256) Gs[i] = 0.f;
    u32x4 A1[5], A2[18], B3[2];
#pragma unroll
    for (int kc = 0; kc < 5; kc++) A1[kc] = nn_wfrag<NN_IN>(prm + NN_OFF_W1, co, 0, 2 * kc + half, NN_IN);
#pragma unroll
    for (int kc = 0; kc < 18; kc++)
        A2[kc] = nn_wfrag<NN_C>(prm + NN_OFF_W2, co, (kc & 1) * 16 + 8 * half, kc >> 1, NN_C);
    B3[0] = nn_w3frag(prm, lane, 0);
    B3[1] = nn_w3frag(prm, lane, 1);
    const float b30 = bf16f(prm[NN_OFF_B3]), b31 = bf16f(prm[NN_OFF_B3 + 1]);
    __syncthreads();
    // thread t = pixel t of the input (row-major 15 x 15): its 7 input
    // channels of a tile, requested one tile ahead
    const bool pin = t < NN_P;
    float xn[NN_IN];
    // FROM_CORRS: offset of this pixel's source in each pair's row (-1: none)
    int psrc[TP_NUM_PAIRS];
#pragma unroll
    for (int p = 0; p < TP_NUM_PAIRS; p++) {
        int sy, sx;
        psrc[p] = (pin && pair_src(p, 7, t / 15 - 7, t % 15 - 7, sy, sx)) ? (sy + 7) * 15 + sx + 7 : -1;
    }
    auto load_x = [&](int tile) {
        if (!FROM_CORRS) {
            const float *xt = in.x + (size_t)tile * NN_IN * NN_P + (pin ? t : 0);
#pragma unroll
            for (int c = 0; c < NN_IN; c++) xn[c] = (tile < n && pin) ? xt[c * NN_P] : 0.f;
        } else {
            // as k_nn_input: pair p's mapped value if the task has the pair
            // and the tile is on the grid, channel 0 their sum in pair order
            int task = 0;
            if (tile < n) {
                gtask *tk = gtasks(in.tasks) + tile;
                const int txy = tk->txy, tx = txy & 0xfff, ty = txy >> 16;
                task = (tx < in.tiles_x && ty >= 0 && ty < in.tiles_y) ? tk->task : 0;
            }
            gfloat *rows = gptr(in.corrs) + (size_t)6 * (tile < n ? tile : 0) * in.corr_stride;
            float sum = 0.f;
#pragma unroll
            for (int p = 0; p < TP_NUM_PAIRS; p++) {
                const bool use = ((task >> (TP_TASK_CORR_BITS + p)) & 1) && psrc[p] >= 0;
                const float v = use ? rows[p * in.corr_stride + psrc[p]] : 0.f;
                xn[1 + p] = v;
                sum += v;
            }
            xn[0] = sum;
        }
    };
    const int xpos = pin ? NN_PW * (t / 15) + t % 15 + NN_POFF : NN_PJUNK;
    const f32x16 zero = {};
    const int s0 = 64 * wave + co, s1 = s0 + 32;  // this lane's slot column in the wave's 2 blocks
    // thread t = slot t in the head
    const int sx = t & 15;
    const bool px = nn_slot_in_tile(t);
    // the head of the previous tile: one thread combines the wavefronts'
    // partial sums in red (after the second barrier of the next tile, which
    // every wavefront reaches only after writing them; red is rewritten only
    // after the third)
    auto finish = [&](int tile) {
        const float M = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        float S = 0.f, SX = 0.f, SC = 0.f;
#pragma unroll
        for (int w = 0; w < 4; w++) {
            const float f = __expf(red[w] - M);  // 0 for a wavefront with no pixel
            S = fmaf(red[4 + w], f, S);
            SX = fmaf(red[8 + w], f, SX);
            SC = fmaf(red[12 + w], f, SC);
        }
        const float inv = 1.0f / S;
        out[2 * tile] = SX * inv;
        out[2 * tile + 1] = 1.0f / (1.0f + __expf(-SC * inv));
    };
    load_x(blockIdx.x);
    int prev = -1;
    for (int tile = blockIdx.x; tile < n; tile += gridDim.x) {
        // x, channel-last on the padded grid (channel 7 and the ring stay 0):
        // one 16-byte store per pixel
        {
            u32x4 v;
            v[0] = bf16pair(xn[0], xn[1]);
            v[1] = bf16pair(xn[2], xn[3]);
            v[2] = bf16pair(xn[4], xn[5]);
            v[3] = bf16pair(xn[6], 0.f);
            *(u32x4 *)(X0 + xpos * 8) = v;
        }
        __syncthreads();
        load_x(tile + gridDim.x);
        {
            const f32x16 b1 = nn_bias_acc(bias[0], lane);
            f32x16 a0 = b1, a1 = b1;
            nn_conv_2blocks<8, 5>(X0, A1, s0, s1, lane, a0, a1);
            nn_store_block(H1, a0, nn_slot_in_tile(s0) ? s0 + NN_POFF : NN_PJUNK, lane);
            nn_store_block(H1, a1, nn_slot_in_tile(s1) ? s1 + NN_POFF : NN_PJUNK, lane);
        }
        __syncthreads();
        // the head of the previous tile, by the lane of the wavefront with
        // the fewest pixels, off the path to the next barrier's other waits
        if (t == 192 && prev >= 0) finish(prev);
        {
            const f32x16 b2 = nn_bias_acc(bias[1], lane);
            f32x16 a0 = b2, a1 = b2;
            nn_conv_2blocks<NN_C, 18>(H1, A2, s0, s1, lane, a0, a1);
            // h2 of the wave's own 64 slots
            nn_store_block(H2, a0, nn_slot_in_tile(s0) ? s0 + NN_POFF : NN_PJUNK, lane);
            nn_store_block(H2, a1, nn_slot_in_tile(s1) ? s1 + NN_POFF : NN_PJUNK, lane);
        }
        wave_lds_sync();
        {
            // G of the wave's two blocks: A = h2 rows (slot = lane % 32)
            f32x16 g0 = zero, g1 = zero;
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const u32x4 a0 = *(const u32x4 *)(H2 + (64 * wave + co + NN_POFF) * NN_CS + 16 * s + 8 * half);
                const u32x4 a1 =
                    *(const u32x4 *)(H2 + (64 * wave + 32 + co + NN_POFF) * NN_CS + 16 * s + 8 * half);
                g0 = nn_mfma(a0, B3[s], g0);
                g1 = nn_mfma(a1, B3[s], g1);
            }
            // lane holds G[slot 8 (v / 4) + 4 half + v % 4 of the block][column co]
            if (co < 18) {
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int q0 = 64 * wave + 8 * (v >> 2) + 4 * half + (v & 3);  // slots past the tile: junk
                    Gs[(q0 + NN_POFF) * NN_GS + co] = g0[v];
                    Gs[(q0 + 32 + NN_POFF) * NN_GS + co] = g1[v];
                }
            }
        }
        __syncthreads();
        // z = b3 + the 9-tap shift-add of G: a neighbour outside the tile is
        // a slot past it (G = 0, its h2 row is the zero ring) or a zero row
        // of Gs, so every tap adds as read (what a masked tap adds: 0)
        float z0 = b30, z1 = b31;
        {
            const float *gt = Gs + (t + NN_POFF) * NN_GS;
#pragma unroll
            for (int r = 0; r < 9; r++) {
                const float *gq = gt + (NN_PW * (r / 3 - 1) + r % 3 - 1) * NN_GS;
                z0 += gq[r];
                z1 += gq[9 + r];
            }
        }
        // softmax statistics of the head (as head_stats) with one barrier:
        // each wavefront reduces relative to its own maximum, thread 0
        // rescales the 4 partial sums to the common one
        {
            const float m = nn_wave_max(px ? z0 : -INFINITY);
            const float e = px ? __expf(z0 - m) : 0.f;
            const float se = nn_wave_sum(e), sxe = nn_wave_sum(e * (float)(sx - 7));
            const float sc = nn_wave_sum(e * z1);
            if (lane == 0) {
                red[wave] = m;
                red[4 + wave] = se;
                red[8 + wave] = sxe;
                red[12 + wave] = sc;
            }
        }
        prev = tile;
        // no barrier here: the next tile rewrites X0 (last read in its layer
        // 1), H1 and H2 (layer 2) and Gs (after its second barrier, which
        // every thread passes only after this shift-add); red is read by one
        // thread after the next second barrier and rewritten only after the
        // third
    }
    __syncthreads();
    if (t == 192 && prev >= 0) finish(prev);
}

// ---------------------------------------------------------------- Adam
__global__ void k_nn_adam(float *w, float *m, float *v, const float *g, int n, float lr, float b1,
                          float b2, float eps, float c1, float c2, uint16_t *wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = fmaf(b1, m[i], (1.f - b1) * gi);
    const float vi = fmaf(b2, v[i], (1.f - b2) * gi * gi);
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    wb[i] = f2bf16(wi);
}

__global__ void k_nn_pack(const float *w, int n, uint16_t *wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = f2bf16(w[i]);
}

__global__ void k_fill(float *p, size_t n, float v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace tp
