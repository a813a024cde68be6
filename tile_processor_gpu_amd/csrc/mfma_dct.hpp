// mfma_dct.hpp - 8-point DCT-IV / DST-IV of a wavefront's 64 lane vectors as
// batched matrix contractions on the matrix cores, with split-precision
// operands (FP16 or BF16 hi + lo, FP32 accumulation).
//
// Layout (the forward CLT's): lane i = 8 u + c holds the 8-vector x of
// column / row c of unit u.  One v_mfma_f32_16x16x16 computes
//   D[r][n] = sum_k A[r][k] B[k][n],  lane i holding B[4 (i/16) + e][i % 16]
//   = x_i[e + 4h] (its own data, half h of the vector), A[r][4 (i/16) + e]
//   and D[4 (i/16) + t][i % 16].
// With A[r][4 g + e] = [g == 2 s + r / 8] T[r % 8][e + 4 h] the 16 rows of D
// are the transforms of the lanes of lane groups 2 s and 2 s + 1, so the
// sum over h = 0, 1 and the two settings s = 0, 1 (4 instructions) gives all
// 64 transformed vectors; 3 instructions per product for the split:
//   T x ~ T_hi x_hi + T_lo x_hi + T_hi x_lo   (T_lo x_lo dropped).
// Lane i ends up with y[m], m = 4 ((i/16) % 2) + t, of the vector of lane
// 16 (2 s + (i/16) / 2) + i % 16; it stores them (one float4) into the LDS
// staging of that unit at [c][m], the layout the FP32 path writes.
//
// PREC 3 splits both operands into three BF16 parts, x = hi + mid + lo
// (each part the BF16 rounding of what the previous ones leave, so the
// parts carry 24 bits and the sum is exact to 2^-26), and keeps the six
// products down to 2^-24: hi hi, hi mid, mid hi, mid mid, hi lo, lo hi
// (6 instructions per product, smallest first).
//
// Accuracy: FP16 hi + lo carries 22 significant bits (relative error of a
// product about 2^-22); BF16 hi + lo 16 bits (2^-16); BF16 hi + mid + lo
// about 24 (2^-24, the FP32 accumulation's own).  For FP16 each lane's
// vector is first scaled by a power of two so that its largest element lies
// in [2^14, 2^15): the hi part cannot overflow and the lo parts of all
// elements that matter (those above 2^-11 of the largest) stay normal FP16
// numbers (subnormal operands lose the lo correction); the constants are
// scaled by 2^8 for the same reason.  The lane storing a result undoes both
// scales of the vector it came from.
#pragma once

#include "tp_common.hpp"

namespace tp {

typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef short i16x4 __attribute__((ext_vector_type(4)));

// PREC 1: FP16 hi + lo, 2: BF16 hi + lo, 3: BF16 hi + mid + lo
template <int PREC>
struct SplitOps;

template <>
struct SplitOps<1> {
    typedef f16x4 vec;
    __device__ static __forceinline__ void split(float v, _Float16 &hi, _Float16 &lo) {
        hi = (_Float16)v;
        lo = (_Float16)(v - (float)hi);
    }
    __device__ static __forceinline__ void split4(const float *v, vec &hi, vec &lo) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            _Float16 h, l;
            split(v[e], h, l);
            hi[e] = h;
            lo[e] = l;
        }
    }
    __device__ static __forceinline__ f32x4 mma(const vec &a, const vec &b, const f32x4 &c) {
        return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
    }
};

__device__ __forceinline__ short bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ __forceinline__ float bf16_widen(short h) { return __uint_as_float((uint32_t)(uint16_t)h << 16); }

template <>
struct SplitOps<2> {
    typedef i16x4 vec;
    __device__ static __forceinline__ void split4(const float *v, vec &hi, vec &lo) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const short h = bf16_rne(v[e]);
            hi[e] = h;
            lo[e] = bf16_rne(v[e] - bf16_widen(h));
        }
    }
    __device__ static __forceinline__ f32x4 mma(const vec &a, const vec &b, const f32x4 &c) {
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
};

template <>
struct SplitOps<3> {
    typedef i16x4 vec;
    __device__ static __forceinline__ void split4(const float *v, vec &hi, vec &mid, vec &lo) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const short h = bf16_rne(v[e]);
            const float r = v[e] - bf16_widen(h);  // exact
            const short m = bf16_rne(r);
            hi[e] = h;
            mid[e] = m;
            lo[e] = bf16_rne(r - bf16_widen(m));  // r - m exact
        }
    }
    __device__ static __forceinline__ void split4(const float *v, vec &hi, vec &lo) {
        vec m;
        split4(v, hi, m, lo);
    }
    __device__ static __forceinline__ f32x4 mma(const vec &a, const vec &b, const f32x4 &c) {
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
};

// Constant operand of one transform for this lane: T[lane % 8][e + 4 h],
// split (mid: PREC 3 only), and the lane's block-diagonal masks for s = 0, 1.
template <int PREC>
struct MfmaConst {
    typename SplitOps<PREC>::vec hi[2], mid[2], lo[2];
    bool on[2];
};

// KIND: 0 DCT-IV, 1 DST-IV (kTab.dtt rows 0 and 1)
template <int PREC, int KIND>
__device__ __forceinline__ MfmaConst<PREC> mfma_const(int lane) {
    MfmaConst<PREC> k;
    const int m = lane & 7;
    float row[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        float v = 0.f;
#pragma unroll
        for (int t = 0; t < 8; t++) v = (m == t) ? kTab.dtt[KIND][t][j] : v;
        row[j] = PREC == 1 ? v * 256.0f : v;  // FP16: constants x 2^8 keep their lo parts normal
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
        if constexpr (PREC == 3)
            SplitOps<PREC>::split4(row + 4 * h, k.hi[h], k.mid[h], k.lo[h]);
        else
            SplitOps<PREC>::split4(row + 4 * h, k.hi[h], k.lo[h]);
    }
    const int g = lane >> 4, u = lane >> 3;
#pragma unroll
    for (int s = 0; s < 2; s++) k.on[s] = (g == 2 * s + (u & 1));
    return k;
}

// y = T x for the 64 lane vectors of the wavefront; lane (u, c) of the
// results is stored at stage + u * ustride + c * 8 + m (m = 0..7).
template <int PREC>
__device__ __forceinline__ void mfma_dct8_to_lds(const float (&x)[8], const MfmaConst<PREC> &k,
                                                 float *stage, int lane, int ustride) {
    typedef typename SplitOps<PREC>::vec vec;
    vec xh[2], xm[2], xl[2];
    float inv = 1.0f;
    if constexpr (PREC == 1) {
        float mx = 0.f;
#pragma unroll
        for (int j = 0; j < 8; j++) mx = fmaxf(mx, fabsf(x[j]));
        // exponent of mx (mx = f 2^e, f in [0.5, 1)); scale 2^(15 - e)
        const int e = min(100, max(-100, (int)((__float_as_uint(mx) >> 23) & 0xff) - 126));
        const float sc = __uint_as_float((uint32_t)(127 + 15 - e) << 23);
        inv = __uint_as_float((uint32_t)(127 - 15 - 8 + e) << 23);  // data and constant scales
        float xs[8];
#pragma unroll
        for (int j = 0; j < 8; j++) xs[j] = x[j] * sc;
#pragma unroll
        for (int h = 0; h < 2; h++) SplitOps<PREC>::split4(xs + 4 * h, xh[h], xl[h]);
    } else if constexpr (PREC == 3) {
#pragma unroll
        for (int h = 0; h < 2; h++) SplitOps<PREC>::split4(x + 4 * h, xh[h], xm[h], xl[h]);
    } else {
#pragma unroll
        for (int h = 0; h < 2; h++) SplitOps<PREC>::split4(x + 4 * h, xh[h], xl[h]);
    }
    const vec zero = {};
#pragma unroll
    for (int s = 0; s < 2; s++) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const vec ah = k.on[s] ? k.hi[h] : zero, al = k.on[s] ? k.lo[h] : zero;
            if constexpr (PREC == 3) {
                const vec am = k.on[s] ? k.mid[h] : zero;
                d = SplitOps<PREC>::mma(al, xh[h], d);
                d = SplitOps<PREC>::mma(ah, xl[h], d);
                d = SplitOps<PREC>::mma(am, xm[h], d);
                d = SplitOps<PREC>::mma(am, xh[h], d);
                d = SplitOps<PREC>::mma(ah, xm[h], d);
                d = SplitOps<PREC>::mma(ah, xh[h], d);
            } else {
                d = SplitOps<PREC>::mma(ah, xh[h], d);
                d = SplitOps<PREC>::mma(al, xh[h], d);
                d = SplitOps<PREC>::mma(ah, xl[h], d);
            }
        }
        const int grp = lane >> 4;
        const int src = 16 * (2 * s + (grp >> 1)) + (lane & 15);  // lane whose vector this is
        if constexpr (PREC == 1) {
            const float iv = __shfl(inv, src);
#pragma unroll
            for (int t = 0; t < 4; t++) d[t] *= iv;
        }
        const int u = src >> 3;
        const int c = lane & 7;
        const int m0 = 4 * (grp & 1);
        *(f32x4 *)(stage + u * ustride + c * 8 + m0) = d;
    }
}

}  // namespace tp
