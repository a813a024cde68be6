// k_dtt.hpp - 2-D 8x8 DCT/DST of every 8x8 tile of an image, the transform
// test entry of the reference (GPU_DTT24_DRV, src/dtt8x8.cuh:103-161).
//
// Mode (as the reference): bit 0 selects the sine transform horizontally,
// bit 1 vertically; modes 0-3 are DCT-IV / DST-IV, modes 4-7 DCT-II /
// DST-II, all orthonormal (dttiv_shared_mem and dttii_shared_mem,
// src/dtt8x8.cuh:299-380, :465-545).  Output tile (ty, tx) replaces the
// input tile at the same place: row = vertical frequency, column =
// horizontal frequency.
//
// One 8-lane unit per tile, 8 tiles per wavefront: lane i loads row i and
// transforms it horizontally, the unit transposes through its LDS slice,
// lane i transforms column i vertically and stores it.  dst may equal src
// (a unit reads all of its tile before it writes any of it).
#pragma once

#include "tp_common.hpp"

namespace tp {

// in-place 8-point transform of x; kind: 0 DCT-IV, 1 DST-IV, 2 DCT-II,
// 3 DST-II (orthonormal: the DCT-II DC term and the DST-II last term carry
// the extra 1 / sqrt 2)
__device__ __forceinline__ void dtt8_ortho(float (&x)[8], int kind) {
    constexpr float kRsqrt2 = 0.70710678118654752f;
    if (kind == 0) {
        dct4_8(x);
    } else if (kind == 1) {
        dst4_8(x);
    } else if (kind == 2) {
        dct2_8(x);
        x[0] *= kRsqrt2;
    } else {
        dst2_8(x);
        x[7] *= kRsqrt2;
    }
}

__global__ __launch_bounds__(256) void k_dtt8x8(const float *src, float *dst, size_t stride, int tiles_x,
                                                 int num_tiles, int mode) {
    __shared__ float slab[32][8][9];  // [unit of the workgroup][row][col], padded
    const int unit = threadIdx.x >> 3, li = threadIdx.x & 7;
    const int tile = blockIdx.x * 32 + unit;
    const bool ok = tile < num_tiles;
    const int ty = ok ? tile / tiles_x : 0, tx = ok ? tile - ty * tiles_x : 0;
    const size_t base = (size_t)(ty * 8) * stride + (size_t)tx * 8;
    const int kh = (mode > 3 ? 2 : 0) + (mode & 1);
    const int kv = (mode > 3 ? 2 : 0) + ((mode >> 1) & 1);
    float x[8];
#pragma unroll
    for (int c = 0; c < 8; c++) x[c] = ok ? src[base + (size_t)li * stride + c] : 0.f;
    dtt8_ortho(x, kh);  // horizontal: row li
#pragma unroll
    for (int c = 0; c < 8; c++) slab[unit][li][c] = x[c];
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 8; r++) x[r] = slab[unit][r][li];
    dtt8_ortho(x, kv);  // vertical: column li
    if (!ok) return;
#pragma unroll
    for (int k = 0; k < 8; k++) dst[base + (size_t)k * stride + li] = x[k];
}

}  // namespace tp
