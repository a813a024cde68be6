// hbm_probe.hip - streaming kernels that measure what this GPU's HBM
// delivers in practice for pure reads, pure writes, copies and the forward
// CLT's mix (one byte read per three written).  The benchmark divides its
// kernels' measured traffic by these rates to tell how close they run to the
// memory system's practical ceiling (the 8 TB/s peak is not reachable by
// any access pattern).  Built into libtp_probe.so, separate from the
// library's own entry points.
//
// Form: one-shot, not grid-stride.  Workgroup b (256 threads) moves the
// float4 elements [256 b, 256 (b + 1)) and exits; the grid covers the whole
// buffer, so the dispatcher streams workgroups over it in address order and
// every CU stays full.  On MI355X this form measures 6.6-6.9 TB/s for reads
// and writes and 6.2-6.3 TB/s for copies, where grid-stride loops over a
// resident grid (2-32 workgroups per CU, 1-8 elements in flight per thread)
// top out at 5.2-5.3 (write), 6.3-6.4 (read) and 5.3-5.8 TB/s (copy).
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) f4 gf4;

template <bool NT>
__device__ __forceinline__ void put(gf4 *p, f4 v) {
    if (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

__global__ __launch_bounds__(256) void k_read(const f4 *src, size_t n, float *sink) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const f4 a = ((const gf4 *)src)[i];
    // data-dependent store the compiler cannot drop (never taken for finite data)
    if (a.x + a.y + a.z + a.w == __builtin_inff()) sink[0] = 1.f;
}

template <bool NT>
__global__ __launch_bounds__(256) void k_write(f4 *dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const f4 v = {1.f, 2.f, 3.f, (float)threadIdx.x};
    put<NT>((gf4 *)dst + i, v);
}

template <bool NT>
__global__ __launch_bounds__(256) void k_copy(const f4 *src, f4 *dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    put<NT>((gf4 *)dst + i, ((const gf4 *)src)[i]);
}

// reads n elements, writes 3 n (three coalesced output streams)
template <bool NT>
__global__ __launch_bounds__(256) void k_mix13(const f4 *src, f4 *dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    gf4 *d = (gf4 *)dst;
    const f4 a = ((const gf4 *)src)[i];
    put<NT>(d + i, a);
    put<NT>(d + n + i, a * 2.f);
    put<NT>(d + 2 * n + i, a * 3.f);
}

}  // namespace

extern "C" {

// kind 0: read `bytes` from src; 1: write `bytes` to dst; 2: copy `bytes`
// src -> dst; 3: read `bytes` from src and write 3 x `bytes` to dst.
// bytes: a multiple of 16; dst must hold 3 x bytes for kind 3.  sink: one
// float of device memory (kind 0).  nontemporal: 1 = stores marked
// non-temporal, 0 = plain stores.  Returns 0, or -1 on a bad argument or
// launch error.
int tp_probe_hbm(int kind, const void *src, void *dst, float *sink, size_t bytes, int nontemporal,
                 void *stream) {
    if (bytes == 0 || bytes % 16 || (nontemporal & ~1)) return -1;
    if ((kind == 0 && (!src || !sink)) || (kind == 1 && !dst) || ((kind == 2 || kind == 3) && (!src || !dst)))
        return -1;
    const size_t n = bytes / 16;
    if ((n + 255) / 256 > 0x7fffffffu) return -1;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    hipStream_t st = (hipStream_t)stream;
    const bool nt = nontemporal != 0;
    switch (kind) {
        case 0: hipLaunchKernelGGL(k_read, grid, block, 0, st, (const f4 *)src, n, sink); break;
        case 1:
            if (nt) hipLaunchKernelGGL(k_write<true>, grid, block, 0, st, (f4 *)dst, n);
            else hipLaunchKernelGGL(k_write<false>, grid, block, 0, st, (f4 *)dst, n);
            break;
        case 2:
            if (nt) hipLaunchKernelGGL(k_copy<true>, grid, block, 0, st, (const f4 *)src, (f4 *)dst, n);
            else hipLaunchKernelGGL(k_copy<false>, grid, block, 0, st, (const f4 *)src, (f4 *)dst, n);
            break;
        case 3:
            if (nt) hipLaunchKernelGGL(k_mix13<true>, grid, block, 0, st, (const f4 *)src, (f4 *)dst, n);
            else hipLaunchKernelGGL(k_mix13<false>, grid, block, 0, st, (const f4 *)src, (f4 *)dst, n);
            break;
        default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
