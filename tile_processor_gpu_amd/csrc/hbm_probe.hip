// hbm_probe.hip - streaming kernels that measure what this GPU's HBM
// delivers in practice for pure reads, pure writes, copies and the forward
// CLT's mix (one byte read per three written).  The benchmark divides its
// kernels' measured traffic by these rates to tell how close they run to the
// memory system's practical ceiling (the 8 TB/s peak is not reachable by
// any access pattern).  Built into libtp_probe.so, separate from the
// library's own entry points.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) f4 gf4;

// Every kernel is a grid-stride loop over float4 elements.
__global__ __launch_bounds__(256) void k_read(const f4 *src, size_t n, float *sink) {
    const gf4 *s = (const gf4 *)src;
    const size_t st = (size_t)gridDim.x * blockDim.x;
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (; i + 3 * st < n; i += 4 * st) {
        const f4 a = s[i], b = s[i + st], c = s[i + 2 * st], d = s[i + 3 * st];
        acc += (a + b) + (c + d);
    }
    for (; i < n; i += st) acc += s[i];
    // data-dependent store the compiler cannot drop (never taken for finite data)
    if (acc.x + acc.y + acc.z + acc.w == __builtin_inff()) sink[0] = 1.f;
}

__global__ __launch_bounds__(256) void k_write(f4 *dst, size_t n) {
    gf4 *d = (gf4 *)dst;
    const size_t st = (size_t)gridDim.x * blockDim.x;
    const f4 v = {1.f, 2.f, 3.f, (float)threadIdx.x};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += st)
        __builtin_nontemporal_store(v, d + i);
}

__global__ __launch_bounds__(256) void k_copy(const f4 *src, f4 *dst, size_t n) {
    const gf4 *s = (const gf4 *)src;
    gf4 *d = (gf4 *)dst;
    const size_t st = (size_t)gridDim.x * blockDim.x;
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    for (; i + st < n; i += 2 * st) {
        const f4 a = s[i], b = s[i + st];
        __builtin_nontemporal_store(a, d + i);
        __builtin_nontemporal_store(b, d + i + st);
    }
    for (; i < n; i += st) __builtin_nontemporal_store(s[i], d + i);
}

// reads n elements, writes 3 n (three coalesced output streams)
__global__ __launch_bounds__(256) void k_mix13(const f4 *src, f4 *dst, size_t n) {
    const gf4 *s = (const gf4 *)src;
    gf4 *d = (gf4 *)dst;
    const size_t st = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += st) {
        const f4 a = s[i];
        __builtin_nontemporal_store(a, d + i);
        __builtin_nontemporal_store(a * 2.f, d + n + i);
        __builtin_nontemporal_store(a * 3.f, d + 2 * n + i);
    }
}

}  // namespace

extern "C" {

// kind 0: read `bytes` from src; 1: write `bytes` to dst; 2: copy `bytes`
// src -> dst; 3: read `bytes` from src and write 3 x `bytes` to dst.
// bytes: a multiple of 16; dst must hold 3 x bytes for kind 3.  sink: one
// float of device memory (kind 0).  blocks_per_cu: workgroups (256 threads)
// per CU.  Returns 0, or -1 on a bad argument or launch error.
int tp_probe_hbm(int kind, const void *src, void *dst, float *sink, size_t bytes, int blocks_per_cu,
                 void *stream) {
    if (bytes == 0 || bytes % 16 || blocks_per_cu < 1 || blocks_per_cu > 64) return -1;
    if ((kind == 0 && (!src || !sink)) || (kind == 1 && !dst) || ((kind == 2 || kind == 3) && (!src || !dst)))
        return -1;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
    const size_t n = bytes / 16;
    const dim3 grid(cus * blocks_per_cu), block(256);
    hipStream_t st = (hipStream_t)stream;
    switch (kind) {
        case 0: hipLaunchKernelGGL(k_read, grid, block, 0, st, (const f4 *)src, n, sink); break;
        case 1: hipLaunchKernelGGL(k_write, grid, block, 0, st, (f4 *)dst, n); break;
        case 2: hipLaunchKernelGGL(k_copy, grid, block, 0, st, (const f4 *)src, (f4 *)dst, n); break;
        case 3: hipLaunchKernelGGL(k_mix13, grid, block, 0, st, (const f4 *)src, (f4 *)dst, n); break;
        default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
