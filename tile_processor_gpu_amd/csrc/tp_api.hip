// tp_api.hip - extern "C" entry points of libtp.so (see include/tp.h).
//
// The library is one translation unit: the kernel headers are included here
// so the __constant__ tables are shared without relocatable device code.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "tp_common.hpp"
#include "k_convert.hpp"
#include "k_correlate.hpp"
#include "k_imclt.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, const char *what = "", long v = 0) {
    char buf[256];
    snprintf(buf, sizeof(buf), fmt, what, v);
    g_err = buf;
    return code;
}

int check_launch(const char *kernel) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        char buf[256];
        snprintf(buf, sizeof(buf), "%s: launch failed: %s", kernel, hipGetErrorString(e));
        g_err = buf;
        return TP_ELAUNCH;
    }
    return TP_OK;
}

int check_geometry(const tp_geometry *g) {
    if (!g) return fail(TP_EINVAL, "geometry is NULL%s", "");
    if (g->img_width <= 0 || g->img_height <= 0 || (g->img_width % 16) || (g->img_height % 16))
        return fail(TP_EINVAL, "image size must be positive multiples of 16%s", "");
    if (g->kernels_hor <= 0 || g->kernels_vert <= 0)
        return fail(TP_EINVAL, "kernel grid must be positive%s", "");
    return TP_OK;
}

inline hipStream_t as_stream(tp_stream_t s) { return (hipStream_t)s; }

}  // namespace

extern "C" {

const char *tp_last_error(void) { return g_err.c_str(); }

const char *tp_version(void) { return "tp-mi355x 0.1 (gfx950)"; }

int tp_set_lpf(int which, const float *host_lpf) {
    if (which < 0 || which > 5) return fail(TP_EINVAL, "tp_set_lpf: which must be 0..5%s", "");
    const float *src = host_lpf ? host_lpf : tp::kTab.lpf[which];  // NULL: built-in default
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(tp::c_lpf), src, 64 * sizeof(float),
                                     (size_t)which * 64 * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(TP_EHIP, "tp_set_lpf: %s", hipGetErrorString(e));
    g_err.clear();
    return TP_OK;
}

int tp_convert_correct_tiles(const tp_geometry *geom, float **gpu_kernel_offsets,
                             float **gpu_kernels, float **gpu_images, const tp_task *gpu_tasks,
                             float **gpu_clt, size_t dstride, int num_tiles, int lpf_mask,
                             tp_stream_t stream) {
    int rc = check_geometry(geom);
    if (rc) return rc;
    if (!gpu_kernel_offsets || !gpu_kernels || !gpu_images || !gpu_tasks || !gpu_clt)
        return fail(TP_EINVAL, "tp_convert_correct_tiles: NULL pointer argument%s", "");
    if (num_tiles < 0) return fail(TP_EINVAL, "tp_convert_correct_tiles: num_tiles < 0%s", "");
    if (dstride < (size_t)geom->img_width)
        return fail(TP_EINVAL, "tp_convert_correct_tiles: dstride < image width%s", "");
    if (lpf_mask & ~7) return fail(TP_EINVAL, "tp_convert_correct_tiles: lpf_mask has bits above 2%s", "");
    g_err.clear();
    if (num_tiles == 0) return TP_OK;
    tp::ConvertArgs a;
    a.g = tp::make_geo(*geom);
    a.koffs = (const float *const *)gpu_kernel_offsets;
    a.kernels = (const float *const *)gpu_kernels;
    a.images = (const float *const *)gpu_images;
    a.tasks = gpu_tasks;
    a.clt = gpu_clt;
    a.dstride = dstride;
    a.num_tiles = num_tiles;
    a.lpf_mask = lpf_mask;
    a.pairs_per_block = 4;
    const int npairs = (num_tiles + 1) / 2;
    const int grid = (npairs + a.pairs_per_block - 1) / a.pairs_per_block;
    hipLaunchKernelGGL(tp::k_convert_correct_tiles, dim3(grid), dim3(192), 0, as_stream(stream), a);
    return check_launch("convert_correct_tiles");
}

int tp_correlate2D(const tp_geometry *geom, float **gpu_clt, int colors, float scale0,
                   float scale1, float scale2, float fat_zero, size_t num_corr_tiles,
                   const int *gpu_corr_indices, size_t corr_stride, int corr_radius,
                   float *gpu_corrs, tp_stream_t stream) {
    int rc = check_geometry(geom);
    if (rc) return rc;
    if (!gpu_clt || !gpu_corr_indices || !gpu_corrs)
        return fail(TP_EINVAL, "tp_correlate2D: NULL pointer argument%s", "");
    if (colors != 1 && colors != 3) return fail(TP_EINVAL, "tp_correlate2D: colors must be 1 or 3%s", "");
    if (corr_radius < 0 || corr_radius > TP_MAX_CORR_RADIUS)
        return fail(TP_EINVAL, "tp_correlate2D: corr_radius must be 0..7%s", "");
    const size_t n = 2 * corr_radius + 1;
    if (corr_stride < n * n) return fail(TP_EINVAL, "tp_correlate2D: corr_stride < (2r+1)^2%s", "");
    if (num_corr_tiles > (size_t)0x7fffffff) return fail(TP_EINVAL, "tp_correlate2D: too many entries%s", "");
    g_err.clear();
    if (num_corr_tiles == 0) return TP_OK;
    tp::CorrArgs a;
    a.clt = gpu_clt;
    a.colors = colors;
    a.scale[0] = scale0;
    a.scale[1] = scale1;
    a.scale[2] = scale2;
    a.fat_zero = fat_zero;
    a.num_corr = (int)num_corr_tiles;
    a.idx = gpu_corr_indices;
    a.stride = corr_stride;
    a.radius = corr_radius;
    a.corrs = gpu_corrs;
    a.ntiles = (geom->img_width / 8) * (geom->img_height / 8);
    a.entries_per_block = 64;
    const int grid = (int)((num_corr_tiles + a.entries_per_block - 1) / a.entries_per_block);
    hipLaunchKernelGGL(tp::k_correlate2D, dim3(grid), dim3(256), 0, as_stream(stream), a);
    return check_launch("correlate2D");
}

int tp_imclt_rbg(const tp_geometry *geom, const float *gpu_clt, float *gpu_rbg, int color,
                 int v_offset, int h_offset, size_t dstride, tp_stream_t stream) {
    int rc = check_geometry(geom);
    if (rc) return rc;
    if (!gpu_clt || !gpu_rbg) return fail(TP_EINVAL, "tp_imclt_rbg: NULL pointer argument%s", "");
    if (color < 0 || color >= TP_NUM_COLORS) return fail(TP_EINVAL, "tp_imclt_rbg: color must be 0..2%s", "");
    if ((v_offset & ~1) || (h_offset & ~1))
        return fail(TP_EINVAL, "tp_imclt_rbg: offsets must be 0 or 1%s", "");
    if (dstride < (size_t)geom->img_width + 8)
        return fail(TP_EINVAL, "tp_imclt_rbg: dstride < image width + 8%s", "");
    g_err.clear();
    tp::Geo g = tp::make_geo(*geom);
    const int th = (g.TX + (h_offset ^ 1)) >> 1;
    const int tv = (g.TY + (v_offset ^ 1)) >> 1;
    const int ntiles = th * tv;
    const int grid = (ntiles + 7) / 8;
    hipLaunchKernelGGL(tp::k_imclt_pass, dim3(grid), dim3(64), 0, as_stream(stream), g, gpu_clt,
                       gpu_rbg, color, v_offset, h_offset, dstride);
    return check_launch("imclt_rbg");
}

int tp_imclt_rbg_all(const tp_geometry *geom, float **gpu_clt, float **gpu_rbg, size_t dstride,
                     tp_stream_t stream) {
    int rc = check_geometry(geom);
    if (rc) return rc;
    if (!gpu_clt || !gpu_rbg) return fail(TP_EINVAL, "tp_imclt_rbg_all: NULL pointer argument%s", "");
    if (dstride < (size_t)geom->img_width + 8)
        return fail(TP_EINVAL, "tp_imclt_rbg_all: dstride < image width + 8%s", "");
    g_err.clear();
    tp::Geo g = tp::make_geo(*geom);
    // output-stationary: one workgroup per 64x64 block of one plane
    const int bx = (g.W + 8 + 63) / 64, by = (g.H + 8 + 63) / 64;
    dim3 grid(bx * by, TP_NUM_COLORS, TP_NUM_CAMS);
    hipLaunchKernelGGL(tp::k_imclt_frame, grid, dim3(256), 0, as_stream(stream), g,
                       (const float *const *)gpu_clt, gpu_rbg, dstride);
    return check_launch("imclt_rbg_all");
}

}  // extern "C"
