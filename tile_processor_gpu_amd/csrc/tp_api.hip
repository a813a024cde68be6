// tp_api.hip - extern "C" entry points of libtp.so (see include/tp.h).
//
// The library is one translation unit: the kernel headers are included here
// so the __constant__ tables are shared without relocatable device code.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdio>
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "tp_common.hpp"
#include "k_convert.hpp"
#include "k_correlate.hpp"
#include "k_imclt.hpp"
#include "k_texture.hpp"
#include "k_combine.hpp"
#include "k_fused.hpp"
#include "k_nn.hpp"
#include "k_dtt.hpp"
#include "k_psf.hpp"

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

static_assert(sizeof(tp_frame) == sizeof(tp::FrameDesc), "tp_frame layout");
static_assert(TP_NN_PARAMS == tp::NN_NPARAMS && TP_NN_ACTS == tp::NN_ACTS &&
              TP_NN_INPUT == tp::NN_IN * tp::NN_P, "CNN sizes");
static_assert(offsetof(tp_frame, num_items) == offsetof(tp::FrameDesc, num_items), "tp_frame layout");

struct FusedCfg {
    int tpi, gt, bpc, dct;
};

int fused_config(const char *fn, const tp_fused_config *c, FusedCfg &o) {
    o = {12, 4, 4, 0};
    if (c) o = {c->tiles_per_item, c->tiles_per_group, c->blocks_per_cu, c->dct};
    if (o.gt != 2 && o.gt != 4) return fail(TP_EINVAL, "%s: tiles_per_group must be 2 or 4", fn);
    if (o.tpi < o.gt || o.tpi % o.gt || o.tpi > 4096)
        return fail(TP_EINVAL, "%s: tiles_per_item must be a multiple of tiles_per_group, <= 4096", fn);
    if (o.bpc < 1 || o.bpc > 8) return fail(TP_EINVAL, "%s: blocks_per_cu must be 1..8", fn);
    if (o.dct < 0 || o.dct > 3) return fail(TP_EINVAL, "%s: dct must be 0..3", fn);
    return TP_OK;
}

int fused_common(const char *fn, const tp_geometry *geom, size_t dstride, int lpf_mask, int colors,
                 size_t corr_stride, int corr_radius) {
    int rc = check_geometry(geom);
    if (rc) return rc;
    if (dstride < (size_t)geom->img_width) return fail(TP_EINVAL, "%s: dstride < image width", fn);
    if (lpf_mask & ~7) return fail(TP_EINVAL, "%s: lpf_mask has bits above 2", fn);
    if (colors != 1 && colors != 3) return fail(TP_EINVAL, "%s: colors must be 1 or 3", fn);
    if (corr_radius < 0 || corr_radius > TP_MAX_CORR_RADIUS)
        return fail(TP_EINVAL, "%s: corr_radius must be 0..7", fn);
    const size_t n = 2 * corr_radius + 1;
    if (corr_stride < n * n) return fail(TP_EINVAL, "%s: corr_stride < (2r+1)^2", fn);
    return TP_OK;
}

tp::FusedArgs fused_args(const tp_geometry *geom, size_t dstride, int lpf_mask, float s0, float s1,
                         float s2, float fat_zero, size_t corr_stride, int corr_radius, int tpi,
                         size_t combined_stride, int pair_mask) {
    tp::FusedArgs a = {};
    a.g = tp::make_geo(*geom);
    a.tiles_per_item = tpi;
    a.dstride = dstride;
    a.lpf_mask = lpf_mask;
    a.scale[0] = s0;
    a.scale[1] = s1;
    a.scale[2] = s2;
    a.fat_zero = fat_zero;
    a.corr_stride = corr_stride;
    a.radius = corr_radius;
    a.combined_stride = combined_stride;
    a.pair_mask = pair_mask;
    return a;
}

template <bool Q, int DCT>
void launch_fused_d(const tp::FusedArgs &a, const FusedCfg &c, int colors, int grid, hipStream_t st) {
    const dim3 gd(grid), bd(64 * TP_FUSED_WG_WAVES);
    if (c.gt == 2) {
        if (colors == 3) hipLaunchKernelGGL((tp::k_fused<2, Q, 3, DCT>), gd, bd, 0, st, a);
        else hipLaunchKernelGGL((tp::k_fused<2, Q, 1, DCT>), gd, bd, 0, st, a);
    } else {
        if (colors == 3) hipLaunchKernelGGL((tp::k_fused<4, Q, 3, DCT>), gd, bd, 0, st, a);
        else hipLaunchKernelGGL((tp::k_fused<4, Q, 1, DCT>), gd, bd, 0, st, a);
    }
}

template <bool Q>
void launch_fused(const tp::FusedArgs &a, const FusedCfg &c, int colors, int grid, hipStream_t st) {
    if (c.dct == 1) launch_fused_d<Q, 1>(a, c, colors, grid, st);
    else if (c.dct == 2) launch_fused_d<Q, 2>(a, c, colors, grid, st);
    else if (c.dct == 3) launch_fused_d<Q, 3>(a, c, colors, grid, st);
    else launch_fused_d<Q, 0>(a, c, colors, grid, st);
}

int num_cus() {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
    return n;
}

// Enqueue the stream form (set-up kernel + persistent kernel) on st.
int enqueue_stream(const tp_geometry *geom, tp_frame *gpu_frames, int nframes, int32_t *gpu_work,
                   uint64_t *gpu_stamps, size_t dstride, int lpf_mask, int colors, float s0,
                   float s1, float s2, float fat_zero, size_t corr_stride, int corr_radius,
                   size_t combined_stride, int pair_mask, const tp_fused_config *config,
                   hipStream_t st, const char *fn) {
    int rc = fused_common(fn, geom, dstride, lpf_mask, colors, corr_stride, corr_radius);
    if (rc) return rc;
    const size_t nn = 2 * corr_radius + 1;
    if (combined_stride < nn * nn) return fail(TP_EINVAL, "%s: combined_stride < (2r+1)^2", fn);
    if (!gpu_frames || !gpu_work) return fail(TP_EINVAL, "%s: NULL pointer argument", fn);
    if (nframes <= 0) return fail(TP_EINVAL, "%s: nframes must be positive", fn);
    FusedCfg c;
    if ((rc = fused_config(fn, config, c))) return rc;
    const int cus = num_cus();
    if (cus <= 0) return fail(TP_EHIP, "%s: cannot query the device", fn);
    tp::FusedArgs a = fused_args(geom, dstride, lpf_mask, s0, s1, s2, fat_zero, corr_stride,
                                 corr_radius, c.tpi, combined_stride, pair_mask);
    a.frames = (const tp::FrameDesc *)gpu_frames;
    a.nframes = nframes;
    a.work = gpu_work;
    a.stamps = (unsigned long long *)gpu_stamps;
    hipLaunchKernelGGL(tp::k_stream_setup, dim3(1), dim3(64), 0, st, (tp::FrameDesc *)gpu_frames,
                       nframes, c.tpi, gpu_work, a.stamps);
    if ((rc = check_launch("stream_setup"))) return rc;
    launch_fused<true>(a, c, colors, cus * c.bpc, st);
    return check_launch("fused (stream)");
}

}  // namespace

// Work buffer of tp_stream_convert_correlate_noalloc: queue counters, item
// count and per-queue item prefixes of up to TP_STREAM_NOALLOC_MAX_FRAMES
// frames (132 KB), one copy per GPU for the life of the process.  In host
// code this name is not a device address; host code uses the address that
// hipGetSymbolAddress returns for the current device (scratch_ptr).
static __device__ __align__(256) int32_t scratch[TP_STREAM_WORK_INTS(TP_STREAM_NOALLOC_MAX_FRAMES)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t *g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call on that device and kept.
int scratch_ptr(const char *fn, int32_t **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(TP_EHIP, "%s: hipGetDevice failed", fn);
    if (dev < 0 || dev >= kMaxDevices) return fail(TP_EHIP, "%s: device index out of range", fn);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            char buf[256];
            snprintf(buf, sizeof(buf), "%s: hipGetSymbolAddress of the work buffer: %s", fn,
                     hipGetErrorString(e));
            g_err = buf;
            return TP_EHIP;
        }
        g_scratch[dev] = (int32_t *)p;
    }
    *out = g_scratch[dev];
    return TP_OK;
}

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
    tp::ConvertArgs a{};
    a.g = tp::make_geo(*geom);
    a.koffs = (const float *const *)gpu_kernel_offsets;
    a.kernels = (const float *const *)gpu_kernels;
    a.images = (const float *const *)gpu_images;
    a.tasks = gpu_tasks;
    a.clt = gpu_clt;
    a.dstride = dstride;
    a.num_tiles = num_tiles;
    a.lpf_mask = lpf_mask;
    // one workgroup (4 wavefronts) per camera and patch of 8 x 4 list
    // entries, camera-major
    const long long grid = 4LL * tp::convert_blocks(num_tiles, a.g.TX);
    if (grid > 0x7fffffffLL) return fail(TP_EINVAL, "tp_convert_correct_tiles: too many tiles%s", "");
    hipLaunchKernelGGL(tp::k_convert_cam, dim3((unsigned)grid), dim3(256), 0, as_stream(stream), a);
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
    a.entries_per_block = TP_CORR_EPB;
    const int grid = (int)((num_corr_tiles + a.entries_per_block - 1) / a.entries_per_block);
    if (colors == 3)
        hipLaunchKernelGGL(tp::k_correlate2D<3>, dim3(grid), dim3(64 * TP_CORR_WG_WAVES), 0,
                           as_stream(stream), a);
    else
        hipLaunchKernelGGL(tp::k_correlate2D<1>, dim3(grid), dim3(64 * TP_CORR_WG_WAVES), 0,
                           as_stream(stream), a);
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
    // strips of 15 column groups x bands of IMS_BAND row groups, each plane
    constexpr int band = tp::IMS_BAND;
    dim3 grid((g.TX + 1 + tp::IMS_OWN - 1) / tp::IMS_OWN, (g.TY + 1 + band - 1) / band,
              TP_NUM_COLORS * TP_NUM_CAMS);
    hipLaunchKernelGGL(tp::k_imclt_strip, grid, dim3(256), 0, as_stream(stream), g,
                       (const float *const *)gpu_clt, gpu_rbg, dstride, band);
    return check_launch("imclt_rbg_all");
}

}  // extern "C"

namespace {

// n: list length or its cap
int launch_textures(const tp::TexArgs &a, int n, hipStream_t st) {
    if (n <= 0) return TP_OK;
    // workgroups loop over the list, TP_TEX_WG_PER_CU per CU
    const int cus = num_cus();
    const int grid = std::min((n + TP_TEX_TPW - 1) / TP_TEX_TPW, (cus > 0 ? cus : 256) * TP_TEX_WG_PER_CU);
    if (a.colors == 3 && a.keep_weights)
        hipLaunchKernelGGL((tp::k_textures_accumulate<3, TP_TEX_TPW, true>), dim3(grid), dim3(256), 0, st, a);
    else if (a.colors == 3)
        hipLaunchKernelGGL((tp::k_textures_accumulate<3, TP_TEX_TPW, false>), dim3(grid), dim3(256), 0, st, a);
    else if (a.keep_weights)
        hipLaunchKernelGGL((tp::k_textures_accumulate<1, TP_TEX_TPW, true>), dim3(grid), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((tp::k_textures_accumulate<1, TP_TEX_TPW, false>), dim3(grid), dim3(256), 0, st, a);
    return check_launch("textures_accumulate");
}

tp::TexArgs tex_args(const tp_geometry *geom, const int *woi, float **clt, const float *ports,
                     int colors, float min_shot, float scale_shot, float diff_sigma,
                     float diff_threshold, float min_agree, float w0, float w1, float w2,
                     int dust_remove, int keep_weights) {
    tp::TexArgs a;
    a.g = tp::make_geo(*geom);
    a.woi = woi;
    a.clt = (const float *const *)clt;
    a.n = 0;
    a.quarter = 0;
    a.num_ptr = nullptr;
    a.idx = nullptr;
    a.ports = ports;
    a.colors = colors;
    a.min_shot = min_shot;
    a.scale_shot = scale_shot;
    a.diff_sigma = diff_sigma;
    a.diff_threshold = diff_threshold;
    a.min_agree = min_agree;
    a.w[0] = w0;
    a.w[1] = w1;
    a.w[2] = w2;
    a.dust_remove = dust_remove;
    a.keep_weights = keep_weights;
    a.first_pass = 0;
    a.rbga_stride = 0;
    a.rbga = nullptr;
    a.tex_stride = 0;
    a.tiles = nullptr;
    return a;
}

int check_tex_params(int colors, float min_shot, float diff_sigma) {
    if (colors != 1 && colors != 3) return fail(TP_EINVAL, "textures: colors must be 1 or 3%s", "");
    if (!(min_shot > 0.f)) return fail(TP_EINVAL, "textures: min_shot must be > 0%s", "");
    if (!(diff_sigma > 0.f)) return fail(TP_EINVAL, "textures: diff_sigma must be > 0%s", "");
    return TP_OK;
}

// the list-building part of prepare_texture_list / generate_RBGA
int enqueue_texture_list(const tp_geometry *geom, tp_task *tasks, int num_tiles, int *indices,
                         int *num, int *woi, int width, int height, hipStream_t st) {
    tp::Geo g = tp::make_geo(*geom);
    const int nclr = std::max(8, std::min(width, g.TX) * height);
    hipLaunchKernelGGL(tp::k_tex_clear, dim3((nclr + 255) / 256), dim3(256), 0, st, indices, width,
                       height, g.TX, g.TY, woi, num);
    int rc = check_launch("clear_texture_list");
    if (rc) return rc;
    if (num_tiles > 0) {
        const int grid = (num_tiles + 255) / 256;
        hipLaunchKernelGGL(tp::k_tex_mark, dim3(grid), dim3(256), 0, st, tasks, num_tiles, indices,
                           g.TX, g.TY);
        if ((rc = check_launch("mark_texture_tiles"))) return rc;
        hipLaunchKernelGGL(tp::k_tex_neighbors,
                           dim3((num_tiles + TP_NEIGHBOR_THREADS - 1) / TP_NEIGHBOR_THREADS),
                           dim3(TP_NEIGHBOR_THREADS), 0, st, tasks, num_tiles, indices, woi, g.TX, g.TY);
        if ((rc = check_launch("mark_texture_neighbor_tiles"))) return rc;
        hipLaunchKernelGGL(tp::k_tex_gen_list,
                           dim3((num_tiles + TP_NEIGHBOR_THREADS - 1) / TP_NEIGHBOR_THREADS),
                           dim3(TP_NEIGHBOR_THREADS), 0, st, tasks, num_tiles, indices, num, woi, g.TX,
                           g.TY, g.TYA);
        if ((rc = check_launch("gen_texture_list"))) return rc;
    }
    hipLaunchKernelGGL(tp::k_tex_finish_woi, dim3(1), dim3(64), 0, st, woi);
    return check_launch("texture woi");
}

int check_list_args(const tp_geometry *geom, const tp_task *tasks, int num_tiles, const int *indices,
                    const int *num, const int *woi, int width, int height) {
    int rc = check_geometry(geom);
    if (rc) return rc;
    if (!tasks || !indices || !num || !woi)
        return fail(TP_EINVAL, "texture list: NULL pointer argument%s", "");
    if (num_tiles < 0) return fail(TP_EINVAL, "texture list: num_tiles < 0%s", "");
    if (width < 0 || height < 0 || height > geom->img_height / 8)
        return fail(TP_EINVAL, "texture list: width/height out of range%s", "");
    return TP_OK;
}

}  // namespace

extern "C" {

int tp_textures_accumulate(const tp_geometry *geom, const int *woi, float **gpu_clt,
                           size_t num_texture_tiles, const int *gpu_texture_indices,
                           const float *gpu_port_offsets, int colors, int is_lwir, float min_shot,
                           float scale_shot, float diff_sigma, float diff_threshold,
                           float min_agree, float weight0, float weight1, float weight2,
                           int dust_remove, int keep_weights, size_t texture_rbg_stride,
                           float *gpu_texture_rbg, size_t texture_stride,
                           float *gpu_texture_tiles, tp_stream_t stream) {
    (void)is_lwir;  // unused by the reference kernel too
    int rc = check_geometry(geom);
    if (rc) return rc;
    if ((rc = check_tex_params(colors, min_shot, diff_sigma))) return rc;
    if (!gpu_clt || !gpu_texture_indices || !gpu_port_offsets)
        return fail(TP_EINVAL, "tp_textures_accumulate: NULL pointer argument%s", "");
    const bool want_rbga = gpu_texture_rbg && texture_rbg_stride;
    const bool want_tiles = gpu_texture_tiles && texture_stride;
    if (want_rbga && !woi) return fail(TP_EINVAL, "tp_textures_accumulate: RGBA output needs woi%s", "");
    const size_t nsl = (size_t)(keep_weights ? 2 * colors + 6 : colors + 1);
    if (want_tiles && texture_stride < nsl * 256)
        return fail(TP_EINVAL, "tp_textures_accumulate: texture_stride too small%s", "");
    if (num_texture_tiles > (size_t)0x7fffffff) return fail(TP_EINVAL, "tp_textures_accumulate: too many tiles%s", "");
    g_err.clear();
    tp::TexArgs a = tex_args(geom, woi, gpu_clt, gpu_port_offsets, colors, min_shot, scale_shot,
                             diff_sigma, diff_threshold, min_agree, weight0, weight1, weight2,
                             dust_remove, keep_weights);
    a.n = (int)num_texture_tiles;
    a.idx = gpu_texture_indices;
    if (want_rbga) {
        a.rbga_stride = texture_rbg_stride;
        a.rbga = gpu_texture_rbg;
    }
    if (want_tiles) {
        a.tex_stride = texture_stride;
        a.tiles = gpu_texture_tiles;
    }
    return launch_textures(a, (int)num_texture_tiles, as_stream(stream));
}

int tp_prepare_texture_list(const tp_geometry *geom, tp_task *gpu_tasks, int num_tiles,
                            int *gpu_texture_indices, int *num_texture_tiles, int *woi, int width,
                            int height, tp_stream_t stream) {
    int rc = check_list_args(geom, gpu_tasks, num_tiles, gpu_texture_indices, num_texture_tiles, woi,
                             width, height);
    if (rc) return rc;
    g_err.clear();
    return enqueue_texture_list(geom, gpu_tasks, num_tiles, gpu_texture_indices, num_texture_tiles,
                                woi, width, height, as_stream(stream));
}

int tp_generate_RBGA(const tp_geometry *geom, tp_task *gpu_tasks, int num_tiles,
                     int *gpu_texture_indices, int *num_texture_tiles, int *woi, int width,
                     int height, float **gpu_clt, const float *gpu_port_offsets, int colors,
                     int is_lwir, float min_shot, float scale_shot, float diff_sigma,
                     float diff_threshold, float min_agree, float weight0, float weight1,
                     float weight2, int dust_remove, int keep_weights, size_t texture_rbga_stride,
                     float *gpu_texture_tiles, tp_stream_t stream) {
    (void)is_lwir;
    (void)keep_weights;  // the reference's passes run with keep_weights = 0
    int rc = check_list_args(geom, gpu_tasks, num_tiles, gpu_texture_indices, num_texture_tiles, woi,
                             width, height);
    if (rc) return rc;
    if ((rc = check_tex_params(colors, min_shot, diff_sigma))) return rc;
    if (!gpu_clt || !gpu_port_offsets || !gpu_texture_tiles)
        return fail(TP_EINVAL, "tp_generate_RBGA: NULL pointer argument%s", "");
    if (texture_rbga_stride < (size_t)(geom->img_width / 8 + 1) * 8)
        return fail(TP_EINVAL, "tp_generate_RBGA: texture_rbga_stride < (tilesx + 1) * 8%s", "");
    g_err.clear();
    hipStream_t st = as_stream(stream);
    if ((rc = enqueue_texture_list(geom, gpu_tasks, num_tiles, gpu_texture_indices,
                                   num_texture_tiles, woi, width, height, st)))
        return rc;
    tp::Geo g = tp::make_geo(*geom);
    const int slices = colors + 1;
    const int max_rows = (g.TY + 1) * 8 * slices;
    const dim3 cgrid((max_rows + TP_CLEAR_ROWS - 1) / TP_CLEAR_ROWS);
    if (texture_rbga_stride % 4 == 0 && (uintptr_t)gpu_texture_tiles % 16 == 0)
        hipLaunchKernelGGL(tp::k_tex_clear_rbga<true>, cgrid, dim3(256), 0, st, gpu_texture_tiles,
                           texture_rbga_stride, woi, slices);
    else
        hipLaunchKernelGGL(tp::k_tex_clear_rbga<false>, cgrid, dim3(256), 0, st, gpu_texture_tiles,
                           texture_rbga_stride, woi, slices);
    if ((rc = check_launch("clear_texture_rbga"))) return rc;
    tp::TexArgs a = tex_args(geom, woi, gpu_clt, gpu_port_offsets, colors, min_shot, scale_shot,
                             diff_sigma, diff_threshold, min_agree, weight0, weight1, weight2,
                             dust_remove, 0);
    a.rbga_stride = texture_rbga_stride;
    a.rbga = gpu_texture_tiles;
    // The reference runs 8 passes, the inner then the border list of each
    // tile parity (:1562-1610).  Tiles of one parity never overlap, inner or
    // border, so one launch per parity takes both of its lists (inner from
    // the start of its quarter, border grown down from the end, counts read
    // on the device); only the order of the overlap-add sums differs.
    const int quarter = g.TX * (g.TYA >> 2);
    for (int par = 0; par < 4; par++) {
        tp::TexArgs p = a;
        p.num_ptr = num_texture_tiles + 2 * par;  // [inner, border]
        p.idx = gpu_texture_indices + par * quarter;
        p.quarter = quarter;
        p.n = quarter;
        p.first_pass = par == 0;
        if ((rc = launch_textures(p, quarter, st))) return rc;
    }
    return TP_OK;
}

}  // extern "C"

extern "C" {

int tp_corr_combine(const tp_geometry *geom, size_t num_corr_tiles, const int *gpu_corr_indices,
                    const float *gpu_corrs, size_t corr_stride, int corr_radius, int pair_mask,
                    float *gpu_combined, size_t combined_stride, tp_stream_t stream) {
    int rc = check_geometry(geom);
    if (rc) return rc;
    if (!gpu_corr_indices || !gpu_corrs || !gpu_combined)
        return fail(TP_EINVAL, "tp_corr_combine: NULL pointer argument%s", "");
    if (corr_radius < 0 || corr_radius > TP_MAX_CORR_RADIUS)
        return fail(TP_EINVAL, "tp_corr_combine: corr_radius must be 0..7%s", "");
    const size_t n = 2 * corr_radius + 1;
    if (corr_stride < n * n || combined_stride < n * n)
        return fail(TP_EINVAL, "tp_corr_combine: stride < (2r+1)^2%s", "");
    if (num_corr_tiles > (size_t)0x7fffffff) return fail(TP_EINVAL, "tp_corr_combine: too many entries%s", "");
    g_err.clear();
    if (!num_corr_tiles) return TP_OK;
    tp::CombineArgs a;
    a.num_corr = (int)num_corr_tiles;
    a.idx = gpu_corr_indices;
    a.corrs = gpu_corrs;
    a.stride = corr_stride;
    a.radius = corr_radius;
    a.pair_mask = pair_mask;
    a.ntiles = (geom->img_width / 8) * (geom->img_height / 8);
    a.out = gpu_combined;
    a.out_stride = combined_stride;
    hipLaunchKernelGGL(tp::k_corr_combine, dim3((a.num_corr + TP_COMBINE_EPW - 1) / TP_COMBINE_EPW), dim3(256), 0,
                       as_stream(stream), a);
    return check_launch("corr_combine");
}

int tp_corr_disparity(const float *gpu_tiles, size_t stride, int num_tiles, int corr_radius,
                      int mode, float *gpu_out, tp_stream_t stream) {
    if (!gpu_tiles || !gpu_out) return fail(TP_EINVAL, "tp_corr_disparity: NULL pointer argument%s", "");
    if (corr_radius < 1 || corr_radius > TP_MAX_CORR_RADIUS)
        return fail(TP_EINVAL, "tp_corr_disparity: corr_radius must be 1..7%s", "");
    const size_t n = 2 * corr_radius + 1;
    if (stride < n * n) return fail(TP_EINVAL, "tp_corr_disparity: stride < (2r+1)^2%s", "");
    if (num_tiles < 0) return fail(TP_EINVAL, "tp_corr_disparity: num_tiles < 0%s", "");
    if (mode != 0 && mode != 1) return fail(TP_EINVAL, "tp_corr_disparity: mode must be 0 or 1%s", "");
    g_err.clear();
    if (!num_tiles) return TP_OK;
    hipLaunchKernelGGL(tp::k_corr_disparity, dim3((num_tiles + 255) / 256), dim3(256), 0,
                       as_stream(stream), gpu_tiles, stride, num_tiles, corr_radius, mode, gpu_out);
    return check_launch("corr_disparity");
}

int tp_convert_correlate(const tp_geometry *geom, float **gpu_kernel_offsets, float **gpu_kernels,
                         float **gpu_images, const tp_task *gpu_tasks, float **gpu_clt,
                         size_t dstride, int num_tiles, int lpf_mask, int colors, float scale0,
                         float scale1, float scale2, float fat_zero, size_t corr_stride,
                         int corr_radius, float *gpu_corrs, float *gpu_combined,
                         size_t combined_stride, int pair_mask, const tp_fused_config *config,
                         tp_stream_t stream) {
    const char *fn = "tp_convert_correlate";
    int rc = fused_common(fn, geom, dstride, lpf_mask, colors, corr_stride, corr_radius);
    if (rc) return rc;
    if (!gpu_kernel_offsets || !gpu_kernels || !gpu_images || !gpu_tasks)
        return fail(TP_EINVAL, "%s: NULL pointer argument", fn);
    const size_t nn = 2 * corr_radius + 1;
    if (gpu_combined && combined_stride < nn * nn)
        return fail(TP_EINVAL, "%s: combined_stride < (2r+1)^2", fn);
    if (num_tiles < 0) return fail(TP_EINVAL, "%s: num_tiles < 0", fn);
    FusedCfg c;
    if ((rc = fused_config(fn, config, c))) return rc;
    g_err.clear();
    if (num_tiles == 0) return TP_OK;
    tp::FusedArgs a = fused_args(geom, dstride, lpf_mask, scale0, scale1, scale2, fat_zero,
                                 corr_stride, corr_radius, c.tpi, combined_stride, pair_mask);
    a.f0.tasks = gpu_tasks;
    a.f0.images = (const float *const *)gpu_images;
    a.f0.kernels = (const float *const *)gpu_kernels;
    a.f0.koffs = (const float *const *)gpu_kernel_offsets;
    a.f0.clt = gpu_clt;
    a.f0.corrs = gpu_corrs;
    a.f0.combined = gpu_combined;
    a.f0.num_tiles = num_tiles;
    a.f0.num_items = 0;
    launch_fused<false>(a, c, colors, (num_tiles + c.tpi - 1) / c.tpi, as_stream(stream));
    return check_launch("fused");
}

int tp_stream_convert_correlate(const tp_geometry *geom, tp_frame *gpu_frames, int nframes,
                                int32_t *gpu_work, uint64_t *gpu_stamps, size_t dstride,
                                int lpf_mask, int colors, float scale0, float scale1, float scale2,
                                float fat_zero, size_t corr_stride, int corr_radius,
                                size_t combined_stride, int pair_mask,
                                const tp_fused_config *config, tp_stream_t stream) {
    g_err.clear();
    return enqueue_stream(geom, gpu_frames, nframes, gpu_work, gpu_stamps, dstride, lpf_mask, colors,
                          scale0, scale1, scale2, fat_zero, corr_stride, corr_radius, combined_stride,
                          pair_mask, config, as_stream(stream), "tp_stream_convert_correlate");
}

int tp_stream_convert_correlate_noalloc(const tp_geometry *geom, tp_frame *gpu_frames, int nframes,
                                        uint64_t *gpu_stamps, size_t dstride, int lpf_mask,
                                        int colors, float scale0, float scale1, float scale2,
                                        float fat_zero, size_t corr_stride, int corr_radius,
                                        size_t combined_stride, int pair_mask,
                                        const tp_fused_config *config, tp_stream_t stream) {
    const char *fn = "tp_stream_convert_correlate_noalloc";
    g_err.clear();
    if (nframes > TP_STREAM_NOALLOC_MAX_FRAMES)
        return fail(TP_EINVAL, "%s: nframes %ld exceeds the built-in work buffer (4096 frames)", fn,
                    (long)nframes);
    if (!gpu_frames) return fail(TP_EINVAL, "%s: NULL pointer argument", fn);
    if (nframes <= 0) return fail(TP_EINVAL, "%s: nframes must be positive", fn);
    int32_t *work = nullptr;
    int rc = scratch_ptr(fn, &work);
    if (rc) return rc;
    return enqueue_stream(geom, gpu_frames, nframes, work, gpu_stamps, dstride, lpf_mask, colors,
                          scale0, scale1, scale2, fat_zero, corr_stride, corr_radius, combined_stride,
                          pair_mask, config, as_stream(stream), fn);
}

int tp_stream_graph_create(const tp_geometry *geom, tp_frame *gpu_frames, int nframes,
                           int32_t *gpu_work, uint64_t *gpu_stamps, size_t dstride, int lpf_mask,
                           int colors, float scale0, float scale1, float scale2, float fat_zero,
                           size_t corr_stride, int corr_radius, size_t combined_stride,
                           int pair_mask, const tp_fused_config *config, tp_graph_t *graph) {
    const char *fn = "tp_stream_graph_create";
    if (!graph) return fail(TP_EINVAL, "%s: graph is NULL", fn);
    *graph = nullptr;
    g_err.clear();
    hipStream_t st;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(TP_EHIP, "tp_stream_graph_create: %s", hipGetErrorString(e));
    e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(st);
        return fail(TP_EHIP, "tp_stream_graph_create: begin capture: %s", hipGetErrorString(e));
    }
    int rc = enqueue_stream(geom, gpu_frames, nframes, gpu_work, gpu_stamps, dstride, lpf_mask, colors,
                            scale0, scale1, scale2, fat_zero, corr_stride, corr_radius, combined_stride,
                            pair_mask, config, st, fn);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(st, &g);
    (void)hipStreamDestroy(st);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(TP_EHIP, "tp_stream_graph_create: end capture: %s", hipGetErrorString(e));
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(TP_EHIP, "tp_stream_graph_create: instantiate: %s", hipGetErrorString(e));
    *graph = (tp_graph_t)ex;
    return TP_OK;
}

int tp_graph_launch(tp_graph_t graph, tp_stream_t stream) {
    if (!graph) return fail(TP_EINVAL, "tp_graph_launch: graph is NULL%s", "");
    hipError_t e = hipGraphLaunch((hipGraphExec_t)graph, as_stream(stream));
    if (e != hipSuccess) return fail(TP_EHIP, "tp_graph_launch: %s", hipGetErrorString(e));
    g_err.clear();
    return TP_OK;
}

int tp_graph_destroy(tp_graph_t graph) {
    if (!graph) return TP_OK;
    hipError_t e = hipGraphExecDestroy((hipGraphExec_t)graph);
    if (e != hipSuccess) return fail(TP_EHIP, "tp_graph_destroy: %s", hipGetErrorString(e));
    g_err.clear();
    return TP_OK;
}

double tp_wall_clock_hz(void) {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0.0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess) return 0.0;
    return 1e3 * khz;
}

int tp_nn_input(const tp_geometry *geom, const tp_task *gpu_tasks, int num_tasks,
                const float *gpu_corrs, size_t corr_stride, float *gpu_x, tp_stream_t stream) {
    int rc = check_geometry(geom);
    if (rc) return rc;
    if (!gpu_tasks || !gpu_corrs || !gpu_x) return fail(TP_EINVAL, "tp_nn_input: NULL pointer argument%s", "");
    if (num_tasks < 0) return fail(TP_EINVAL, "tp_nn_input: num_tasks < 0%s", "");
    if (corr_stride < 225) return fail(TP_EINVAL, "tp_nn_input: corr_stride < 225 (radius 7 rows)%s", "");
    g_err.clear();
    if (!num_tasks) return TP_OK;
    hipLaunchKernelGGL(tp::k_nn_input, dim3(num_tasks), dim3(256), 0, as_stream(stream), gpu_tasks,
                       num_tasks, geom->img_width / 8, geom->img_height / 8, gpu_corrs, corr_stride, gpu_x);
    return check_launch("nn_input");
}

namespace {
int nn_grid(int n) {
    const int cus = num_cus();
    const int g = (cus > 0 ? cus : 256) * 4;
    return std::max(1, std::min(n, g));
}
}  // namespace

int tp_nn_forward(const uint16_t *gpu_params_bf16, const float *gpu_x, int n, float *gpu_acts,
                  float *gpu_out, tp_stream_t stream) {
    if (!gpu_params_bf16 || !gpu_x || !gpu_out) return fail(TP_EINVAL, "tp_nn_forward: NULL pointer argument%s", "");
    if (n < 0) return fail(TP_EINVAL, "tp_nn_forward: n < 0%s", "");
    g_err.clear();
    if (!n) return TP_OK;
    hipLaunchKernelGGL(tp::k_nn_forward, dim3(nn_grid(n)), dim3(256), 0, as_stream(stream),
                       gpu_params_bf16, gpu_x, n, gpu_acts, gpu_out);
    return check_launch("nn_forward");
}

int tp_nn_infer(const uint16_t *gpu_params_bf16, const float *gpu_x, int n, float *gpu_out,
                tp_stream_t stream) {
    if (!gpu_params_bf16 || !gpu_x || !gpu_out) return fail(TP_EINVAL, "tp_nn_infer: NULL pointer argument%s", "");
    if (n < 0) return fail(TP_EINVAL, "tp_nn_infer: n < 0%s", "");
    g_err.clear();
    if (!n) return TP_OK;
    tp::NNInferIn in{};
    in.x = gpu_x;
    hipLaunchKernelGGL(tp::k_nn_infer<false>, dim3(nn_grid(n)), dim3(256), 0, as_stream(stream),
                       gpu_params_bf16, in, n, gpu_out);
    return check_launch("nn_infer");
}

int tp_nn_infer_corrs(const tp_geometry *geom, const uint16_t *gpu_params_bf16, const tp_task *gpu_tasks,
                      int num_tasks, const float *gpu_corrs, size_t corr_stride, float *gpu_out,
                      tp_stream_t stream) {
    int rc = check_geometry(geom);
    if (rc) return rc;
    if (!gpu_params_bf16 || !gpu_tasks || !gpu_corrs || !gpu_out)
        return fail(TP_EINVAL, "tp_nn_infer_corrs: NULL pointer argument%s", "");
    if (num_tasks < 0) return fail(TP_EINVAL, "tp_nn_infer_corrs: num_tasks < 0%s", "");
    if (corr_stride < 225) return fail(TP_EINVAL, "tp_nn_infer_corrs: corr_stride < 225 (radius 7 rows)%s", "");
    g_err.clear();
    if (!num_tasks) return TP_OK;
    tp::NNInferIn in{};
    in.tasks = gpu_tasks;
    in.corrs = gpu_corrs;
    in.corr_stride = corr_stride;
    in.tiles_x = geom->img_width / 8;
    in.tiles_y = geom->img_height / 8;
    hipLaunchKernelGGL(tp::k_nn_infer<true>, dim3(nn_grid(num_tasks)), dim3(256), 0, as_stream(stream),
                       gpu_params_bf16, in, num_tasks, gpu_out);
    return check_launch("nn_infer_corrs");
}

int tp_nn_backward(const uint16_t *gpu_params_bf16, const float *gpu_x, const float *gpu_acts,
                   const float *gpu_target, int n, float conf_weight, float huber_delta,
                   float loss_scale, float *gpu_grads, float *gpu_loss, tp_stream_t stream) {
    if (!gpu_params_bf16 || !gpu_x || !gpu_acts || !gpu_target || !gpu_grads || !gpu_loss)
        return fail(TP_EINVAL, "tp_nn_backward: NULL pointer argument%s", "");
    if (n < 0) return fail(TP_EINVAL, "tp_nn_backward: n < 0%s", "");
    if (!(huber_delta > 0.f)) return fail(TP_EINVAL, "tp_nn_backward: huber_delta must be positive%s", "");
    g_err.clear();
    if (!n) return TP_OK;
    tp::NNBackArgs a;
    a.prm = gpu_params_bf16;
    a.x = gpu_x;
    a.acts = gpu_acts;
    a.target = gpu_target;
    a.n = n;
    a.conf_weight = conf_weight;
    a.huber_delta = huber_delta;
    a.scale = loss_scale;
    a.grads = gpu_grads;
    a.loss = gpu_loss;
    hipLaunchKernelGGL(tp::k_nn_backward, dim3(nn_grid(n)), dim3(256), 0, as_stream(stream), a);
    return check_launch("nn_backward");
}

int tp_nn_adam(float *gpu_params, float *gpu_m, float *gpu_v, const float *gpu_grads, int n,
               float lr, float beta1, float beta2, float eps, int step, uint16_t *gpu_params_bf16,
               tp_stream_t stream) {
    if (!gpu_params || !gpu_m || !gpu_v || !gpu_grads || !gpu_params_bf16)
        return fail(TP_EINVAL, "tp_nn_adam: NULL pointer argument%s", "");
    if (n < 0 || step < 1) return fail(TP_EINVAL, "tp_nn_adam: n < 0 or step < 1%s", "");
    g_err.clear();
    if (!n) return TP_OK;
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)beta1, step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)beta2, step)));
    hipLaunchKernelGGL(tp::k_nn_adam, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream), gpu_params,
                       gpu_m, gpu_v, gpu_grads, n, lr, beta1, beta2, eps, c1, c2, gpu_params_bf16);
    return check_launch("nn_adam");
}

int tp_nn_pack_bf16(const float *gpu_src, int n, uint16_t *gpu_dst, tp_stream_t stream) {
    if (!gpu_src || !gpu_dst) return fail(TP_EINVAL, "tp_nn_pack_bf16: NULL pointer argument%s", "");
    if (n < 0) return fail(TP_EINVAL, "tp_nn_pack_bf16: n < 0%s", "");
    g_err.clear();
    if (!n) return TP_OK;
    hipLaunchKernelGGL(tp::k_nn_pack, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream), gpu_src, n,
                       gpu_dst);
    return check_launch("nn_pack_bf16");
}

int tp_psf_to_clt(const tp_geometry *geom, float **gpu_psfs, int psf_height, int psf_width,
                  int center, float **gpu_kernels, float **gpu_kernel_offsets, tp_stream_t stream) {
    int rc = check_geometry(geom);
    if (rc) return rc;
    if (!gpu_psfs || !gpu_kernels || !gpu_kernel_offsets)
        return fail(TP_EINVAL, "tp_psf_to_clt: NULL pointer argument%s", "");
    if (psf_height < 1 || psf_width < 1 || psf_height > tp::PSF_MAX || psf_width > tp::PSF_MAX)
        return fail(TP_EINVAL, "tp_psf_to_clt: PSF sides must be 1..32%s", "");
    if (center != 0 && center != 1) return fail(TP_EINVAL, "tp_psf_to_clt: center must be 0 or 1%s", "");
    g_err.clear();
    tp::PsfArgs a;
    a.psfs = (const float *const *)gpu_psfs;
    a.kernels = gpu_kernels;
    a.koffs = gpu_kernel_offsets;
    a.KV = geom->kernels_vert;
    a.KH = geom->kernels_hor;
    a.h = psf_height;
    a.w = psf_width;
    a.center = center;
    const int units = TP_NUM_CAMS * a.KV * a.KH * TP_NUM_COLORS;
    hipLaunchKernelGGL(tp::k_psf_to_clt, dim3((units + 3) / 4), dim3(256), 0, as_stream(stream), a);
    if ((rc = check_launch("psf_to_clt"))) return rc;
    hipLaunchKernelGGL(tp::k_psf_slopes, dim3((units + 255) / 256), dim3(256), 0, as_stream(stream),
                       gpu_kernel_offsets, a.KV, a.KH);
    return check_launch("psf_slopes");
}

int tp_dtt8x8(float *gpu_dst, const float *gpu_src, int src_stride, int dtt_mode, int width,
              int height, tp_stream_t stream) {
    if (!gpu_dst || !gpu_src) return fail(TP_EINVAL, "tp_dtt8x8: NULL pointer argument%s", "");
    if (dtt_mode < 0 || dtt_mode > 7) return fail(TP_EINVAL, "tp_dtt8x8: dtt_mode must be 0..7%s", "");
    if (width <= 0 || height <= 0 || (width % 8) || (height % 8))
        return fail(TP_EINVAL, "tp_dtt8x8: width and height must be positive multiples of 8%s", "");
    if (src_stride < width) return fail(TP_EINVAL, "tp_dtt8x8: src_stride < width%s", "");
    g_err.clear();
    const int tiles_x = width / 8, ntiles = tiles_x * (height / 8);
    hipLaunchKernelGGL(tp::k_dtt8x8, dim3((ntiles + 31) / 32), dim3(256), 0, as_stream(stream), gpu_src,
                       gpu_dst, (size_t)src_stride, tiles_x, ntiles, dtt_mode);
    return check_launch("dtt8x8");
}

int tp_fill(float *gpu_dst, size_t n, float value, tp_stream_t stream) {
    if (!gpu_dst) return fail(TP_EINVAL, "tp_fill: NULL pointer argument%s", "");
    g_err.clear();
    if (!n) return TP_OK;
    hipLaunchKernelGGL(tp::k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream),
                       gpu_dst, n, value);
    return check_launch("fill");
}

}  // extern "C"
