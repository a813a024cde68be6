"""Host-side mirror of the reference kernels' calling convention over libtp.so.

Each function here has the name, argument order and argument meaning of the
reference kernel it replaces (src/TileProcessor.cuh), plus the run-time
geometry and an optional stream.  Arguments the reference passes as
`float **` (device arrays of per-camera device pointers) are torch int64
device tensors built by `ptr_array`.  Work is enqueued on `stream` (default:
torch's current stream) and nothing synchronises.

All device memory here is torch's; the kernels are the HIP code of libtp.so.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .geometry import Geometry, NUM_CAMS, NUM_COLORS, TASK_DTYPE

_VP = ctypes.c_void_p


def _p(t):
    return _VP(t.data_ptr()) if t is not None else _VP(0)


def _stream(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    if isinstance(stream, torch.cuda.Stream):
        return _VP(stream.cuda_stream)
    return _VP(int(stream))


def _check_dev(*ts):
    for t in ts:
        if t is not None and not t.is_cuda:
            raise ValueError("tensor arguments must be on the GPU")


def ptr_array(tensors):
    """Device array of device pointers (the reference's `float **`)."""
    _check_dev(*tensors)
    return torch.tensor([t.data_ptr() for t in tensors], dtype=torch.int64,
                        device=tensors[0].device)


def version():
    return _lib.lib().tp_version().decode()


def set_lpf(which, table):
    """Replace a low-pass table (reference: cudaMemcpyToSymbol on lpf_data,
    lpf_rb_corr, lpf_corr).  which: 0..3 lpf_data[R,B,G,mono], 4 lpf_rb_corr,
    5 lpf_corr; table: 64 floats [vertical][horizontal], or None for the
    built-in table."""
    if table is None:
        _lib.call("tp_set_lpf", int(which), _VP(0))
        return
    a = np.ascontiguousarray(table, dtype=np.float32).reshape(64)
    _lib.call("tp_set_lpf", int(which), a.ctypes.data_as(_VP))


def convert_correct_tiles(geom, gpu_kernel_offsets, gpu_kernels, gpu_images, gpu_tasks, gpu_clt,
                          dstride, num_tiles, lpf_mask=0, stream=None):
    """convert_correct_tiles (src/TileProcessor.cuh:1859-1937)."""
    _check_dev(gpu_kernel_offsets, gpu_kernels, gpu_images, gpu_tasks, gpu_clt)
    g = geom.c()
    _lib.call("tp_convert_correct_tiles", ctypes.byref(g), _p(gpu_kernel_offsets), _p(gpu_kernels),
              _p(gpu_images), _p(gpu_tasks), _p(gpu_clt), int(dstride), int(num_tiles),
              int(lpf_mask), _stream(stream))


def correlate2D(geom, gpu_clt, colors, scale0, scale1, scale2, fat_zero, num_corr_tiles,
                gpu_corr_indices, corr_stride, corr_radius, gpu_corrs, stream=None):
    """correlate2D (src/TileProcessor.cuh:1176-1443)."""
    _check_dev(gpu_clt, gpu_corr_indices, gpu_corrs)
    g = geom.c()
    _lib.call("tp_correlate2D", ctypes.byref(g), _p(gpu_clt), int(colors), float(scale0),
              float(scale1), float(scale2), float(fat_zero), int(num_corr_tiles),
              _p(gpu_corr_indices), int(corr_stride), int(corr_radius), _p(gpu_corrs),
              _stream(stream))


def imclt_rbg(geom, gpu_clt, gpu_rbg, color, v_offset, h_offset, dstride, stream=None):
    """imclt_rbg (src/TileProcessor.cuh:2588-2712): one camera, one color,
    one of the four tile-parity passes."""
    _check_dev(gpu_clt, gpu_rbg)
    g = geom.c()
    _lib.call("tp_imclt_rbg", ctypes.byref(g), _p(gpu_clt), _p(gpu_rbg), int(color),
              int(v_offset), int(h_offset), int(dstride), _stream(stream))


def imclt_rbg_all(geom, gpu_clt, gpu_rbg, dstride, stream=None):
    """All cameras, colors and passes in one call (gpu_clt, gpu_rbg: pointer arrays)."""
    _check_dev(gpu_clt, gpu_rbg)
    g = geom.c()
    _lib.call("tp_imclt_rbg_all", ctypes.byref(g), _p(gpu_clt), _p(gpu_rbg), int(dstride),
              _stream(stream))


class FusedConfig(ctypes.Structure):
    """tp_fused_config: work-item size, tiles held in LDS at once, resident
    workgroups per CU (stream form) and the forward-DCT method."""
    _fields_ = [("tiles_per_item", ctypes.c_int32), ("tiles_per_group", ctypes.c_int32),
                ("blocks_per_cu", ctypes.c_int32), ("dct", ctypes.c_int32)]

    def __init__(self, tiles_per_item=12, tiles_per_group=4, blocks_per_cu=4, dct=0):
        super().__init__(tiles_per_item, tiles_per_group, blocks_per_cu, dct)

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _cfg(config):
    if config is None:
        return _VP(0)
    if isinstance(config, dict):
        config = FusedConfig(**config)
    return ctypes.byref(config)


def convert_correlate(geom, gpu_kernel_offsets, gpu_kernels, gpu_images, gpu_tasks, gpu_clt,
                      dstride, num_tiles, lpf_mask, colors, scale0, scale1, scale2, fat_zero,
                      corr_stride, corr_radius, gpu_corrs, gpu_combined=None, combined_stride=0,
                      pair_mask=0x3F, config=None, stream=None):
    """Forward CLT + pair correlation of one frame in one pass (rows 6 task + pair);
    gpu_clt, gpu_corrs and gpu_combined (six-pair combination, row = task)
    may each be None."""
    _check_dev(gpu_kernel_offsets, gpu_kernels, gpu_images, gpu_tasks, gpu_clt, gpu_corrs,
               gpu_combined)
    g = geom.c()
    _lib.call("tp_convert_correlate", ctypes.byref(g), _p(gpu_kernel_offsets), _p(gpu_kernels),
              _p(gpu_images), _p(gpu_tasks), _p(gpu_clt), int(dstride), int(num_tiles),
              int(lpf_mask), int(colors), float(scale0), float(scale1), float(scale2),
              float(fat_zero), int(corr_stride), int(corr_radius), _p(gpu_corrs),
              _p(gpu_combined), int(combined_stride), int(pair_mask), _cfg(config),
              _stream(stream))


# tp_frame (include/tp.h): one frame of the stream form
def stream_work_ints(nframes):
    """TP_STREAM_WORK_INTS(nframes): int32 words of the stream work buffer."""
    return 296 + 8 * nframes


FRAME_DTYPE = np.dtype([("tasks", "<u8"), ("images", "<u8"), ("kernels", "<u8"),
                        ("kernel_offsets", "<u8"), ("clt", "<u8"), ("corrs", "<u8"),
                        ("combined", "<u8"), ("num_tiles", "<i4"), ("num_items", "<i4")])


def wall_clock_hz():
    return _lib.lib().tp_wall_clock_hz()


def frame_descriptors(frames, store_clt=False, store_pairs=True, combined=False):
    """tp_frame array of the stream form for a list of FrameBuffers, after
    checking that each frame has a task list and that its outputs hold the
    rows the kernel writes: 6 rows per task (row 6 task + pair) for the pair
    tiles and one per task for the six-pair combination."""
    desc = np.zeros(len(frames), dtype=FRAME_DTYPE)
    for i, fb in enumerate(frames):
        if fb.tasks is None:
            raise ValueError(f"frame {i}: no task list uploaded")
        if store_pairs and fb.corrs.shape[0] < 6 * fb.num_tiles:
            raise ValueError(f"frame {i}: corrs has {fb.corrs.shape[0]} rows, the stream form "
                             f"writes 6 per task ({6 * fb.num_tiles})")
        comb = fb.combined_buffer() if combined else None
        if comb is not None and comb.shape[0] < fb.num_tiles:
            raise ValueError(f"frame {i}: combined buffer has fewer rows than tasks")
        desc[i] = (fb.tasks.data_ptr(), fb.p_images.data_ptr(), fb.p_kernels.data_ptr(),
                   fb.p_koffs.data_ptr(), fb.p_clt.data_ptr() if store_clt else 0,
                   fb.corrs.data_ptr() if store_pairs else 0,
                   comb.data_ptr() if comb is not None else 0, fb.num_tiles, 0)
    return desc


class StreamRun:
    """The stream form over a list of FrameBuffers: one persistent launch
    (tp_stream_convert_correlate) per call, or a captured graph of it."""

    def drained(self):
        """True when every per-XCD queue has handed out all its items (call
        after synchronising)."""
        w = self.work.cpu().numpy().astype(np.int64)
        nf = len(self.frames)
        lengths = w[288 + 8 * nf:288 + 8 * nf + 8]  # per-queue item counts
        return int(lengths.sum()) == int(w[256]) and all(w[32 * x] >= lengths[x] for x in range(8))

    def __init__(self, frames, colors=3, scales=(0.25, 0.25, 0.5), fat_zero=0.05, lpf_mask=0,
                 store_clt=False, config=None, store_pairs=True, combined=False, pair_mask=0x3F,
                 stamps=True):
        if not frames:
            raise ValueError("no frames")
        self.frames = frames
        geom = frames[0].geom
        dev = frames[0].device
        for fb in frames:
            if fb.geom != geom or fb.dstride != frames[0].dstride or \
                    fb.corr_radius != frames[0].corr_radius or \
                    fb.corr_stride != frames[0].corr_stride:
                raise ValueError("frames of one stream share geometry, dstride and correlation shape")
        self.geom, self.device = geom, dev
        self.args = (colors, *scales, fat_zero)
        self.lpf_mask = lpf_mask
        self.config = FusedConfig(**config) if isinstance(config, dict) else (config or FusedConfig())
        self.store_clt, self.store_pairs, self.combined = store_clt, store_pairs, combined
        self.pair_mask = pair_mask
        self._desc_host = frame_descriptors(frames, store_clt, store_pairs, combined)
        self.desc = torch.from_numpy(self._desc_host.view(np.uint8).copy()).to(dev)
        # per-XCD queue counters (one 128-byte line each) and the item count
        self.work = torch.zeros(stream_work_ints(len(frames)), dtype=torch.int32, device=dev)
        # per-frame GPU time stamps (two atomics per work item on the frame's words)
        self.stamps = torch.zeros((len(frames), 2), dtype=torch.int64, device=dev) if stamps else None
        self.graph = None

    def refresh(self):
        """Re-check the frames and rewrite the device frame descriptors in
        place if a frame's buffers or task count changed since the last
        launch (a captured graph reads the descriptors at replay time, so it
        follows).  Raises ValueError if a frame's output rows no longer fit."""
        d = frame_descriptors(self.frames, self.store_clt, self.store_pairs, self.combined)
        if d.tobytes() != self._desc_host.tobytes():
            self.desc.copy_(torch.from_numpy(d.view(np.uint8).copy()))
            self._desc_host = d
            return True
        return False

    def _common(self):
        f0 = self.frames[0]
        colors, s0, s1, s2, fz = self.args
        return (ctypes.byref(self.geom.c()), _p(self.desc), len(self.frames), _p(self.work),
                _p(self.stamps), int(f0.dstride), int(self.lpf_mask), int(colors), float(s0),
                float(s1), float(s2), float(fz), int(f0.corr_stride), int(f0.corr_radius),
                int(f0.combined_stride), int(self.pair_mask), ctypes.byref(self.config))

    def launch(self, stream=None):
        self.refresh()
        _lib.call("tp_stream_convert_correlate", *self._common(), _stream(stream))

    def launch_noalloc(self, stream=None):
        """The same launch through tp_stream_convert_correlate_noalloc (the
        library's own work buffer instead of self.work)."""
        self.refresh()
        args = self._common()
        _lib.call("tp_stream_convert_correlate_noalloc", *args[:3], *args[4:], _stream(stream))

    def capture(self):
        self.release()
        self.refresh()
        h = _VP(0)
        _lib.call("tp_stream_graph_create", *self._common(), ctypes.byref(h))
        self.graph = h

    def replay(self, stream=None):
        if self.graph is None:
            self.capture()
        self.refresh()
        _lib.call("tp_graph_launch", self.graph, _stream(stream))

    def release(self):
        if self.graph is not None:
            _lib.call("tp_graph_destroy", self.graph)
            self.graph = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def frame_times_s(self):
        """Per frame (start, end) of its work items in seconds after the
        first start, from the GPU wall clock (call after synchronising)."""
        if self.stamps is None:
            raise ValueError("this StreamRun keeps no time stamps")
        raw = self.stamps.cpu().numpy()
        hz = wall_clock_hz()
        done = raw[:, 0] != -1  # frames without work items keep the initial stamps
        out = np.full(raw.shape, np.nan)
        if done.any():
            t0 = raw[done, 0].min()
            out[done] = (raw[done] - t0).astype(np.float64) / hz
        return out


class FrameBuffers:
    """Device buffers of one 4-camera frame in the reference layout, with the
    pointer arrays the `float **` arguments need."""

    def __init__(self, geom: Geometry, device="cuda:0", dstride=None, corr_radius=7,
                 max_corr=None, corr_stride=None):
        self.geom = geom
        self.device = torch.device(device)
        W, H = geom.img_width, geom.img_height
        self.dstride = dstride or W
        f32 = dict(dtype=torch.float32, device=self.device)
        self.images = [torch.zeros((H, self.dstride), **f32) for _ in range(NUM_CAMS)]
        self.kernels = [torch.zeros(geom.kernel_floats, **f32) for _ in range(NUM_CAMS)]
        self.koffs = [torch.zeros(geom.kernels_vert * geom.kernels_hor * NUM_COLORS * 8, **f32)
                      for _ in range(NUM_CAMS)]
        self.clt = [torch.zeros(geom.clt_floats, **f32) for _ in range(NUM_CAMS)]
        self.p_images = ptr_array(self.images)
        self.p_kernels = ptr_array(self.kernels)
        self.p_koffs = ptr_array(self.koffs)
        self.p_clt = ptr_array(self.clt)
        self.tasks = None
        self.num_tiles = 0
        self.corr_radius = corr_radius
        n = 2 * corr_radius + 1
        self.corr_stride = corr_stride or n * n
        self.max_corr = max_corr if max_corr is not None else geom.num_tiles * 6
        self.corr_idx = torch.zeros(max(1, self.max_corr), dtype=torch.int32, device=self.device)
        self.corrs = torch.zeros((max(1, self.max_corr), self.corr_stride), **f32)
        self.num_corr = 0

    def upload(self, images=None, kernels=None, koffs=None, tasks=None, corr_idx=None):
        if images is not None:
            for c in range(NUM_CAMS):
                im = torch.from_numpy(np.ascontiguousarray(images[c], dtype=np.float32))
                self.images[c][:, :im.shape[1]].copy_(im)
        if kernels is not None:
            for c in range(NUM_CAMS):
                self.kernels[c].copy_(torch.from_numpy(
                    np.ascontiguousarray(kernels[c], dtype=np.float32).reshape(-1)))
        if koffs is not None:
            for c in range(NUM_CAMS):
                self.koffs[c].copy_(torch.from_numpy(
                    np.ascontiguousarray(koffs[c], dtype=np.float32).reshape(-1)))
        if tasks is not None:
            # one task buffer per frame, sized for a full task list; task
            # lists are copied into it, so its address (which a stream's
            # frame descriptors hold) stays the same while they fit
            t = np.ascontiguousarray(tasks, dtype=TASK_DTYPE)
            nbytes = t.nbytes
            if self.tasks is None or self.tasks.numel() < nbytes:
                cap = max(nbytes, self.geom.num_tiles * TASK_DTYPE.itemsize)
                self.tasks = torch.zeros(cap, dtype=torch.uint8, device=self.device)
            if nbytes:
                self.tasks[:nbytes].copy_(torch.from_numpy(t.view(np.uint8).copy()))
            self.num_tiles = len(t)
        if corr_idx is not None:
            ci = np.ascontiguousarray(corr_idx, dtype=np.int32)
            if len(ci) > self.max_corr:
                raise ValueError("correlation list longer than max_corr")
            self.corr_idx[:len(ci)].copy_(torch.from_numpy(ci))
            self.num_corr = len(ci)

    def kernels_from_psfs(self, psfs, center=True, stream=None):
        """Fill self.kernels and self.koffs from sampled PSFs, a device tensor
        [cams][KV][KH][3][h][w] (tp_psf_to_clt)."""
        if psfs.dim() != 6 or psfs.shape[0] != NUM_CAMS:
            raise ValueError("psfs must be [cams][KV][KH][3][h][w]")
        psfs = psfs.contiguous()
        p_psfs = ptr_array([psfs[c] for c in range(NUM_CAMS)])
        psf_to_clt(self.geom, p_psfs, psfs.shape[4], psfs.shape[5], int(center), self.p_kernels,
                   self.p_koffs, stream)
        return p_psfs  # keep alive until the work is done

    def convert(self, lpf_mask=0, stream=None):
        convert_correct_tiles(self.geom, self.p_koffs, self.p_kernels, self.p_images, self.tasks,
                              self.p_clt, self.dstride, self.num_tiles, lpf_mask, stream)

    def correlate(self, colors=3, scales=(0.25, 0.25, 0.5), fat_zero=0.05, stream=None):
        correlate2D(self.geom, self.p_clt, colors, scales[0], scales[1], scales[2], fat_zero,
                    self.num_corr, self.corr_idx, self.corr_stride, self.corr_radius, self.corrs,
                    stream)

    def combined_buffer(self):
        """[tiles][(2r+1)^2] six-pair combination rows (row = task), made on first use."""
        n = 2 * self.corr_radius + 1
        if getattr(self, "combined", None) is None or self.combined.shape[0] < max(1, self.num_tiles):
            rows = max(1, self.num_tiles, self.geom.num_tiles)
            self.combined = torch.zeros((rows, n * n), dtype=torch.float32, device=self.device)
        return self.combined

    @property
    def combined_stride(self):
        n = 2 * self.corr_radius + 1
        return n * n

    def fused(self, colors=3, scales=(0.25, 0.25, 0.5), fat_zero=0.05, lpf_mask=0,
              store_clt=False, config=None, stream=None, store_pairs=True, combined=False,
              pair_mask=0x3F):
        """Forward CLT + correlation in one pass; rows of self.corrs are
        6 task + pair (self.corrs needs 6 num_tiles rows); with combined=True
        also the six-pair combination into self.combined (row = task)."""
        if store_pairs and self.corrs.shape[0] < 6 * self.num_tiles:
            raise ValueError("corrs needs 6 rows per task for the fused form")
        comb = self.combined_buffer() if combined else None
        convert_correlate(self.geom, self.p_koffs, self.p_kernels, self.p_images, self.tasks,
                          self.p_clt if store_clt else None, self.dstride, self.num_tiles,
                          lpf_mask, colors, scales[0], scales[1], scales[2], fat_zero,
                          self.corr_stride, self.corr_radius, self.corrs if store_pairs else None,
                          comb, self.combined_stride, pair_mask, config, stream)


# texture parameters of the reference harness (src/dtt8x8.cu:791-814)
TEXTURE_DEFAULTS = dict(colors=3, is_lwir=0, min_shot=10.0, scale_shot=3.0, diff_sigma=1.5,
                        diff_threshold=10.0, min_agree=3.0, weight0=0.294118, weight1=0.117647,
                        weight2=0.588235, dust_remove=1, keep_weights=0)


def textures_accumulate(geom, woi, gpu_clt, num_texture_tiles, gpu_texture_indices,
                        gpu_port_offsets, colors, is_lwir, min_shot, scale_shot, diff_sigma,
                        diff_threshold, min_agree, weight0, weight1, weight2, dust_remove,
                        keep_weights, texture_rbg_stride, gpu_texture_rbg, texture_stride,
                        gpu_texture_tiles, stream=None):
    """textures_accumulate (src/TileProcessor.cuh:2207-2585)."""
    g = geom.c()
    f = float
    _lib.call("tp_textures_accumulate", ctypes.byref(g), _p(woi), _p(gpu_clt),
              int(num_texture_tiles), _p(gpu_texture_indices), _p(gpu_port_offsets), int(colors),
              int(is_lwir), f(min_shot), f(scale_shot), f(diff_sigma), f(diff_threshold),
              f(min_agree), f(weight0), f(weight1), f(weight2), int(dust_remove), int(keep_weights),
              int(texture_rbg_stride), _p(gpu_texture_rbg), int(texture_stride),
              _p(gpu_texture_tiles), _stream(stream))


def prepare_texture_list(geom, gpu_tasks, num_tiles, gpu_texture_indices, num_texture_tiles, woi,
                         width, height, stream=None):
    """prepare_texture_list (src/TileProcessor.cuh:1643-1713)."""
    g = geom.c()
    _lib.call("tp_prepare_texture_list", ctypes.byref(g), _p(gpu_tasks), int(num_tiles),
              _p(gpu_texture_indices), _p(num_texture_tiles), _p(woi), int(width), int(height),
              _stream(stream))


def generate_RBGA(geom, gpu_tasks, num_tiles, gpu_texture_indices, num_texture_tiles, woi, width,
                  height, gpu_clt, gpu_port_offsets, colors, is_lwir, min_shot, scale_shot,
                  diff_sigma, diff_threshold, min_agree, weight0, weight1, weight2, dust_remove,
                  keep_weights, texture_rbga_stride, gpu_texture_tiles, stream=None):
    """generate_RBGA (src/TileProcessor.cuh:1446-1613), without dynamic parallelism."""
    g = geom.c()
    f = float
    _lib.call("tp_generate_RBGA", ctypes.byref(g), _p(gpu_tasks), int(num_tiles),
              _p(gpu_texture_indices), _p(num_texture_tiles), _p(woi), int(width), int(height),
              _p(gpu_clt), _p(gpu_port_offsets), int(colors), int(is_lwir), f(min_shot),
              f(scale_shot), f(diff_sigma), f(diff_threshold), f(min_agree), f(weight0),
              f(weight1), f(weight2), int(dust_remove), int(keep_weights),
              int(texture_rbga_stride), _p(gpu_texture_tiles), _stream(stream))


def corr_combine(geom, num_corr_tiles, gpu_corr_indices, gpu_corrs, corr_stride, corr_radius,
                 pair_mask, gpu_combined, combined_stride, stream=None):
    """Sum the 6 pair tiles of each tile mapped onto the horizontal disparity
    axis (no counterpart in the reference kernels)."""
    g = geom.c()
    _lib.call("tp_corr_combine", ctypes.byref(g), int(num_corr_tiles), _p(gpu_corr_indices),
              _p(gpu_corrs), int(corr_stride), int(corr_radius), int(pair_mask),
              _p(gpu_combined), int(combined_stride), _stream(stream))


def corr_disparity(gpu_tiles, stride, num_tiles, corr_radius, mode, gpu_out, stream=None):
    """(disparity, strength) per combined tile; mode 0 parabola, 1 centroid."""
    _lib.call("tp_corr_disparity", _p(gpu_tiles), int(stride), int(num_tiles), int(corr_radius),
              int(mode), _p(gpu_out), _stream(stream))


def dtt8x8(gpu_dst, gpu_src, src_stride, dtt_mode, width, height, stream=None):
    """GPU_DTT24_DRV (src/dtt8x8.cuh:103-161): 2-D 8x8 DCT/DST of every 8x8
    tile (mode bit 0: horizontal sine, bit 1: vertical sine, modes 4-7:
    type II instead of IV; orthonormal)."""
    _check_dev(gpu_dst, gpu_src)
    _lib.call("tp_dtt8x8", _p(gpu_dst), _p(gpu_src), int(src_stride), int(dtt_mode), int(width),
              int(height), _stream(stream))


def psf_to_clt(geom, gpu_psfs, psf_height, psf_width, center, gpu_kernels, gpu_kernel_offsets,
               stream=None):
    """CLT-domain kernels [cam][KV][KH][3][4][8][8] and CltExtra records of
    sampled PSFs [cam][KV][KH][3][h][w] (pointer arrays), the inputs of
    convert_correct_tiles (the reference's .kernel / .kernel_offsets files,
    src/dtt8x8.cu:228-238)."""
    _check_dev(gpu_psfs, gpu_kernels, gpu_kernel_offsets)
    g = geom.c()
    _lib.call("tp_psf_to_clt", ctypes.byref(g), _p(gpu_psfs), int(psf_height), int(psf_width),
              int(center), _p(gpu_kernels), _p(gpu_kernel_offsets), _stream(stream))
