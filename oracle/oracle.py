"""ctypes binding of the CPU oracle (liboracle.so built from tp_oracle.cpp).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the
benchmark's CPU-baseline leg as the checker and the timed CPU reference,
never by the product.  All arrays are host numpy arrays.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# TP_ORACLE_LIB: another build of the same source (the sanitizer build)
LIB_PATH = os.environ.get("TP_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

_VP = ctypes.c_void_p


class CGeometry(ctypes.Structure):
    _fields_ = [("img_width", ctypes.c_int32), ("img_height", ctypes.c_int32),
                ("kernels_hor", ctypes.c_int32), ("kernels_vert", ctypes.c_int32)]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.oracle_num_threads.restype = ctypes.c_int
        _lib.oracle_table.restype = ctypes.c_int
    return _lib


def _cgeo(geom):
    return CGeometry(geom.img_width, geom.img_height, geom.kernels_hor, geom.kernels_vert)


def _ptrs(arrs):
    return (_VP * len(arrs))(*[a.ctypes.data for a in arrs])


def num_threads(n=0):
    return lib().oracle_num_threads(int(n))


def table(kind):
    """0: dtt matrices [4][8][8]; 1: fold rind/rcs/rss [3][16]; 2: windows
    hw, hw2, hwsq, hwsqi [4][8]; 3: lpf [6][64]; 4: alpha fade [16][256]."""
    sizes = {0: (4, 8, 8), 1: (3, 16), 2: (4, 8), 3: (6, 64), 4: (16, 256)}
    out = np.zeros(int(np.prod(sizes[kind])), np.float64)
    lib().oracle_table(int(kind), out.ctypes.data_as(_VP))
    return out.reshape(sizes[kind])


def set_lpf(which, tab=None):
    a = None if tab is None else np.ascontiguousarray(tab, np.float32).reshape(64)
    lib().oracle_set_lpf(int(which), None if a is None else a.ctypes.data_as(_VP))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def convert_tile(geom, koffs_cam, kernels_cam, image, dstride, color, cx, cy, lpf_mask=0,
                 dtype=np.float64):
    """One tile, one camera, one color -> [4][8][8]."""
    q = np.zeros((4, 8, 8), dtype)
    ko, ke, im = _f32(koffs_cam), _f32(kernels_cam), _f32(image)
    fn = lib().oracle_convert_tile_f64 if dtype == np.float64 else lib().oracle_convert_tile_f32
    g = _cgeo(geom)
    fn(ctypes.byref(g), ko.ctypes.data_as(_VP), ke.ctypes.data_as(_VP), im.ctypes.data_as(_VP),
       ctypes.c_size_t(dstride), ctypes.c_int(color), ctypes.c_int(lpf_mask), ctypes.c_float(cx),
       ctypes.c_float(cy), q.ctypes.data_as(_VP))
    return q


def convert_correct_tiles(geom, koffs, kernels, images, tasks, dstride=None, lpf_mask=0,
                          dtype=np.float64, clt=None):
    """Oracle of convert_correct_tiles.  koffs/kernels/images: per-camera
    arrays; tasks: TASK_DTYPE array.  Returns 4 arrays [tilesy*tilesx*3*256]."""
    dstride = dstride or geom.img_width
    ko = [_f32(a) for a in koffs]
    ke = [_f32(a) for a in kernels]
    im = [_f32(a) for a in images]
    if clt is None:
        clt = [np.zeros(geom.clt_floats, dtype) for _ in range(4)]
    t = np.ascontiguousarray(tasks)
    fn = (lib().oracle_convert_correct_tiles_f64 if dtype == np.float64
          else lib().oracle_convert_correct_tiles_f32)
    g = _cgeo(geom)
    fn(ctypes.byref(g), _ptrs(ko), _ptrs(ke), _ptrs(im), t.ctypes.data_as(_VP), _ptrs(clt),
       ctypes.c_size_t(dstride), ctypes.c_int(len(t)), ctypes.c_int(lpf_mask))
    return clt


def correlate2D(geom, clt, colors, scales, fat_zero, idx, radius, stride=None, dtype=None):
    """Oracle of correlate2D; clt: 4 arrays (float32 or float64, selects the
    variant).  Returns [len(idx)][stride]."""
    dtype = dtype or clt[0].dtype
    clt = [np.ascontiguousarray(c, dtype) for c in clt]
    n = 2 * radius + 1
    stride = stride or n * n
    idx = np.ascontiguousarray(idx, np.int32)
    out = np.zeros((len(idx), stride), dtype)
    fn = lib().oracle_correlate2D_f64 if dtype == np.float64 else lib().oracle_correlate2D_f32
    g = _cgeo(geom)
    fn(ctypes.byref(g), _ptrs(clt), ctypes.c_int(colors), ctypes.c_float(scales[0]),
       ctypes.c_float(scales[1]), ctypes.c_float(scales[2]), ctypes.c_float(fat_zero),
       ctypes.c_size_t(len(idx)), idx.ctypes.data_as(_VP), ctypes.c_size_t(stride),
       ctypes.c_int(radius), out.ctypes.data_as(_VP))
    return out


def imclt_tile(q, horizontal_first=1):
    """Inverse CLT of one [4][8][8] tile -> [16][16] (no accumulation)."""
    q = np.ascontiguousarray(q, np.float64)
    out = np.zeros((16, 16))
    lib().oracle_imclt_tile_f64(q.ctypes.data_as(_VP), out.ctypes.data_as(_VP),
                                ctypes.c_int(horizontal_first))
    return out


def imclt_rbg(geom, clt_cam, rbg, color, v_offset, h_offset, dstride):
    """Oracle of one imclt_rbg pass, in place on rbg (3 planes of
    (H + 8) x dstride, dtype of rbg selects the variant)."""
    clt_cam = np.ascontiguousarray(clt_cam, rbg.dtype)
    fn = lib().oracle_imclt_rbg_f64 if rbg.dtype == np.float64 else lib().oracle_imclt_rbg_f32
    g = _cgeo(geom)
    fn(ctypes.byref(g), clt_cam.ctypes.data_as(_VP), rbg.ctypes.data_as(_VP), ctypes.c_int(color),
       ctypes.c_int(v_offset), ctypes.c_int(h_offset), ctypes.c_size_t(dstride))
    return rbg


def imclt_frame(geom, clt, dstride=None, dtype=np.float64):
    """All cameras, colors and the 4 passes in reference order, planes
    zero-initialised -> list of 4 arrays [3][H + 8][dstride]."""
    dstride = dstride or geom.img_width + 8
    out = []
    for cam in range(4):
        rbg = np.zeros((3, geom.img_height + 8, dstride), dtype)
        for color in range(3):
            for v in range(2):
                for h in range(2):
                    imclt_rbg(geom, clt[cam], rbg, color, v, h, dstride)
        out.append(rbg)
    return out


class TexParams:
    """Texture parameters of textures_accumulate (reference harness values
    by default, src/dtt8x8.cu:791-814)."""

    def __init__(self, colors=3, min_shot=10.0, scale_shot=3.0, diff_sigma=1.5,
                 diff_threshold=10.0, min_agree=3.0, weights=(0.294118, 0.117647, 0.588235),
                 dust_remove=1, keep_weights=0):
        self.colors, self.min_shot, self.scale_shot = colors, min_shot, scale_shot
        self.diff_sigma, self.diff_threshold, self.min_agree = diff_sigma, diff_threshold, min_agree
        self.weights, self.dust_remove, self.keep_weights = weights, dust_remove, keep_weights

    def cargs(self):
        f = ctypes.c_float
        return [ctypes.c_int(self.colors), f(self.min_shot), f(self.scale_shot),
                f(self.diff_sigma), f(self.diff_threshold), f(self.min_agree),
                f(self.weights[0]), f(self.weights[1]), f(self.weights[2]),
                ctypes.c_int(self.dust_remove), ctypes.c_int(self.keep_weights)]

    def slices(self, tiles=True):
        full = self.colors + 1 + 4 + self.colors + 1
        return full if (self.keep_weights or not tiles) else self.colors + 1


PORTS = np.array([[-0.5, -0.5], [0.5, -0.5], [-0.5, 0.5], [0.5, 0.5]], np.float32)


def textures_accumulate(geom, woi, clt, idx, params, rbga=None, rbga_stride=0, tiles=None,
                        tex_stride=0, ports=PORTS):
    """Oracle of textures_accumulate; rbga / tiles are updated in place (their
    dtype must match clt's)."""
    dtype = clt[0].dtype
    clt = [np.ascontiguousarray(c, dtype) for c in clt]
    idx = np.ascontiguousarray(idx, np.int32)
    w = np.ascontiguousarray(woi if woi is not None else [0, 0, 0, 0], np.int32)
    po = np.ascontiguousarray(ports, np.float32)
    fn = lib().oracle_textures_accumulate_f64 if dtype == np.float64 else lib().oracle_textures_accumulate_f32
    g = _cgeo(geom)
    fn(ctypes.byref(g), w.ctypes.data_as(_VP), _ptrs(clt), ctypes.c_size_t(len(idx)),
       idx.ctypes.data_as(_VP), po.ctypes.data_as(_VP), *params.cargs(),
       ctypes.c_size_t(rbga_stride), _VP(rbga.ctypes.data if rbga is not None else 0),
       ctypes.c_size_t(tex_stride), _VP(tiles.ctypes.data if tiles is not None else 0))


def prepare_texture_list(geom, tasks, width=None, height=None):
    """Returns (mutated tasks, indices [tilesx * tilesya], num[8], woi[4])."""
    t = np.array(tasks, copy=True)
    ind = np.zeros(geom.tilesx * geom.tilesya, np.int32)
    num = np.zeros(8, np.int32)
    woi = np.zeros(4, np.int32)
    g = _cgeo(geom)
    lib().oracle_prepare_texture_list(ctypes.byref(g), t.ctypes.data_as(_VP), ctypes.c_int(len(t)),
                                      ind.ctypes.data_as(_VP), num.ctypes.data_as(_VP),
                                      woi.ctypes.data_as(_VP),
                                      ctypes.c_int(width or geom.tilesx),
                                      ctypes.c_int(height or geom.tilesy))
    return t, ind, num, woi


def sub_lists(geom, ind, num):
    """The 8 (parity x inner/border) texture lists as in generate_RBGA's passes."""
    q = geom.tilesx * (geom.tilesya >> 2)
    out = []
    for p in range(8):
        border = p >> 2
        ntt = int(num[((p & 3) << 1) + border])
        off = (p & 3) * q + (q - ntt if border else 0)
        out.append(ind[off:off + ntt].copy())
    return out


def generate_rbga(geom, tasks, clt, params, rbga_stride=None, ports=PORTS):
    """Oracle of generate_RBGA: texture list, then the 8 passes into a zeroed
    RGBA image of (woi_h + 1) * 8 rows per slice.  Returns (tasks, woi, num,
    lists, rbga [colors + 1][rows][rbga_stride])."""
    t, ind, num, woi = prepare_texture_list(geom, tasks)
    rbga_stride = rbga_stride or (geom.tilesx + 1) * 8
    rows = (woi[3] + 1) * 8
    rbga = np.zeros((params.colors + 1, rows, rbga_stride), clt[0].dtype)
    lists = sub_lists(geom, ind, num)
    for lst in lists:
        textures_accumulate(geom, woi, clt, lst, params, rbga=rbga, rbga_stride=rbga_stride,
                            ports=ports)
    return t, woi, num, lists, rbga


def debayer_shot(src, rb_mode, min_shot=10.0, scale_shot=3.0):
    src = np.ascontiguousarray(src, np.float64)
    dst = np.zeros((16, 16))
    lib().oracle_debayer_shot_f64(ctypes.c_int(rb_mode), ctypes.c_float(min_shot),
                                  ctypes.c_float(scale_shot), src.ctypes.data_as(_VP),
                                  dst.ctypes.data_as(_VP))
    return dst


def combine_pixel(D, M, r, c, colors=3, diff_sigma=1.5, weights=(0.294118, 0.117647, 0.588235),
                  dust_remove=1, keep_weights=1, ports=PORTS, alt=0, eta=None):
    """tile_combine_rgba of one pixel: out[colors + 1 + 4 + colors + 1].
    alt overrides the discrete choices (bit 0: blend with the third-best
    port, bit 1: the second-lowest port is the worst); with eta, returns
    (out, ties[2]) where ties are the tie ratios of those two choices."""
    D = np.zeros((4, 3)) + np.asarray(D, np.float64)
    M = np.zeros((4, 3)) + np.asarray(M, np.float64)
    out = np.zeros(12)
    ties = np.zeros(2)
    po = np.ascontiguousarray(ports, np.float32)
    f = ctypes.c_float
    lib().oracle_combine_pixel_f64(ctypes.c_int(colors), f(diff_sigma), f(weights[0]), f(weights[1]),
                                   f(weights[2]), ctypes.c_int(dust_remove), ctypes.c_int(keep_weights),
                                   po.ctypes.data_as(_VP), D.ctypes.data_as(_VP), M.ctypes.data_as(_VP),
                                   ctypes.c_int(r), ctypes.c_int(c), out.ctypes.data_as(_VP),
                                   ctypes.c_int(alt), ctypes.c_double(eta or 0.0),
                                   _VP(ties.ctypes.data if eta is not None else 0))
    return (out, ties) if eta is not None else out


def texture_ties(geom, clt, idx, params, alt=0, eta=1e-6, ports=PORTS):
    """Per valid list entry: all slices [n][colors + 1 + 4 + colors + 1][256]
    of its texture tile (FP64) with the consensus choices overridden by alt,
    and the tie ratios [n][2][256] of every pixel (combine_pixel); rows of
    invalid entries are NaN."""
    clt = [np.ascontiguousarray(c, np.float64) for c in clt]
    idx = np.ascontiguousarray(idx, np.int32)
    nsl = 2 * params.colors + 6
    sl = np.full((len(idx), nsl, 256), np.nan)
    ties = np.full((len(idx), 2, 256), np.nan)
    po = np.ascontiguousarray(ports, np.float32)
    f = ctypes.c_float
    p = params
    g = _cgeo(geom)
    lib().oracle_texture_ties_f64(ctypes.byref(g), _ptrs(clt), ctypes.c_size_t(len(idx)),
                                  idx.ctypes.data_as(_VP), po.ctypes.data_as(_VP),
                                  ctypes.c_int(p.colors), f(p.min_shot), f(p.scale_shot),
                                  f(p.diff_sigma), f(p.diff_threshold), f(p.min_agree),
                                  f(p.weights[0]), f(p.weights[1]), f(p.weights[2]),
                                  ctypes.c_int(p.dust_remove), ctypes.c_int(alt),
                                  ctypes.c_double(eta), sl.ctypes.data_as(_VP),
                                  ties.ctypes.data_as(_VP))
    return sl, ties


def corr_combine(corrs, radius, pair_mask=0x3F):
    """corrs: [tiles * 6][>= (2r+1)^2] rows of a tile-major, pair-minor list
    -> [tiles][(2r+1)^2] combined tiles."""
    corrs = np.ascontiguousarray(corrs, np.float64)
    n = 2 * radius + 1
    nt = corrs.shape[0] // 6
    out = np.zeros((nt, n * n))
    lib().oracle_corr_combine_f64(corrs.ctypes.data_as(_VP), ctypes.c_size_t(corrs.shape[1]),
                                  ctypes.c_int(nt), ctypes.c_int(pair_mask), ctypes.c_int(radius),
                                  out.ctypes.data_as(_VP))
    return out


def corr_disparity(tiles, radius, mode=0):
    """tiles: [n][(2r+1)^2] -> [n][2] (disparity, strength); mode 0 parabola,
    1 centroid."""
    tiles = np.ascontiguousarray(tiles, np.float64)
    out = np.zeros((tiles.shape[0], 2))
    lib().oracle_corr_disparity_f64(tiles.ctypes.data_as(_VP), ctypes.c_size_t(tiles.shape[1]),
                                    ctypes.c_int(tiles.shape[0]), ctypes.c_int(radius),
                                    ctypes.c_int(mode), out.ctypes.data_as(_VP))
    return out


NN_PARAMS = 11874
NN_ACTS = 14850


def bf16_round(a):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def nn_forward(prm, x, with_acts=False, bf16_acts=False):
    """prm: [11874] parameters (float64; pass the BF16-rounded values to
    mirror the GPU); x: [n][7][225] -> out [n][2] (disparity, confidence)
    and optionally acts [n][14850]."""
    prm = np.ascontiguousarray(prm, np.float64)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 7 * 225)
    n = x.shape[0]
    out = np.zeros((n, 2))
    acts = np.zeros((n, NN_ACTS)) if with_acts else None
    lib().oracle_nn_forward_ex(prm.ctypes.data_as(_VP), x.ctypes.data_as(_VP), ctypes.c_int(n),
                               acts.ctypes.data_as(_VP) if with_acts else _VP(0),
                               out.ctypes.data_as(_VP), ctypes.c_int(int(bf16_acts)))
    return (out, acts) if with_acts else out


def nn_backward(prm, x, target, conf_weight=0.1, delta=1.0, scale=None):
    """Gradient (float64[11874]) and value of the mean training loss over the
    tiles with a finite target."""
    prm = np.ascontiguousarray(prm, np.float64)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 7 * 225)
    target = np.ascontiguousarray(target, np.float32)
    n = x.shape[0]
    if scale is None:
        scale = 1.0 / max(1, int(np.isfinite(target).sum()))
    g = np.zeros(NN_PARAMS)
    loss = np.zeros(1)
    lib().oracle_nn_backward(prm.ctypes.data_as(_VP), x.ctypes.data_as(_VP), ctypes.c_int(n),
                             target.ctypes.data_as(_VP), ctypes.c_double(conf_weight),
                             ctypes.c_double(delta), ctypes.c_double(scale), g.ctypes.data_as(_VP),
                             loss.ctypes.data_as(_VP))
    return g, float(loss[0])
