"""Aberration kernels in the CLT domain from spatial kernels, and the
reference's on-disk formats.

The reference loads per-camera kernel files built by its (absent) Java host
(src/dtt8x8.cu:228-238, :364-378):
  *.kernel          float32 [KV][KH][3 colors][4 quadrants][8][8]
  *.kernel_offsets  float32 [KV][KH][3] x CltExtra (data_x, data_y,
                    center_x, center_y, dxc_dx, dxc_dy, dyc_dx, dyc_dy),
                    src/TileProcessor.cuh:187-196
  *.bayer           float32 [H][W] image
  *.portsxy         float32 [tiles][2] tile centres of one camera
all raw little-endian floats (readFloatsFromFile, src/dtt8x8.cu:144-155).

CLT-domain kernel of a spatial kernel p (p sampled on the integer grid, its
element [i][j] at x = j - ox, y = i - oy):
    K(k, l) = sum_{x,y} p(x, y) e^{i u_k x} e^{j v_l y},
    u_k = pi (2k + 1) / 16 (horizontal bin k), v_l likewise (vertical bin l)
in the tessarine algebra of the kernel multiply (i: horizontal sine, j:
vertical sine, ij = ji), i.e. the quadrants
    CC = sum p cos(u x) cos(v y)      SC = sum p sin(u x) cos(v y)
    CS = sum p cos(u x) sin(v y)      SS = sum p sin(u x) sin(v y).
The kernel of a unit impulse at t is the phase rotation that
convert_correct_tiles applies for a residual shift t (shiftTileHor /
shiftTileVert, src/TileProcessor.cuh:2715-2776), which moves the content by
+t; by linearity, multiplying a tile's CLT by K(p) approximates the CLT of
the image convolved with p (up to the windowing of the lapped transform,
a few percent for a PSF of about 1 px).  With centre=True the kernel is
taken about the centroid (cx, cy) of p and the centroid goes to CltExtra
data_x/data_y, which convert_correct_tiles applies by moving the window
(the kernel then stays near the identity); the slopes dxc_dx ... come from
central differences of the centroids over the kernel grid (16 px apart).
"""
import os

import numpy as np

from .geometry import Geometry, NUM_CAMS, NUM_COLORS

U = np.pi * (2 * np.arange(8) + 1) / 16.0


def clt_kernel(p, ox=None, oy=None, center=False):
    """[4][8][8] CLT-domain kernel of the spatial kernel p ([h][w]); origin
    (ox, oy) in element coordinates (default: the central element).  With
    center=True the origin is the centroid of p instead.  Returns
    (kernel, (cx, cy)): the origin used, relative to the central element."""
    p = np.asarray(p, np.float64)
    h, w = p.shape
    ox = (w - 1) / 2.0 if ox is None else float(ox)
    oy = (h - 1) / 2.0 if oy is None else float(oy)
    ys, xs = np.mgrid[0:h, 0:w]
    if center:
        s = p.sum()
        if s == 0:
            raise ValueError("kernel with zero sum has no centroid")
        ox, oy = (p * xs).sum() / s, (p * ys).sum() / s
    x = (xs - ox).ravel()
    y = (ys - oy).ravel()
    pv = p.ravel()
    cu, su = np.cos(np.outer(U, x)), np.sin(np.outer(U, x))   # [k][n]
    cv, sv = np.cos(np.outer(U, y)), np.sin(np.outer(U, y))   # [l][n]
    k = np.empty((4, 8, 8))
    k[0] = (cv * pv) @ cu.T           # [l][k]
    k[1] = (cv * pv) @ su.T
    k[2] = (sv * pv) @ cu.T
    k[3] = (sv * pv) @ su.T
    return k, (ox - (w - 1) / 2.0, oy - (h - 1) / 2.0)


def kernel_grid(psfs, center=True):
    """Kernels and CltExtra of a kernel grid.

    psfs: float [cams][KV][KH][3][h][w] spatial kernels per cell and color.
    Returns kernels float32 [cams][KV][KH][3][4][8][8] and kernel offsets
    float32 [cams][KV][KH][3][8]; cell (kx, ky) is centred at pixel
    (16 kx - 8, 16 ky - 8) as the reference's kernel grid."""
    psfs = np.asarray(psfs, np.float64)
    cams, kv, kh, nc = psfs.shape[:4]
    kern = np.zeros((cams, kv, kh, nc, 4, 8, 8), np.float32)
    offs = np.zeros((cams, kv, kh, nc, 8), np.float32)
    cen = np.zeros((cams, kv, kh, nc, 2))
    for idx in np.ndindex(cams, kv, kh, nc):
        k, c = clt_kernel(psfs[idx], center=center)
        kern[idx] = k
        cen[idx] = c
    ky, kx = np.mgrid[0:kv, 0:kh]
    offs[..., 0] = cen[..., 0]
    offs[..., 1] = cen[..., 1]
    offs[..., 2] = (16 * kx - 8)[None, :, :, None]
    offs[..., 3] = (16 * ky - 8)[None, :, :, None]
    # d(centroid)/d(pixel) by central differences over the grid (one-sided at the edges)
    offs[..., 4] = np.gradient(cen[..., 0], 16.0, axis=2)
    offs[..., 5] = np.gradient(cen[..., 0], 16.0, axis=1)
    offs[..., 6] = np.gradient(cen[..., 1], 16.0, axis=2)
    offs[..., 7] = np.gradient(cen[..., 1], 16.0, axis=1)
    return kern, offs


def gaussian_psf(sx, sy=None, theta=0.0, cx=0.0, cy=0.0, size=15):
    """Normalised anisotropic Gaussian sampled on a size x size grid, centred
    (cx, cy) px from the central element."""
    sy = sx if sy is None else sy
    r = (size - 1) / 2.0
    ys, xs = np.mgrid[0:size, 0:size] - r
    xs, ys = xs - cx, ys - cy
    c, s = np.cos(theta), np.sin(theta)
    a = c * xs + s * ys
    b = -s * xs + c * ys
    g = np.exp(-0.5 * ((a / sx) ** 2 + (b / sy) ** 2))
    return g / g.sum()


# ------------------------------------------------------------ file formats
def write_floats(path, a):
    np.ascontiguousarray(a, np.float32).astype("<f4").tofile(path)


def read_floats(path, shape=None):
    a = np.fromfile(path, dtype="<f4")
    return a.reshape(shape) if shape is not None else a


def save_kernels(prefix, kernels, offsets):
    """prefix_chn{c}_transposed.kernel / .kernel_offsets for every camera."""
    for c in range(len(kernels)):
        write_floats(f"{prefix}_chn{c}_transposed.kernel", kernels[c])
        write_floats(f"{prefix}_chn{c}_transposed.kernel_offsets", offsets[c])


def load_kernels(prefix, geom: Geometry, cams=NUM_CAMS):
    kv, kh = geom.kernels_vert, geom.kernels_hor
    k = np.stack([read_floats(f"{prefix}_chn{c}_transposed.kernel", (kv, kh, NUM_COLORS, 4, 8, 8))
                  for c in range(cams)])
    o = np.stack([read_floats(f"{prefix}_chn{c}_transposed.kernel_offsets", (kv, kh, NUM_COLORS, 8))
                  for c in range(cams)])
    return k, o


def save_frame(prefix, images, tasks):
    """prefix_chn{c}.bayer and prefix_chn{c}.portsxy (tile centres of camera c
    from a full task list)."""
    for c in range(len(images)):
        write_floats(f"{prefix}_chn{c}.bayer", images[c])
        write_floats(f"{prefix}_chn{c}.portsxy", np.asarray(tasks["xy"])[:, c, :])


def load_frame(prefix, geom: Geometry, cams=NUM_CAMS):
    from .geometry import full_tasks
    imgs = np.stack([read_floats(f"{prefix}_chn{c}.bayer", (geom.img_height, geom.img_width))
                     for c in range(cams)])
    xy = np.stack([read_floats(f"{prefix}_chn{c}.portsxy", (geom.num_tiles, 2))
                   for c in range(cams)], 1)
    return imgs, full_tasks(geom, xy)
