"""Seeded synthetic inputs for the tile processor (host numpy arrays).

  images      4 Bayer RGGB-layout (GR/BG, red at row 0 col 1) FP32 frames
  kernels     CLT-domain aberration kernels [4][KV][KH][3][4][8][8] derived
              from Gaussian point-spread functions
  koffs       per kernel cell/color offsets (CltExtra) [4][KV][KH][3][8]
  tasks       all-tiles task list with tile centres shifted by a disparity
              map along the camera port offsets, plus jitter

The CLT-domain kernel of a point-spread function p is its bicomplex
frequency response at the transform's bin centres w_k = pi (2k+1) / 16:
  K = sum_{x,y} p(x, y) e^{i w_k x} e^{j w_l y},   i^2 = j^2 = -1, ij = ji
so that multiplying a tile's CLT by K (the reference's convolveTiles) acts
as convolution with p.  For a centred Gaussian with covariance S the sum is
taken in closed form, P(u, v) = exp(-[u v] S [u v]^T / 2):
  K_cc = (P(w_k, w_l) + P(w_k, -w_l)) / 2,   K_ss = (P(w_k, -w_l) - P(w_k, w_l)) / 2,
  K_sc = K_cs = 0.
"""
import numpy as np

from .geometry import Geometry, NUM_CAMS, NUM_COLORS, PORT_OFFSETS, full_tasks


def _blur(img, sigma):
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(img, sigma, mode="wrap")


def make_images(geom, rng, blur_sigma=2.0, hdr=False):
    """4 frames: uniform [0, 1] plus a Gaussian-blurred random texture.  With
    hdr=True pixel values are log-uniform over 1e-3 .. 1e3 instead."""
    H, W = geom.img_height, geom.img_width
    out = np.empty((NUM_CAMS, H, W), np.float32)
    for c in range(NUM_CAMS):
        if hdr:
            out[c] = np.exp(rng.uniform(np.log(1e-3), np.log(1e3), (H, W)))
        else:
            tex = _blur(rng.standard_normal((H, W)).astype(np.float32), blur_sigma)
            tex = (tex - tex.min()) / max(1e-12, float(tex.max() - tex.min()))
            out[c] = rng.uniform(0.0, 1.0, (H, W)) + tex
    return out


def psf_kernels(sx, sy, theta):
    """CLT kernels [..., 4, 8, 8] of centred Gaussian PSFs with principal
    sigmas sx, sy (px) rotated by theta (arrays of equal shape)."""
    sx, sy, theta = np.broadcast_arrays(np.asarray(sx, float), np.asarray(sy, float),
                                        np.asarray(theta, float))
    c, s = np.cos(theta), np.sin(theta)
    sxx = (c * sx) ** 2 + (s * sy) ** 2
    syy = (s * sx) ** 2 + (c * sy) ** 2
    sxy = c * s * (sx ** 2 - sy ** 2)
    w = np.pi * (2 * np.arange(8) + 1) / 16.0
    u = w[None, :]    # horizontal frequency (column k)
    v = w[:, None]    # vertical frequency (row l)
    e = lambda a: a[..., None, None]
    base = -0.5 * (e(sxx) * u ** 2 + e(syy) * v ** 2)
    pp = np.exp(base - e(sxy) * u * v)   # P(w_k, w_l)
    pm = np.exp(base + e(sxy) * u * v)   # P(w_k, -w_l)
    k = np.zeros(sx.shape + (4, 8, 8))
    k[..., 0, :, :] = 0.5 * (pp + pm)
    k[..., 3, :, :] = 0.5 * (pm - pp)
    return k


def psf_grid(geom, rng, sigma=(0.5, 2.0), size=15, offset=0.25):
    """Sampled anisotropic Gaussian PSFs of every cell and color,
    float32 [cams][KV][KH][3][size][size]: principal sigmas U[sigma], random
    orientation, centre U[-offset, offset] px from the central element,
    normalised to sum 1 (the input of tp_psf_to_clt)."""
    shape = (NUM_CAMS, geom.kernels_vert, geom.kernels_hor, NUM_COLORS)
    sx = rng.uniform(sigma[0], sigma[1], shape)
    sy = rng.uniform(sigma[0], sigma[1], shape)
    th = rng.uniform(0, np.pi, shape)
    cx = rng.uniform(-offset, offset, shape)
    cy = rng.uniform(-offset, offset, shape)
    r = (size - 1) / 2.0
    g = np.arange(size, dtype=np.float64) - r
    out = np.empty(shape + (size, size), np.float32)
    for cam in range(NUM_CAMS):  # one camera at a time keeps the temporaries small
        e = lambda a: a[cam][..., None, None]  # noqa: E731
        xs = g[None, None, None, None, :] - e(cx)
        ys = g[None, None, None, :, None] - e(cy)
        c, s_ = np.cos(e(th)), np.sin(e(th))
        a = c * xs + s_ * ys
        b = -s_ * xs + c * ys
        p = np.exp(-0.5 * ((a / e(sx)) ** 2 + (b / e(sy)) ** 2))
        out[cam] = p / p.sum(axis=(-2, -1), keepdims=True)
    return out


def identity_kernels(geom):
    """K_cc = 1, others 0: the aberration step becomes the identity."""
    k = np.zeros((NUM_CAMS, geom.kernels_vert, geom.kernels_hor, NUM_COLORS, 4, 8, 8), np.float32)
    k[..., 0, :, :] = 1.0
    return k


def make_kernels(geom, rng, sigma=(0.5, 1.5), anisotropic=False, offsets=0.25, slopes=1e-3):
    """Random PSF kernels and CltExtra offsets for every cell and color."""
    shape = (NUM_CAMS, geom.kernels_vert, geom.kernels_hor, NUM_COLORS)
    sx = rng.uniform(sigma[0], sigma[1], shape)
    if anisotropic:
        sy = rng.uniform(sigma[0], sigma[1], shape)
        th = rng.uniform(0, np.pi, shape)
    else:
        sy, th = sx, np.zeros(shape)
    kern = psf_kernels(sx, sy, th).astype(np.float32)
    koffs = np.zeros(shape + (8,), np.float32)
    ky, kx = np.meshgrid(np.arange(geom.kernels_vert), np.arange(geom.kernels_hor), indexing="ij")
    koffs[..., 0] = rng.uniform(-offsets, offsets, shape)       # data_x
    koffs[..., 1] = rng.uniform(-offsets, offsets, shape)       # data_y
    koffs[..., 2] = (16 * kx - 8)[None, :, :, None]              # center_x
    koffs[..., 3] = (16 * ky - 8)[None, :, :, None]              # center_y
    koffs[..., 4:8] = rng.uniform(-slopes, slopes, shape + (4,))  # d(data)/d(x, y)
    return kern, koffs


def zero_offsets(geom):
    return np.zeros((NUM_CAMS, geom.kernels_vert, geom.kernels_hor, NUM_COLORS, 8), np.float32)


def tile_centres(geom, disparity=0.0, jitter=0.0, rng=None, ports=PORT_OFFSETS):
    """Tile centres [tiles][4][2]: (8 tx + 4, 8 ty + 4) + d * port + jitter.
    disparity: scalar or per-tile array."""
    n = geom.num_tiles
    ty, tx = np.divmod(np.arange(n), geom.tilesx)
    d = np.broadcast_to(np.asarray(disparity, np.float64), (n,))
    xy = np.empty((n, NUM_CAMS, 2))
    for c in range(NUM_CAMS):
        xy[:, c, 0] = 8 * tx + 4 + d * ports[c][0]
        xy[:, c, 1] = 8 * ty + 4 + d * ports[c][1]
    if jitter:
        xy += rng.uniform(-jitter, jitter, xy.shape)
    return xy.astype(np.float32)


def planar_disparity(geom, rng, dmin=0.0, dmax=20.0):
    """Disparity of a random plane over the tile grid, within [dmin, dmax]."""
    ty, tx = np.divmod(np.arange(geom.num_tiles), geom.tilesx)
    a, b = rng.uniform(-1, 1, 2)
    p = a * tx / max(1, geom.tilesx - 1) + b * ty / max(1, geom.tilesy - 1)
    p = (p - p.min()) / max(1e-12, p.max() - p.min())
    return dmin + (dmax - dmin) * p


def config_small(seed=1, geom=None):
    """4 x 256 x 256 Bayer frames, 32 x 32 tiles, isotropic PSF kernels
    (sigma 0.5-1.5 px), per-tile fractional shifts U[-0.5, 0.5] px."""
    geom = geom or Geometry(256, 256)
    rng = np.random.default_rng(seed)
    images = make_images(geom, rng)
    kern, koffs = make_kernels(geom, rng, sigma=(0.5, 1.5))
    xy = tile_centres(geom, 0.0, 0.5, rng)
    return dict(geom=geom, images=images, kernels=kern, koffs=koffs, tasks=full_tasks(geom, xy))


def config_full(seed=2):
    """4 x 2592 x 1936 frames, 324 x 242 tiles, anisotropic PSF kernels (sigma
    0.5-2 px, random orientation), planar disparity 0-20 px + U[-0.5, 0.5]."""
    geom = Geometry(2592, 1936)
    rng = np.random.default_rng(seed)
    images = make_images(geom, rng)
    kern, koffs = make_kernels(geom, rng, sigma=(0.5, 2.0), anisotropic=True)
    d = planar_disparity(geom, rng, 0.0, 20.0)
    xy = tile_centres(geom, d, 0.5, rng)
    return dict(geom=geom, images=images, kernels=kern, koffs=koffs, tasks=full_tasks(geom, xy),
                disparity=d)


def cluster_mask(geom, rng, frac, sigma=3.0):
    """Clustered tile mask: Gaussian-smoothed white noise over the tile grid,
    thresholded so that a fraction frac of the tiles is active."""
    noise = _blur(rng.standard_normal((geom.tilesy, geom.tilesx)), sigma)
    k = int(round(frac * geom.num_tiles))
    if k <= 0:
        return np.zeros(geom.num_tiles, bool)
    thr = np.sort(noise.ravel())[::-1][k - 1]
    return (noise.ravel() >= thr)


def sparse_tasks(tasks, mask):
    """Task list of the active tiles only (list order = tile order)."""
    return np.ascontiguousarray(tasks[np.asarray(mask, bool)])


def config_stream(nframes=64, seed=4, geom=None, frac=(0.1, 1.0)):
    """Stream of small-config frames (4 x 256 x 256), each with its own
    images, isotropic PSF kernels (sigma 0.5-1.5), fractional shifts
    U[-0.5, 0.5] and a clustered sparse task list whose active fraction is
    uniform over frac."""
    geom = geom or Geometry(256, 256)
    rng = np.random.default_rng(seed)
    frames = []
    for _ in range(nframes):
        images = make_images(geom, rng)
        kern, koffs = make_kernels(geom, rng, sigma=(0.5, 1.5))
        xy = tile_centres(geom, 0.0, 0.5, rng)
        f = rng.uniform(*frac)
        tasks = sparse_tasks(full_tasks(geom, xy), cluster_mask(geom, rng, f))
        frames.append(dict(geom=geom, images=images, kernels=kern, koffs=koffs, tasks=tasks,
                           active=f))
    return frames
