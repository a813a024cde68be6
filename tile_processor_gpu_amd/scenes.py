"""Synthetic scenes with known disparity, for training and judging the
disparity-refinement network.

A scene is a random texture seen by the 4 cameras (port offsets of
geometry.PORT_OFFSETS) through a piecewise-planar disparity map: Voronoi
regions, each a plane d = a + gx (x - sx) + gy (y - sy) clipped to
[0, dmax] px.  Camera c sees the texture at (x - d px_c, y - d py_c)
(bicubic resampling of a band-limited texture), plus Gaussian noise with a
per-frame sigma drawn from `noise` (fraction of the signal range).  Tasks
are built at a prior disparity d0 = d_true(tile centre) + U[-residual,
residual], and the target of each tile is the residual d_true - d0 that the
correlation should recover.  Tiles within `margin` tiles of the frame edge
(whose windows are clamped) get a NaN target.

Rendering uses torch on the GPU (data generation, not part of the pipeline).
"""
import numpy as np
import torch
import torch.nn.functional as F

from . import synth
from .geometry import Geometry, PORT_OFFSETS, full_tasks


def _gauss1d(sigma, device):
    r = max(1, int(np.ceil(3 * sigma)))
    x = torch.arange(-r, r + 1, device=device, dtype=torch.float32)
    k = torch.exp(-0.5 * (x / sigma) ** 2)
    return k / k.sum()


def _texture(h, w, sigma, gen, device):
    """Band-limited texture in [0, 1]: Gaussian-smoothed white noise."""
    t = torch.randn((1, 1, h, w), generator=gen, device=device)
    k = _gauss1d(sigma, device)
    t = F.conv2d(F.pad(t, (len(k) // 2, len(k) // 2, 0, 0), mode="reflect"), k.view(1, 1, 1, -1))
    t = F.conv2d(F.pad(t, (0, 0, len(k) // 2, len(k) // 2), mode="reflect"), k.view(1, 1, -1, 1))
    t = t - t.min()
    return (t / t.max().clamp_min(1e-12))[0, 0]


def disparity_map(geom, rng, dmax=20.0, regions=(2, 6), slope=0.03):
    """Piecewise-planar disparity [H][W] (float64) and its plane parameters."""
    H, W = geom.img_height, geom.img_width
    k = int(rng.integers(regions[0], regions[1] + 1))
    seeds = rng.uniform([0, 0], [W, H], (k, 2))
    planes = np.stack([rng.uniform(0, dmax, k), rng.uniform(-slope, slope, k),
                       rng.uniform(-slope, slope, k)], 1)
    return seeds, planes


def eval_disparity(seeds, planes, x, y, dmax=20.0):
    """Disparity at pixel coordinates (x, y) (arrays) of a Voronoi plane map."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    d2 = (x[..., None] - seeds[:, 0]) ** 2 + (y[..., None] - seeds[:, 1]) ** 2
    lab = np.argmin(d2, -1)
    a, gx, gy = planes[lab, 0], planes[lab, 1], planes[lab, 2]
    return np.clip(a + gx * (x - seeds[lab, 0]) + gy * (y - seeds[lab, 1]), 0.0, dmax)


def render_frame(geom, rng, gen, device, dmax=20.0, noise=(0.005, 0.05), residual=2.0,
                 margin=2, psf_sigma=(0.5, 1.0), tex_sigma=(1.0, 2.5)):
    """One frame: dict(images [4][H][W] torch on device, kernels, koffs,
    tasks (numpy), target float32[tiles], d_true float32[tiles])."""
    H, W = geom.img_height, geom.img_width
    pad = int(np.ceil(dmax * 0.5)) + 8
    tex = _texture(H + 2 * pad, W + 2 * pad, float(rng.uniform(*tex_sigma)), gen, device)
    seeds, planes = disparity_map(geom, rng, dmax)
    yy, xx = np.mgrid[0:H, 0:W]
    d = torch.from_numpy(eval_disparity(seeds, planes, xx, yy, dmax).astype(np.float32)).to(device)
    gx = torch.arange(W, device=device, dtype=torch.float32)[None, :].expand(H, W)
    gy = torch.arange(H, device=device, dtype=torch.float32)[:, None].expand(H, W)
    sig = float(rng.uniform(*noise))
    imgs = torch.empty((4, H, W), device=device)
    for c in range(4):
        px, py = PORT_OFFSETS[c]
        sx = gx - d * float(px) + pad
        sy = gy - d * float(py) + pad
        grid = torch.stack([2 * sx / (W + 2 * pad - 1) - 1, 2 * sy / (H + 2 * pad - 1) - 1], -1)
        v = F.grid_sample(tex[None, None], grid[None], mode="bicubic", padding_mode="border",
                          align_corners=True)[0, 0]
        imgs[c] = v + sig * torch.randn((H, W), generator=gen, device=device)
    kern, _ = synth.make_kernels(geom, rng, sigma=psf_sigma)
    koffs = synth.zero_offsets(geom)
    ty, tx = np.divmod(np.arange(geom.num_tiles), geom.tilesx)
    cxs, cys = 8.0 * tx + 4.0, 8.0 * ty + 4.0
    d_true = eval_disparity(seeds, planes, cxs, cys, dmax)
    d0 = d_true + rng.uniform(-residual, residual, geom.num_tiles)
    xy = np.zeros((geom.num_tiles, 4, 2), np.float32)
    for c in range(4):
        xy[:, c, 0] = cxs + d0 * PORT_OFFSETS[c][0]
        xy[:, c, 1] = cys + d0 * PORT_OFFSETS[c][1]
    target = (d_true - d0).astype(np.float32)
    inner = (tx >= margin) & (tx < geom.tilesx - margin) & (ty >= margin) & (ty < geom.tilesy - margin)
    target[~inner] = np.nan
    return dict(images=imgs, kernels=kern, koffs=koffs, tasks=full_tasks(geom, xy), target=target,
                d_true=d_true.astype(np.float32), noise=sig)


def generator(seed, device):
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    return g


SMALL = Geometry(256, 256)
