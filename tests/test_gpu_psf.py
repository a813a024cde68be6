"""GPU parity of tp_psf_to_clt (CLT-domain kernels and CltExtra records from
sampled PSFs, the inputs the reference reads from its .kernel /
.kernel_offsets files, src/dtt8x8.cu:228-238) against the host generator
kernelgen.kernel_grid (FP64), which tests/test_kernelgen.py pins by impulse
phasors, the closed-form Gaussian response and the convolution invariant.

The GPU sums in FP64 and stores FP32: kernels within 1e-6 of the host
result (max |x - x64| / max |x64|), centroids within 1e-6 px, slopes (from
the stored FP32 centroids) within 1e-8 px/px."""
import numpy as np
import pytest
import torch

from tile_processor_gpu_amd import kernelgen as kg, tp
from tile_processor_gpu_amd.geometry import Geometry
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _psfs(geom, h, w, seed):
    rng = np.random.default_rng(seed)
    kv, kh = geom.kernels_vert, geom.kernels_hor
    out = np.zeros((4, kv, kh, 3, h, w))
    for idx in np.ndindex(4, kv, kh, 3):
        sx, sy = rng.uniform(0.5, 2.0, 2)
        p = kg.gaussian_psf(sx, sy, rng.uniform(0, np.pi), *rng.uniform(-0.4, 0.4, 2), size=max(h, w))
        r0, c0 = (max(h, w) - h) // 2, (max(h, w) - w) // 2
        p = p[r0:r0 + h, c0:c0 + w]
        out[idx] = p + rng.normal(0, 1e-3, p.shape)  # not symmetric about any axis
    return out


def _gpu(geom, psfs, center):
    fb = tp.FrameBuffers(geom, "cuda:0", max_corr=1)
    for c in range(4):
        fb.kernels[c].fill_(np.nan)
        fb.koffs[c].fill_(np.nan)
    keep = fb.kernels_from_psfs(torch.from_numpy(psfs.astype(np.float32)).cuda(), center=center)
    torch.cuda.synchronize()
    del keep
    kv, kh = geom.kernels_vert, geom.kernels_hor
    k = np.stack([fb.kernels[c].cpu().numpy().reshape(kv, kh, 3, 4, 8, 8) for c in range(4)])
    o = np.stack([fb.koffs[c].cpu().numpy().reshape(kv, kh, 3, 8) for c in range(4)])
    return k, o


@pytest.mark.parametrize("h,w,center", [(15, 15, 1), (15, 15, 0), (9, 13, 1), (32, 32, 1)])
def test_psf_to_clt_matches_host_generator(h, w, center):
    g = Geometry(256, 192)
    print(f"[gpu psf] {g} psf {h}x{w} center {center}", flush=True)
    psfs = _psfs(g, h, w, seed=h * 100 + w + center)
    # the host generator takes the same FP32 samples the GPU reads
    p32 = psfs.astype(np.float32).astype(np.float64)
    k_ref, o_ref = kg.kernel_grid(p32, center=bool(center))
    k, o = _gpu(g, psfs, center)
    assert np.isfinite(k).all() and np.isfinite(o).all()
    assert rel_err(k, k_ref) <= 1e-6, rel_err(k, k_ref)
    assert np.abs(o[..., :2] - o_ref[..., :2]).max() <= 1e-6
    assert np.array_equal(o[..., 2:4], o_ref[..., 2:4])
    assert np.abs(o[..., 4:] - o_ref[..., 4:]).max() <= 1e-8
    if center:
        assert np.abs(o_ref[..., 4:]).max() > 1e-4  # the slopes are exercised
    else:
        assert (o[..., :2] == 0).all() and (o[..., 4:] == 0).all()


def test_psf_impulse_gives_identity_kernel():
    """A 1 x 1 unit PSF is the identity kernel (CC = 1, other quadrants 0)."""
    g = Geometry(64, 64)
    psfs = np.ones((4, g.kernels_vert, g.kernels_hor, 3, 1, 1))
    for center in (0, 1):
        k, o = _gpu(g, psfs, center)
        assert np.abs(k[..., 0, :, :] - 1).max() < 1e-7 and np.abs(k[..., 1:, :, :]).max() < 1e-7
        assert (o[..., :2] == 0).all() and (o[..., 4:] == 0).all()
