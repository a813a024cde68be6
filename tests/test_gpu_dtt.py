"""GPU parity of tp_dtt8x8, the 8-mode 2-D transform test entry
(GPU_DTT24_DRV, src/dtt8x8.cuh:103-161), against scipy's orthonormal
DCT/DST of types II and IV: the GPU butterflies on their own."""
import numpy as np
import pytest
import scipy.fft as sf
import torch

from tile_processor_gpu_amd import tp

pytestmark = pytest.mark.gpu
TOL = 1e-6  # max |x - x64| / max |x64|: FP32 butterflies of 8 points


def reference(img, mode):
    """scipy: horizontal transform of each tile row, then vertical."""
    t = 2 if mode > 3 else 4
    fh = sf.dst if mode & 1 else sf.dct
    fv = sf.dst if mode & 2 else sf.dct
    h, w = img.shape
    x = img.reshape(h // 8, 8, w // 8, 8).astype(np.float64)
    x = fh(x, type=t, axis=3, norm="ortho")
    x = fv(x, type=t, axis=1, norm="ortho")
    return x.reshape(h, w)


@pytest.mark.parametrize("mode", range(8))
def test_dtt8x8_matches_scipy(mode):
    rng = np.random.default_rng(mode)
    h, w, stride = 48, 72, 80
    img = rng.standard_normal((h, w)).astype(np.float32)
    print(f"[gpu dtt] mode {mode} {w}x{h} stride {stride}", flush=True)
    src = torch.full((h, stride), 5.0, device="cuda:0")
    src[:, :w] = torch.from_numpy(img).cuda()
    dst = torch.full((h, stride), -3.0, device="cuda:0")
    tp.dtt8x8(dst, src, stride, mode, w, h)
    torch.cuda.synchronize()
    got = dst.cpu().numpy()
    ref = reference(img, mode)
    err = np.abs(got[:, :w] - ref).max() / np.abs(ref).max()
    assert err <= TOL, (mode, err)
    assert (got[:, w:] == -3.0).all()  # past the width: untouched


def test_dtt8x8_in_place_and_inverse():
    """In place (dst = src); DCT-IV and DST-IV are their own inverses, so
    mode 0 twice gives back the image (and mode 3 twice)."""
    rng = np.random.default_rng(11)
    img = rng.standard_normal((32, 64)).astype(np.float32)
    for mode in (0, 3):
        buf = torch.from_numpy(img).cuda()
        tp.dtt8x8(buf, buf, 64, mode, 64, 32)
        once = buf.cpu().numpy()
        tp.dtt8x8(buf, buf, 64, mode, 64, 32)
        torch.cuda.synchronize()
        assert np.abs(once - reference(img, mode)).max() <= TOL * np.abs(once).max()
        assert np.abs(buf.cpu().numpy() - img).max() <= 1e-6 * np.abs(img).max()
