"""Split-precision matrix-core form of the forward DCT-IV/DST-IV (fused path,
tp_fused_config.dct = 1: FP16 hi + lo, 2: BF16 hi + lo, 3: BF16 hi + mid +
lo) against the FP64 oracle, on the small configuration and its
high-dynamic-range variant (pixels log-uniform over 1e-3 .. 1e3).

Tolerances (max |x - x64| / max |x64| per camera and color, and per
correlation list):  FP16 hi + lo and BF16 hi + mid + lo 1e-5, as the FP32
path;  BF16 hi + lo 1e-4 (16 significant bits: each product carries ~2^-17
relative error; outside the 1e-5 contract, kept as the faster BF16 form).
"""
import numpy as np
import pytest
import torch

from tile_processor_gpu_amd import synth, tp
from tile_processor_gpu_amd.geometry import corr_indices
from conftest import rel_err

pytestmark = pytest.mark.gpu

FZ = 0.05
SCALES = (0.25, 0.25, 0.5)
TOL = {1: 1e-5, 2: 1e-4, 3: 1e-5}


def _case(hdr, seed):
    cfg = synth.config_small(seed=seed)
    g = cfg["geom"]
    if hdr:
        cfg["images"] = synth.make_images(g, np.random.default_rng(seed + 100), hdr=True)
    fb = tp.FrameBuffers(g, "cuda:0", corr_radius=7)
    fb.upload(cfg["images"], cfg["kernels"], cfg["koffs"], cfg["tasks"], corr_indices(g, cfg["tasks"]))
    return cfg, fb


@pytest.fixture(scope="module", params=[False, True], ids=["ldr", "hdr"])
def case(request, orc):
    cfg, fb = _case(request.param, 41)
    g = cfg["geom"]
    print(f"[gpu mfma] {'hdr' if request.param else 'ldr'} {g}", flush=True)
    clt64 = orc.convert_correct_tiles(g, cfg["koffs"], cfg["kernels"], cfg["images"], cfg["tasks"])
    corr64 = orc.correlate2D(g, clt64, 3, SCALES, FZ, corr_indices(g, cfg["tasks"]), 7)
    return dict(cfg=cfg, fb=fb, clt64=clt64, corr64=corr64)


@pytest.mark.parametrize("dct", [1, 2, 3])
def test_mfma_dct_matches_fp64(case, dct):
    fb = case["fb"]
    for c in fb.clt:
        c.fill_(np.nan)
    fb.fused(3, SCALES, FZ, store_clt=True, config=dict(tiles_per_item=8, dct=dct))
    torch.cuda.synchronize()
    worst = 0.0
    for cam in range(4):
        got = fb.clt[cam].cpu().numpy().reshape(-1, 3, 256)
        ref = case["clt64"][cam].reshape(-1, 3, 256)
        for color in range(3):
            e = rel_err(got[:, color], ref[:, color])
            worst = max(worst, e)
            assert e <= TOL[dct], (dct, cam, color, e)
    corr = fb.corrs[:6 * fb.num_tiles].cpu().numpy()
    ec = rel_err(corr, case["corr64"])
    print(f"[gpu mfma] dct={dct}: clt {worst:.2e}, corr {ec:.2e}", flush=True)
    assert ec <= TOL[dct], (dct, ec)


def test_mfma_stream_form(case):
    """Persistent stream launch with the FP16 split: same as the single-frame
    fused launch with it."""
    fb = case["fb"]
    fb.fused(3, SCALES, FZ, config=dict(tiles_per_item=8, dct=1))
    torch.cuda.synchronize()
    ref = fb.corrs[:6 * fb.num_tiles].clone()
    fb.corrs.fill_(-7.0)
    run = tp.StreamRun([fb], 3, SCALES, FZ, config=dict(tiles_per_item=8, dct=1))
    run.launch()
    torch.cuda.synchronize()
    assert torch.equal(fb.corrs[:6 * fb.num_tiles], ref)
