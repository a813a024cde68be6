"""GPU kernels of the disparity-refinement network against the CNN oracle
(FP64), and what the trained network gains over the plain centroid.

Tolerances: forward outputs and activations max|x - x64| / max|x64| <= 1e-5
(FP32 arithmetic on the same BF16 parameters); gradients per parameter
block <= 1e-4 (FP32 sums over up to 225 pixels x 32 channels per tile and
atomically accumulated workgroup partials); the input assembly is a copy of
correlation values (exact) and their sum (<= 1e-6).
"""
import os

import numpy as np
import pytest
import torch

from tile_processor_gpu_amd import nn, synth, tp, _lib
from tile_processor_gpu_amd.geometry import corr_indices
from conftest import rel_err

pytestmark = pytest.mark.gpu

BLOCKS = [(0, 2016), (2016, 2048), (2048, 11264), (11264, 11296), (11296, 11872), (11872, 11874)]


def _inputs(seed=3, ntiles=None):
    cfg = synth.config_small(seed=seed)
    g = cfg["geom"]
    fb = tp.FrameBuffers(g, "cuda:0", corr_radius=7)
    fb.upload(cfg["images"], cfg["kernels"], cfg["koffs"], cfg["tasks"], corr_indices(g, cfg["tasks"]))
    fb.fused(3, (0.25, 0.25, 0.5), 0.05)
    n = g.num_tiles
    x = torch.zeros((n, 7, 225), device="cuda:0")
    import ctypes
    _lib.call("tp_nn_input", ctypes.byref(g.c()), ctypes.c_void_p(fb.tasks.data_ptr()), n,
              ctypes.c_void_p(fb.corrs.data_ptr()), fb.corr_stride, ctypes.c_void_p(x.data_ptr()),
              tp._stream(None))
    torch.cuda.synchronize()
    return g, fb, x


def test_nn_input_matches_oracle(orc):
    g, fb, x = _inputs()
    corrs = fb.corrs[:6 * g.num_tiles].cpu().numpy().astype(np.float64)
    xs = x.cpu().numpy()
    for p in range(6):
        ref = orc.corr_combine(corrs, 7, pair_mask=1 << p)
        assert np.array_equal(xs[:, 1 + p], ref), p
    ref = orc.corr_combine(corrs, 7)
    assert rel_err(xs[:, 0], ref) <= 1e-6


def test_nn_infer_corrs_equals_input_then_infer():
    """tp_nn_infer_corrs (input assembled while reading the correlation rows)
    gives exactly tp_nn_input + tp_nn_infer, also for tasks with pairs off,
    a task off the grid and a list shorter than the frame."""
    import ctypes
    g, fb, _ = _inputs(seed=8)
    model = nn.DisparityCNN(seed=11)
    tasks = fb.tasks[:40 * g.num_tiles].clone()
    tv = tasks.view(-1, 40)
    rng = np.random.default_rng(2)
    words = tv[:, :4].contiguous().view(torch.int32).cpu().numpy()[:, 0].copy()
    for i in rng.choice(g.num_tiles, 200, replace=False):
        words[i] &= ~(int(rng.integers(1, 63)) << 4)  # some pairs off
    tv[:, :4] = torch.from_numpy(words.astype(np.int32).view(np.uint8).reshape(-1, 4)).cuda()
    txy = tv[:, 4:8].contiguous().view(torch.int32).cpu().numpy()[:, 0].copy()
    txy[5] = (g.tilesy + 3) << 16  # off the grid
    tv[:, 4:8] = torch.from_numpy(txy.astype(np.int32).view(np.uint8).reshape(-1, 4)).cuda()
    for n in (g.num_tiles, 517):
        x = torch.zeros((n, 7, 225), device="cuda:0")
        _lib.call("tp_nn_input", ctypes.byref(g.c()), ctypes.c_void_p(tasks.data_ptr()), n,
                  ctypes.c_void_p(fb.corrs.data_ptr()), fb.corr_stride, ctypes.c_void_p(x.data_ptr()),
                  tp._stream(None))
        two = model.infer(x, n)
        one = model.infer_corrs(g, tasks, n, fb.corrs, fb.corr_stride)
        torch.cuda.synchronize()
        print(f"[gpu nn] infer_corrs n={n}", flush=True)
        assert torch.equal(one, two), n
        assert torch.isfinite(one).all()


@pytest.fixture(scope="module")
def net_case(orc):
    g, fb, x = _inputs(seed=4)
    n = 48
    xs = x[:n].contiguous()
    model = nn.DisparityCNN(seed=7)
    prm = model.params_bf16().astype(np.float64)
    return dict(x=xs, n=n, model=model, prm=prm)


def test_nn_forward_matches_oracle(orc, net_case):
    x, n, model, prm = net_case["x"], net_case["n"], net_case["model"], net_case["prm"]
    acts = torch.zeros((n, nn.NACTS), device="cuda:0")
    out = model.forward(x, n, acts=acts)
    torch.cuda.synchronize()
    ref, ref_acts = orc.nn_forward(prm, x.cpu().numpy(), with_acts=True)
    got = out.cpu().numpy()
    assert np.abs(got[:, 0] - ref[:, 0]).max() <= 1e-5 * 7, np.abs(got[:, 0] - ref[:, 0]).max()
    assert rel_err(got[:, 1], ref[:, 1]) <= 1e-5
    a = acts.cpu().numpy()
    for lo, hi in [(0, 7200), (7200, 14400), (14400, 14850)]:
        assert rel_err(a[:, lo:hi], ref_acts[:, lo:hi]) <= 1e-5, (lo, hi)
    # inference without saved activations gives the same outputs
    out2 = model.forward(x, n)
    torch.cuda.synchronize()
    assert torch.equal(out, out2)


def test_nn_infer_matrix_core_matches_oracle(orc, net_case):
    """Matrix-core inference (BF16 activations) against the oracle with the
    activations rounded to BF16 at the same points: disparity within 0.02 px
    and confidence within 0.01 (a BF16 rounding that falls on the other side
    of a tie moves one activation by 2^-8 relative)."""
    x, n, model, prm = net_case["x"], net_case["n"], net_case["model"], net_case["prm"]
    out = model.infer(x, n)
    torch.cuda.synchronize()
    ref = orc.nn_forward(prm, x.cpu().numpy(), bf16_acts=True)
    got = out.cpu().numpy()
    assert np.abs(got[:, 0] - ref[:, 0]).max() <= 0.02, np.abs(got[:, 0] - ref[:, 0]).max()
    assert np.abs(got[:, 1] - ref[:, 1]).max() <= 0.01, np.abs(got[:, 1] - ref[:, 1]).max()
    full = orc.nn_forward(prm, x.cpu().numpy())  # FP32-activation network, for scale
    print(f"[gpu nn] infer vs bf16 oracle {np.abs(got[:, 0] - ref[:, 0]).max():.2e} px, "
          f"bf16 vs fp activations {np.abs(ref[:, 0] - full[:, 0]).max():.2e} px", flush=True)


def test_nn_backward_matches_oracle(orc, net_case):
    x, n, model, prm = net_case["x"], net_case["n"], net_case["model"], net_case["prm"]
    disp = orc.nn_forward(prm, x.cpu().numpy())[:, 0]
    rng = np.random.default_rng(5)
    target = (disp + rng.choice([-1.7, -0.25, 0.2, 1.4], n)).astype(np.float32)
    target[::7] = np.nan
    tgt = torch.from_numpy(target).cuda()
    acts = torch.zeros((n, nn.NACTS), device="cuda:0")
    loss = model.gradient(x, tgt, n, acts, conf_weight=0.3, huber_delta=1.0)
    torch.cuda.synchronize()
    g_ref, loss_ref = orc.nn_backward(prm, x.cpu().numpy(), target, conf_weight=0.3, delta=1.0)
    g = model.grads.cpu().numpy()
    assert abs(float(loss.item()) - loss_ref) <= 1e-5 * abs(loss_ref)
    for lo, hi in BLOCKS:
        assert rel_err(g[lo:hi], g_ref[lo:hi]) <= 1e-4, (lo, hi, rel_err(g[lo:hi], g_ref[lo:hi]))


def test_nn_adam_matches_numpy(net_case):
    model = nn.DisparityCNN(seed=9)
    rng = np.random.default_rng(1)
    w = model.params().astype(np.float64)
    m = np.zeros_like(w); v = np.zeros_like(w)
    for step in range(1, 4):
        gr = rng.standard_normal(nn.NPARAMS).astype(np.float32)
        model.grads.copy_(torch.from_numpy(gr))
        model.adam(1e-2)
        m = 0.9 * m + 0.1 * gr
        v = 0.999 * v + 0.001 * gr.astype(np.float64) ** 2
        w = w - 1e-2 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
    torch.cuda.synchronize()
    assert np.abs(model.params() - w).max() <= 1e-5
    from oracle import oracle as orc_mod
    assert np.array_equal(model.params_bf16(), orc_mod.bf16_round(model.params()))


def test_nn_training_reduces_loss():
    """A short training run on rendered scenes lowers the loss."""
    model, hist = nn.train(tiles=100_000, frames_per_batch=4, lr=3e-3, log_every=1, log=lambda s: None)
    first, last = hist[0][1], np.mean([h[1] for h in hist[-2:]])
    print(f"[gpu nn] loss {first:.4f} -> {last:.4f}", flush=True)
    assert last < 0.8 * first


@pytest.mark.skipif(not os.path.exists(nn.DEFAULT_WEIGHTS), reason="no trained weights")
def test_trained_network_beats_centroid():
    """The committed network (trained on the GPU by these kernels) against
    the plain centroid on freshly rendered frames it has not seen."""
    model = nn.DisparityCNN.load(nn.DEFAULT_WEIGHTS)
    res = nn.evaluate(model, nn.TileSource(4, seed=777), batches=2)
    print(f"[gpu nn] {res}", flush=True)
    assert res["nn"]["rmse"] <= res["centroid"]["rmse"]
    assert res["nn_matrix_core_bf16"]["rmse"] <= res["centroid"]["rmse"]
    assert res["nn_with_fallback"]["rmse"] <= res["centroid"]["rmse"]
