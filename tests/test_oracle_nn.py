"""Pins of the CNN oracle (forward and gradient) used to check the GPU
network kernels.

* forward: an independent restatement with scipy.signal.correlate2d
  (zero-padded 'same' cross-correlation per channel pair), softmax and
  soft-argmax in numpy;
* gradient: central finite differences of the oracle's own loss along random
  directions in parameter space (float64 parameters), away from the loss's
  kinks (ReLU zeros, the Huber knee, the confidence label's 0.5 px switch);
* BF16 rounding: ties-to-even on hand-made bit patterns.
"""
import numpy as np
import pytest
from scipy.signal import correlate2d

NP = 11874


def _net(rng, scale=0.15):
    p = rng.standard_normal(NP) * scale
    p[2016:2048] = 0.05   # b1
    p[11264:11296] = 0.05  # b2
    return p


def _np_forward(prm, x):
    W1 = prm[0:2016].reshape(32, 7, 3, 3); b1 = prm[2016:2048]
    W2 = prm[2048:11264].reshape(32, 32, 3, 3); b2 = prm[11264:11296]
    W3 = prm[11296:11872].reshape(2, 32, 3, 3); b3 = prm[11872:11874]

    def conv(inp, W, b, relu):
        out = np.zeros((W.shape[0], 15, 15))
        for co in range(W.shape[0]):
            acc = np.full((15, 15), b[co])
            for ci in range(W.shape[1]):
                acc += correlate2d(inp[ci], W[co, ci], mode="same", boundary="fill")
            out[co] = np.maximum(acc, 0) if relu else acc
        return out

    res = []
    for t in range(x.shape[0]):
        h = conv(x[t].reshape(7, 15, 15).astype(np.float64), W1, b1, True)
        h = conv(h, W2, b2, True)
        z = conv(h, W3, b3, False)
        e = np.exp(z[0] - z[0].max())
        p = e / e.sum()
        xs = np.arange(15)[None, :] - 7.0
        res.append(((p * xs).sum(), 1 / (1 + np.exp(-(p * z[1]).sum()))))
    return np.array(res)


def test_forward_matches_scipy(orc):
    rng = np.random.default_rng(1)
    prm = _net(rng)
    x = rng.standard_normal((4, 7, 225)).astype(np.float32)
    got = orc.nn_forward(prm, x)
    ref = _np_forward(prm, x)
    assert np.allclose(got, ref, rtol=1e-11, atol=1e-12)


def test_backward_matches_finite_differences(orc):
    rng = np.random.default_rng(2)
    prm = _net(rng)
    x = rng.standard_normal((3, 7, 225)).astype(np.float32)
    disp = orc.nn_forward(prm, x)[:, 0]
    # errors of 0.2 and 1.6 px: away from the 0.5 px label switch and the
    # Huber knee at 1 px; one tile without a target
    target = np.array([disp[0] - 0.2, disp[1] + 1.6, np.nan], np.float32)
    g, loss = orc.nn_backward(prm, x, target, conf_weight=0.3, delta=1.0)
    assert loss > 0 and np.isfinite(g).all()
    h = 1e-6
    for k in range(6):
        v = rng.standard_normal(NP)
        v /= np.linalg.norm(v)
        lp = orc.nn_backward(prm + h * v, x, target, conf_weight=0.3, delta=1.0)[1]
        lm = orc.nn_backward(prm - h * v, x, target, conf_weight=0.3, delta=1.0)[1]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g @ v) <= 1e-5 * max(1.0, abs(fd)), (k, fd, g @ v)
    # per-block sanity: every parameter block receives gradient
    for a, b in [(0, 2016), (2016, 2048), (2048, 11264), (11264, 11296), (11296, 11872), (11872, NP)]:
        assert np.abs(g[a:b]).max() > 0


def test_backward_ignores_nan_targets(orc):
    rng = np.random.default_rng(3)
    prm = _net(rng)
    x = rng.standard_normal((2, 7, 225)).astype(np.float32)
    g, loss = orc.nn_backward(prm, x, np.array([np.nan, np.nan], np.float32))
    assert loss == 0 and not g.any()


def test_bf16_round(orc):
    bits = np.array([0x3F800000, 0x3F808000, 0x3F818000, 0x3F807FFF, 0x3F80FFFF], np.uint32)
    want = np.array([0x3F800000, 0x3F800000, 0x3F820000, 0x3F800000, 0x3F810000], np.uint32)
    got = orc.bf16_round(bits.view(np.float32)).view(np.uint32)
    assert np.array_equal(got, want)
