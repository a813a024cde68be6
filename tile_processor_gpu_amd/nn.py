"""Disparity-refinement network: parameters, training and evaluation over
the HIP kernels of libtp.so (tp_nn_*).

The network reads, per tile, the combined 15 x 15 correlation and the six
pair tiles mapped onto the horizontal axis (tp_nn_input) and predicts the
residual disparity (soft-argmax column) and a confidence.  Training data
comes from scenes.render_frame: synthetic piecewise-planar scenes with known
disparity, rendered to 4 Bayer cameras and passed through this pipeline
(fused forward CLT + correlation, tp_convert_correlate).  The plain centroid
of the combined tile (tp_corr_disparity, mode 1) is the baseline, the
fallback for tiles the network is not confident about, and the yardstick
its accuracy is compared with.

    python -m tile_processor_gpu_amd.nn train --tiles 2000000 --out weights.npz
"""
import argparse
import ctypes
import json
import math
import os
import time

import numpy as np
import torch

from . import _lib, scenes, tp

NPARAMS = 11874
NINPUT = 1575
NACTS = 14850
LAYERS = (("W1", 0, (32, 7, 3, 3)), ("b1", 2016, (32,)), ("W2", 2048, (32, 32, 3, 3)),
          ("b2", 11264, (32,)), ("W3", 11296, (2, 32, 3, 3)), ("b3", 11872, (2,)))
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights",
                               "disparity_cnn.npz")
_VP = ctypes.c_void_p


def _p(t):
    return _VP(t.data_ptr()) if t is not None else _VP(0)


def init_params(seed=0):
    """He-normal weights, zero biases (float32[NPARAMS])."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, np.float32)
    for name, off, shape in LAYERS:
        n = int(np.prod(shape))
        if name.startswith("W"):
            fan_in = shape[1] * 9
            std = math.sqrt(2.0 / fan_in) * (0.5 if name == "W3" else 1.0)
            p[off:off + n] = rng.standard_normal(n) * std
    return p


class DisparityCNN:
    """FP32 master parameters + Adam state on the GPU, BF16 copy for the
    kernels."""

    def __init__(self, params=None, device="cuda:0", seed=0):
        self.device = torch.device(device)
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        if p.shape != (NPARAMS,):
            raise ValueError("parameter vector has the wrong size")
        f32 = dict(dtype=torch.float32, device=self.device)
        self.master = torch.from_numpy(p.copy()).to(self.device)
        self.m = torch.zeros(NPARAMS, **f32)
        self.v = torch.zeros(NPARAMS, **f32)
        self.grads = torch.zeros(NPARAMS, **f32)
        self.loss = torch.zeros(1, **f32)
        self.bf16 = torch.zeros(NPARAMS, dtype=torch.int16, device=self.device)
        self.step = 0
        _lib.call("tp_nn_pack_bf16", _p(self.master), NPARAMS, _p(self.bf16), tp._stream(None))

    # ---------------------------------------------------------------- I/O
    def params(self):
        return self.master.cpu().numpy()

    def params_bf16(self):
        """The BF16 values the kernels read, as float32."""
        b = self.bf16.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
        return b.view(np.float32)

    def save(self, path, meta=None):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, params=self.params(), meta=json.dumps(meta or {}))

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS, device="cuda:0"):
        with np.load(path, allow_pickle=False) as z:
            return cls(z["params"], device=device)

    # --------------------------------------------------------- kernels
    def forward(self, x, n, out=None, acts=None, stream=None):
        """x: float32 [n][7][225] device tensor -> out [n][2] (disparity, confidence)."""
        if out is None:
            out = torch.empty((n, 2), dtype=torch.float32, device=self.device)
        _lib.call("tp_nn_forward", _p(self.bf16), _p(x), int(n), _p(acts), _p(out), tp._stream(stream))
        return out

    def infer(self, x, n, out=None, stream=None):
        """Inference on the matrix cores (BF16 activations): x [n][7][225] ->
        out [n][2] (disparity, confidence)."""
        if out is None:
            out = torch.empty((n, 2), dtype=torch.float32, device=self.device)
        _lib.call("tp_nn_infer", _p(self.bf16), _p(x), int(n), _p(out), tp._stream(stream))
        return out

    def infer_corrs(self, geom, tasks, n, corrs, corr_stride, out=None, stream=None):
        """infer() on the input that tp_nn_input would assemble, read straight
        from the correlation rows (6 task + pair, radius 7): tasks is the
        device task list, n the number of tasks."""
        if out is None:
            out = torch.empty((n, 2), dtype=torch.float32, device=self.device)
        g = geom.c()
        _lib.call("tp_nn_infer_corrs", ctypes.byref(g), _p(self.bf16), _p(tasks), int(n), _p(corrs),
                  int(corr_stride), _p(out), tp._stream(stream))
        return out

    def gradient(self, x, target, n, acts, conf_weight=0.1, huber_delta=1.0, stream=None):
        """Gradient of the mean loss into self.grads (overwritten); returns
        the loss tensor (device, not synchronised)."""
        st = tp._stream(stream)
        nvalid = int(torch.isfinite(target[:n]).sum().item())
        _lib.call("tp_fill", _p(self.grads), NPARAMS, ctypes.c_float(0.0), st)
        _lib.call("tp_fill", _p(self.loss), 1, ctypes.c_float(0.0), st)
        self.forward(x, n, acts=acts, stream=stream, out=torch.empty((n, 2), device=self.device))
        _lib.call("tp_nn_backward", _p(self.bf16), _p(x), _p(acts), _p(target), int(n),
                  float(conf_weight), float(huber_delta), 1.0 / max(1, nvalid), _p(self.grads),
                  _p(self.loss), st)
        return self.loss

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        self.step += 1
        _lib.call("tp_nn_adam", _p(self.master), _p(self.m), _p(self.v), _p(self.grads), NPARAMS,
                  float(lr), float(beta1), float(beta2), float(eps), self.step, _p(self.bf16),
                  tp._stream(stream))


# --------------------------------------------------------------- data
class TileSource:
    """Renders synthetic frames and turns them into network inputs with the
    pipeline: fused forward CLT + correlation, then tp_nn_input."""

    def __init__(self, frames_per_batch=4, geom=scenes.SMALL, device="cuda:0", seed=1,
                 noise=(0.005, 0.05), residual=2.0):
        self.geom, self.device = geom, torch.device(device)
        self.rng = np.random.default_rng(seed)
        self.gen = scenes.generator(seed, self.device)
        self.noise, self.residual = noise, residual
        self.fbs = [tp.FrameBuffers(geom, self.device, corr_radius=7) for _ in range(frames_per_batch)]
        n = frames_per_batch * geom.num_tiles
        self.n = n
        self.x = torch.zeros((n, 7, 225), dtype=torch.float32, device=self.device)
        self.target = torch.zeros(n, dtype=torch.float32, device=self.device)
        self.baseline = torch.zeros((n, 2), dtype=torch.float32, device=self.device)

    def next(self):
        """Fill self.x / self.target / self.baseline (centroid) for a new batch."""
        g = self.geom
        nt = g.num_tiles
        tgt = np.empty(self.n, np.float32)
        for i, fb in enumerate(self.fbs):
            fr = scenes.render_frame(g, self.rng, self.gen, self.device, noise=self.noise,
                                     residual=self.residual)
            for c in range(4):
                fb.images[c][:, :g.img_width].copy_(fr["images"][c])
            fb.upload(None, fr["kernels"], fr["koffs"], fr["tasks"])
            fb.fused(3, (0.25, 0.25, 0.5), 0.05)
            xs = self.x[i * nt:(i + 1) * nt]
            _lib.call("tp_nn_input", ctypes.byref(g.c()), _p(fb.tasks), nt, _p(fb.corrs),
                      fb.corr_stride, _p(xs), tp._stream(None))
            tgt[i * nt:(i + 1) * nt] = fr["target"]
        self.target.copy_(torch.from_numpy(tgt))
        _lib.call("tp_corr_disparity", _p(self.x), 7 * 225, self.n, 7, 1, _p(self.baseline),
                  tp._stream(None))
        return self.x, self.target, self.n


def metrics(est, target, conf=None, fallback=None, threshold=0.5):
    """RMSE and outlier rate (> 0.5 px) over tiles with a finite target."""
    est = np.asarray(est, np.float64)
    t = np.asarray(target, np.float64)
    ok = np.isfinite(t) & np.isfinite(est)
    if conf is not None and fallback is not None:
        est = np.where(np.asarray(conf) >= threshold, est, np.asarray(fallback))
    e = est[ok] - t[ok]
    return {"rmse": float(np.sqrt(np.mean(e * e))), "outliers": float(np.mean(np.abs(e) > 0.5)),
            "tiles": int(ok.sum())}


def evaluate(model, source, batches=2):
    """Network vs centroid on fresh frames from `source`."""
    nn_est, nn_conf, nn_bf, cen, tg = [], [], [], [], []
    for _ in range(batches):
        x, target, n = source.next()
        out = model.forward(x, n)
        out_bf = model.infer(x, n)
        torch.cuda.synchronize()
        o = out.cpu().numpy()
        nn_est.append(o[:, 0]); nn_conf.append(o[:, 1]); nn_bf.append(out_bf[:, 0].cpu().numpy())
        cen.append(source.baseline[:, 0].cpu().numpy()); tg.append(target.cpu().numpy())
    nn_est, nn_conf, nn_bf, cen, tg = map(np.concatenate, (nn_est, nn_conf, nn_bf, cen, tg))
    return {"nn": metrics(nn_est, tg), "nn_matrix_core_bf16": metrics(nn_bf, tg),
            "centroid": metrics(cen, tg),
            "nn_with_fallback": metrics(nn_est, tg, nn_conf, cen),
            "mean_confidence": float(np.nanmean(nn_conf))}


def train(tiles=2_000_000, frames_per_batch=4, lr=2e-3, seed=0, data_seed=1, log_every=50,
          conf_weight=0.1, huber_delta=1.0, model=None, log=print):
    """Train on freshly rendered frames until `tiles` tiles have been seen
    (Adam, cosine learning-rate decay).  Returns (model, history)."""
    model = model or DisparityCNN(seed=seed)
    src = TileSource(frames_per_batch, seed=data_seed)
    acts = torch.empty((src.n, NACTS), dtype=torch.float32, device=model.device)
    steps = max(1, tiles // src.n)
    hist = []
    t0 = time.time()
    for s in range(steps):
        x, target, n = src.next()
        lr_s = lr * 0.5 * (1 + math.cos(math.pi * s / steps))
        loss = model.gradient(x, target, n, acts, conf_weight, huber_delta)
        model.adam(max(lr_s, lr * 0.02))
        if (s + 1) % log_every == 0 or s == steps - 1:
            hist.append((s + 1, float(loss.item())))
            log(f"[nn] step {s + 1}/{steps} tiles {(s + 1) * n} loss {hist[-1][1]:.4f} "
                f"({time.time() - t0:.1f} s)")
    return model, hist


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--tiles", type=int, default=2_000_000)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--frames-per-batch", type=int, default=4)
    t.add_argument("--out", default=DEFAULT_WEIGHTS)
    t.add_argument("--eval-batches", type=int, default=4)
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=DEFAULT_WEIGHTS)
    e.add_argument("--eval-batches", type=int, default=4)
    a = ap.parse_args()
    if a.cmd == "train":
        t0 = time.time()
        model, hist = train(a.tiles, a.frames_per_batch, a.lr)
        secs = time.time() - t0
        res = evaluate(model, TileSource(a.frames_per_batch, seed=12345), a.eval_batches)
        meta = {"tiles": a.tiles, "train_s": round(secs, 1), "final_loss": hist[-1][1], "eval": res}
        model.save(a.out, meta)
        print(json.dumps(meta), flush=True)
    else:
        model = DisparityCNN.load(a.weights)
        print(json.dumps(evaluate(model, TileSource(4, seed=12345), a.eval_batches)), flush=True)


if __name__ == "__main__":
    main()
