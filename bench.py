#!/usr/bin/env python3
"""Benchmark of the tile processor hot path on one MI355X.

One step = one 4-camera 2592 x 1936 Bayer frame through
  convert_correct_tiles (forward CLT + aberration kernel + phase shift,
  all 78 408 tiles x 4 cameras x 3 colors) and
  correlate2D (6 camera pairs per tile, 15 x 15 output tiles),
the drop-in staged path (the CLT is written to and read back from HBM).
Inputs are synthetic (seeded), already resident in HBM when timing starts.

Prints one JSON line: frames/s, the HBM roofline of the dominant kernel
(algorithmic bytes / its HIP-event duration), the error against the FP64
oracle on a sample of tiles, and the CPU oracle timed on a bounded sample.

  python bench.py [--gpus 1] [--steps 200] [--warmup 300] [--settle-ms 100]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix-core peak
NN_FLOP_PER_TILE = 2 * 11808 * 225  # disparity network: MAC per pixel x pixels x 2
FAT_ZERO = 0.05
SCALES = (0.25, 0.25, 0.5)
RADIUS = 7
# per-kernel HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE runs of
# this benchmark, calibrated on a 1 GiB device copy (written by
# tile_processor_gpu_amd.rocprof_summary; the matching .md gives the commands)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "clt_corr_full_frame_traffic.json")
KERNEL_SYMBOLS = {"convert_correct_tiles": "tp::k_convert_cam",
                  "correlate2D": "void tp::k_correlate2D<3>"}


def measured_traffic(kernel):
    """HBM bytes per launch of one kernel from the committed counter summary,
    or None when there is none."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)[KERNEL_SYMBOLS[kernel]]
    except (OSError, KeyError, ValueError):
        return None
    return {"bytes": int(t["fetch_bytes"] + t["write_bytes"]), "read": int(t["fetch_bytes"]),
            "written": int(t["write_bytes"]), "profiled_avg_us": round(t["avg_us"], 2),
            "source": os.path.relpath(TRAFFIC_FILE, ROOT)}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def hbm_ceilings(reps=10, mib=1024):
    """Practical HBM rates (GB/s, median of reps) of one-shot streaming
    kernels (libtp_probe.so, one 4 KB chunk per workgroup over a grid that
    covers the buffer, the fastest form found): pure read, pure write, copy,
    and the forward CLT's mix of one byte read per three written, each with
    plain and non-temporal stores (the better one counts).  No access
    pattern reaches the 8 TB/s peak; these say how far the measured kernels
    are from what the memory system delivers."""
    import ctypes
    import torch
    from tile_processor_gpu_amd import build as tpb
    lib = ctypes.CDLL(tpb.PROBE_LIB)
    lib.tp_probe_hbm.restype = ctypes.c_int
    lib.tp_probe_hbm.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    nbytes = mib << 20
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty(3 * nbytes // 4, dtype=torch.float32, device="cuda:0")
    sink = torch.zeros(1, dtype=torch.float32, device="cuda:0")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {}
    for name, kind, moved in (("read", 0, 1), ("write", 1, 1), ("copy", 2, 2), ("mix_1r3w", 3, 4)):
        best = None
        for nt in ((0,) if kind == 0 else (0, 1)):
            def run():
                rc = lib.tp_probe_hbm(kind, ctypes.c_void_p(src.data_ptr()),
                                      ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(sink.data_ptr()),
                                      nbytes, nt, st)
                if rc != 0:
                    raise RuntimeError("tp_probe_hbm failed")
            ms = median_ms(run, reps, 2)
            gbs = moved * nbytes / (ms * 1e-3) / 1e9
            best = gbs if best is None else max(best, gbs)
        out[name + "_GBps"] = round(best, 1)
    del src, dst
    torch.cuda.empty_cache()
    return out


def algorithmic_bytes(g, num_corr):
    """Compulsory unique bytes at the reference's buffer boundaries."""
    cams, colors = 4, 3
    images = cams * g.img_width * g.img_height * 4
    kernels = cams * g.kernels_vert * g.kernels_hor * colors * 256 * 4
    koffs = cams * g.kernels_vert * g.kernels_hor * colors * 32
    tasks = g.num_tiles * 40
    clt = cams * g.num_tiles * colors * 256 * 4
    n = 2 * RADIUS + 1
    convert = images + kernels + koffs + tasks + clt
    correlate = clt + num_corr * 4 + num_corr * n * n * 4
    return convert, correlate


def err_vs_fp64(cfg, fb, idx, sample_tiles):
    """Max relative error (max|x - x64| / max|x64|) of the GPU outputs on a
    sample of tiles against the FP64 oracle."""
    import torch
    from oracle import oracle as orc
    g = cfg["geom"]
    tasks = cfg["tasks"][sample_tiles]
    clt64 = orc.convert_correct_tiles(g, cfg["koffs"], cfg["kernels"], cfg["images"], tasks)
    torch.cuda.synchronize()
    worst_clt, rms_clt = 0.0, 0.0
    for cam in range(4):
        got = fb.clt[cam].view(-1, 3, 256)[torch.as_tensor(sample_tiles, device=fb.device)]
        got = got.cpu().numpy().astype(np.float64)
        ref = clt64[cam].reshape(-1, 3, 256)[sample_tiles]
        for color in range(3):
            den = np.abs(ref[:, color]).max()
            d = got[:, color] - ref[:, color]
            worst_clt = max(worst_clt, float(np.abs(d).max() / den))
            rms_clt = max(rms_clt, float(np.sqrt(np.mean(d * d)) / den))
    sel = np.nonzero(np.isin(idx >> 8, sample_tiles))[0]
    ref = orc.correlate2D(g, clt64, 3, SCALES, FAT_ZERO, idx[sel], RADIUS)
    got = fb.corrs[torch.as_tensor(sel, device=fb.device)].cpu().numpy().astype(np.float64)
    n = 2 * RADIUS + 1
    got = got[:, :n * n]
    den = np.abs(ref).max()
    e_corr = float(np.abs(got - ref).max() / den)
    rms_corr = float(np.sqrt(np.mean((got - ref) ** 2)) / den)
    pk_got = got.argmax(1)
    pk_ref = ref.argmax(1)
    # sub-pixel disparity of the combined tiles (centroid), GPU vs FP64
    out = {"clt": worst_clt, "clt_rms": rms_clt, "corr": e_corr, "corr_rms": rms_corr,
           "argmax_mismatch": int((pk_got != pk_ref).sum()), "sample_tiles": int(len(sample_tiles))}
    if len(sel) % 6 == 0 and len(sel):
        d_got = orc.corr_disparity(orc.corr_combine(got, RADIUS), RADIUS, mode=1)[:, 0]
        d_ref = orc.corr_disparity(orc.corr_combine(ref, RADIUS), RADIUS, mode=1)[:, 0]
        out["disparity_max_abs_px"] = float(np.abs(d_got - d_ref).max())
    return out


def cpu_baseline(cfg, idx, rows, min_s=10.0, max_reps=200, threads=0):
    """FP32 oracle (OpenMP over tiles) on the first `rows` tile rows, run
    repeatedly for at least min_s seconds (at least 3 runs); the median run
    is reported.  threads: OpenMP threads (0: all of the host's)."""
    from oracle import oracle as orc
    g = cfg["geom"]
    prev = orc.num_threads()
    if threads:
        orc.num_threads(threads)
    try:
        return _cpu_baseline(orc, cfg, idx, rows, min_s, max_reps)
    finally:
        orc.num_threads(prev)


def _cpu_baseline(orc, cfg, idx, rows, min_s, max_reps):
    g = cfg["geom"]
    cores = orc.num_threads()
    n = rows * g.tilesx
    tasks = cfg["tasks"][:n]
    sel = idx[(idx >> 8) < n]
    # warm-up on a few tiles (page in the inputs)
    orc.convert_correct_tiles(g, cfg["koffs"], cfg["kernels"], cfg["images"], tasks[:g.tilesx],
                              dtype=np.float32)
    times = []
    t_all = time.perf_counter()
    while len(times) < 3 or (time.perf_counter() - t_all < min_s and len(times) < max_reps):
        t0 = time.perf_counter()
        clt = orc.convert_correct_tiles(g, cfg["koffs"], cfg["kernels"], cfg["images"], tasks,
                                        dtype=np.float32)
        orc.correlate2D(g, clt, 3, SCALES, FAT_ZERO, sel, RADIUS)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    frac = n / g.num_tiles
    return {"value": frac / dt, "unit": "frames/s", "cores": int(cores), "kind": "port",
            "sample": f"{rows} of {g.tilesy} tile rows ({n} tiles, {len(sel)} pairs), "
                      f"convert + correlate, FP32 oracle, median of {len(times)} runs "
                      f"({sum(times):.1f} s of CPU work): {dt:.3f} s per run"}


def median_ms(fn, reps, warmup=3):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    for a, b in ev:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    return float(np.median([a.elapsed_time(b) for a, b in ev]))


METHODS = (("fp32", 0), ("mfma_fp16x2", 1), ("mfma_bf16x2", 2), ("mfma_bf16x3", 3))


def measure_fused(cfg, fb, idx, reps, staged_ms, frame_bytes_fused, sample):
    """Fused forward CLT + correlation of the same frame (CLT kept on chip)
    for each forward-DCT method (FP32 butterflies; matrix-core FP16 and BF16
    hi + lo), autotuned; also as one persistent launch replayed from a
    captured graph; errors against the FP64 oracle on the standard frame and
    on a high-dynamic-range frame (pixels log-uniform over 1e-3 .. 1e3)."""
    import torch
    from tile_processor_gpu_amd import autotune, synth, tp
    ref = fb.corrs[:fb.num_corr].clone()
    hdr = dict(cfg, images=synth.make_images(cfg["geom"], np.random.default_rng(99), hdr=True))
    out = {}
    for name, dct in METHODS:
        report = []
        c = autotune.tune_fused(fb, 3, SCALES, FAT_ZERO, cache_path=None, report=report, dct=dct)
        ms = median_ms(lambda: fb.fused(3, SCALES, FAT_ZERO, config=c), reps)
        run = tp.StreamRun([fb], 3, SCALES, FAT_ZERO, config=c)
        run.capture()
        ms_graph = median_ms(run.replay, reps)
        run.release()
        dev = float((fb.corrs[:fb.num_corr] - ref).abs().max() / ref.abs().max())
        fb.fused(3, SCALES, FAT_ZERO, store_clt=True, config=c)
        err = err_vs_fp64(cfg, fb, idx, sample)
        fb.upload(hdr["images"])
        fb.fused(3, SCALES, FAT_ZERO, store_clt=True, config=c)
        err_hdr = err_vs_fp64(hdr, fb, idx, sample)
        fb.upload(cfg["images"])
        torch.cuda.synchronize()
        best = min(ms, ms_graph)
        out[name] = {"frames_per_s": round(1e3 / best, 2), "ms": round(ms, 4),
                     "ms_persistent_graph": round(ms_graph, 4),
                     "speedup_vs_plain": round(staged_ms / best, 3),
                     "hbm_GBps_algorithmic": round(frame_bytes_fused / (best * 1e-3) / 1e9, 1),
                     "config": c, "tuned_over": len(report), "max_rel_dev_vs_plain": dev,
                     "err_vs_fp64": err, "err_vs_fp64_hdr": err_hdr}
    return out


def measure_stream(nframes, reps, seed=4):
    """Batched stream: nframes small-config frames with clustered sparse task
    lists, one persistent launch per batch replayed from a captured graph,
    autotuned; per-frame latency from the kernel's GPU wall-clock stamps."""
    import torch
    from tile_processor_gpu_amd import autotune, synth, tp
    from tile_processor_gpu_amd.geometry import corr_indices
    frames = synth.config_stream(nframes, seed=seed)
    fbs = []
    for f in frames:
        fb = tp.FrameBuffers(f["geom"], "cuda:0", corr_radius=RADIUS, max_corr=6 * len(f["tasks"]))
        fb.upload(f["images"], f["kernels"], f["koffs"], f["tasks"], corr_indices(f["geom"], f["tasks"]))
        fbs.append(fb)
    report = []
    cfg = autotune.tune_stream(fbs, 3, SCALES, FAT_ZERO, cache_path=None, report=report)
    run = tp.StreamRun(fbs, 3, SCALES, FAT_ZERO, config=cfg)
    run.capture()
    ms = median_ms(run.replay, reps)
    run.replay()
    torch.cuda.synchronize()
    t = run.frame_times_s()
    service = (t[:, 1] - t[:, 0]) * 1e3
    done = t[:, 1] * 1e3
    tiles = sum(fb.num_tiles for fb in fbs)
    # staged per-frame launches of the same stream, for comparison
    def staged():
        for fb in fbs:
            fb.convert()
            fb.correlate(3, SCALES, FAT_ZERO)
    staged_ms = median_ms(staged, max(3, reps // 4), warmup=1)
    run.release()
    # the same stream with the forward DCT on the matrix cores (FP16 hi + lo
    # split), autotuned and graph-replayed; deviation from the FP32 stream
    ref = [fb.corrs[:fb.num_corr].clone() for fb in fbs]
    cfg16 = autotune.tune_stream(fbs, 3, SCALES, FAT_ZERO, cache_path=None, dct=1)
    run16 = tp.StreamRun(fbs, 3, SCALES, FAT_ZERO, config=cfg16)
    run16.capture()
    ms16 = median_ms(run16.replay, reps)
    run16.replay()
    torch.cuda.synchronize()
    dev16 = max(float((fb.corrs[:fb.num_corr] - r).abs().max() / r.abs().max()) for fb, r in zip(fbs, ref))
    run16.release()
    return {"frames": nframes, "tiles": int(tiles),
            "active_fraction": [round(min(f["active"] for f in frames), 3),
                                round(max(f["active"] for f in frames), 3)],
            "frames_per_s": round(nframes * 1e3 / ms, 1), "ms_per_batch": round(ms, 4),
            "staged_ms_per_batch": round(staged_ms, 4),
            "speedup_vs_staged": round(staged_ms / ms, 3),
            "latency_ms": {"service_p50": round(float(np.percentile(service, 50)), 4),
                           "service_p99": round(float(np.percentile(service, 99)), 4),
                           "completion_p50": round(float(np.percentile(done, 50)), 4),
                           "completion_p99": round(float(np.percentile(done, 99)), 4)},
            "config": cfg, "tuned_over": len(report), "graph": True,
            "mfma_fp16x2": {"ms_per_batch": round(ms16, 4), "frames_per_s": round(nframes * 1e3 / ms16, 1),
                            "speedup_vs_staged": round(staged_ms / ms16, 3), "config": cfg16,
                            "max_rel_dev_vs_fp32_stream": dev16}}


def make_texture(cfg, fb):
    """generate_RBGA of the whole frame from fb's CLT into an RGBA image
    (the task words, which it rewrites, restored first)."""
    import torch
    from tile_processor_gpu_amd import tp
    g = cfg["geom"]
    dev = torch.device("cuda:0")
    t = np.ascontiguousarray(cfg["tasks"])
    tasks0 = torch.from_numpy(t.view(np.uint8).copy()).to(dev)
    tasks = tasks0.clone()
    ind = torch.zeros(g.tilesx * g.tilesya, dtype=torch.int32, device=dev)
    num = torch.zeros(8, dtype=torch.int32, device=dev)
    woi = torch.zeros(4, dtype=torch.int32, device=dev)
    stride = (g.tilesx + 1) * 8
    rbga = torch.zeros((4 * (g.tilesy + 1) * 8, stride), device=dev)
    ports = torch.tensor([[-0.5, -0.5], [0.5, -0.5], [-0.5, 0.5], [0.5, 0.5]], device=dev)
    kw = tp.TEXTURE_DEFAULTS

    def texture():
        tasks.copy_(tasks0)
        tp.generate_RBGA(g, tasks, len(t), ind, num, woi, g.tilesx, g.tilesy, fb.p_clt, ports,
                         3, 0, kw["min_shot"], kw["scale_shot"], kw["diff_sigma"], kw["diff_threshold"],
                         kw["min_agree"], kw["weight0"], kw["weight1"], kw["weight2"],
                         kw["dust_remove"], 0, stride, rbga)
    return texture


def measure_stages(cfg, fb, idx, reps):
    """The rest of the full-frame pipeline, timed separately (not part of
    `value`): six-pair combination to 15 x 15 tiles, texture output
    (generate_RBGA: list, texture tiles, RGBA overlap-add) and the
    disparity-refinement network (input assembly + forward)."""
    import ctypes
    import torch
    from tile_processor_gpu_amd import _lib, nn, tp
    from tile_processor_gpu_amd.geometry import TASK_DTYPE
    g = cfg["geom"]
    out = {}
    dev = torch.device("cuda:0")
    # staged CLT + correlation list in fb
    fb.convert()
    fb.correlate(3, SCALES, FAT_ZERO)
    comb = torch.zeros((g.num_tiles, 225), device=dev)
    ms = median_ms(lambda: tp.corr_combine(g, fb.num_corr, fb.corr_idx, fb.corrs, fb.corr_stride,
                                           RADIUS, 0x3F, comb, 225), reps)
    out["combine_ms"] = round(ms, 4)
    # algorithmic bytes: the pair tiles read once, the combined tiles written
    comb_bytes = fb.num_corr * (225 * 4 + 4) + g.num_tiles * 225 * 4
    out["combine_GBps"] = round(comb_bytes / (ms * 1e-3) / 1e9, 1)
    disp = torch.zeros((g.num_tiles, 2), device=dev)
    out["centroid_ms"] = round(median_ms(lambda: tp.corr_disparity(comb, 225, g.num_tiles, RADIUS, 1,
                                                                    disp), reps), 4)
    # inverse CLT of the whole frame (4 cameras x 3 colors, all four tile
    # parities) into (W + 8) x (H + 8) planes: one launch (column strips,
    # every pixel written once), and the reference's 48 per-pass launches
    # (read-modify-write) of the same planes
    dstride = g.img_width + 8
    planes = [torch.zeros(3 * (g.img_height + 8) * dstride, device=dev) for _ in range(4)]
    p_planes = tp.ptr_array(planes)
    ms = median_ms(lambda: tp.imclt_rbg_all(g, fb.p_clt, p_planes, dstride), reps)
    imclt_bytes = 4 * g.num_tiles * 3 * 1024 + 4 * 3 * (g.img_height + 8) * dstride * 4
    out["imclt_frame_ms"] = round(ms, 4)
    out["imclt_frame_GBps"] = round(imclt_bytes / (ms * 1e-3) / 1e9, 1)

    def imclt_passes():
        for cam in range(4):
            for color in range(3):
                for v in range(2):
                    for h in range(2):
                        tp.imclt_rbg(g, fb.clt[cam], planes[cam], color, v, h, dstride)
    out["imclt_48_launches_ms"] = round(median_ms(imclt_passes, max(3, reps // 4), warmup=1), 4)
    del planes
    # texture output
    texture = make_texture(cfg, fb)
    out["texture_rgba_ms"] = round(median_ms(texture, max(3, reps // 2)), 4)

    # the whole per-frame pipeline (forward CLT, 6-pair correlation, 15 x 15
    # combination, texture output), staged kernels against the fused kernel
    # (CLT still stored for the texture step)
    def staged_pipeline():
        fb.convert()
        fb.correlate(3, SCALES, FAT_ZERO)
        tp.corr_combine(g, fb.num_corr, fb.corr_idx, fb.corrs, fb.corr_stride, RADIUS, 0x3F, comb, 225)
        texture()

    def fused_pipeline():
        fb.fused(3, SCALES, FAT_ZERO, store_clt=True, combined=True)
        texture()
    out["pipeline_staged_ms"] = round(median_ms(staged_pipeline, max(3, reps // 2)), 4)
    out["pipeline_fused_ms"] = round(median_ms(fused_pipeline, max(3, reps // 2)), 4)
    out["fused_combined_only_ms"] = round(median_ms(
        lambda: fb.fused(3, SCALES, FAT_ZERO, store_pairs=False, combined=True), reps), 4)
    # network on the full frame
    if os.path.exists(nn.DEFAULT_WEIGHTS):
        model = nn.DisparityCNN.load(nn.DEFAULT_WEIGHTS)
        fb.fused(3, SCALES, FAT_ZERO)
        x = torch.empty((g.num_tiles, 7, 225), device=dev)

        def nn_input():
            _lib.call("tp_nn_input", ctypes.byref(g.c()), ctypes.c_void_p(fb.tasks.data_ptr()),
                      g.num_tiles, ctypes.c_void_p(fb.corrs.data_ptr()), fb.corr_stride,
                      ctypes.c_void_p(x.data_ptr()), tp._stream(None))
        out["nn_input_ms"] = round(median_ms(nn_input, reps, warmup=5), 4)
        # a warm-up long enough for the clocks to settle (the kernel draws
        # more power than the memory-bound stages timed before it)
        ms_nn = median_ms(lambda: model.infer(x, g.num_tiles), reps, warmup=20)
        out["nn_frame_ms_matrix_core_bf16"] = round(ms_nn, 4)
        # roofline of the inference kernel in FLOP terms: 11808 MAC per
        # pixel (7 -> 32, 32 -> 32 and 32 -> 2 channels, 3 x 3) x 225 pixels
        # per tile, against the 2.5 PFLOP/s dense BF16 matrix-core peak
        flops = NN_FLOP_PER_TILE * g.num_tiles
        out["nn_infer_TFLOPs"] = round(flops / (ms_nn * 1e-3) / 1e12, 1)
        out["nn_infer_frac_of_bf16_peak"] = round(flops / (ms_nn * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4)
        # the same network reading the correlation rows directly (no
        # [tiles][7][225] input buffer): the input assembly and inference of
        # a frame in one launch
        out["nn_from_corrs_ms"] = round(median_ms(
            lambda: model.infer_corrs(g, fb.tasks, g.num_tiles, fb.corrs, fb.corr_stride), reps, warmup=20), 4)
        out["nn_frame_ms_fp32"] = round(
            median_ms(lambda: model.forward(x, g.num_tiles), 3, warmup=1), 4)
        del x
        res = nn.evaluate(model, nn.TileSource(4, seed=777), batches=2)
        out["nn_accuracy"] = {"data": "rendered piecewise-planar scenes, 4 x 256 x 256, "
                                      "residual U[-2, 2] px, noise 0.5-5%, unseen seeds",
                              **res}
    return out


def setup_frame(seed):
    """The full 4 x 2592 x 1936 frame in HBM.  Its aberration kernels and
    CltExtra records are made on the GPU by tp_psf_to_clt from sampled
    anisotropic Gaussian PSFs (15 x 15, sigma 0.5-2 px, random orientation,
    centroid U[-0.25, 0.25] px), then read back so that the oracle and the CPU
    baseline use the same kernels.  Returns (cfg, fb, idx, kernel_gen)."""
    import torch
    from tile_processor_gpu_amd import synth, tp
    from tile_processor_gpu_amd.geometry import corr_indices
    cfg = synth.config_full(seed=seed)
    g = cfg["geom"]
    idx = corr_indices(g, cfg["tasks"])
    fb = tp.FrameBuffers(g, "cuda:0", corr_radius=RADIUS, max_corr=len(idx))
    fb.upload(cfg["images"], None, None, cfg["tasks"], idx)
    psfs = torch.from_numpy(synth.psf_grid(g, np.random.default_rng(seed + 1000))).cuda()
    keep = fb.kernels_from_psfs(psfs)
    ms = median_ms(lambda: fb.kernels_from_psfs(psfs), 5, 1)
    torch.cuda.synchronize()
    del keep
    cfg["kernels"] = np.stack([k.cpu().numpy() for k in fb.kernels])
    cfg["koffs"] = np.stack([k.cpu().numpy() for k in fb.koffs])
    gen = {"ms": round(ms, 4), "psfs": int(psfs.shape[0] * psfs.shape[1] * psfs.shape[2] * psfs.shape[3]),
           "psf_size": int(psfs.shape[-1])}
    del psfs
    return cfg, fb, idx, gen


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=300,
                    help="untimed frames before the timed ones")
    ap.add_argument("--settle-ms", type=float, default=100.0,
                    help="frames run for this long before the warm-up: the clocks take about 50 "
                         "frames to settle after the host's set-up")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-rows", type=int, default=242, help="tile rows for the CPU baseline (0: skip)")
    ap.add_argument("--err-tiles", type=int, default=512,
                    help="tiles checked against the FP64 oracle after timing (0: skip)")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the plain staged path (no HBM ceilings, fused or stream measurements)")
    ap.add_argument("--stream-frames", type=int, default=64)
    ap.add_argument("--dct", type=int, choices=[0, 1, 2, 3], default=0,
                    help="forward-DCT method of --mode fused: 0 FP32, 1 FP16 hi+lo, 2 BF16 hi+lo, "
                         "3 BF16 hi+mid+lo on the matrix cores")
    ap.add_argument("--mode", choices=["frame", "copy", "fused", "texture", "nn"], default="frame",
                    help="copy: device-to-device copies of 1 GiB (calibrates rocprof "
                         "FETCH_SIZE/WRITE_SIZE against known bytes); fused: the fused kernel "
                         "alone; texture: generate_RBGA, then the whole-frame inverse CLT, each "
                         "--steps times on the full frame; nn: network inference on the full "
                         "frame's tiles (for profiling)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build runs on a single MI355X "
              "(use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from tile_processor_gpu_amd import synth, tp
    from tile_processor_gpu_amd.geometry import corr_indices

    if args.mode == "copy":
        n = 1 << 28  # floats: 1 GiB
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(max(1, args.steps)):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"mode": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n,
                          "copies": max(1, args.steps)}), flush=True)
        return

    if args.mode == "fused":  # the fused kernel alone (default launch config), for profiling
        cfg, fb, idx, _ = setup_frame(args.seed)
        c = tp.FusedConfig(dct=args.dct)
        ms = median_ms(lambda: fb.fused(3, SCALES, FAT_ZERO, config=c), max(1, args.steps), args.warmup)
        print(json.dumps({"mode": "fused", "dct": args.dct, "ms": round(ms, 4), "steps": args.steps}),
              flush=True)
        return

    if args.mode == "nn":  # network inference alone on the full frame's tiles, for profiling
        import ctypes
        from tile_processor_gpu_amd import _lib, nn
        cfg, fb, idx, _ = setup_frame(args.seed)
        g = cfg["geom"]
        model = nn.DisparityCNN.load(nn.DEFAULT_WEIGHTS)
        fb.fused(3, SCALES, FAT_ZERO)
        x = torch.empty((g.num_tiles, 7, 225), device="cuda:0")
        _lib.call("tp_nn_input", ctypes.byref(g.c()), ctypes.c_void_p(fb.tasks.data_ptr()), g.num_tiles,
                  ctypes.c_void_p(fb.corrs.data_ptr()), fb.corr_stride, ctypes.c_void_p(x.data_ptr()),
                  tp._stream(None))
        ms = median_ms(lambda: model.infer(x, g.num_tiles), max(1, args.steps), args.warmup)
        tf = NN_FLOP_PER_TILE * g.num_tiles / (ms * 1e-3) / 1e12
        print(json.dumps({"mode": "nn", "ms": round(ms, 4), "TFLOPs": round(tf, 1),
                          "frac_of_bf16_peak": round(tf / BF16_PEAK_TFLOPS, 4), "steps": args.steps}),
              flush=True)
        return

    if args.mode == "texture":
        cfg, fb, idx, _ = setup_frame(args.seed)
        g = cfg["geom"]
        fb.convert()
        ms_tex = median_ms(make_texture(cfg, fb), max(1, args.steps), args.warmup)
        dstride = g.img_width + 8
        planes = [torch.zeros(3 * (g.img_height + 8) * dstride, device="cuda:0") for _ in range(4)]
        p_planes = tp.ptr_array(planes)
        ms_imclt = median_ms(lambda: tp.imclt_rbg_all(g, fb.p_clt, p_planes, dstride),
                             max(1, args.steps), args.warmup)
        print(json.dumps({"mode": "texture", "texture_rgba_ms": round(ms_tex, 4),
                          "imclt_frame_ms": round(ms_imclt, 4), "steps": args.steps}), flush=True)
        return

    t0 = time.time()
    cfg, fb, idx, kernel_gen = setup_frame(args.seed)
    g = cfg["geom"]
    log(f"inputs {g.img_width}x{g.img_height}, {g.num_tiles} tiles, {len(idx)} pairs, "
        f"kernels from {kernel_gen['psfs']} PSFs in {kernel_gen['ms']} ms ({time.time() - t0:.1f} s)")
    torch.cuda.synchronize()

    def step():
        fb.convert()
        fb.correlate(3, SCALES, FAT_ZERO)

    # After the host has been idle (the frame's set-up above), the GPU needs
    # about 50 frames (30 ms) to reach its steady clocks: the first 25 run
    # about 5% slower.  Frames are run for settle_ms before the warm-up
    # steps, so that a short --warmup still times the steady state.
    t0 = time.perf_counter()
    settle_steps = 0
    while 1e3 * (time.perf_counter() - t0) < args.settle_ms:
        for _ in range(8):
            step()
        settle_steps += 8
        torch.cuda.synchronize()
    settle = {"ms": round(1e3 * (time.perf_counter() - t0), 1), "steps": settle_steps}
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"settle {settle}, warmup {args.warmup} done")

    # Events: s[k] before step k's convert, m[k] between its convert and its
    # correlate, s[K] after the last step (two records per step).
    K = args.steps
    s_ev = [torch.cuda.Event(enable_timing=True) for _ in range(K + 1)]
    m_ev = [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(K):
        s_ev[k].record()
        fb.convert()
        m_ev[k].record()
        fb.correlate(3, SCALES, FAT_ZERO)
    s_ev[K].record()
    torch.cuda.synchronize()
    t_end = time.perf_counter()
    total_s = t_end - t_start
    conv_ms = float(np.mean([s_ev[k].elapsed_time(m_ev[k]) for k in range(K)]))
    corr_ms = float(np.mean([m_ev[k].elapsed_time(s_ev[k + 1]) for k in range(K)]))
    log(f"timed {K} steps: {1e3 * total_s / K:.3f} ms/step; convert {conv_ms:.3f} ms, "
        f"correlate {corr_ms:.3f} ms")

    conv_b, corr_b = algorithmic_bytes(g, len(idx))
    kernels = {
        "convert_correct_tiles": {"ms": conv_ms, "bytes": conv_b,
                                  "GBps": conv_b / (conv_ms * 1e-3) / 1e9},
        "correlate2D": {"ms": corr_ms, "bytes": corr_b, "GBps": corr_b / (corr_ms * 1e-3) / 1e9},
    }
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    d = kernels[dom]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(d["GBps"], 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(d["GBps"] / HBM_PEAK_GBS, 4),
                "traffic": None}
    tr = measured_traffic(dom)
    if tr is not None:
        roofline["traffic"] = tr["bytes"]
        roofline["traffic_detail"] = tr
    ceilings = None
    if not args.no_extra:
        ceilings = hbm_ceilings()
        log(f"hbm ceilings: {ceilings}")
        # the dominant kernel's own read/write mix: convert reads about one
        # byte per three written, correlate about one per two thirds
        practical = ceilings["mix_1r3w_GBps" if dom == "convert_correct_tiles" else "copy_GBps"]
        roofline["practical_ceiling_GBps"] = practical
        roofline["frac_of_practical"] = round(d["GBps"] / practical, 4)
    if tr is not None:
        roofline["traffic_GBps"] = round(tr["bytes"] / (d["ms"] * 1e-3) / 1e9, 1)
    frame_bytes = conv_b + corr_b
    stage = {k: {"ms": round(v["ms"], 4), "GBps": round(v["GBps"], 1)} for k, v in kernels.items()}

    err = None
    if args.err_tiles > 0:
        rng = np.random.default_rng(7)
        sample = np.sort(rng.choice(g.num_tiles, size=min(args.err_tiles, g.num_tiles),
                                    replace=False))
        err = err_vs_fp64(cfg, fb, idx, sample)
        log(f"err vs fp64: {err}")

    extra = {}
    if not args.no_extra:
        fused_bytes = (conv_b - 4 * g.num_tiles * 3 * 1024) + (corr_b - 4 * g.num_tiles * 3 * 1024)
        rng = np.random.default_rng(11)
        sample = np.sort(rng.choice(g.num_tiles, size=min(max(args.err_tiles, 64), g.num_tiles),
                                    replace=False))
        extra["fused"] = measure_fused(cfg, fb, idx, max(5, K // 2), 1e3 * total_s / K, fused_bytes,
                                       sample)
        log(f"fused: {extra['fused']}")
        extra["stream"] = measure_stream(args.stream_frames, max(5, K // 2))
        log(f"stream: {extra['stream']}")
        extra["pipeline_stages"] = measure_stages(cfg, fb, idx, max(5, K // 2))
        log(f"stages: {extra['pipeline_stages']}")

    cpu = cpu1 = None
    if args.cpu_rows > 0:
        cpu = cpu_baseline(cfg, idx, args.cpu_rows)
        log(f"cpu baseline: {cpu}")
        # one thread, on 8 tile rows (about 5 s)
        cpu1 = cpu_baseline(cfg, idx, min(8, args.cpu_rows), min_s=5.0, threads=1)
        log(f"cpu baseline, one thread: {cpu1}")

    value = K / total_s
    out = {
        "metric": "quad-frame CLT+corr frames/s, % HBM roofline, err vs FP64 (plain vs MFMA/NN)",
        "value": round(value, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "settle": settle,
        "ms_per_step": round(1e3 * total_s / K, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Bayer frames; aberration kernels made on the GPU from sampled "
                "anisotropic Gaussian PSFs; planar disparity 0-20 px)",
        "config": {"workload": "plain FP32 staged CLT + aberration + 6-pair correlation (r=7)",
                   "cameras": 4, "width": g.img_width, "height": g.img_height,
                   "tiles": g.num_tiles, "pairs": int(len(idx)), "fat_zero": FAT_ZERO,
                   "method": "plain"},
        "roofline": roofline,
        "hbm_ceilings": ceilings,
        "frame_hbm_GBps": round(frame_bytes / total_s * K / 1e9, 1),
        "stages": stage,
        "err_vs_fp64": err,
        "cpu_baseline": cpu,
        "cpu_baseline_1thread": cpu1,
        "kernel_generation": kernel_gen,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
