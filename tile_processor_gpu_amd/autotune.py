"""Autotuner for the fused and stream launches.

Times every candidate launch configuration (tp_fused_config: tiles per work
item, tiles held in LDS per group, resident workgroups per CU) on the
caller's own buffers with HIP events, and keeps the fastest.  Results are
cached in a JSON file keyed by the library version, the frame geometry and
the workload's shape (number of frames, tiles per frame rounded to a power of
two), so a serving process tunes once per shape.

    best = autotune.tune_stream(frame_buffers)        # list of FrameBuffers
    run = tp.StreamRun(frame_buffers, config=best)
"""
import json
import math
import os

import numpy as np
import torch

from . import tp

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "tile_processor_gpu_amd",
                             "tuning.json")

TILES_PER_ITEM = (2, 4, 6, 8, 12, 16)
TILES_PER_GROUP = (2, 4)
BLOCKS_PER_CU = (2, 3, 4, 6)


def candidates(stream=True):
    out = []
    for gt in TILES_PER_GROUP:
        for tpi in TILES_PER_ITEM:
            if tpi % gt:
                continue
            for bpc in (BLOCKS_PER_CU if stream else (4,)):
                out.append(dict(tiles_per_item=tpi, tiles_per_group=gt, blocks_per_cu=bpc))
    return out


def _median_ms(fn, reps, warmup):
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


def _key(kind, frames):
    g = frames[0].geom
    mean_tiles = max(1, int(np.mean([fb.num_tiles for fb in frames])))
    return (f"{tp.version()}|{kind}|{g.img_width}x{g.img_height}|frames={len(frames)}|"
            f"tiles~2^{int(round(math.log2(mean_tiles)))}")


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, cache):
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune_stream(frames, colors=3, scales=(0.25, 0.25, 0.5), fat_zero=0.05, reps=5, warmup=2,
                cache_path=DEFAULT_CACHE, retune=False, report=None, dct=0):
    """Fastest tp_fused_config for one persistent launch over `frames`
    (a list of tp.FrameBuffers), with the forward-DCT method dct (0 FP32,
    1 FP16 hi + lo, 2 BF16 hi + lo on the matrix cores).  Writes the frames'
    correlation buffers.  report: optional list that receives (config,
    median ms) per candidate."""
    key = _key(f"stream|dct={dct}", frames)
    cache = _load(cache_path) if cache_path else {}
    if not retune and key in cache:
        return dict(cache[key]["config"])
    best, best_ms = None, float("inf")
    for c in candidates(stream=True):
        c = dict(c, dct=dct)
        run = tp.StreamRun(frames, colors, scales, fat_zero, config=c)
        ms = _median_ms(run.launch, reps, warmup)
        if report is not None:
            report.append((c, ms))
        if ms < best_ms:
            best, best_ms = c, ms
    if cache_path:
        cache[key] = {"config": best, "ms": best_ms}
        _store(cache_path, cache)
    return dict(best)


def tune_fused(fb, colors=3, scales=(0.25, 0.25, 0.5), fat_zero=0.05, store_clt=False, reps=5,
               warmup=2, cache_path=DEFAULT_CACHE, retune=False, report=None, dct=0):
    """Fastest tp_fused_config (for the given forward-DCT method) for the
    single-frame fused launch on fb."""
    key = _key(f"fused{'+clt' if store_clt else ''}|dct={dct}", [fb])
    cache = _load(cache_path) if cache_path else {}
    if not retune and key in cache:
        return dict(cache[key]["config"])
    best, best_ms = None, float("inf")
    for c in candidates(stream=False):
        c = dict(c, dct=dct)
        ms = _median_ms(lambda: fb.fused(colors, scales, fat_zero, store_clt=store_clt, config=c),
                        reps, warmup)
        if report is not None:
            report.append((c, ms))
        if ms < best_ms:
            best, best_ms = c, ms
    if cache_path:
        cache[key] = {"config": best, "ms": best_ms}
        _store(cache_path, cache)
    return dict(best)
