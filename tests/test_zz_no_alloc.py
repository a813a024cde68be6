"""tp_stream_convert_correlate_noalloc (the stream form with the library's
own work buffer, for callers that cannot allocate) against
tp_stream_convert_correlate with a caller-allocated work buffer, on two
calls in a row, the second with a smaller input (fewer frames, shorter task
lists), so the second call runs on the work buffer the first one left."""
import numpy as np
import pytest
import torch

from tile_processor_gpu_amd import synth, tp
from tile_processor_gpu_amd.geometry import corr_indices

pytestmark = pytest.mark.gpu
SCALES = (0.25, 0.25, 0.5)
FZ = 0.05


def _frames(nframes, seed, shrink=1.0):
    out = []
    for f in synth.config_stream(nframes=nframes, seed=seed):
        t = f["tasks"][:max(1, int(len(f["tasks"]) * shrink))]
        g = f["geom"]
        fb = tp.FrameBuffers(g, "cuda:0", corr_radius=7, max_corr=6 * len(t))
        fb.upload(f["images"], f["kernels"], f["koffs"], t, corr_indices(g, t))
        out.append(fb)
    return out


def _outputs(fbs, launch):
    for fb in fbs:
        fb.corrs.fill_(-7.0)
        fb.combined_buffer().fill_(-7.0)
    launch()
    torch.cuda.synchronize()
    return [(fb.corrs[:6 * fb.num_tiles].clone(), fb.combined[:fb.num_tiles].clone()) for fb in fbs]


@pytest.mark.parametrize("cfg", [dict(tiles_per_item=8), dict(tiles_per_item=4, tiles_per_group=2,
                                                              blocks_per_cu=2)])
def test_noalloc_matches_work_buffer_form(cfg):
    for call, (nframes, seed, shrink) in enumerate([(12, 51, 1.0), (5, 52, 0.5)]):
        print(f"[noalloc] call {call}: {nframes} frames, seed {seed}, cfg {cfg}", flush=True)
        fbs = _frames(nframes, seed, shrink)
        run = tp.StreamRun(fbs, 3, SCALES, FZ, config=cfg, combined=True)
        ref = _outputs(fbs, run.launch)
        got = _outputs(fbs, run.launch_noalloc)
        for i, ((rc, rb), (gc, gb)) in enumerate(zip(ref, got)):
            assert not (rc == -7.0).any(), (call, i)
            # the same kernels on the same frames: bit-identical rows
            assert torch.equal(rc, gc), (call, i)
            assert torch.equal(rb, gb), (call, i)
        times = run.frame_times_s()  # stamps written by the no-alloc launch
        assert np.isfinite(times).all() and (times[:, 1] >= times[:, 0]).all()
