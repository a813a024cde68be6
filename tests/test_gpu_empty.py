"""Zero-length inputs on the GPU: no tasks, an empty correlation, combination
or texture list.  Every entry point returns success and writes nothing to
its outputs, and a stream batch may hold frames without tiles next to full
ones (their rows stay untouched, the others match the single-frame form).

For an empty texture list the reference's list build leaves the window of
interest at its cleared value (TX, TY, 0, 0) (clear_texture_list,
src/TileProcessor.cuh:1716-1728) and generate_RBGA then turns x1, y1 into a
width and height (:1534-1535): (TX, TY, 1 - TX, 1 - TY), with all eight
sub-list counts zero.
"""
import numpy as np
import pytest
import torch

from tile_processor_gpu_amd import synth, tp
from tile_processor_gpu_amd.geometry import TASK_DTYPE, corr_indices
from conftest import rel_err

pytestmark = pytest.mark.gpu
S = -5.0
SCALES = (0.25, 0.25, 0.5)
FZ = 0.05
PORTS = [[-0.5, -0.5], [0.5, -0.5], [-0.5, 0.5], [0.5, 0.5]]


@pytest.fixture(scope="module")
def frame():
    cfg = synth.config_small(seed=8)
    g = cfg["geom"]
    print(f"[gpu empty] {g}", flush=True)
    fb = tp.FrameBuffers(g, "cuda:0", corr_radius=7, max_corr=6 * len(cfg["tasks"]))
    fb.upload(cfg["images"], cfg["kernels"], cfg["koffs"], cfg["tasks"], corr_indices(g, cfg["tasks"]))
    return cfg, fb


def test_empty_lists_write_nothing(frame):
    cfg, fb = frame
    g = fb.geom
    for c in fb.clt:
        c.fill_(S)
    fb.corrs.fill_(S)
    tp.convert_correct_tiles(g, fb.p_koffs, fb.p_kernels, fb.p_images, fb.tasks, fb.p_clt,
                             fb.dstride, 0)
    tp.correlate2D(g, fb.p_clt, 3, *SCALES, FZ, 0, fb.corr_idx, fb.corr_stride, 7, fb.corrs)
    comb = torch.full((4, 225), S, device="cuda:0")
    tp.convert_correlate(g, fb.p_koffs, fb.p_kernels, fb.p_images, fb.tasks, fb.p_clt, fb.dstride,
                         0, 0, 3, *SCALES, FZ, fb.corr_stride, 7, fb.corrs, comb, 225)
    tp.corr_combine(g, 0, fb.corr_idx, fb.corrs, fb.corr_stride, 7, 0x3F, comb, 225)
    ports = torch.tensor(PORTS, device="cuda:0")
    tiles = torch.full((4, 1024), S, device="cuda:0")
    kw = tp.TEXTURE_DEFAULTS
    tp.textures_accumulate(g, None, fb.p_clt, 0, fb.corr_idx, ports, 3, 0, kw["min_shot"],
                           kw["scale_shot"], kw["diff_sigma"], kw["diff_threshold"], kw["min_agree"],
                           kw["weight0"], kw["weight1"], kw["weight2"], kw["dust_remove"], 0, 0,
                           None, 1024, tiles)
    torch.cuda.synchronize()
    assert all(bool((c == S).all()) for c in fb.clt)
    assert bool((fb.corrs == S).all())
    assert bool((comb == S).all()) and bool((tiles == S).all())


def test_generate_rbga_without_tiles(frame):
    cfg, fb = frame
    g = fb.geom
    tasks = torch.zeros(40, dtype=torch.uint8, device="cuda:0")
    ind = torch.full((g.tilesx * g.tilesya,), 7, dtype=torch.int32, device="cuda:0")
    num = torch.full((8,), 7, dtype=torch.int32, device="cuda:0")
    woi = torch.full((4,), 7, dtype=torch.int32, device="cuda:0")
    stride = (g.tilesx + 1) * 8
    rbga = torch.full((4 * (g.tilesy + 1) * 8, stride), S, device="cuda:0")
    ports = torch.tensor(PORTS, device="cuda:0")
    kw = tp.TEXTURE_DEFAULTS
    tp.generate_RBGA(g, tasks, 0, ind, num, woi, g.tilesx, g.tilesy, fb.p_clt, ports, 3, 0,
                     kw["min_shot"], kw["scale_shot"], kw["diff_sigma"], kw["diff_threshold"],
                     kw["min_agree"], kw["weight0"], kw["weight1"], kw["weight2"],
                     kw["dust_remove"], 0, stride, rbga)
    torch.cuda.synchronize()
    assert num.cpu().tolist() == [0] * 8
    assert woi.cpu().tolist() == [g.tilesx, g.tilesy, 1 - g.tilesx, 1 - g.tilesy]
    assert bool((rbga == S).all())


def test_stream_with_empty_frames(frame):
    """Frames 1 and 3 of a five-frame batch have no tiles; the persistent
    launch and a graph replay of it fill the other frames exactly as the
    single-frame fused form does and leave the empty frames' rows alone."""
    cfg, fb = frame
    g = fb.geom
    full = synth.config_stream(nframes=3, seed=44)
    fbs = []
    for i in range(5):
        f = tp.FrameBuffers(g, "cuda:0", corr_radius=7, max_corr=6 * g.num_tiles)
        if i in (1, 3):
            f.upload(cfg["images"], cfg["kernels"], cfg["koffs"], np.zeros(0, dtype=TASK_DTYPE))
        else:
            src = full[[0, 2, 4].index(i)]
            f.upload(src["images"], src["kernels"], src["koffs"], src["tasks"])
        fbs.append(f)
    assert [f.num_tiles for f in fbs][1] == 0 and fbs[3].num_tiles == 0
    refs = {}
    for i in (0, 2, 4):
        fbs[i].fused(3, SCALES, FZ)
        torch.cuda.synchronize()
        refs[i] = fbs[i].corrs[:6 * fbs[i].num_tiles].cpu().numpy()
    run = tp.StreamRun(fbs, 3, SCALES, FZ, config=dict(tiles_per_item=4))
    for launch in (run.launch, run.replay):
        for f in fbs:
            f.corrs.fill_(S)
        launch()
        torch.cuda.synchronize()
        for i, f in enumerate(fbs):
            got = f.corrs.cpu().numpy()
            if i in refs:
                n = 6 * f.num_tiles
                assert rel_err(got[:n], refs[i]) <= 1e-6, i
                assert bool((got[n:] == S).all()), i
            else:
                assert bool((got == S).all()), i
        assert run.drained()
    times = run.frame_times_s()
    assert np.isnan(times[1]).all() and np.isnan(times[3]).all()
    run.release()
