"""GPU parity of the fused (forward CLT + correlation in one pass) and stream
(persistent launch over many frames, optionally replayed from a captured
graph) forms against the staged kernels and the FP64 oracle.

Tolerance: max |x - x64| / max |x64| <= 1e-5 against the oracle, as for the
staged path; against the staged GPU kernels (the same per-tile arithmetic,
differently scheduled) <= 1e-6.
"""
import numpy as np
import pytest
import torch

from tile_processor_gpu_amd import synth, tp
from tile_processor_gpu_amd.geometry import corr_indices
from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5
TOL_STAGED = 1e-6
FZ = 0.05
SCALES = (0.25, 0.25, 0.5)
SENTINEL = -7.0


def _frame(cfg, radius=7):
    g = cfg["geom"]
    fb = tp.FrameBuffers(g, "cuda:0", corr_radius=radius, max_corr=6 * len(cfg["tasks"]))
    fb.upload(cfg["images"], cfg["kernels"], cfg["koffs"], cfg["tasks"], corr_indices(g, cfg["tasks"]))
    return fb


def _staged(fb, colors=3):
    """Staged results: CLT and the correlation list output."""
    fb.convert()
    fb.correlate(colors, SCALES, FZ)
    torch.cuda.synchronize()
    return [c.clone() for c in fb.clt], fb.corrs[:fb.num_corr].clone()


def _dense_rows(cfg):
    """Row (6 task + pair) of every correlation-list entry."""
    t = cfg["tasks"]
    cm = (t["task"] >> 4) & 63
    bits = (cm[:, None] >> np.arange(6)[None, :]) & 1
    ti, pi = np.nonzero(bits)
    return 6 * ti + pi


@pytest.fixture(scope="module")
def small():
    cfg = synth.config_small(seed=21)
    print(f"[gpu fused] small config {cfg['geom']}", flush=True)
    fb = _frame(cfg)
    clt, corr = _staged(fb)
    return dict(cfg=cfg, fb=fb, clt=clt, corr=corr)


def test_fused_matches_staged_and_oracle(small, orc):
    fb, cfg = small["fb"], small["cfg"]
    for c in fb.clt:
        c.fill_(np.nan)
    fb.corrs.fill_(SENTINEL)
    fb.fused(3, SCALES, FZ, store_clt=True)
    torch.cuda.synchronize()
    for cam in range(4):
        a, b = fb.clt[cam].cpu().numpy(), small["clt"][cam].cpu().numpy()
        assert rel_err(a, b) <= TOL_STAGED, (cam, rel_err(a, b))
    got = fb.corrs[:6 * fb.num_tiles].cpu().numpy()
    ref = small["corr"].cpu().numpy()
    assert rel_err(got, ref) <= TOL_STAGED, rel_err(got, ref)
    g = cfg["geom"]
    clt64 = orc.convert_correct_tiles(g, cfg["koffs"], cfg["kernels"], cfg["images"], cfg["tasks"])
    o = orc.correlate2D(g, clt64, 3, SCALES, FZ, corr_indices(g, cfg["tasks"]), 7)
    assert rel_err(got, o) < TOL, rel_err(got, o)


@pytest.mark.parametrize("gt,tpi", [(2, 2), (2, 8), (4, 4), (4, 16), (2, 64)])
def test_fused_configs(small, gt, tpi):
    fb = small["fb"]
    fb.corrs.fill_(SENTINEL)
    fb.fused(3, SCALES, FZ, config=dict(tiles_per_item=tpi, tiles_per_group=gt))
    torch.cuda.synchronize()
    got = fb.corrs[:6 * fb.num_tiles].cpu().numpy()
    assert rel_err(got, small["corr"].cpu().numpy()) <= TOL_STAGED, (gt, tpi)


def test_fused_mono_and_radius(orc):
    cfg = synth.config_small(seed=22)
    g = cfg["geom"]
    fb = _frame(cfg, radius=4)
    _, ref = _staged(fb, colors=1)
    fb.corrs.fill_(SENTINEL)
    fb.fused(1, SCALES, FZ, config=dict(tiles_per_item=4, tiles_per_group=4))
    torch.cuda.synchronize()
    got = fb.corrs[:6 * fb.num_tiles].cpu().numpy()
    assert rel_err(got[:, :81], ref.cpu().numpy()[:, :81]) <= TOL_STAGED
    assert (got[:, 81:] == SENTINEL).all()  # past (2r+1)^2: untouched


def test_fused_sparse_bits_and_nop(orc):
    """Tasks with some pairs off, NOP tasks and out-of-grid tasks: only the
    rows of set pairs are written, and they match the staged kernels."""
    cfg = synth.config_small(seed=23)
    rng = np.random.default_rng(3)
    t = cfg["tasks"].copy()
    n = len(t)
    bits = rng.integers(0, 64, n)
    t["task"] = (t["task"] & 0xF) | (bits << 4)
    t["task"][rng.choice(n, 40, replace=False)] = 0
    bad = rng.choice(n, 5, replace=False)
    t["txy"][bad] = (1000 << 16) | 3  # ty outside the grid
    cfg = dict(cfg, tasks=t)
    fb = _frame(cfg)
    _, ref = _staged(fb)
    fb.corrs.fill_(SENTINEL)
    fb.fused(3, SCALES, FZ)
    torch.cuda.synchronize()
    got = fb.corrs[:6 * n].cpu().numpy()
    idx = corr_indices(cfg["geom"], t)
    rows = _dense_rows(cfg)
    ok = np.ones(len(idx), bool)
    ok[np.isin(rows // 6, bad)] = False  # staged list entries of out-of-grid tiles
    r = ref.cpu().numpy()
    assert rel_err(got[rows[ok]], r[ok]) <= TOL_STAGED
    written = np.zeros(6 * n, bool)
    written[rows[ok]] = True
    assert (got[~written] == SENTINEL).all()


@pytest.fixture(scope="module")
def stream():
    frames = synth.config_stream(nframes=12, seed=31)
    print(f"[gpu fused] stream of {len(frames)} frames, tiles "
          f"{[len(f['tasks']) for f in frames]}", flush=True)
    fbs = [_frame(f) for f in frames]
    refs = []
    for fb in fbs:
        fb.fused(3, SCALES, FZ)
        torch.cuda.synchronize()
        refs.append(fb.corrs[:6 * fb.num_tiles].clone())
    return dict(frames=frames, fbs=fbs, refs=refs)


@pytest.mark.parametrize("cfg", [dict(tiles_per_item=8, tiles_per_group=2, blocks_per_cu=4),
                                 dict(tiles_per_item=2, tiles_per_group=2, blocks_per_cu=1),
                                 dict(tiles_per_item=16, tiles_per_group=4, blocks_per_cu=2)])
def test_stream_matches_single_frames(stream, cfg):
    fbs = stream["fbs"]
    for fb in fbs:
        fb.corrs.fill_(SENTINEL)
    run = tp.StreamRun(fbs, 3, SCALES, FZ, config=cfg)
    run.launch()
    torch.cuda.synchronize()
    for i, fb in enumerate(fbs):
        got = fb.corrs[:6 * fb.num_tiles].cpu().numpy()
        assert not (got == SENTINEL).any(), (cfg, i)
        assert rel_err(got, stream["refs"][i].cpu().numpy()) <= TOL_STAGED, (cfg, i)
    times = run.frame_times_s()
    assert (times[:, 1] >= times[:, 0]).all() and times[:, 0].min() == 0.0
    assert run.drained()


@pytest.mark.parametrize("dct,tol", [(1, 1e-5), (2, 1e-4), (3, 1e-5)])
def test_stream_mfma_matches_fp32_stream(stream, dct, tol):
    """The persistent stream with the split-precision matrix-core forward DCT
    (1: FP16 hi + lo, 2: BF16 hi + lo, 3: BF16 hi + mid + lo), launched and
    replayed from a graph,
    against the FP32 stream on the same frames (FP32 is within 1e-6 of the
    FP64 oracle, so the tolerance is the split form's own bound)."""
    fbs = stream["fbs"]
    ref_run = tp.StreamRun(fbs, 3, SCALES, FZ, config=dict(tiles_per_item=8))
    ref_run.launch()
    torch.cuda.synchronize()
    refs = [fb.corrs[:6 * fb.num_tiles].cpu().numpy() for fb in fbs]
    for fb in fbs:
        fb.corrs.fill_(SENTINEL)
    run = tp.StreamRun(fbs, 3, SCALES, FZ, config=dict(tiles_per_item=8, dct=dct))
    run.capture()
    run.replay()
    torch.cuda.synchronize()
    for i, fb in enumerate(fbs):
        got = fb.corrs[:6 * fb.num_tiles].cpu().numpy()
        assert not (got == SENTINEL).any(), (dct, i)
        assert rel_err(got, refs[i]) <= tol, (dct, i)
    run.release()


def test_stream_graph_replays(stream):
    """A captured graph replayed several times, with new image data uploaded
    between replays and eager work in between: every replay matches the
    eager stream launch on the same data."""
    fbs = stream["fbs"][:5]
    run = tp.StreamRun(fbs, 3, SCALES, FZ, config=dict(tiles_per_item=4))
    run.capture()
    rng = np.random.default_rng(9)
    for rep in range(3):
        k = rep % len(fbs)
        fbs[k].images[1].mul_(float(rng.uniform(0.5, 2.0)))
        for fb in fbs:
            fb.corrs.fill_(SENTINEL)
        run.replay()
        torch.cuda.synchronize()
        got = [fb.corrs[:6 * fb.num_tiles].clone() for fb in fbs]
        for fb in fbs:
            fb.corrs.fill_(SENTINEL)
        run.launch()
        torch.cuda.synchronize()
        for i, fb in enumerate(fbs):
            assert torch.equal(got[i], fb.corrs[:6 * fb.num_tiles]), (rep, i)
    run.release()


@pytest.mark.parametrize("pair_mask", [0x3F, 0x13])
def test_fused_combined_matches_combine(small, orc, pair_mask):
    """The six-pair combination written by the fused kernel (pair rows not
    stored) against tp_corr_combine on the staged correlations and against
    the oracle's combination of the FP64 correlations."""
    fb, cfg = small["fb"], small["cfg"]
    g = cfg["geom"]
    ref_rows = small["corr"]
    comb_ref = torch.zeros((g.num_tiles, 225), device="cuda:0")
    tp.corr_combine(g, fb.num_corr, fb.corr_idx, ref_rows, 225, 7, pair_mask, comb_ref, 225)
    fb.corrs.fill_(SENTINEL)
    fb.combined_buffer().fill_(SENTINEL)
    fb.fused(3, SCALES, FZ, store_pairs=False, combined=True, pair_mask=pair_mask)
    torch.cuda.synchronize()
    got = fb.combined[:g.num_tiles].cpu().numpy()
    assert (fb.corrs.cpu().numpy() == SENTINEL).all()  # pair rows not written
    assert rel_err(got, comb_ref.cpu().numpy()) <= TOL_STAGED
    clt64 = orc.convert_correct_tiles(g, cfg["koffs"], cfg["kernels"], cfg["images"], cfg["tasks"])
    c64 = orc.correlate2D(g, clt64, 3, SCALES, FZ, corr_indices(g, cfg["tasks"]), 7)
    o = orc.corr_combine(c64, 7, pair_mask=pair_mask)
    assert rel_err(got, o) < TOL


def test_stream_combined(stream):
    fbs = stream["fbs"][:4]
    refs = []
    for fb in fbs:
        fb.fused(3, SCALES, FZ, store_pairs=False, combined=True)
        torch.cuda.synchronize()
        refs.append(fb.combined[:fb.num_tiles].clone())
        fb.combined.fill_(SENTINEL)
    run = tp.StreamRun(fbs, 3, SCALES, FZ, config=dict(tiles_per_item=8), store_pairs=False,
                       combined=True)
    run.launch()
    torch.cuda.synchronize()
    for i, fb in enumerate(fbs):
        assert rel_err(fb.combined[:fb.num_tiles].cpu().numpy(), refs[i].cpu().numpy()) <= TOL_STAGED


@pytest.fixture(scope="module")
def stream64(orc):
    """The 64-frame sparse stream (clustered masks, 10-100 % of the tiles
    active, per-frame images, kernels and shifts; the bench's batch) and the
    FP64 oracle's correlation rows of every frame."""
    frames = synth.config_stream(nframes=64, seed=4)
    act = [f["active"] for f in frames]
    print(f"[gpu fused] C4 stream: 64 frames, active {min(act):.3f}..{max(act):.3f}, "
          f"{sum(len(f['tasks']) for f in frames)} tiles", flush=True)
    fbs = [_frame(f) for f in frames]
    refs = []
    for f in frames:
        g = f["geom"]
        clt64 = orc.convert_correct_tiles(g, f["koffs"], f["kernels"], f["images"], f["tasks"])
        refs.append(orc.correlate2D(g, clt64, 3, SCALES, FZ, corr_indices(g, f["tasks"]), 7))
    return dict(frames=frames, fbs=fbs, refs=refs)


@pytest.mark.parametrize("dct", [0, 1])
def test_stream64_matches_oracle(stream64, dct):
    """Every row (6 task + pair) of every frame of the 64-frame stream, run
    as one persistent launch replayed from a captured graph (FP32 forward DCT,
    or the FP16 hi + lo matrix-core form), within 1e-5 of the FP64 oracle;
    rows past each frame's tasks untouched."""
    fbs = stream64["fbs"]
    for fb in fbs:
        fb.corrs.fill_(SENTINEL)
    run = tp.StreamRun(fbs, 3, SCALES, FZ, config=dict(tiles_per_item=8, dct=dct))
    run.capture()
    run.replay()
    torch.cuda.synchronize()
    assert run.drained()
    worst = []
    for i, (f, fb) in enumerate(zip(stream64["frames"], fbs)):
        rows = _dense_rows(f)
        allr = fb.corrs.cpu().numpy()
        got = allr[rows]
        e = rel_err(got, stream64["refs"][i])
        worst.append(e)
        assert e <= TOL, (dct, i, f["active"], e)
        assert (allr[6 * fb.num_tiles:] == SENTINEL).all(), (dct, i)
    i_min = int(np.argmin([f["active"] for f in stream64["frames"]]))
    i_max = int(np.argmax([f["active"] for f in stream64["frames"]]))
    print(f"dct {dct}: max rel err {max(worst):.2e}; sparsest frame ({stream64['frames'][i_min]['active']:.2f}) "
          f"{worst[i_min]:.2e}, densest ({stream64['frames'][i_max]['active']:.2f}) {worst[i_max]:.2e}")
    run.release()


def test_stream64_new_task_lists_between_replays(stream64, orc):
    """New sparse task lists uploaded into a frame after the graph was
    captured (another mask, so another task count): the replay follows it,
    and the frame's rows match the FP64 oracle on its new list."""
    fbs = stream64["fbs"][:6]
    frames = stream64["frames"][:6]
    run = tp.StreamRun(fbs, 3, SCALES, FZ, config=dict(tiles_per_item=8))
    run.capture()
    rng = np.random.default_rng(17)
    k = 2
    f = frames[k]
    g = f["geom"]
    from tile_processor_gpu_amd.geometry import full_tasks
    full = full_tasks(g, synth.tile_centres(g, 0.0, 0.5, rng))
    # half the frame's own active fraction: the rows fit its corrs buffer
    new_tasks = synth.sparse_tasks(full, synth.cluster_mask(g, rng, 0.5 * f["active"]))
    assert 0 < len(new_tasks) < len(f["tasks"])
    fbs[k].upload(tasks=new_tasks)
    fbs[k].corrs.fill_(SENTINEL)
    run.replay()
    torch.cuda.synchronize()
    clt64 = orc.convert_correct_tiles(g, f["koffs"], f["kernels"], f["images"], new_tasks)
    ref = orc.correlate2D(g, clt64, 3, SCALES, FZ, corr_indices(g, new_tasks), 7)
    nf = dict(f, tasks=new_tasks)
    allr = fbs[k].corrs.cpu().numpy()
    assert rel_err(allr[_dense_rows(nf)], ref) <= TOL
    assert (allr[6 * len(new_tasks):] == SENTINEL).all()
    # restore the fixture's frame
    fbs[k].upload(tasks=f["tasks"])
    run.release()
