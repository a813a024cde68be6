"""GPU parity: the HIP kernels of libtp.so against the CPU oracle (FP64 and
FP32) on the small 4 x 256 x 256 configuration, plus edge cases.

Tolerance (FP32 outputs vs FP64): max |x - x64| / max |x64| <= 1e-5 per
array (per camera and color for the CLT, per list for correlations, per
plane for imclt), with fat_zero = 0.05 for the correlation normalisation.
"""
import numpy as np
import pytest
import torch

from tile_processor_gpu_amd import synth, tp
from tile_processor_gpu_amd.geometry import Geometry, corr_indices, full_tasks
from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5
FAT_ZERO = 0.05
SCALES = (0.25, 0.25, 0.5)


@pytest.fixture(scope="module")
def small(orc):
    cfg = synth.config_small(seed=11)
    g = cfg["geom"]
    print(f"[gpu parity] small config {g}", flush=True)
    fb = tp.FrameBuffers(g, "cuda:0", corr_radius=7)
    idx = corr_indices(g, cfg["tasks"])
    fb.upload(cfg["images"], cfg["kernels"], cfg["koffs"], cfg["tasks"], idx)
    clt64 = orc.convert_correct_tiles(g, cfg["koffs"], cfg["kernels"], cfg["images"], cfg["tasks"])
    clt32 = orc.convert_correct_tiles(g, cfg["koffs"], cfg["kernels"], cfg["images"], cfg["tasks"],
                                      dtype=np.float32)
    return dict(cfg=cfg, g=g, fb=fb, idx=idx, clt64=clt64, clt32=clt32)


def gpu_clt(fb):
    torch.cuda.synchronize()
    return [c.cpu().numpy().astype(np.float64) for c in fb.clt]


def test_convert_matches_oracle(small):
    fb = small["fb"]
    for c in fb.clt:
        c.fill_(np.nan)
    fb.convert()
    got = gpu_clt(fb)
    for cam in range(4):
        for color in range(3):
            a = got[cam].reshape(-1, 3, 256)[:, color]
            b = small["clt64"][cam].reshape(-1, 3, 256)[:, color]
            c = small["clt32"][cam].reshape(-1, 3, 256)[:, color]
            assert np.isfinite(a).all()
            assert rel_err(a, b) < TOL, (cam, color, rel_err(a, b))
            assert rel_err(a, c) < TOL, (cam, color, rel_err(a, c))


def test_convert_other_grid(orc):
    """A 34 x 26 tile grid (tile rows not a multiple of 4, tile columns not a
    multiple of 4): the forward CLT's visiting order must still cover every
    tile exactly once."""
    g = Geometry(272, 208)
    cfg = synth.config_small(seed=5, geom=g)
    print(f"[gpu parity] odd grid {g}", flush=True)
    fb = tp.FrameBuffers(g, "cuda:0", corr_radius=7)
    fb.upload(cfg["images"], cfg["kernels"], cfg["koffs"], cfg["tasks"],
              corr_indices(g, cfg["tasks"]))
    for c in fb.clt:
        c.fill_(np.nan)
    fb.convert()
    got = gpu_clt(fb)
    ref = orc.convert_correct_tiles(g, cfg["koffs"], cfg["kernels"], cfg["images"], cfg["tasks"])
    for cam in range(4):
        assert np.isfinite(got[cam]).all(), cam
        assert rel_err(got[cam], ref[cam]) < TOL, (cam, rel_err(got[cam], ref[cam]))


def test_convert_is_deterministic(small):
    fb = small["fb"]
    fb.convert()
    a = gpu_clt(fb)
    fb.convert()
    b = gpu_clt(fb)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_correlate_matches_oracle(small, orc):
    fb, g, idx = small["fb"], small["g"], small["idx"]
    fb.convert()
    fb.correlate(3, SCALES, FAT_ZERO)
    torch.cuda.synchronize()
    got = fb.corrs[:len(idx)].cpu().numpy()
    ref = orc.correlate2D(g, small["clt64"], 3, SCALES, FAT_ZERO, idx, 7)
    assert rel_err(got, ref) < TOL, rel_err(got, ref)
    # isolated: oracle CLT in, GPU correlation out
    for cam in range(4):
        fb.clt[cam].copy_(torch.from_numpy(small["clt64"][cam].astype(np.float32)))
    fb.correlate(3, SCALES, FAT_ZERO)
    torch.cuda.synchronize()
    got2 = fb.corrs[:len(idx)].cpu().numpy()
    assert rel_err(got2, ref) < TOL, rel_err(got2, ref)


@pytest.mark.parametrize("radius", [0, 1, 4, 7])
def test_correlate_radius_and_stride(small, orc, radius):
    g = small["g"]
    n = 2 * radius + 1
    idx = small["idx"][:600]
    stride = n * n + 3
    fb = tp.FrameBuffers(g, "cuda:0", corr_radius=radius, max_corr=len(idx), corr_stride=stride)
    for cam in range(4):
        fb.clt[cam].copy_(torch.from_numpy(small["clt64"][cam].astype(np.float32)))
    fb.upload(corr_idx=idx)
    fb.corrs.fill_(-7.0)
    fb.correlate(3, SCALES, FAT_ZERO)
    torch.cuda.synchronize()
    got = fb.corrs.cpu().numpy()
    ref = orc.correlate2D(g, small["clt64"], 3, SCALES, FAT_ZERO, idx, radius)
    assert rel_err(got[:, :n * n], ref) < TOL
    assert (got[:, n * n:] == -7.0).all()  # padding untouched


def test_correlate_mono_and_bad_pairs(small, orc):
    g = small["g"]
    idx = small["idx"][:300].copy()
    idx[5] = (idx[5] & ~0xFF) | 6       # pair 6: skipped
    idx[17] = (idx[17] & ~0xFF) | 0x3F  # skipped
    fb = tp.FrameBuffers(g, "cuda:0", corr_radius=7, max_corr=len(idx))
    for cam in range(4):
        fb.clt[cam].copy_(torch.from_numpy(small["clt64"][cam].astype(np.float32)))
    fb.upload(corr_idx=idx)
    fb.corrs.fill_(3.0)
    fb.correlate(1, (1.0, 0.0, 0.0), FAT_ZERO)
    torch.cuda.synchronize()
    got = fb.corrs.cpu().numpy()
    ref = orc.correlate2D(g, small["clt64"], 1, (1.0, 0.0, 0.0), FAT_ZERO, idx, 7)
    keep = np.ones(len(idx), bool)
    keep[[5, 17]] = False
    assert rel_err(got[keep], ref[keep]) < TOL
    assert (got[~keep] == 3.0).all()


def test_convert_skips_nop_and_odd_count(small, orc):
    cfg, g = small["cfg"], small["g"]
    tasks = cfg["tasks"][:101].copy()          # odd count, partial last pair
    tasks["task"][[3, 50, 100]] = 0             # NOP tiles
    fb = tp.FrameBuffers(g, "cuda:0", max_corr=1)
    fb.upload(cfg["images"], cfg["kernels"], cfg["koffs"], tasks)
    for c in fb.clt:
        c.fill_(-1.0)
    fb.convert()
    got = gpu_clt(fb)
    ref = orc.convert_correct_tiles(g, cfg["koffs"], cfg["kernels"], cfg["images"], tasks)
    done = np.zeros(g.num_tiles, bool)
    done[np.nonzero(tasks["task"])[0]] = True   # task i is tile i here
    for cam in range(4):
        a = got[cam].reshape(g.num_tiles, -1)
        b = ref[cam].reshape(g.num_tiles, -1)
        assert rel_err(a[done], b[done]) < TOL
        assert (a[~done] == -1.0).all()
    fb2 = tp.FrameBuffers(g, "cuda:0", max_corr=1)
    fb2.upload(tasks=tasks[:0])
    fb2.tasks = torch.zeros(40, dtype=torch.uint8, device="cuda:0")
    fb2.convert()  # num_tiles = 0: no-op


def test_convert_lpf_mask_and_custom_lpf(small, orc):
    cfg, g = small["cfg"], small["g"]
    rng = np.random.default_rng(4)
    custom = rng.uniform(0.2, 1.0, 64).astype(np.float32)
    tp.set_lpf(2, custom)
    orc.set_lpf(2, custom)
    try:
        fb = small["fb"]
        fb.convert(lpf_mask=7)
        got = gpu_clt(fb)
        ref = orc.convert_correct_tiles(g, cfg["koffs"], cfg["kernels"], cfg["images"], cfg["tasks"],
                                        lpf_mask=7)
        for cam in range(4):
            assert rel_err(got[cam], ref[cam]) < TOL
    finally:
        tp.set_lpf(2, None)
        orc.set_lpf(2, None)


def test_imclt_pass_matches_oracle(small, orc):
    g = small["g"]
    dstride = g.img_width + 8 + 4
    clt0 = small["clt64"][0]
    dev_clt = torch.from_numpy(clt0.astype(np.float32)).cuda()
    rbg = torch.zeros((3, g.img_height + 8, dstride), dtype=torch.float32, device="cuda:0")
    ref = np.zeros((3, g.img_height + 8, dstride))
    for color in range(3):
        for v in range(2):
            for h in range(2):
                tp.imclt_rbg(g, dev_clt, rbg, color, v, h, dstride)
                orc.imclt_rbg(g, clt0, ref, color, v, h, dstride)
    torch.cuda.synchronize()
    got = rbg.cpu().numpy()
    for color in range(3):
        assert rel_err(got[color], ref[color]) < TOL


def test_imclt_frame_matches_oracle(small, orc):
    g = small["g"]
    dstride = g.img_width + 8
    planes = [torch.full((3, g.img_height + 8, dstride), 5.0, device="cuda:0") for _ in range(4)]
    clts = [torch.from_numpy(c.astype(np.float32)).cuda() for c in small["clt64"]]
    tp.imclt_rbg_all(g, tp.ptr_array(clts), tp.ptr_array(planes), dstride)
    torch.cuda.synchronize()
    ref = orc.imclt_frame(g, small["clt64"], dstride)
    for cam in range(4):
        got = planes[cam].cpu().numpy()
        for color in range(3):
            assert rel_err(got[color], ref[cam][color]) < TOL


def test_imclt_frame_other_grid(orc):
    """Whole-frame inverse CLT on a 34 x 26 tile grid: 35 x 27 groups of 8
    output columns / rows, so the last strip of 15 column groups and the
    last band of row groups are partial.  Every plane pixel is written
    (NaN before) and matches the oracle."""
    g = Geometry(272, 208)
    cfg = synth.config_small(seed=6, geom=g)
    print(f"[gpu parity] imclt odd grid {g}", flush=True)
    clt64 = orc.convert_correct_tiles(g, cfg["koffs"], cfg["kernels"], cfg["images"], cfg["tasks"])
    dstride = g.img_width + 8 + 3
    planes = [torch.full((3, g.img_height + 8, dstride), float("nan"), device="cuda:0") for _ in range(4)]
    clts = [torch.from_numpy(c.astype(np.float32)).cuda() for c in clt64]
    tp.imclt_rbg_all(g, tp.ptr_array(clts), tp.ptr_array(planes), dstride)
    torch.cuda.synchronize()
    ref = orc.imclt_frame(g, clt64, dstride)
    for cam in range(4):
        got = planes[cam].cpu().numpy()[:, :, :g.img_width + 8]
        assert np.isfinite(got).all(), cam
        for color in range(3):
            assert rel_err(got[color], ref[cam][color][:, :g.img_width + 8]) < TOL, (cam, color)


def test_full_frame_tdac_and_correlation_peak(orc):
    """Full 2592 x 1936 frame, size-independent properties: identity kernels
    on the 8-px grid reconstruct every Bayer plane (TDAC) and an all-zero
    disparity gives the correlation peak at the centre."""
    g = Geometry(2592, 1936)
    rng = np.random.default_rng(9)
    imgs = synth.make_images(g, rng)
    tasks = full_tasks(g, synth.tile_centres(g))
    fb = tp.FrameBuffers(g, "cuda:0", corr_radius=7)
    fb.upload(imgs, synth.identity_kernels(g), synth.zero_offsets(g), tasks, corr_indices(g, tasks))
    fb.convert()
    dstride = g.img_width + 8
    planes = [torch.zeros((3, g.img_height + 8, dstride), device="cuda:0") for _ in range(4)]
    tp.imclt_rbg_all(g, fb.p_clt, tp.ptr_array(planes), dstride)
    from direct import bayer_mask
    H, W = g.img_height, g.img_width
    for cam in (0, 3):
        got = planes[cam].cpu().numpy()
        for color in range(3):
            exp = imgs[cam] * bayer_mask(color, np.arange(H), np.arange(W))
            assert rel_err(got[color][12:H - 4, 12:W - 4], exp[8:-8, 8:-8]) < TOL
    # same image in every camera -> peak at the centre for every pair
    for cam in range(1, 4):
        fb.clt[cam].copy_(fb.clt[0])
    fb.correlate(3, SCALES, FAT_ZERO)
    c = fb.corrs[:fb.num_corr].view(-1, 225)
    assert (c.argmax(1) == 112).all()


def test_combine_and_disparity_match_oracle(small, orc):
    g, idx = small["g"], small["idx"]
    fb = small["fb"]
    for cam in range(4):
        fb.clt[cam].copy_(torch.from_numpy(small["clt64"][cam].astype(np.float32)))
    fb.correlate(3, SCALES, FAT_ZERO)
    n = 15
    comb = torch.full((g.num_tiles, n * n + 1), -5.0, device="cuda:0")
    tp.corr_combine(g, len(idx), fb.corr_idx, fb.corrs, fb.corr_stride, 7, 0x3F, comb, n * n + 1)
    disp = torch.zeros((g.num_tiles, 2), device="cuda:0")
    tp.corr_disparity(comb, n * n + 1, g.num_tiles, 7, 0, disp)
    torch.cuda.synchronize()
    ref_c = orc.correlate2D(g, small["clt64"], 3, SCALES, FAT_ZERO, idx, 7)
    ref = orc.corr_combine(ref_c, 7)
    got = comb.cpu().numpy()
    assert rel_err(got[:, :n * n], ref) < TOL
    assert (got[:, n * n] == -5.0).all()
    d_ref = orc.corr_disparity(ref, 7, 0)
    d = disp.cpu().numpy()
    same_peak = np.abs(d[:, 1] - d_ref[:, 1]) <= 1e-5 * np.abs(d_ref[:, 1]).max()
    assert same_peak.all()
    assert np.abs(d[:, 0] - d_ref[:, 0]).max() < 1e-3


@pytest.mark.parametrize("r,stride", [(5, 130), (5, 121), (7, 225)])
def test_combine_irregular_list(small, orc, r, stride):
    """Combine over a list that is not the full tile-major one: random pair
    subsets, tiles with duplicated entries (more than the 6 rows staged at a
    time), tiles with no entry; a padded row stride, and contiguous rows
    (stride (2r+1)^2, the software-pipelined path).  Expected output: per
    tile, the sum in list order of each entry's pair tile mapped by the
    oracle's combination (one entry at a time)."""
    g = small["g"]
    rng = np.random.default_rng(21)
    entries = []
    for t in range(g.num_tiles):
        u = rng.random()
        if u < 0.1:
            continue                                        # no entry
        if u < 0.2:
            pairs = list(range(6)) + list(range(4))         # 10 entries
        else:
            pairs = sorted(rng.choice(6, size=rng.integers(1, 7), replace=False).tolist())
        entries += [(t << 8) | p for p in pairs]
    idx = np.array(entries, np.int32)
    n = 2 * r + 1
    fb = tp.FrameBuffers(g, "cuda:0", corr_radius=r, max_corr=len(idx), corr_stride=stride)
    for cam in range(4):
        fb.clt[cam].copy_(torch.from_numpy(small["clt64"][cam].astype(np.float32)))
    fb.upload(corr_idx=idx)
    fb.correlate(3, SCALES, FAT_ZERO)
    out = torch.full((g.num_tiles, n * n), -5.0, device="cuda:0")
    tp.corr_combine(g, len(idx), fb.corr_idx, fb.corrs, stride, r, 0x3F, out, n * n)
    torch.cuda.synchronize()
    rows = fb.corrs[:len(idx)].cpu().numpy().astype(np.float64)
    got = out.cpu().numpy()
    ref = np.full((g.num_tiles, n * n), -5.0)
    for k, e in enumerate(idx):
        t, p = int(e) >> 8, int(e) & 0xFF
        six = np.zeros((6, n * n))
        six[p] = rows[k, :n * n]
        mapped = orc.corr_combine(six, r, pair_mask=1 << p)[0]
        first = k == 0 or (int(idx[k - 1]) >> 8) != t    # the list is grouped by tile
        ref[t] = mapped if first else ref[t] + mapped
    has = np.zeros(g.num_tiles, bool)
    has[idx >> 8] = True
    assert rel_err(got[has], ref[has]) < TOL
    assert (got[~has] == -5.0).all()


@pytest.mark.parametrize("W,H,seed", [(144, 112, 21), (208, 160, 22), (96, 240, 23)])
def test_frame_stages_other_geometries(orc, W, H, seed):
    """Forward CLT, correlation (all pairs, r = 7) and whole-frame inverse CLT
    against the FP64 oracle on tile grids of other shapes (18 x 14, 26 x 20,
    12 x 30; image sizes are multiples of 16): list lengths, strips and bands
    that are not multiples of the kernels' work sizes."""
    g = Geometry(W, H)
    cfg = synth.config_small(seed=seed, geom=g)
    idx = corr_indices(g, cfg["tasks"])
    print(f"[gpu parity] geometry {g}, {len(idx)} pairs", flush=True)
    fb = tp.FrameBuffers(g, "cuda:0", corr_radius=7, max_corr=len(idx))
    fb.upload(cfg["images"], cfg["kernels"], cfg["koffs"], cfg["tasks"], idx)
    fb.convert()
    fb.correlate(3, SCALES, FAT_ZERO)
    dstride = g.img_width + 8
    planes = [torch.zeros((3, g.img_height + 8, dstride), device="cuda:0") for _ in range(4)]
    tp.imclt_rbg_all(g, fb.p_clt, tp.ptr_array(planes), dstride)
    torch.cuda.synchronize()
    clt64 = orc.convert_correct_tiles(g, cfg["koffs"], cfg["kernels"], cfg["images"], cfg["tasks"])
    got = gpu_clt(fb)
    for cam in range(4):
        assert rel_err(got[cam], clt64[cam]) < TOL, cam
    ref = orc.correlate2D(g, clt64, 3, SCALES, FAT_ZERO, idx, 7)
    assert rel_err(fb.corrs[:len(idx), :225].cpu().numpy(), ref) < TOL
    rplanes = orc.imclt_frame(g, clt64, dstride)
    for cam in range(4):
        p = planes[cam].cpu().numpy()
        for color in range(3):
            assert rel_err(p[color], rplanes[cam][color]) < TOL, (cam, color)
