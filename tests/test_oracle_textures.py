"""Pins of the oracle's texture path (textures_accumulate, debayer_shot,
tile_combine_rgba, prepare_texture_list) by known answers and invariants."""
import numpy as np

from tile_processor_gpu_amd.geometry import Geometry, full_tasks
from tile_processor_gpu_amd import synth

W = np.sin(np.pi * (np.arange(16) + 0.5) / 16)  # full 16-sample window
W2 = np.outer(W ** 2, W ** 2)


def test_debayer_of_constant_is_constant_then_companded(orc):
    for rb in (0, 1):
        for c, scale in ((25.0, 3.0), (4.0, 3.0), (7.0, 0.0)):
            out = orc.debayer_shot(c * W2, rb, 10.0, scale)
            if scale > 0:
                exp = scale * (np.sqrt(c) if c > 10.0 else c / np.sqrt(10.0))
            else:
                exp = c
            np.testing.assert_allclose(out, exp * W2, rtol=1e-12)


def test_debayer_kernels_sum_and_shape(orc):
    """Interior: R/B use [1 2 1]x[1 2 1]/16, G the cross /8 + centre /2;
    the border rows/columns pass through (no companding: scale_shot 0)."""
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 1, (16, 16))
    out_rb = orc.debayer_shot(x * W2, 1, 10.0, 0.0) / W2
    out_g = orc.debayer_shot(x * W2, 0, 10.0, 0.0) / W2
    kr = np.outer([1, 2, 1], [1, 2, 1]) / 16.0
    kg = np.array([[0, 1, 0], [1, 4, 1], [0, 1, 0]]) / 8.0
    for r in range(1, 15):
        for c in range(1, 15):
            np.testing.assert_allclose(out_rb[r, c], (x[r - 1:r + 2, c - 1:c + 2] * kr).sum(), rtol=1e-12)
            np.testing.assert_allclose(out_g[r, c], (x[r - 1:r + 2, c - 1:c + 2] * kg).sum(), rtol=1e-12)
    for o in (out_rb, out_g):
        np.testing.assert_allclose(o[0], x[0], rtol=1e-12)
        np.testing.assert_allclose(o[:, 15], x[:, 15], rtol=1e-12)


def test_combine_identical_cameras(orc):
    """4 identical cameras: the colour is the common (non-debayered) value and
    alpha = w^2 (1 + 1e-4); each port weight is 1 + 1e-4, rms is 0."""
    D = np.array([0.3, 0.7, 0.5])
    M = np.array([0.31, 0.69, 0.52])
    for r, c in ((0, 0), (7, 8), (15, 3)):
        out = orc.combine_pixel(np.tile(D, (4, 1)), np.tile(M, (4, 1)), r, c)
        np.testing.assert_allclose(out[:3], M, rtol=1e-12)
        np.testing.assert_allclose(out[3], W2[r, c] * (1 + 1e-4), rtol=1e-12)
        np.testing.assert_allclose(out[4:8], 1 + 1e-4, rtol=1e-12)
        np.testing.assert_allclose(out[8:12], 0, atol=1e-12)


def test_combine_rejects_outlier_camera(orc):
    """One camera far from the other three gets ~no weight; with dust_remove
    the three agreeing cameras keep ~full weight."""
    D = np.tile([0.3, 0.7, 0.5], (4, 1))
    M = np.tile([0.3, 0.7, 0.5], (4, 1))
    D[2] += 5.0
    M[2] += 5.0
    out = orc.combine_pixel(D, M, 8, 8)
    np.testing.assert_allclose(out[:3], [0.3, 0.7, 0.5], atol=1e-3)
    assert out[4 + 2] < 1e-3
    assert min(out[4], out[5], out[7]) > 0.9


def test_texture_list_known_answers(orc):
    """3x2 block of textured tiles at (5..7, 4..5) of a 16 x 8 tile grid."""
    g = Geometry(128, 64)
    tasks = full_tasks(g, np.zeros((g.num_tiles, 4, 2), np.float32), task=0)
    on = [(x, y) for y in (4, 5) for x in (5, 6, 7)]
    for x, y in on:
        tasks["task"][y * g.tilesx + x] = 0xF | (0x3F << 4)
    t, ind, num, woi = orc.prepare_texture_list(g, tasks)
    assert list(woi) == [5, 4, 3, 2]           # x, y, width, height
    nb = {(x, y): t["task"][y * g.tilesx + x] & 0xF for x, y in on}
    # N=1 E=2 S=4 W=8
    assert nb[(5, 4)] == 2 | 4 and nb[(6, 4)] == 2 | 4 | 8 and nb[(7, 4)] == 4 | 8
    assert nb[(5, 5)] == 1 | 2 and nb[(6, 5)] == 1 | 2 | 8 and nb[(7, 5)] == 1 | 8
    assert all((t["task"][i] >> 4) == 0x3F for i in [y * g.tilesx + x for x, y in on])
    # all 6 are border tiles (on the woi edge): sub-lists by (x & 1, y & 1)
    lists = orc.sub_lists(g, ind, num)
    assert [len(l) for l in lists[:4]] == [0, 0, 0, 0]
    got = {p: sorted(int(v) >> 8 for v in lists[4 + p]) for p in range(4)}
    exp = {p: sorted(y * g.tilesx + x for x, y in on if (x & 1) + 2 * (y & 1) == p) for p in range(4)}
    assert got == exp
    for lst in lists:
        for v in lst:
            assert v & 0x80
            x, y = (v >> 8) % g.tilesx, (v >> 8) // g.tilesx
            assert (v & 0xF) == nb[(x, y)]


def test_isolated_texture_tile_gets_no_entry(orc):
    """A textured tile with no textured neighbour ends with texture bits 0
    (mark_texture_neighbor_tiles replaces them) and is not listed."""
    g = Geometry(64, 64)
    tasks = full_tasks(g, np.zeros((g.num_tiles, 4, 2), np.float32), task=0)
    tasks["task"][3 * g.tilesx + 3] = 0xF
    t, ind, num, woi = orc.prepare_texture_list(g, tasks)
    assert t["task"][3 * g.tilesx + 3] == 0 and num.sum() == 0


def test_rbga_alpha_is_one_in_interior(orc):
    """All tiles textured, identical cameras: interior alpha = sum of the
    overlapping w^2 windows = 1 (TDAC), times (1 + 1e-4)."""
    g = Geometry(96, 64)
    rng = np.random.default_rng(1)
    img = rng.uniform(20, 200, (g.img_height, g.img_width)).astype(np.float32)
    imgs = np.stack([img] * 4)
    tasks = full_tasks(g, synth.tile_centres(g))
    clt = orc.convert_correct_tiles(g, synth.zero_offsets(g), synth.identity_kernels(g), imgs, tasks)
    t, woi, num, lists, rbga = orc.generate_rbga(g, tasks, clt, orc.TexParams())
    assert list(woi) == [0, 0, g.tilesx, g.tilesy]
    alpha = rbga[3]
    # pixels whose 4 covering tiles are all inner (texture mask 0xF, no fade)
    inner = alpha[16:8 * (g.tilesy - 1), 16:8 * (g.tilesx - 1)]
    np.testing.assert_allclose(inner, 1 + 1e-4, rtol=1e-10)
    # the colour planes there are the plain (unfaded) overlap-add
    assert np.isfinite(rbga).all()
