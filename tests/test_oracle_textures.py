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


# ---- tile_combine_rgba known answers (src/TileProcessor.cuh:4202-4560) ----
#
# Port offsets (-1/2,-1/2), (1/2,-1/2), (-1/2,1/2), (1/2,1/2): squared pair
# distances are 1 for (0,1), (0,2), (1,3), (2,3) and 2 for the diagonals
# (0,3), (1,2).  Colours 0 and 1 are equal in all cameras, so a colour
# distance is w2 * x^2 for a difference x of colour 2 divided by the window
# squared.  The cases below differ in colour 2 only and are written with
# scaled differences v (D = base + v * wnd2), so the window drops out of the
# exponents.  Each expected value is the closed form of that case, derived
# by hand (below, per case), not a restatement of the general loops.

WEIGHTS = (0.294118, 0.117647, 0.588235)
KSIG = 1.0 / (2.0 * 1.5 ** 2)
FZ = 1e-4
W2C = float(np.float32(WEIGHTS[2]))  # the oracle receives the weights as float


def _wnd2(orc, r, c):
    hwsq = orc.table(2)[2]
    sym = lambda i: i if i < 8 else 15 - i  # noqa: E731
    return hwsq[sym(r)] * hwsq[sym(c)]


def _pixel(orc, v, r, c):
    """D with colour 2 offset by v * wnd2 (colours 0, 1 equal), and M with
    distinct values per camera so that the output colour shows the weights."""
    wnd2 = _wnd2(orc, r, c)
    D = np.tile([40.0, 55.0, 70.0], (4, 1))
    D[:, 2] += np.asarray(v, float) * wnd2
    M = np.array([[41.0, 52.0, 71.0], [39.0, 58.0, 66.0], [47.0, 50.0, 80.0], [36.0, 57.0, 69.0]])
    return D, M, wnd2


def _expect(q, M, wnd2):
    """Output of the final weights q: colour = sum q M / sum q, alpha =
    wnd2 * sum q / 4, port weights q."""
    q = np.asarray(q)
    return np.concatenate([(q[:, None] * M).sum(0) / q.sum(), [wnd2 * q.sum() / 4], q])


def test_combine_moderate_outlier_known_answer(orc):
    """Cameras 0, 1, 3 equal, camera 2 off by v = 3 (worst weight ~0.3).
    With u = w2 v^2, A = exp(-k u), B = exp(-k u / 2):
      pairwise: e01 = e03 = e13 = 1 + f, e02 = e23 = A + f, e12 = B + f, so
      w = (2 + A + 3f, 2 + B + 3f, 2A + B + 3f, 2 + A + 3f): best ports 1
      and 0 (camera 3 ties camera 0 and does not replace it), blend = base;
      re-computed q = (1 + f, 1 + f, A + f, 1 + f); worst = 2;
      avg = 1 + f (the other three over NUM_CAMS - 1 = 3),
      scale = 1 + q2 (avg - q2) / (3 avg^2), q_i *= scale (i != 2),
      q2 *= q2 / avg."""
    r, c = 8, 5
    v = 3.0
    D, M, wnd2 = _pixel(orc, [0, 0, v, 0], r, c)
    A = np.exp(-KSIG * W2C * v * v)
    assert 0.25 < A < 0.35
    avg = 1 + FZ
    q2 = A + FZ
    scale = 1 + q2 * (avg - q2) / (3 * avg * avg)
    q = [(1 + FZ) * scale, (1 + FZ) * scale, q2 * q2 / avg, (1 + FZ) * scale]
    out = orc.combine_pixel(D, M, r, c, keep_weights=1)
    np.testing.assert_allclose(out[:8], _expect(q, M, wnd2), rtol=1e-12, atol=0)
    # rms of colour 2 (values base, base, base + v wnd2, base): mean
    # square deviation 3 (v wnd2)^2 / 16; colours 0, 1: 0
    d = v * wnd2
    np.testing.assert_allclose(out[8:12], [0, 0, np.sqrt(3) * d / 4, np.sqrt(W2C * 3 * d * d / 16)],
                               rtol=1e-12, atol=1e-12)
    # without dust_remove the re-computed weights are the output weights
    out0 = orc.combine_pixel(D, M, r, c, dust_remove=0)
    np.testing.assert_allclose(out0[:8], _expect([1 + FZ, 1 + FZ, A + FZ, 1 + FZ], M, wnd2),
                               rtol=1e-12, atol=0)


def _split_case():
    """2+1+1 split: v = (0, 0.8, 2.2, -1.6)."""
    return np.array([0.0, 0.8, 2.2, -1.6])


def _split_expect(orc, third=False):
    """Known answer of the split case.  Pairwise weights (squared distances
    in brackets): e01 = E(0.8^2 [1]), e02 = E(2.2^2 [1]), e03 = E(1.6^2 [2]),
    e12 = E(1.4^2 [2]), e13 = E(2.4^2 [1]), e23 = E(3.8^2 [1]) with
    E(x [d]) = exp(-k w2 x / d) + f;  w0 = e01 + e02 + e03, w1 = e01 + e12 +
    e13, w2 = e02 + e12 + e23, w3 = e03 + e13 + e23.  Best ports 0 and 1 (w0 >
    w1 > w2 > w3, asserted), blend = (w0 v0 + w1 v1) / (w0 + w1); with
    third=True the blend takes port 2 instead of port 1.  Re-computed q_i =
    exp(-k w2 (v_i - blend)^2) + f; worst = 3 (asserted); dust_remove as in
    the outlier case over the other three."""
    v = _split_case()
    E = lambda x, d: np.exp(-KSIG * W2C * x * x / d) + FZ  # noqa: E731
    e01, e02, e03 = E(0.8, 1), E(2.2, 1), E(1.6, 2)
    e12, e13, e23 = E(1.4, 2), E(2.4, 1), E(3.8, 1)
    w = np.array([e01 + e02 + e03, e01 + e12 + e13, e02 + e12 + e23, e03 + e13 + e23])
    assert w[0] > w[1] > w[2] > w[3]
    b2 = 2 if third else 1
    blend = (w[0] * v[0] + w[b2] * v[b2]) / (w[0] + w[b2])
    q = np.exp(-KSIG * W2C * (v - blend) ** 2) + FZ
    assert np.argmin(q) == 3
    avg = (q[0] + q[1] + q[2]) / 3
    scale = 1 + q[3] * (avg - q[3]) / (3 * avg * avg)
    return np.array([q[0] * scale, q[1] * scale, q[2] * scale, q[3] * q[3] / avg])


def test_combine_split_known_answer(orc):
    """The blend of the two best ports matters: ports 0 and 1 differ, and the
    third port (2) would give another blend, other weights and colour."""
    r, c = 3, 12
    D, M, wnd2 = _pixel(orc, _split_case(), r, c)
    out = orc.combine_pixel(D, M, r, c)
    np.testing.assert_allclose(out[:8], _expect(_split_expect(orc), M, wnd2), rtol=1e-12, atol=0)
    # alt bit 0 = blend with the third-best port: the same closed form with port 2
    alt = orc.combine_pixel(D, M, r, c, alt=1)
    np.testing.assert_allclose(alt[:8], _expect(_split_expect(orc, third=True), M, wnd2), rtol=1e-12,
                               atol=0)
    assert np.abs(alt[:3] - out[:3]).max() > 1e-3


def test_combine_alt_worst_and_tie_ratios(orc):
    """alt bit 1 treats the second-lowest re-computed weight as the worst;
    tie ratios are large for clear choices and ~0 (summation-order rounding)
    for exact ties."""
    r, c = 8, 8
    D, M, wnd2 = _pixel(orc, _split_case(), r, c)
    out, ties = orc.combine_pixel(D, M, r, c, eta=1e-6)
    assert ties[0] > 10 and ties[1] > 10
    # swapping the worst: port 2 (second lowest, asserted) takes q2^2 / avg'
    # with avg' over ports 0, 1, 3; the blend is that of the split case
    v = _split_case()
    E = lambda x, d: np.exp(-KSIG * W2C * x * x / d) + FZ  # noqa: E731
    wv = np.array([E(0.8, 1) + E(2.2, 1) + E(1.6, 2), E(0.8, 1) + E(1.4, 2) + E(2.4, 1)])
    blend = (wv[0] * v[0] + wv[1] * v[1]) / wv.sum()
    qa = np.exp(-KSIG * W2C * (v - blend) ** 2) + FZ
    assert np.argsort(qa)[1] == 2
    avg = (qa[0] + qa[1] + qa[3]) / 3
    scale = 1 + qa[2] * (avg - qa[2]) / (3 * avg * avg)
    exp_alt = np.array([qa[0] * scale, qa[1] * scale, qa[2] * qa[2] / avg, qa[3] * scale])
    alt = orc.combine_pixel(D, M, r, c, alt=2)
    np.testing.assert_allclose(alt[:8], _expect(exp_alt, M, wnd2), rtol=1e-12, atol=0)
    # exact ties: cameras 0, 1, 3 equal make ports 0 and 3 tie for second
    # best; cameras 0 and 2 equal, 1 and 3 equal (v = 0, 1, 0, 1) make the
    # re-computed weights tie two by two
    _, t1 = orc.combine_pixel(*_pixel(orc, [0, 0, 3.0, 0], r, c)[:2], r, c, eta=1e-6)
    assert t1[0] < 1e-6
    _, t2 = orc.combine_pixel(*_pixel(orc, [0, 1.0, 0, 1.0], r, c)[:2], r, c, eta=1e-6)
    assert t2[1] < 1e-6


# Mutations of the oracle's consensus that a wrong restatement could make;
# each must fail one of the known-answer tests above.
MUTATIONS = {
    "avg over 4 ports": ("avg /= (T)3;", "avg /= (T)4;"),
    "no rescale of the other ports": ("if (cam != worst) pw[cam] *= scale;", "if (cam != worst) pw[cam] *= (T)1;"),
    "no worst-port factor": ("pw[worst] *= wv / avg;", "pw[worst] *= (T)1;"),
    "best-port weights swapped": ("const T w1 = pw[b1] / (pw[b1] + pw[b2]);",
                                  "const T w1 = pw[b2] / (pw[b1] + pw[b2]);"),
    "second port = the lowest": ("if (cam != b1 && pw[cam] > best) { b2 = cam; best = pw[cam]; }",
                                 "if (cam != b1 && pw[cam] < best) { b2 = cam; best = pw[cam]; }"),
    "scale over 4 ports": ("T scale = (T)1 + wv * (avg - wv) / (avg * avg * (T)3);",
                           "T scale = (T)1 + wv * (avg - wv) / (avg * avg * (T)4);"),
}


def test_known_answers_catch_oracle_mutations(orc, tmp_path):
    """Each mutation of combine_pixel, compiled into a copy of the oracle,
    fails the moderate-outlier or split known answers."""
    import ctypes
    import os
    import shutil
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = open(os.path.join(root, "oracle", "tp_oracle.cpp")).read()
    cxx = shutil.which("g++") or shutil.which("c++")
    (tmp_path / "include").mkdir()
    shutil.copy(os.path.join(root, "include", "tp.h"), tmp_path / "include" / "tp.h")
    (tmp_path / "oracle").mkdir()
    caught = {}
    for name, (a, b) in MUTATIONS.items():
        assert src.count(a) == 1, name
        d = tmp_path / "oracle" / ("m%d" % len(caught))
        d.mkdir()
        (tmp_path / "oracle" / (d.name + ".cpp")).write_text(src.replace(a, b))
        so = str(d / "lib.so")
        subprocess.check_call([cxx, "-O1", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-o", so,
                               str(tmp_path / "oracle" / (d.name + ".cpp"))])
        L = ctypes.CDLL(so)

        def combine(D, M, r, c, dust_remove=1, alt=0):
            D = np.ascontiguousarray(D, np.float64)
            M = np.ascontiguousarray(M, np.float64)
            out = np.zeros(12)
            po = np.ascontiguousarray([[-0.5, -0.5], [0.5, -0.5], [-0.5, 0.5], [0.5, 0.5]], np.float32)
            f = ctypes.c_float
            L.oracle_combine_pixel_f64(ctypes.c_int(3), f(1.5), f(WEIGHTS[0]), f(WEIGHTS[1]),
                                       f(WEIGHTS[2]), ctypes.c_int(dust_remove), ctypes.c_int(1),
                                       po.ctypes.data_as(ctypes.c_void_p), D.ctypes.data_as(ctypes.c_void_p),
                                       M.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(r), ctypes.c_int(c),
                                       out.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(alt),
                                       ctypes.c_double(0.0), ctypes.c_void_p(0))
            return out

        ok = True
        D, M, wnd2 = _pixel(orc, [0, 0, 3.0, 0], 8, 5)
        A = np.exp(-KSIG * W2C * 9.0)
        avg, q2 = 1 + FZ, A + FZ
        sc = 1 + q2 * (avg - q2) / (3 * avg * avg)
        exp1 = _expect([(1 + FZ) * sc, (1 + FZ) * sc, q2 * q2 / avg, (1 + FZ) * sc], M, wnd2)
        ok &= np.allclose(combine(D, M, 8, 5)[:8], exp1, rtol=1e-12, atol=0)
        D, M, wnd2 = _pixel(orc, _split_case(), 3, 12)
        ok &= np.allclose(combine(D, M, 3, 12)[:8], _expect(_split_expect(orc), M, wnd2), rtol=1e-12, atol=0)
        caught[name] = not ok
    assert all(caught.values()), caught


def test_fp32_oracle_mismatches_are_certified_ties(orc):
    """The certification the GPU texture tests use, exercised on the CPU with
    the FP32 oracle in the role of an FP32 implementation: every texture-tile
    pixel where it misses the FP64 result by more than 1e-5 is a pixel whose
    tie ratio is <= 1, and the FP64 result with that choice overridden
    matches it."""
    from texcert import scene_cfg, certify_tiles
    g = Geometry(128, 128)
    imgs, kern, koffs, tasks = scene_cfg(3, g)
    clt64 = orc.convert_correct_tiles(g, koffs, kern, imgs, tasks)
    clt32 = [c.astype(np.float32) for c in clt64]
    from tile_processor_gpu_amd.geometry import texture_indices
    idx = texture_indices(g, tasks)
    for colors in (3, 1):
        prm = orc.TexParams(colors=colors, keep_weights=1)
        stride = prm.slices() * 256
        r32 = np.zeros((len(idx), stride), np.float32)
        orc.textures_accumulate(g, None, clt32, idx, prm, tiles=r32, tex_stride=stride)
        # the FP32 oracle keeps the one-pass rms of the reference, which
        # cancels in FP32 (the GPU kernel's is two-pass): colours, alpha and
        # port weights only
        got = r32.reshape(len(idx), -1, 256)[:, :colors + 5]
        n_bad, n_tied = certify_tiles(orc, g, clt64, idx, prm, got, colors + 5)
        print(f"colors {colors}: {n_bad} pixels beyond 1e-5, all certified ties "
              f"({n_tied} pixels of {got.shape[0] * 256} have a tie ratio <= 1)")


def test_fp32_oracle_rbga_within_certified_slack(orc):
    """The same for the overlap-added RGBA image of generate_RBGA: the FP32
    oracle's image is within 1e-5 of the FP64 one except where certified tie
    flips of the overlapping tiles account for the difference."""
    from texcert import scene_cfg, rbga_slack, certify_rbga
    g = Geometry(128, 128)
    imgs, kern, koffs, tasks = scene_cfg(4, g)
    clt64 = orc.convert_correct_tiles(g, koffs, kern, imgs, tasks)
    clt32 = [c.astype(np.float32) for c in clt64]
    prm = orc.TexParams()
    _, woi, _, lists, ref = orc.generate_rbga(g, tasks, clt64, prm)
    _, woi32, _, _, got = orc.generate_rbga(g, tasks, clt32, prm)
    assert list(woi) == list(woi32)
    rows, cols = (woi[3] + 1) * 8, (woi[2] + 1) * 8
    ref, got = ref[:, :rows, :cols], got[:, :rows, :cols]
    slack = rbga_slack(orc, g, clt64, lists, woi, prm, ref.shape)
    n = certify_rbga(got, ref, slack)
    print(f"{n} RGBA pixels beyond 1e-5, all within the certified-tie slack")


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
