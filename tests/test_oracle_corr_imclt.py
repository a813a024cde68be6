"""Pins of the oracle's correlate2D and imclt restatements: bicomplex
algebra + textbook transforms, known-answer correlation peaks, the adjoint
identity imclt = F^T / 4 and perfect reconstruction (TDAC)."""
import numpy as np
import scipy.fft as sf

from tile_processor_gpu_amd.geometry import Geometry, full_tasks, corr_indices, PAIRS
from tile_processor_gpu_amd import synth
from direct import bicomplex_mul, bicomplex_conj, lapped_basis, bayer_mask, smooth_scene

LPF_RB, LPF_C = 4, 5


def _dtt2(sin):
    E = np.eye(8)
    if sin:
        m = sf.dst(E, type=2, norm="ortho", axis=0)
        m[7] *= np.sqrt(2)
    else:
        m = sf.dct(E, type=2, norm="ortho", axis=0)
        m[0] *= np.sqrt(2)
    return m


def corr_direct(A, B, scales, fat_zero, lpf, radius):
    """A, B: [colors][4][8][8] CLT of the two cameras."""
    C = np.zeros((4, 8, 8))
    for c in range(len(A)):
        C += scales[c] * bicomplex_mul(bicomplex_conj(A[c]), B[c])
        if c == 1:
            C *= lpf[LPF_RB].reshape(8, 8)
    C /= np.sqrt(fat_zero ** 2 + (C ** 2).sum(0))
    C *= lpf[LPF_C].reshape(8, 8)
    P = np.zeros((4, 8, 8))
    for q in range(4):
        P[q] = _dtt2(q >> 1) @ C[q] @ _dtt2(q & 1).T
    # even/odd reconstruction: lag (i, j) = even part (CC) + odd-in-x (SC, lag j -> j-1)
    # + odd-in-y (CS) + odd-odd (SS), all / 4
    n = 2 * radius + 1
    out = np.zeros((n, n))
    for y in range(n):
        for x in range(n):
            i, j = y - radius, x - radius
            ai, aj = abs(i), abs(j)
            v = P[0][ai][aj]
            if aj:
                v += np.sign(j) * P[1][ai][aj - 1]
            if ai:
                v += np.sign(i) * P[2][ai - 1][aj]
            if ai and aj:
                v += np.sign(i) * np.sign(j) * P[3][ai - 1][aj - 1]
            out[y, x] = 0.25 * v
    return out


def test_correlation_equals_algebra_and_textbook_transforms(orc):
    g = Geometry(32, 32)
    rng = np.random.default_rng(1)
    clt = [rng.standard_normal(g.clt_floats) for _ in range(4)]
    lpf = orc.table(3)
    tiles = [0, 5, 11]
    idx = np.array([(t << 8) | p for t in tiles for p in range(6)], np.int32)
    for radius in (7, 4, 0):
        out = orc.correlate2D(g, clt, 3, (0.25, 0.25, 0.5), 0.25, idx, radius)
        for e, v in enumerate(idx):
            t, p = v >> 8, v & 0xFF
            A = [clt[PAIRS[p][0]][(t * 3 + c) * 256:(t * 3 + c + 1) * 256].reshape(4, 8, 8) for c in range(3)]
            B = [clt[PAIRS[p][1]][(t * 3 + c) * 256:(t * 3 + c + 1) * 256].reshape(4, 8, 8) for c in range(3)]
            d = corr_direct(A, B, (0.25, 0.25, 0.5), 0.25, lpf, radius)
            np.testing.assert_allclose(out[e], d.ravel(), atol=1e-13)


def test_mono_correlation_uses_color0_only(orc):
    g = Geometry(32, 32)
    rng = np.random.default_rng(2)
    clt = [rng.standard_normal(g.clt_floats) for _ in range(4)]
    idx = np.array([(3 << 8) | 4], np.int32)
    out = orc.correlate2D(g, clt, 1, (0.75, 0.0, 0.0), 0.125, idx, 7)
    A = [clt[0][9 * 256:10 * 256].reshape(4, 8, 8)]
    B = [clt[3][9 * 256:10 * 256].reshape(4, 8, 8)]
    np.testing.assert_allclose(out[0], corr_direct(A, B, (0.75,), 0.125, orc.table(3), 7).ravel(), atol=1e-13)


def _shifted_pair(dx, dy, seed=0, size=128, fmax=0.12):
    """Gray scene seen by 4 cameras; camera 1's content is displaced by
    (+dx, +dy) (fractional allowed)."""
    g = Geometry(size, size)
    f = smooth_scene(seed, n=64, fmax=fmax)
    y, x = np.mgrid[0:size, 0:size].astype(np.float64)
    imgs = np.zeros((4, size, size), np.float32)
    for c in range(4):
        imgs[c] = f(x, y)
    imgs[1] = f(x - dx, y - dy)
    return g, imgs


def _pair0_peaks(orc, g, imgs, xy):
    tasks = full_tasks(g, xy)
    clt = orc.convert_correct_tiles(g, synth.zero_offsets(g), synth.identity_kernels(g), imgs, tasks)
    tiles = [(ty * g.tilesx + tx) for ty in range(4, 12) for tx in range(4, 12)]
    idx = np.array([(t << 8) | 0 for t in tiles], np.int32)
    c = orc.correlate2D(g, clt, 3, (0.25, 0.25, 0.5), 0.01, idx, 7)
    return [np.unravel_index(np.argmax(row), (15, 15)) for row in c]


def test_correlation_peak_known_shift(orc):
    """Pair (0, 1), camera 1's content displaced by (+dx, +dy) px and tile
    centres unchanged: the 15x15 correlation peaks at (r + dy, r + dx)."""
    for dx, dy in [(2, 0), (0, -2), (-2, 1), (1, 1)]:
        g, imgs = _shifted_pair(dx, dy, fmax=0.35)
        peaks = _pair0_peaks(orc, g, imgs, synth.tile_centres(g))
        hits = sum((y - 7, x - 7) == (dy, dx) for y, x in peaks)
        assert hits >= 0.95 * len(peaks), (dx, dy, hits, len(peaks))


def test_task_centres_compensate_fractional_shift(orc):
    """Camera 1's content displaced by a fractional (dx, dy) and its tile
    centres moved by the same amount: window placement + phase rotation
    re-align it, so the correlation peaks at the centre."""
    for dx, dy in [(2.5, -1.25), (-3.75, 0.5), (0.3, 4.6)]:
        g, imgs = _shifted_pair(dx, dy, seed=1)
        xy = synth.tile_centres(g)
        xy[:, 1, 0] += dx
        xy[:, 1, 1] += dy
        peaks = _pair0_peaks(orc, g, imgs, xy)
        hits = sum((y, x) == (7, 7) for y, x in peaks)
        assert hits >= 0.95 * len(peaks), (dx, dy, hits, len(peaks))


def test_autocorrelation_is_point_symmetric(orc):
    g, imgs = _shifted_pair(0, 0)
    koffs, kid = synth.zero_offsets(g), synth.identity_kernels(g)
    tasks = full_tasks(g, synth.tile_centres(g, 0.0, 0.3, np.random.default_rng(4)))
    clt = orc.convert_correct_tiles(g, koffs, kid, imgs, tasks)
    clt[1] = clt[0].copy()
    idx = np.array([(t << 8) | 0 for t in range(100, 140)], np.int32)
    c = orc.correlate2D(g, clt, 3, (0.25, 0.25, 0.5), 0.01, idx, 7).reshape(-1, 15, 15)
    np.testing.assert_allclose(c, c[:, ::-1, ::-1], atol=1e-12)


def test_imclt_is_quarter_adjoint_of_forward(orc):
    bh = lapped_basis(0.0)
    rng = np.random.default_rng(3)
    for _ in range(4):
        q = rng.standard_normal((4, 8, 8))
        adj = sum(bh[qq >> 1].T @ q[qq] @ bh[qq & 1] for qq in range(4))
        for hf in (0, 1):
            np.testing.assert_allclose(orc.imclt_tile(q, hf), 0.25 * adj, atol=1e-13)


def test_tdac_reconstructs_each_bayer_plane(orc):
    """Identity kernels, tile centres on the 8-px grid: imclt overlap-add of
    the forward CLT returns each color's masked image (shifted by the 4-px
    plane margin) away from the frame border."""
    g = Geometry(96, 64)
    rng = np.random.default_rng(5)
    imgs = rng.uniform(0, 1, (4, g.img_height, g.img_width)).astype(np.float32)
    tasks = full_tasks(g, synth.tile_centres(g))
    clt = orc.convert_correct_tiles(g, synth.zero_offsets(g), synth.identity_kernels(g), imgs, tasks)
    planes = orc.imclt_frame(g, clt)
    H, W = g.img_height, g.img_width
    for cam in range(4):
        for color in range(3):
            m = bayer_mask(color, np.arange(H), np.arange(W))
            exp = imgs[cam] * m
            got = planes[cam][color][4:H + 4, 4:W + 4]
            np.testing.assert_allclose(got[8:-8, 8:-8], exp[8:-8, 8:-8], atol=1e-12)
