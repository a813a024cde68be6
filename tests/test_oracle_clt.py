"""Pins of the oracle's forward CLT (convert_correct_tiles restatement)
against the direct lapped-transform sum, the bicomplex algebra and known
answers."""
import numpy as np
import pytest

from tile_processor_gpu_amd.geometry import Geometry
from tile_processor_gpu_amd import synth
from direct import forward_direct, bicomplex_mul

G = Geometry(64, 64)


def _inputs(seed=0):
    rng = np.random.default_rng(seed)
    img = rng.uniform(0, 1, (G.img_height, G.img_width)).astype(np.float32)
    koffs = np.zeros((G.kernels_vert, G.kernels_hor, 3, 8), np.float32)
    return rng, img, koffs


@pytest.mark.parametrize("color", [0, 1, 2])
@pytest.mark.parametrize("parity", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_fold_dct_replication_equals_direct_sum(orc, color, parity):
    """Bayer fold + DCT-IV/DST-IV + quadrant replication + phase rotation
    == direct shifted-window lapped transform of the color-masked patch,
    times e^{i th_k} e^{j th_l}, for every Bayer parity of the window."""
    rng, img, koffs = _inputs(color * 4 + parity[0] * 2 + parity[1])
    kid = synth.identity_kernels(G)[0]
    for _ in range(3):
        itlx, itly = 20 + parity[0], 30 + parity[1]
        cx = np.float32(itlx + 8 - rng.uniform(-0.49, 0.49))
        cy = np.float32(itly + 8 - rng.uniform(-0.49, 0.49))
        sh = np.float64(itlx + 8) - np.float64(cx)
        sv = np.float64(itly + 8) - np.float64(cy)
        q = orc.convert_tile(G, koffs, kid, img, G.img_width, color, cx, cy)
        d = forward_direct(img, color, itlx, itly, sh, sv)
        np.testing.assert_allclose(q, d, atol=1e-13 * np.abs(d).max())


def test_kernel_multiply_is_bicomplex_product(orc):
    rng, img, koffs = _inputs(5)
    kern = rng.standard_normal((G.kernels_vert, G.kernels_hor, 3, 4, 8, 8)).astype(np.float32)
    cx, cy = np.float32(33.3), np.float32(27.8)
    itlx, itly = int(np.floor(cx - 8 + 0.5)), int(np.floor(cy - 8 + 0.5))
    sh = np.float64(itlx + 8) - np.float64(cx)
    sv = np.float64(itly + 8) - np.float64(cy)
    ktx, kty = int(np.rint(cx / 16 + 1)), int(np.rint(cy / 16 + 1))
    for color in range(3):
        q = orc.convert_tile(G, koffs, kern, img, G.img_width, color, cx, cy)
        d = forward_direct(img, color, itlx, itly, sh, sv, kern[kty, ktx, color])
        np.testing.assert_allclose(q, d, atol=1e-12 * np.abs(d).max())


def test_kernel_offset_equals_centre_shift(orc):
    """CltExtra data_x/data_y move the window like moving the tile centre the
    other way (TileProcessor.cuh:3185, :3226)."""
    rng, img, koffs = _inputs(6)
    kid = synth.identity_kernels(G)[0]
    ko = koffs.copy()
    ko[..., 0] = 0.375   # data_x
    ko[..., 1] = -0.25   # data_y
    for color in range(3):
        a = orc.convert_tile(G, ko, kid, img, G.img_width, color, 33.125, 29.5)
        b = orc.convert_tile(G, koffs, kid, img, G.img_width, color, 33.125 - 0.375, 29.5 + 0.25)
        np.testing.assert_allclose(a, b, atol=1e-12)


def test_kernel_cell_selection_known_answers(orc):
    """ktile = rint(c / 16 + 1) with round-half-even, clamped to the grid
    (TileProcessor.cuh:3129-3130): each cell's kernel is a distinct gain."""
    rng, img, koffs = _inputs(7)
    kern = synth.identity_kernels(G)[0].astype(np.float64)
    gain = 1 + np.arange(G.kernels_vert * G.kernels_hor).reshape(G.kernels_vert, G.kernels_hor)
    kern = (kern * gain[:, :, None, None, None, None]).astype(np.float32)
    base = orc.convert_tile(G, koffs, synth.identity_kernels(G)[0], img, G.img_width, 0, 24.0, 40.0)
    # 24/16 + 1 = 2.5 -> 2 (half-even), 40/16 + 1 = 3.5 -> 4
    q = orc.convert_tile(G, koffs, kern, img, G.img_width, 0, 24.0, 40.0)
    np.testing.assert_allclose(q, base * gain[4, 2], rtol=1e-12, atol=1e-12)
    # far outside: clamped to the last cell
    q = orc.convert_tile(G, koffs, kern, img, G.img_width, 0, 60.0, 4.0)
    b2 = orc.convert_tile(G, koffs, synth.identity_kernels(G)[0], img, G.img_width, 0, 60.0, 4.0)
    np.testing.assert_allclose(q, b2 * gain[1, 5], rtol=1e-12, atol=1e-12)


def test_border_window_uses_bayer_clamp(orc):
    """Right/bottom border: pixels beyond the image come from the nearest pixel
    of the same color; the window is snapped like the reference's
    (itl & 1) + size - 10."""
    rng, img, koffs = _inputs(8)
    kid = synth.identity_kernels(G)[0]
    cx, cy = np.float32(G.img_width - 1.7), np.float32(G.img_height + 0.2)
    px, py = cx - 8, cy - 8
    itlx, itly = int(np.floor(px + 0.5)), int(np.floor(py + 0.5))
    itlx = (itlx & 1) + G.img_width - 10 if itlx >= G.img_width - 8 else itlx
    itly = (itly & 1) + G.img_height - 10 if itly >= G.img_height - 8 else itly
    sh = np.float64(itlx + 8) - np.float64(cx)
    sv = np.float64(itly + 8) - np.float64(cy)
    for color in range(3):
        q = orc.convert_tile(G, koffs, kid, img, G.img_width, color, cx, cy)
        d = forward_direct(img, color, itlx, itly, sh, sv)
        np.testing.assert_allclose(q, d, atol=1e-12 * np.abs(d).max())


def test_fp32_oracle_close_to_fp64(orc):
    cfg = synth.config_small(seed=3, geom=Geometry(64, 64))
    g = cfg["geom"]
    c64 = orc.convert_correct_tiles(g, cfg["koffs"], cfg["kernels"], cfg["images"], cfg["tasks"])
    c32 = orc.convert_correct_tiles(g, cfg["koffs"], cfg["kernels"], cfg["images"], cfg["tasks"],
                                    dtype=np.float32)
    for a, b in zip(c32, c64):
        assert np.abs(a - b).max() / np.abs(b).max() < 2e-6
