"""Pins of the pair combination and sub-pixel disparity (not in the
reference; its Java host did this): known disparities rendered into 4
cameras are recovered from the combined 15x15 tile."""
import numpy as np
import pytest

from tile_processor_gpu_amd.geometry import Geometry, full_tasks, corr_indices
from tile_processor_gpu_amd import synth
from direct import render_cameras

R = 7


def _combined(orc, g, imgs, task_disparity=0.0, pair_mask=0x3F):
    tasks = full_tasks(g, synth.tile_centres(g, task_disparity))
    clt = orc.convert_correct_tiles(g, synth.zero_offsets(g), synth.identity_kernels(g), imgs, tasks)
    idx = corr_indices(g, tasks)
    c = orc.correlate2D(g, clt, 3, (0.25, 0.25, 0.5), 0.01, idx, R)
    return orc.corr_combine(c, R, pair_mask)


def _interior(g, margin=3):
    ty, tx = np.divmod(np.arange(g.num_tiles), g.tilesx)
    return (tx >= margin) & (tx < g.tilesx - margin) & (ty >= margin) & (ty < g.tilesy - margin)


@pytest.mark.parametrize("d", [0.0, 2.0, -2.0, 4.0])
def test_combined_peak_recovers_disparity(orc, d):
    """Whole-pixel camera displacements (d * 0.5 integer): both estimators
    recover d to a few hundredths of a pixel."""
    g = Geometry(128, 128)
    imgs = render_cameras(g, d, seed=2)
    comb = _combined(orc, g, imgs)
    sel = _interior(g)
    tol = 0.05 + 0.02 * abs(d)  # the 16-px window attenuates larger offsets
    for mode in (0, 1):
        est = orc.corr_disparity(comb, R, mode)[sel, 0]
        assert np.median(np.abs(est - d)) < tol, (mode, np.median(est))
        assert np.mean(np.abs(est - d) < 2 * tol) > 0.9, (mode, est)


@pytest.mark.parametrize("d", [0.5, 1.0, -0.8, 1.6])
def test_fractional_disparity_within_bayer_bias(orc, d):
    """Sub-pixel camera displacements: the per-color CLT of a Bayer mosaic
    carries spectral images that the phase rotation does not move, which
    biases the sub-pixel peak (up to ~0.3 px at quarter-pixel offsets).
    The learned refinement targets exactly this residual."""
    g = Geometry(128, 128)
    imgs = render_cameras(g, d, seed=2)
    comb = _combined(orc, g, imgs)
    est = orc.corr_disparity(comb, R, 1)[_interior(g), 0]
    assert np.median(np.abs(est - d)) < 0.3, np.median(est)


def test_each_pair_maps_onto_the_horizontal_axis(orc):
    """Every pair alone puts the peak at the same disparity after mapping."""
    g = Geometry(128, 128)
    d = 2.0
    imgs = render_cameras(g, d, seed=3)
    sel = _interior(g)
    for p in range(6):
        comb = _combined(orc, g, imgs, pair_mask=1 << p)
        est = orc.corr_disparity(comb, R, 0)[sel, 0]
        assert np.median(np.abs(est - d)) < 0.2, (p, np.median(est))


def test_task_disparity_is_subtracted(orc):
    """Tile centres placed at the true disparity leave a residual of 0, and
    (true 6, task 4) measures what (true 2, task 0) measures."""
    g = Geometry(128, 128)
    imgs = render_cameras(g, 4.0, seed=4)
    est = orc.corr_disparity(_combined(orc, g, imgs, task_disparity=4.0), R, 0)[_interior(g), 0]
    assert np.abs(est).max() < 1e-6
    a = orc.corr_disparity(_combined(orc, g, render_cameras(g, 6.0, seed=4), 4.0), R, 0)
    b = orc.corr_disparity(_combined(orc, g, render_cameras(g, 2.0, seed=4), 0.0), R, 0)
    sel = _interior(g)
    assert np.median(np.abs(a[sel, 0] - b[sel, 0])) < 0.02


def test_mapping_geometry_known_answer(orc):
    """A single delta at (y, x) of one pair tile lands where the mapping says."""
    n = 2 * R + 1
    rows = np.zeros((6, n * n))
    rows[4][(R + 2) * n + (R + 2)] = 1.0   # pair 4: lag (2, 2) -> (0, +2)
    rows[5][(R - 3) * n + (R + 3)] = 1.0   # pair 5: lag (-3, +3) -> (0, +3)
    rows[2][(R + 1) * n + R] = 1.0         # pair 2: lag (1, 0) -> (0, +1)
    out = orc.corr_combine(rows, R)[0].reshape(n, n)
    assert out[R, R + 2] == 1.0 and out[R, R + 3] == 1.0 and out[R, R + 1] == 1.0
    assert out.sum() == 3.0
