"""PSF -> CLT-domain kernel generator and the reference's file formats.

Pins (no reference generator exists here: the reference's kernels came from
its Java host, absent from the repository):
* the kernel of a unit impulse at t equals the phase rotation the forward
  CLT applies for a residual shift t (bicomplex phasor of tests/direct.py);
  an impulse at the origin gives the identity kernel;
* a Gaussian kernel equals the closed-form Gaussian frequency response
  sampled at the bin centres (wide enough that sampling aliasing is below
  1e-4);
* convolution invariant on a band-limited periodic image (sub-pixel shifts
  exact through the Fourier shift theorem): the lapped transform of the
  convolved image matches kernel x transform of the original within the
  windowing error, and the opposite sign convention is several times worse;
* the CltExtra offset of an impulse kernel's centroid moves the oracle's
  window the same way as the kernel;
* centroids and their slopes over the kernel grid; file round trips.
"""
import numpy as np
import pytest

from tile_processor_gpu_amd import kernelgen as kg, synth
from tile_processor_gpu_amd.geometry import Geometry
import direct as d


def test_impulse_kernel_is_shift_phasor():
    K0, c = kg.clt_kernel(np.array([[1.0]]))
    ident = np.zeros((4, 8, 8)); ident[0] = 1.0
    assert np.abs(K0 - ident).max() < 1e-15 and c == (0.0, 0.0)
    for t in [(0.3, -0.2), (-1.25, 0.5), (2.0, 3.0)]:
        K, _ = kg.clt_kernel(np.array([[1.0]]), ox=-t[0], oy=-t[1])
        assert np.abs(K - d.phasor(*t)).max() < 1e-12, t


def test_gaussian_kernel_matches_closed_form():
    for sx, sy, th in [(1.3, 1.3, 0.0), (1.6, 1.2, 0.7), (2.0, 1.4, 2.0)]:
        p = kg.gaussian_psf(sx, sy, th, size=31)
        K, c = kg.clt_kernel(p)
        A = synth.psf_kernels(sx, sy, th)
        assert np.abs(K - A).max() < 1e-4, (sx, sy, th, np.abs(K - A).max())
        assert abs(c[0]) < 1e-12 and abs(c[1]) < 1e-12


def _periodic_image(seed=1, n=96, fsig=0.05):
    rng = np.random.default_rng(seed)
    F = np.fft.fft2(rng.standard_normal((n, n)))
    fy = np.fft.fftfreq(n)[:, None]
    fx = np.fft.fftfreq(n)[None, :]
    F *= np.exp(-(fx ** 2 + fy ** 2) / (2 * fsig ** 2))
    return F, fx, fy


def test_convolution_invariant():
    F, fx, fy = _periodic_image()
    I = np.real(np.fft.ifft2(F))
    corners = [(30, 30), (40, 52), (50, 20), (60, 64)]
    # sub-pixel content shift +t: window moved by +t, then the kernel of an impulse at t
    for t in [(0.3, -0.2), (-0.45, 0.35)]:
        J = np.real(np.fft.ifft2(F * np.exp(-2j * np.pi * (fx * t[0] + fy * t[1]))))
        K, _ = kg.clt_kernel(np.array([[1.0]]), ox=-t[0], oy=-t[1])
        Km, _ = kg.clt_kernel(np.array([[1.0]]), ox=t[0], oy=t[1])
        for y0, x0 in corners:
            b = d.mclt_direct(J[y0:y0 + 16, x0:x0 + 16], 0.0, 0.0)
            a = d.mclt_direct(I[y0:y0 + 16, x0:x0 + 16], t[0], t[1])
            e = np.abs(d.bicomplex_mul(a, K) - b).max() / np.abs(b).max()
            em = np.abs(d.bicomplex_mul(a, Km) - b).max() / np.abs(b).max()
            assert e < 0.03 and em > 4 * e, (t, e, em)
    # a centred anisotropic PSF of about 1 px on a sharper image: convolution
    # (FFT) against kernel x transform, which must account for most of the blur
    F, fx, fy = _periodic_image(fsig=0.15)
    I = np.real(np.fft.ifft2(F))
    p = kg.gaussian_psf(1.0, 0.7, 0.4, size=15)
    P = np.fft.fft2(np.roll(np.roll(np.pad(p, ((0, 96 - 15), (0, 96 - 15))), -7, 0), -7, 1))
    J = np.real(np.fft.ifft2(F * P))
    K, _ = kg.clt_kernel(p)
    for y0, x0 in corners:
        b = d.mclt_direct(J[y0:y0 + 16, x0:x0 + 16], 0.0, 0.0)
        a = d.mclt_direct(I[y0:y0 + 16, x0:x0 + 16], 0.0, 0.0)
        e = np.abs(d.bicomplex_mul(a, K) - b).max() / np.abs(b).max()
        e_none = np.abs(a - b).max() / np.abs(b).max()  # no kernel at all
        assert e < 0.15 and e < 0.6 * e_none, (e, e_none)


def test_centroid_offset_moves_window_like_kernel(orc):
    g = Geometry(256, 256)
    img = synth.make_images(g, np.random.default_rng(0))[0]
    ident = synth.identity_kernels(g)
    zoff = synth.zero_offsets(g)[0]
    t = (0.3, -0.2)
    p = np.zeros((5, 5)); p[2, 2] = 1.0
    K, c = kg.clt_kernel(np.array([[1.0]]), ox=-t[0], oy=-t[1])
    grid = np.broadcast_to(K, (g.kernels_vert, g.kernels_hor, 3, 4, 8, 8)).astype(np.float32)
    for color in range(3):
        a = orc.convert_tile(g, zoff, grid, img, g.img_width, color, 100.0, 120.0)
        for sgn, ok in ((1, True), (-1, False)):
            off = zoff.copy()
            off[..., 0], off[..., 1] = sgn * t[0], sgn * t[1]
            b = orc.convert_tile(g, off, ident, img, g.img_width, color, 100.0, 120.0)
            e = np.abs(a - b).max() / np.abs(b).max()
            assert (e < 0.1) if ok else (e > 0.5), (color, sgn, e)


def test_kernel_grid_centroids_and_slopes():
    kv, kh = 4, 5
    psfs = np.zeros((1, kv, kh, 3, 15, 15))
    cx = lambda kx, ky: 0.1 * kx - 0.05 * ky
    cy = lambda kx, ky: 0.02 * kx + 0.2
    for ky in range(kv):
        for kx in range(kh):
            for c in range(3):
                psfs[0, ky, kx, c] = kg.gaussian_psf(1.0, 1.0, 0.0, cx(kx, ky), cy(kx, ky))
    kern, offs = kg.kernel_grid(psfs)
    assert kern.shape == (1, kv, kh, 3, 4, 8, 8) and offs.shape == (1, kv, kh, 3, 8)
    ky, kx = np.mgrid[0:kv, 0:kh]
    assert np.allclose(offs[0, ..., 0, 0], cx(kx, ky), atol=1e-5)
    assert np.allclose(offs[0, ..., 0, 1], cy(kx, ky), atol=1e-5)
    assert np.allclose(offs[0, ..., 0, 2], 16 * kx - 8) and np.allclose(offs[0, ..., 0, 3], 16 * ky - 8)
    assert np.allclose(offs[0, ..., 0, 4], 0.1 / 16, atol=1e-6)
    assert np.allclose(offs[0, ..., 0, 5], -0.05 / 16, atol=1e-6)
    assert np.allclose(offs[0, ..., 0, 6], 0.02 / 16, atol=1e-6)
    assert np.allclose(offs[0, ..., 0, 7], 0.0, atol=1e-6)
    # centred kernels of displaced Gaussians: nearly symmetric about the
    # centroid (sampling a shifted Gaussian leaves a small odd part)
    assert np.abs(kern[..., 1:3, :, :]).max() < 1e-2


def test_files_round_trip(tmp_path):
    g = Geometry(64, 48)
    rng = np.random.default_rng(2)
    kern, offs = synth.make_kernels(g, rng)
    pre = str(tmp_path / "main")
    kg.save_kernels(pre, kern, offs)
    k2, o2 = kg.load_kernels(pre, g)
    assert np.array_equal(k2, kern) and np.array_equal(o2, offs)
    assert (tmp_path / "main_chn0_transposed.kernel").stat().st_size == kern[0].size * 4
    cfg = synth.config_small(seed=3, geom=g)
    kg.save_frame(pre, cfg["images"], cfg["tasks"])
    imgs, tasks = kg.load_frame(pre, g)
    assert np.array_equal(imgs, cfg["images"])
    assert np.array_equal(tasks["xy"], cfg["tasks"]["xy"]) and np.array_equal(tasks["txy"], cfg["tasks"]["txy"])
