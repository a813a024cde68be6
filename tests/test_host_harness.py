"""The C++ host harness (tile_processor_gpu_amd/host/tp_harness), the
counterpart of the reference's C++ test harness (src/dtt8x8.cu main): it
reads the harness's input files (*_transposed.kernel, *.kernel_offsets,
*.bayer, *.portsxy), builds the task, correlation and texture lists on the
host and runs the path through libtp.so's C ABI.

CPU tests: the host-built lists against geometry.py's builders, argument and
file errors.  GPU test: every result file against the FP64 oracle (CLT,
correlation, imclt planes at 1e-5) and against the same library called
through the Python binding on the same inputs (texture tiles and RGBA image,
whose oracle comparison needs the tie certification of
tests/test_gpu_textures.py: bit-identical here).
"""
import os
import subprocess

import numpy as np
import pytest

from tile_processor_gpu_amd import kernelgen, synth
from tile_processor_gpu_amd.build import HARNESS
from tile_processor_gpu_amd.geometry import corr_indices, texture_indices
from conftest import rel_err

FZ = 0.05
TOL = 1e-5


def _harness_built():
    if not os.path.exists(HARNESS):  # built with the library; build it alone if only that exists
        from tile_processor_gpu_amd.build import build_harness
        build_harness(verbose=False)


def _run(args, timeout=240):
    _harness_built()
    return subprocess.run([HARNESS] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout)


def _geom_args(g):
    return ["--width", g.img_width, "--height", g.img_height, "--kernels-hor", g.kernels_hor,
            "--kernels-vert", g.kernels_vert]


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    """The small config in the reference harness's file formats."""
    cfg = synth.config_small(seed=12)
    d = tmp_path_factory.mktemp("harness")
    prefix = str(d / "main")
    kernelgen.save_kernels(prefix, cfg["kernels"], cfg["koffs"])
    kernelgen.save_frame(prefix, cfg["images"], cfg["tasks"])
    return cfg, d, prefix


def test_help_runs_without_gpu():
    r = _run(["--help"])
    assert r.returncode == 0 and "--prefix" in r.stdout


def test_lists_match_python_builders(files):
    cfg, d, prefix = files
    g = cfg["geom"]
    r = _run(["--prefix", prefix, "--out", d, "--lists-only"] + _geom_args(g))
    assert r.returncode == 0, r.stderr
    lines = (d / "lists.txt").read_text().split("\n")
    n = int(lines[0].split()[1])
    tw = np.array([list(map(int, s.split())) for s in lines[1:1 + n]])
    assert n == g.num_tiles
    np.testing.assert_array_equal(tw[:, 0], cfg["tasks"]["task"])
    np.testing.assert_array_equal(tw[:, 1], cfg["tasks"]["txy"])
    k = 1 + n
    nc = int(lines[k].split()[1])
    corr = np.array(list(map(int, lines[k + 1:k + 1 + nc])))
    np.testing.assert_array_equal(corr, corr_indices(g, cfg["tasks"]))
    k += 1 + nc
    nt = int(lines[k].split()[1])
    tex = np.array(list(map(int, lines[k + 1:k + 1 + nt])))
    np.testing.assert_array_equal(tex, texture_indices(g, cfg["tasks"]))


def test_errors_are_reported(files, tmp_path):
    cfg, d, prefix = files
    g = cfg["geom"]
    r = _run(["--out", tmp_path])
    assert r.returncode == 1 and "--prefix is required" in r.stderr
    r = _run(["--prefix", prefix, "--bogus"])
    assert r.returncode == 1 and "unknown argument" in r.stderr
    r = _run(["--prefix", str(tmp_path / "none"), "--out", tmp_path, "--lists-only"] + _geom_args(g))
    assert r.returncode == 1 and "cannot open" in r.stderr
    # a .portsxy file of the wrong length for the geometry
    r = _run(["--prefix", prefix, "--out", tmp_path, "--lists-only", "--width", 2 * g.img_width,
              "--height", g.img_height])
    assert r.returncode == 1 and "expected" in r.stderr


@pytest.mark.gpu
def test_harness_results_match_oracle_and_binding(files, orc):
    import torch
    from tile_processor_gpu_amd import tp
    cfg, d, prefix = files
    g = cfg["geom"]
    out = d / "out"
    out.mkdir(exist_ok=True)
    print(f"[gpu harness] {g}", flush=True)
    r = _run(["--prefix", prefix, "--out", out, "--iterations", 2, "--save-clt", "--fat-zero", FZ,
              "--radius", 7, "--keep-weights", 1, "--fused"] + _geom_args(g))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Average correlate2D run time" in r.stdout

    # FP64 oracle: CLT, correlation, imclt planes
    clt64 = orc.convert_correct_tiles(g, cfg["koffs"], cfg["kernels"], cfg["images"], cfg["tasks"])
    for c in range(4):
        got = kernelgen.read_floats(str(out / f"main_chn{c}.clt"))
        assert rel_err(got, clt64[c].reshape(-1)) < TOL, c
    idx = corr_indices(g, cfg["tasks"])
    ref = orc.correlate2D(g, clt64, 3, (0.25, 0.25, 0.5), FZ, idx, 7)
    got = kernelgen.read_floats(str(out / "main_corr.corr"), (len(idx), 225))
    assert rel_err(got, ref) < TOL
    fused = kernelgen.read_floats(str(out / "main_corr_fused.corr"), (len(idx), 225))
    assert rel_err(fused, got) <= 1e-6 and rel_err(fused, ref) < TOL
    planes = orc.imclt_frame(g, clt64)
    for c in range(4):
        got = kernelgen.read_floats(str(out / f"main_chn{c}.rbg"), (3, g.img_height + 8, g.img_width + 8))
        for color in range(3):
            assert rel_err(got[color], planes[c][color]) < TOL, (c, color)
    # the whole-frame inverse CLT (one launch) through the harness
    out2 = d / "out_all"
    out2.mkdir(exist_ok=True)
    r2 = _run(["--prefix", prefix, "--out", out2, "--iterations", 1, "--fat-zero", FZ, "--imclt-all"]
              + _geom_args(g))
    assert r2.returncode == 0, r2.stdout[-2000:] + r2.stderr[-2000:]
    assert "Average imclt_rbg_all run time" in r2.stdout
    for c in range(4):
        got = kernelgen.read_floats(str(out2 / f"main_chn{c}.rbg"), (3, g.img_height + 8, g.img_width + 8))
        for color in range(3):
            assert rel_err(got[color], planes[c][color]) < TOL, (c, color)

    # the same library through the Python binding: texture tiles, RGBA image
    fb = tp.FrameBuffers(g, "cuda:0", corr_radius=7, max_corr=len(idx))
    fb.upload(cfg["images"], cfg["kernels"], cfg["koffs"], cfg["tasks"], idx)
    fb.convert()
    kw = tp.TEXTURE_DEFAULTS
    tidx = texture_indices(g, cfg["tasks"])
    ports = torch.tensor([[-0.5, -0.5], [0.5, -0.5], [-0.5, 0.5], [0.5, 0.5]], device="cuda:0")
    nsl = 3 + 1 + 4 + 3 + 1
    tiles = torch.zeros((len(tidx), nsl * 256), device="cuda:0")
    tp.textures_accumulate(g, None, fb.p_clt, len(tidx), torch.from_numpy(tidx).cuda(), ports, 3, 0,
                           kw["min_shot"], kw["scale_shot"], kw["diff_sigma"], kw["diff_threshold"],
                           kw["min_agree"], kw["weight0"], kw["weight1"], kw["weight2"], kw["dust_remove"],
                           1, 0, None, nsl * 256, tiles)
    tasks = torch.from_numpy(np.ascontiguousarray(cfg["tasks"]).view(np.uint8).copy()).cuda()
    ind = torch.zeros(g.tilesx * g.tilesya, dtype=torch.int32, device="cuda:0")
    num = torch.zeros(8, dtype=torch.int32, device="cuda:0")
    woi = torch.zeros(4, dtype=torch.int32, device="cuda:0")
    stride = (g.tilesx + 1) * 8
    rbga = torch.zeros((4 * (g.tilesy + 1) * 8, stride), device="cuda:0")
    tp.generate_RBGA(g, tasks, len(cfg["tasks"]), ind, num, woi, g.tilesx, g.tilesy, fb.p_clt, ports, 3, 0,
                     kw["min_shot"], kw["scale_shot"], kw["diff_sigma"], kw["diff_threshold"],
                     kw["min_agree"], kw["weight0"], kw["weight1"], kw["weight2"], kw["dust_remove"], 0,
                     stride, rbga)
    torch.cuda.synchronize()
    got = kernelgen.read_floats(str(out / "texture.rgba"), (len(tidx), nsl * 256))
    np.testing.assert_array_equal(got, tiles.cpu().numpy())
    got = kernelgen.read_floats(str(out / "texture_rgba.rgba"), tuple(rbga.shape))
    np.testing.assert_array_equal(got, rbga.cpu().numpy())
    assert f"WOI x=0, y=0, width={g.tilesx}, height={g.tilesy}" in r.stdout
