"""The C ABI library loads, exports every entry point include/tp.h declares,
and rejects bad arguments before touching the GPU (no GPU needed)."""
import ctypes
import os
import re

import pytest

from tile_processor_gpu_amd import _lib
from tile_processor_gpu_amd.geometry import Geometry, CGeometry

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "tp.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tp_\w+)\s*\(", src)))


def test_header_declarations_exported():
    names = declared()
    assert "tp_convert_correct_tiles" in names and "tp_correlate2D" in names
    L = _lib.lib()
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.PROTOTYPES, f"{n} has no ctypes prototype"


def test_version():
    assert b"gfx950" in _lib.lib().tp_version()


VP = ctypes.c_void_p
NZ = VP(16)  # non-NULL dummy; never dereferenced when validation fails


def rc_err(name, *args):
    L = _lib.lib()
    rc = getattr(L, name)(*args)
    return rc, L.tp_last_error().decode()


def test_convert_rejects_bad_arguments():
    g = Geometry(256, 256).c()
    bad = CGeometry(250, 256, 18, 18)
    assert rc_err("tp_convert_correct_tiles", None, NZ, NZ, NZ, NZ, NZ, 256, 1, 0, None)[0] == -1
    rc, msg = rc_err("tp_convert_correct_tiles", ctypes.byref(bad), NZ, NZ, NZ, NZ, NZ, 256, 1, 0, None)
    assert rc == -1 and "multiples of 16" in msg
    assert rc_err("tp_convert_correct_tiles", ctypes.byref(g), None, NZ, NZ, NZ, NZ, 256, 1, 0, None)[0] == -1
    assert rc_err("tp_convert_correct_tiles", ctypes.byref(g), NZ, NZ, NZ, NZ, NZ, 256, -1, 0, None)[0] == -1
    rc, msg = rc_err("tp_convert_correct_tiles", ctypes.byref(g), NZ, NZ, NZ, NZ, NZ, 100, 1, 0, None)
    assert rc == -1 and "dstride" in msg
    assert rc_err("tp_convert_correct_tiles", ctypes.byref(g), NZ, NZ, NZ, NZ, NZ, 256, 1, 8, None)[0] == -1


def test_correlate_rejects_bad_arguments():
    g = Geometry(256, 256).c()
    F = ctypes.c_float
    args = lambda colors=3, n=10, stride=225, r=7, p=NZ: (
        ctypes.byref(g), p, colors, F(.25), F(.25), F(.5), F(.05), n, NZ, stride, r, NZ, None)
    assert rc_err("tp_correlate2D", *args(colors=2))[0] == -1
    assert rc_err("tp_correlate2D", *args(r=8))[0] == -1
    assert rc_err("tp_correlate2D", *args(r=-1))[0] == -1
    rc, msg = rc_err("tp_correlate2D", *args(stride=224))
    assert rc == -1 and "corr_stride" in msg
    assert rc_err("tp_correlate2D", *args(p=None))[0] == -1
    # nothing to do is not an error and launches nothing
    assert rc_err("tp_correlate2D", *args(n=0))[0] == 0


def test_imclt_and_lpf_reject_bad_arguments():
    g = Geometry(256, 256).c()
    assert rc_err("tp_imclt_rbg", ctypes.byref(g), NZ, NZ, 3, 0, 0, 264, None)[0] == -1
    assert rc_err("tp_imclt_rbg", ctypes.byref(g), NZ, NZ, 0, 2, 0, 264, None)[0] == -1
    assert rc_err("tp_imclt_rbg", ctypes.byref(g), NZ, NZ, 0, 0, 0, 256, None)[0] == -1
    assert rc_err("tp_imclt_rbg_all", ctypes.byref(g), NZ, None, 264, None)[0] == -1
    assert rc_err("tp_set_lpf", 6, NZ)[0] == -1
    assert rc_err("tp_set_lpf", -1, None)[0] == -1


def test_python_wrapper_raises_on_error():
    with pytest.raises(_lib.TPError):
        _lib.call("tp_set_lpf", 9, NZ)


def test_hbm_probe_library():
    """The benchmark's bandwidth probe loads and rejects bad arguments
    without launching anything."""
    from tile_processor_gpu_amd import build as tpb
    P = ctypes.CDLL(tpb.PROBE_LIB)
    P.tp_probe_hbm.restype = ctypes.c_int
    P.tp_probe_hbm.argtypes = [ctypes.c_int, VP, VP, VP, ctypes.c_size_t, ctypes.c_int, VP]
    assert P.tp_probe_hbm(0, NZ, NZ, NZ, 0, 4, None) == -1        # no bytes
    assert P.tp_probe_hbm(0, NZ, NZ, NZ, 17, 4, None) == -1       # not a multiple of 16
    assert P.tp_probe_hbm(0, NZ, NZ, NZ, 1024, 0, None) == -1     # bad blocks per CU
    assert P.tp_probe_hbm(0, None, NZ, NZ, 1024, 4, None) == -1   # read without a source
    assert P.tp_probe_hbm(3, NZ, None, NZ, 1024, 4, None) == -1   # mix without a destination


def test_dtt8x8_rejects_bad_arguments():
    ok = (NZ, NZ, 64, 0, 64, 32, None)
    assert rc_err("tp_dtt8x8", None, *ok[1:])[0] == -1
    rc, msg = rc_err("tp_dtt8x8", NZ, NZ, 64, 8, 64, 32, None)
    assert rc == -1 and "dtt_mode" in msg
    rc, msg = rc_err("tp_dtt8x8", NZ, NZ, 64, 0, 60, 32, None)
    assert rc == -1 and "multiples of 8" in msg
    rc, msg = rc_err("tp_dtt8x8", NZ, NZ, 40, 0, 64, 32, None)
    assert rc == -1 and "src_stride" in msg


def test_stream_noalloc_rejects_too_many_frames():
    g = Geometry(256, 256).c()
    F = ctypes.c_float
    args = lambda n: (ctypes.byref(g), NZ, n, None, 256, 0, 3, F(0.25), F(0.25), F(0.5),  # noqa: E731
                      F(0.05), 225, 7, 225, 0x3F, None, None)
    rc, msg = rc_err("tp_stream_convert_correlate_noalloc", *args(4097))
    assert rc == -1 and "4096" in msg
    assert rc_err("tp_stream_convert_correlate_noalloc", *args(0))[0] == -1


def test_header_compiles_as_c_and_cpp(tmp_path):
    """include/tp.h is plain C (and C++): its struct-size assertions hold
    for both compilers."""
    import shutil
    import subprocess
    src = tmp_path / "t.c"
    src.write_text('#include "tp.h"\nint main(void) { tp_task t; (void)t; return 0; }\n')
    inc = os.path.join(ROOT, "include")
    for cc, args in (("gcc", ["-std=c11"]), ("g++", ["-std=c++17", "-x", "c++"])):
        if not shutil.which(cc):
            pytest.skip(f"{cc} not found")
        subprocess.check_call([cc, *args, "-Wall", "-Werror", "-I", inc, "-fsyntax-only", str(src)])
