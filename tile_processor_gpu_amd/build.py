"""Build libtp.so (gfx950) in-tree with hipcc.

Usage: python -m tile_processor_gpu_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtp.so")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    # no packed FP32 instructions anywhere (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fconstexpr-steps=100000000", "-Wall", "-Wno-unused-function",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


PROBE_LIB = os.path.join(HERE, "libtp_probe.so")


def _compile(src, out, defines, verbose):
    tmp = out + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-D" + d for d in defines] + ["-o", tmp, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, out)
    return out


HOST = os.path.join(HERE, "host")
HARNESS = os.path.join(HOST, "tp_harness")


def build_harness(verbose=True):
    """The C++ host harness (host/tp_harness.cpp, host code only) linked
    against the in-tree libtp.so, which it finds next to its own folder."""
    tmp = HARNESS + ".tmp"
    cmd = [hipcc(), "-O2", "-std=c++17", "-Wall", "-I", os.path.join(os.path.dirname(HERE), "include"),
           os.path.join(HOST, "tp_harness.cpp"), "-L", HERE, "-ltp", "-Wl,-rpath,$ORIGIN/..", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, HARNESS)
    return HARNESS


def build(verbose=True, out=None, defines=(), probe=True):
    """Compile csrc/tp_api.hip (the whole library is one translation unit)
    and, with probe=True, the HBM bandwidth probe csrc/hbm_probe.hip that the
    benchmark uses (libtp_probe.so) and the C++ host harness.  defines: extra
    -D definitions for the compiler; out: library path (default: the in-tree
    libtp.so)."""
    out = _compile(os.path.join(CSRC, "tp_api.hip"), out or LIB, defines, verbose)
    if probe:
        _compile(os.path.join(CSRC, "hbm_probe.hip"), PROBE_LIB, (), verbose)
        build_harness(verbose)
    return out


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("-o", "--out", default=None)
    ap.add_argument("-D", "--define", action="append", default=[])
    a = ap.parse_args()
    build(out=a.out, defines=a.define, probe=a.out is None)
    sys.exit(0)
