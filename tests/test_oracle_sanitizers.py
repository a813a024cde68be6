"""The oracle's own tests with the oracle built under AddressSanitizer and
UndefinedBehaviorSanitizer (oracle/Makefile target liboracle_san.so, host
code only): no out-of-bounds access, use after free or undefined behaviour
on any of their inputs.  Runs them in a child process with the ASan runtime
preloaded (Python itself is not instrumented)."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_TESTS = ["tests/test_oracle_clt.py", "tests/test_oracle_corr_imclt.py",
                "tests/test_oracle_textures.py", "tests/test_oracle_tables.py",
                "tests/test_oracle_combine.py", "tests/test_oracle_nn.py", "tests/test_kernelgen.py"]


def test_oracle_tests_clean_under_asan_ubsan():
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("g++ not found")
    asan = subprocess.run([cxx, "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("libasan not found")
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle_san.so"])
    lib = os.path.join(ROOT, "oracle", "liboracle_san.so")
    env = dict(os.environ, LD_PRELOAD=asan, TP_ORACLE_LIB=lib,
               ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    # the child really loads the sanitized build
    probe = subprocess.run([sys.executable, "-c", "from oracle import oracle; oracle.lib(); print(oracle.LIB_PATH)"],
                           cwd=ROOT, env=env, capture_output=True, text=True)
    assert probe.returncode == 0, probe.stderr[-2000:]
    assert probe.stdout.strip() == lib
    # the mutation test compiles its own unsanitized copies: left out
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", *ORACLE_TESTS,
                        "-k", "not mutations"], cwd=ROOT, env=env, capture_output=True, text=True)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in tail and "runtime error" not in tail, tail
