"""CPU tests under AddressSanitizer and UndefinedBehaviorSanitizer: the library's
host-only request preparation (csrc/mr_tiles.cpp: task validation, tile enumeration,
height order, the request cache, COR shift groups), its host copy pool
(csrc/mr_hostpool.cpp, also under ThreadSanitizer) and the C oracle
(oracle/oracle_c.c), each built from source with g++/gcc -fsanitize=... by this test
and run on inputs that reach their edge cases (tests/sanitize/*)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "tomography-local-demo_amd", "csrc")
SAN = os.path.join(ROOT, "tests", "sanitize")
FLAGS = ["-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=all"]


def _run(cmd, exe, tmp_path):
    subprocess.run(cmd, check=True, capture_output=True, text=True, cwd=tmp_path)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="2")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
    return r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_tile_preparation_under_sanitizers(tmp_path):
    exe = str(tmp_path / "tiles_check")
    out = _run(["g++", "-std=c++17", *FLAGS, "-I", CSRC, os.path.join(SAN, "tiles_check.cpp"),
                os.path.join(CSRC, "mr_tiles.cpp"), "-o", exe], exe, tmp_path)
    assert "all checks passed" in out


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("san", ["address,undefined", "thread"])
def test_host_copy_pool_under_sanitizers(tmp_path, san):
    exe = str(tmp_path / "hostpool_check")
    flags = [f if not f.startswith("-fsanitize=") else "-fsanitize=" + san for f in FLAGS]
    out = _run(["g++", "-std=c++17", *flags, "-pthread", "-I", CSRC,
                os.path.join(SAN, "hostpool_check.cpp"), os.path.join(CSRC, "mr_hostpool.cpp"),
                "-o", exe], exe, tmp_path)
    assert "all checks passed" in out


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_c_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_check")
    out = _run(["gcc", *FLAGS, "-fopenmp", os.path.join(SAN, "oracle_check.c"),
                os.path.join(ROOT, "oracle", "oracle_c.c"), "-o", exe, "-lm"], exe, tmp_path)
    assert "all checks passed" in out
