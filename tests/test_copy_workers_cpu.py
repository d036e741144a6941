"""Host copy workers of the staging path (maxmulsk_amd/csrc/copy_workers.hpp):
tests/c_abi/copy_workers_stress.cpp is compiled with g++ and run from three
caller threads at once; a second build with ThreadSanitizer checks the pool's
synchronisation (no data race may be reported)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c_abi", "copy_workers_stress.cpp")
INC = os.path.join(ROOT, "maxmulsk_amd", "csrc")


def _build(tmp_path, name, extra):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / name)
    r = subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", *extra, SRC, "-I", INC, "-pthread", "-o", exe],
                       capture_output=True, text=True)
    return exe, r


def test_copy_workers_copy_every_byte(tmp_path):
    exe, r = _build(tmp_path, "stress", ["-O2"])
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe, "150"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, (out.stdout, out.stderr)
    assert "bad 0" in out.stdout


def test_copy_workers_no_data_race(tmp_path):
    exe, r = _build(tmp_path, "stress_tsan", ["-O1", "-g", "-fsanitize=thread"])
    if r.returncode != 0:
        pytest.skip(f"ThreadSanitizer build unavailable: {r.stderr[:200]}")
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    out = subprocess.run([exe, "12"], capture_output=True, text=True, timeout=300, env=env)
    if out.returncode != 0 and "ThreadSanitizer" in out.stderr and "data race" not in out.stderr:
        pytest.skip(f"ThreadSanitizer cannot run here: {out.stderr[:200]}")
    assert out.returncode == 0, out.stderr[-2000:]
    assert "bad 0" in out.stdout
