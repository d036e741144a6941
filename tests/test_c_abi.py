"""The reference ABI from C: tests/c_abi/drop_in.c is compiled with gcc against
include/maxmulsk.h only and linked with libmaxmulsk.so, as a caller of the
reference's capi/maxmulsk.h would be."""
import os
import shutil
import subprocess

import pytest

import maxmulsk_amd as mm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c_abi", "drop_in.c")


def _build(tmp_path) -> str:
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    lib_dir = os.path.dirname(mm.LIB_PATH)
    exe = str(tmp_path / "drop_in")
    cmd = [gcc, "-std=c99", "-O2", "-Wall", "-Werror", SRC, "-I", os.path.join(ROOT, "include"),
           "-L", lib_dir, "-lmaxmulsk", f"-Wl,-rpath,{lib_dir}", "-lm", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_caller_contract(tmp_path):
    mm.lib()  # the library must be built
    r = subprocess.run([_build(tmp_path), "contract"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)


@pytest.mark.gpu
def test_c_caller_gemm_on_host_arrays(tmp_path):
    from tests.gpu_util import torch_cuda
    torch_cuda()
    r = subprocess.run([_build(tmp_path), "gemm"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
