"""Sanitizer builds of the host code (CPU): the oracle under AddressSanitizer
and UndefinedBehaviorSanitizer over tree, assembly, products, GMRES and the
degenerate inputs (empty boxes, coincident points, a zero kernel); the C ABI's
host code (libhodlr3d_asan.so) driven through its argument checks and error
paths.  The same driver runs a whole workload on a GPU (test_gpu_sanitized.py)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_ENV = dict(os.environ, UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_oracle_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "oracle_san"
    plugin = os.path.join(ROOT, "include", "h3d_plugin_plummer.hpp")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-g", "-fopenmp",
                           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                           "-fno-omit-frame-pointer", "-ffp-contract=off",
                           "-I", os.path.join(ROOT, "include"), f'-DH3D_KERNEL_PLUGIN="{plugin}"',
                           os.path.join(ROOT, "oracle", "h3d_oracle.cpp"),
                           os.path.join(ROOT, "tests", "cpp", "oracle_san_main.cpp"),
                           "-o", str(exe)])
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600,
                       env=dict(SAN_ENV, ASAN_OPTIONS="detect_leaks=1"))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


def capi_driver():
    exe = os.path.join(ROOT, "tests", "cpp", "capi_san")
    if not os.path.exists(exe):
        if not os.path.exists(os.path.join(ROOT, "hodlr3d_amd", "libhodlr3d_asan.so")):
            pytest.skip("libhodlr3d_asan.so not built (make -C hodlr3d_amd/csrc asan)")
        import __graft_entry__ as g
        g.build_capi_san()
    return exe


def test_capi_host_code_under_asan_and_ubsan(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("the GPU run is test_gpu_sanitized.py")
    r = subprocess.run([capi_driver(), str(tmp_path)], capture_output=True, text=True,
                       timeout=600, env=dict(SAN_ENV, ASAN_OPTIONS="detect_leaks=0"))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "argument checks only (0 failures)" in r.stdout
