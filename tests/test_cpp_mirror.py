"""The header-only C++ mirror (include/hodlr3d.hpp) compiles against the C ABI,
links libhodlr3d.so and turns an error status into an exception (CPU)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r"""
#include "hodlr3d.hpp"
#include <cstdio>
int main() {
  std::vector<pts3D> p(8, pts3D{0.1, 0.2, 0.3});
  try {
    HODLR3D H(4, 1.0, 7, 7, p);   // N < leaf1d^3: nLevels < 1
  } catch (const std::exception& e) {
    std::printf("%s\n", e.what());
    return 3;
  }
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_cpp_mirror_builds_and_reports_errors(tmp_path):
    lib = os.path.join(ROOT, "hodlr3d_amd")
    src = tmp_path / "t.cpp"
    src.write_text(SRC)
    exe = tmp_path / "t"
    subprocess.check_call(["g++", "-std=c++17", str(src), "-I", os.path.join(ROOT, "include"),
                           "-L", lib, "-lhodlr3d", f"-Wl,-rpath,{lib}", "-o", str(exe)])
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 3 and "nLevels" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_reference_driver_compiles_against_the_mirror(tmp_path):
    """tests/cpp/ref_driver.cpp makes the reference driver's calls (H->A->assTime,
    H->A->findMemory(sum), H->mykernel->getCol(k), MatVecProduct on an
    Eigen-shaped vector); without a GPU it must fail with the library's error."""
    exe = tmp_path / "drv"
    lib = os.path.join(ROOT, "hodlr3d_amd")
    subprocess.check_call(["g++", "-std=c++17", "-O1", os.path.join(ROOT, "tests", "cpp",
                                                                    "ref_driver.cpp"),
                           "-I", os.path.join(ROOT, "include"), "-L", lib, "-lhodlr3d",
                           f"-Wl,-rpath,{lib}", "-o", str(exe)])
    r = subprocess.run([str(exe), "4", "2", "1.0", "7", "7", "1", str(tmp_path / "o.bin")],
                       capture_output=True, text=True)
    import torch
    if not torch.cuda.is_available():
        assert r.returncode == 1 and "hodlr3d:" in r.stdout
