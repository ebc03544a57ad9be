"""The C ABI's host code under AddressSanitizer and UBSan (libhodlr3d_asan.so)
through a whole workload on the GPU: products in every precision, several
right-hand sides, transposes, entries, the mixed copy, solves, save / load,
tuning, timers, an injected allocation failure and moves (GPU)."""
import os
import subprocess

import pytest

from tests.test_sanitizers import SAN_ENV, capi_driver

pytestmark = pytest.mark.gpu


def test_capi_workload_under_asan_and_ubsan(tmp_path):
    print("[sanitized driver] workload", flush=True)
    r = subprocess.run([capi_driver(), str(tmp_path)], capture_output=True, text=True,
                       timeout=600, env=dict(SAN_ENV, ASAN_OPTIONS="detect_leaks=0"))
    print(r.stdout[-2000:], r.stderr[-4000:], flush=True)
    assert r.returncode == 0
    assert "GPU workload done (0 failures)" in r.stdout
