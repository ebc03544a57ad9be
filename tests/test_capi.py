"""The C ABI library loads and exports every entry point of include/hodlr3d.h;
host-side argument checks answer without a GPU (CPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "hodlr3d.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|void|const char\*)\s+(h3d_\w+)\s*\(",
                                 text, flags=re.M)))


def test_header_lists_entry_points():
    names = declared_functions()
    assert "h3d_create" in names and "h3d_matvec" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    from hodlr3d_amd import _native as nat
    L = nat.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    assert set(declared_functions()) <= set(nat.SIGNATURES)


def test_nlevels_matches_oracle():
    from hodlr3d_amd import _native as nat
    L = nat.lib()
    for N in [1, 7, 64, 65, 4096, 8192, 32768, 32769, 99999, 262144, 2 ** 20, 2 ** 21]:
        for leaf in (1, 2, 3, 4, 6, 8):
            assert L.h3d_nlevels(N, leaf) == O.nlevels(N, leaf), (N, leaf)


def test_invalid_arguments_are_rejected_before_the_gpu():
    from hodlr3d_amd import _native as nat
    L = nat.lib()
    h = ctypes.c_void_p()
    xyz = np.zeros((4, 3))
    p = xyz.ctypes.data_as(ctypes.c_void_p)
    assert L.h3d_create(None, 4, 1, 1.0, 7, None, None, ctypes.byref(h)) == nat.H3D_ERR_INVALID
    assert L.h3d_create(p, 0, 1, 1.0, 7, None, None, ctypes.byref(h)) == nat.H3D_ERR_INVALID
    assert L.h3d_create(p, 4, 0, 1.0, 7, None, None, ctypes.byref(h)) == nat.H3D_ERR_INVALID
    assert L.h3d_create(p, 4, 1, -1.0, 7, None, None, ctypes.byref(h)) == nat.H3D_ERR_INVALID
    bad = xyz.copy()
    bad[1, 2] = np.nan
    assert L.h3d_create(bad.ctypes.data_as(ctypes.c_void_p), 4, 1, 1.0, 7, None, None,
                        ctypes.byref(h)) == nat.H3D_ERR_NONFINITE
    # N < leaf^3 gives nLevels < 1 (the reference would return zeros)
    assert L.h3d_create(p, 4, 4, 1.0, 7, None, None, ctypes.byref(h)) == nat.H3D_ERR_INVALID
    k = nat.Kernel(17, 1.0, 0.0, None)
    many = np.random.default_rng(0).uniform(-1, 1, (100, 3))
    assert L.h3d_create(many.ctypes.data_as(ctypes.c_void_p), 100, 2, 1.0, 7, ctypes.byref(k),
                        None, ctypes.byref(h)) == nat.H3D_ERR_INVALID
    assert not h.value
    assert L.h3d_assemble(None) == nat.H3D_ERR_INVALID
    assert L.h3d_status_string(nat.H3D_ERR_STATE) == b"call out of order"


def test_load_refuses_missing_and_foreign_files_before_the_gpu(tmp_path):
    from hodlr3d_amd import _native as nat
    L = nat.lib()
    h = ctypes.c_void_p()
    missing = str(tmp_path / "absent.h3d").encode()
    assert L.h3d_load(missing, None, ctypes.byref(h)) == nat.H3D_ERR_IO
    for blob in (b"", b"H3DF", b"not a saved operator at all, just text" * 4,
                 b"H3DFACT1" + b"\x03\x00\x00\x00" + b"\0" * 200):
        p = tmp_path / "foreign.h3d"
        p.write_bytes(blob)
        assert L.h3d_load(str(p).encode(), None, ctypes.byref(h)) == nat.H3D_ERR_IO, blob[:12]
    assert not h.value
    assert L.h3d_load(None, None, ctypes.byref(h)) == nat.H3D_ERR_INVALID
    assert L.h3d_save(None, missing) == nat.H3D_ERR_INVALID
    assert L.h3d_get_config(None, None) == nat.H3D_ERR_INVALID
    assert L.h3d_status_string(nat.H3D_ERR_IO) == b"file error"


def test_python_mirror_rejects_bad_shapes():
    import hodlr3d_amd as H3
    with pytest.raises(ValueError):
        H3.HODLR3D(4, 1.0, 7, 7, np.zeros((10, 2)))
    assert H3.nlevels(4096, 4) == 2
