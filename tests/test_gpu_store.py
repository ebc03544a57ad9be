"""Saved assembled representation (h3d_save / h3d_load): a loaded handle has
the saved tree, ranks and factors, and its products equal the saved handle's
bit for bit; damaged files are refused (GPU).

The reference keeps nothing on disk, so the bar is identity with the handle
that was saved."""
import os

import numpy as np
import pytest

from tests.conftest import uniform

pytestmark = pytest.mark.gpu

CASES = {
    # name: (points, leaf1d, tol_pow, Qchoice, centre, per-point diagonal)
    "1/r-20k": (lambda: uniform(20000, 41), 4, 8, 7, None, False),
    "q3-12k-centre-diag": (lambda: uniform(12000, 42) + 0.25, 4, 9, 3, (0.25, 0.25, 0.25), True),
}


@pytest.fixture(scope="module", params=list(CASES))
def saved(request, tmp_path_factory):
    import hodlr3d_amd as H3
    mk, leaf, tol_pow, q, ctr, with_diag = CASES[request.param]
    x = mk()
    print(f"[case {request.param}] N={len(x)}", flush=True)
    kern = None
    if with_diag:
        dv = 1.0 + np.random.default_rng(3).uniform(0, 0.1, len(x))
        kern = H3.userkernel(q, 0.5, 0.0, dv)
    h = H3.HODLR3D(leaf, 1.0, tol_pow, q, x, kernel=kern, center=ctr)
    h.assemble()
    h.set_operator(alpha=0.75, diag=0.0)
    path = str(tmp_path_factory.mktemp("store") / "op.h3d")
    h.save(path)
    g = H3.HODLR3D.load(path)
    yield request.param, x, h, g, path
    g.close()
    h.close()


def test_structure_and_config_round_trip(saved):
    name, x, h, g, path = saved
    assert g.config() == h.config()
    assert (g.N, g.nLevels, g.TOL_POW, g.L) == (h.N, h.nLevels, h.TOL_POW, h.L)
    assert np.array_equal(g.perm(), h.perm())
    for j in range(1, h.nLevels + 1):
        rh, sh = h.ranks(j)
        rg, sg = g.ranks(j)
        assert np.array_equal(rh, rg) and np.array_equal(sh, sg), j
    sh, sg = h.stats(), g.stats()
    for k in ("max_rank", "n_blocks", "rank_sum", "factor_elems", "find_memory", "near_pairs",
              "factor_bytes", "u_elems", "v_elems"):
        assert sh[k] == sg[k], k


def test_products_are_bit_identical(saved):
    name, x, h, g, path = saved
    q = np.random.default_rng(11).standard_normal((len(x), 3))
    for transpose in (False, True):
        assert np.array_equal(h.matvec(q, transpose=transpose), g.matvec(q, transpose=transpose))
    for prec in ("fp32", "bf16"):
        h.prepare_precision(prec)
        g.prepare_precision(prec)
        assert np.array_equal(h.matvec(q[:, 0], prec=prec), g.matvec(q[:, 0], prec=prec)), prec
        h.release_precision(prec)
        g.release_precision(prec)
    # the mixed copy is rebuilt from the loaded factors, deterministically
    sh, sg = h.compress(1e-6), g.compress(1e-6)
    for k in ("bytes", "dropped", "rank_fp32", "rank_fp16", "predicted_rel_err"):
        assert sh[k] == sg[k], k
    assert np.array_equal(h.matvec(q[:, 1], prec="mixed"), g.matvec(q[:, 1], prec="mixed"))
    h.release_precision("mixed")
    g.release_precision("mixed")


def test_factor_blocks_and_pivots_round_trip(saved):
    name, x, h, g, path = saved
    for j in range(1, h.nLevels + 1):
        off, idx = h.il(j)
        _, c = h.boxes(j)
        rk, _ = h.ranks(j)
        for k in np.random.default_rng(j).choice(8 ** j, min(8 ** j, 12), replace=False):
            for b in range(off[k], off[k + 1], 7):
                r, n, m = int(rk[b]), int(c[k]), int(c[idx[b]])
                if r == 0:
                    continue
                Bh, Bg = h.block(j, b, r, n, m, uv=True), g.block(j, b, r, n, m, uv=True)
                for a, d in zip(Bh, Bg):
                    assert np.array_equal(a, d), (j, b)


def test_fp32_only_handle_saves_fp32(tmp_path):
    import hodlr3d_amd as H3
    x = uniform(9000, 43)
    h = H3.HODLR3D(4, 1.0, 7, 7, x)
    h.assemble()
    h.move_precision("fp32")
    path = str(tmp_path / "fp32.h3d")
    h.save(path)
    g = H3.HODLR3D.load(path)
    st = g.stats()
    assert st["factor_bytes"][0] == 0 and st["factor_bytes"][1] == h.stats()["factor_bytes"][1]
    q = np.random.default_rng(12).standard_normal(len(x))
    assert np.array_equal(h.matvec(q, prec="fp32"), g.matvec(q, prec="fp32"))
    g.close()
    h.close()


def test_damaged_files_are_refused(saved, tmp_path):
    import hodlr3d_amd as H3
    from hodlr3d_amd import _native as nat
    name, x, h, g, path = saved
    data = open(path, "rb").read()
    cases = {
        "truncated": data[: len(data) // 2],
        "flipped factor byte": data[:-4096] + bytes([data[-4096] ^ 0x10]) + data[-4095:],
        "flipped header": data[:20] + bytes([data[20] ^ 0x01]) + data[21:],
        "trailing byte": data + b"\0",
    }
    for what, blob in cases.items():
        p = tmp_path / "bad.h3d"
        p.write_bytes(blob)
        with pytest.raises(nat.H3DError) as ei:
            H3.HODLR3D.load(str(p))
        assert ei.value.code == nat.H3D_ERR_IO, what
    os.remove(tmp_path / "bad.h3d")
