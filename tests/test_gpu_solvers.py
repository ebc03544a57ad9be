"""GMRES on the GPU vs the oracle's restatement of src/gmres.cpp, and the
refined GMRES(30) solve (GPU)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import lattice, uniform

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def second_kind():
    import hodlr3d_amd as H3
    x = lattice(12)
    N = len(x)
    h = H3.HODLR3D(3, 1.0, 10, 16, x)  # IE_CUBE_3D: h3 = 1/N, Kii = 1 + self-cell integral
    h.assemble()
    o = O.Oracle(x, 3, tol_pow=10, qchoice=16, scale=1.0 / N, diag=O.ie_cube_kii(N))
    o.assemble()
    yield x, h, o
    h.close()


def test_ie_cube_operator_matches_oracle(second_kind):
    x, h, o = second_kind
    q = np.random.default_rng(0).standard_normal(len(x))
    yg, yo = h.MatVecProduct(q), o.matvec(q)
    assert np.linalg.norm(yg - yo) / np.linalg.norm(yo) < 1e-12


def test_reference_gmres_matches_oracle(second_kind):
    import hodlr3d_amd as H3
    x, h, o = second_kind
    rng = np.random.default_rng(1)
    xs = (rng.uniform(size=len(x)) < 0.05).astype(float)  # planted 0/1 solution
    b = o.matvec(xs)
    xg, eg, ig, hg = H3.classGMRES().gmres(h, b, 400, 1e-10)
    xo, eo, io, ho = o.gmres(b, 400, 1e-10)
    assert ig == io and eg <= 1e-10
    assert np.allclose(hg, ho, rtol=1e-6, atol=1e-14)
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-10
    assert np.linalg.norm(xg - xs) / np.linalg.norm(xs) < 1e-8


@pytest.mark.parametrize("inner", ["fp32", "bf16", "fp64"])
def test_refined_solve_reaches_1e10(inner):
    import hodlr3d_amd as H3
    from hodlr3d_amd import configs
    x = uniform(20000, 2)
    h = H3.HODLR3D(4, 1.0, 8, 7, x)
    h.assemble()
    h.set_operator(alpha=configs.second_kind_scale(27), diag=1.0)
    b = np.random.default_rng(3).standard_normal(len(x))
    sol, info = h.solve(b, tol=1e-10, restart=30, inner_prec=inner, inner_tol=1e-4)
    r = b - h.matvec(sol)
    assert info["rel_residual"] <= 1e-10
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 2e-10
    # transpose solve
    solt, it = h.solve(b, tol=1e-10, transpose=True, inner_prec=inner)
    rt = b - h.matvec_transpose(solt)
    assert np.linalg.norm(rt) / np.linalg.norm(b) <= 2e-10
    h.close()
