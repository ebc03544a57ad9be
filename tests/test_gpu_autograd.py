"""Adjoint products and differentiable solves (hodlr3d_amd.autograd) against a
dense adjoint and central finite differences in fp64 (GPU)."""
import numpy as np
import pytest

from tests.conftest import lattice, uniform

pytestmark = pytest.mark.gpu


def test_transpose_equals_the_dense_adjoint():
    """N = 4096 (the reference's 16^3 lattice): A and A^T formed column by
    column through the ABI; A^T must be the exact transpose, and autograd's
    gradient of <w, A b> must be A^T w."""
    import torch
    import hodlr3d_amd as H3
    from hodlr3d_amd import autograd as AG
    x = lattice(16)
    N = len(x)
    h = H3.HODLR3D(4, 1.0, 10, 7, x)
    h.assemble()
    I = torch.eye(N, dtype=torch.float64, device="cuda")
    A = h.matvec(I)
    AT = h.matvec(I, transpose=True)
    assert torch.linalg.norm(AT - A.T) <= 1e-13 * torch.linalg.norm(A)
    gen = torch.Generator(device="cuda").manual_seed(0)
    b = torch.randn(N, dtype=torch.float64, device="cuda", generator=gen, requires_grad=True)
    w = torch.randn(N, dtype=torch.float64, device="cuda", generator=gen)
    y = AG.matvec(h, b)
    assert torch.linalg.norm(y - A @ b.detach()) <= 1e-13 * torch.linalg.norm(y)
    (w * y).sum().backward()
    assert torch.linalg.norm(b.grad - A.T @ w) <= 1e-13 * torch.linalg.norm(b.grad)
    h.close()


@pytest.fixture(scope="module")
def second_kind():
    """A = diag(c) + h^3/(4 pi r) with c = 1 + 0.1 U[0,1] (the C5 operator, smaller N)."""
    import torch
    import hodlr3d_amd as H3
    from hodlr3d_amd import configs
    x = uniform(8000, 5)
    h = H3.HODLR3D(4, 1.0, 9, 7, x)
    h.assemble()
    h.set_operator(alpha=configs.second_kind_scale(20), diag=0.0)
    gen = torch.Generator(device="cuda").manual_seed(1)
    c = 1.0 + 0.1 * torch.rand(len(x), dtype=torch.float64, device="cuda", generator=gen)
    b = torch.randn(len(x), dtype=torch.float64, device="cuda", generator=gen)
    w = torch.randn(len(x), dtype=torch.float64, device="cuda", generator=gen)
    yield h, c, b, w, gen
    h.close()


def test_solve_solves(second_kind):
    import torch
    from hodlr3d_amd import autograd as AG
    h, c, b, w, gen = second_kind
    xs = AG.solve(h, b, c)
    h.set_diag(c)
    r = b - h.matvec(xs)
    assert torch.linalg.norm(r) <= 1e-11 * torch.linalg.norm(b)


def test_solve_gradients_match_finite_differences(second_kind):
    import torch
    from hodlr3d_amd import autograd as AG
    h, c, b, w, gen = second_kind
    bb = b.clone().requires_grad_(True)
    cc = c.clone().requires_grad_(True)
    xs = AG.solve(h, bb, cc)
    (w * xs).sum().backward()
    vb = torch.randn(len(b), dtype=torch.float64, device="cuda", generator=gen)
    vc = torch.randn(len(b), dtype=torch.float64, device="cuda", generator=gen)

    def L(bv, cv):
        with torch.no_grad():
            return float((w * AG.solve(h, bv, cv)).sum())

    # in b the map is linear: the difference quotient is exact up to the solve tolerance
    eps = 1e-3
    fd_b = (L(b + eps * vb, c) - L(b - eps * vb, c)) / (2 * eps)
    assert abs(fd_b - float((bb.grad * vb).sum())) <= 1e-8 * abs(fd_b)
    # in c: central differences, O(eps^2) truncation
    eps = 1e-5
    fd_c = (L(b, c + eps * vc) - L(b, c - eps * vc)) / (2 * eps)
    assert abs(fd_c - float((cc.grad * vc).sum())) <= 1e-6 * abs(fd_c)
