"""Differentiation at the differentiation workload (GPU): N = 262,144 points iid
uniform in [-1,1]^3, leaf1d 4, ACA tolerance 1e-8, A = diag(c) + h^3 K with
K = 1/(4 pi r), h = 2/64, c_i = 1 + 0.1 u_i (u_i iid U[0,1]).

x = A^{-1} b for b iid N(0,1), loss L = ||x - x*||^2 with x* iid N(0,1).  The
adjoint solve A^T lambda = 2 (x - x*) gives dL/db = lambda and dL/dc =
-lambda * x (hodlr3d_amd.autograd.solve); the product's backward gives
d(g^T K q)/dq = K^T g (hodlr3d_amd.autograd.matvec).  Checked against central
finite differences in fp64 along random directions, with the loss differences
formed as (x+ - x-).(x+ + x- - 2 x*) so that they do not cancel; and, at
N = 4096, against a dense fp64 adjoint of the same operator formed column by
column and of the exact kernel matrix.  The forward solve differentiated is
the caller side of classGMRES::gmres (src/gmres.cpp:40-94).
"""
import math

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

SOLVE = dict(tol=1e-12, inner_prec="fp64", inner_tol=1e-6, restart=30, max_inner=600)


def _problem(N, seed, tol_pow):
    import torch
    import hodlr3d_amd as H3
    from hodlr3d_amd import configs
    x = configs.uniform(N, seed=seed)
    h = H3.HODLR3D(4, 1.0, tol_pow, 7, x)
    h.assemble()
    n1d = round(N ** (1 / 3))
    alpha = (2.0 / n1d) ** 3 / (4 * math.pi)
    h.set_operator(alpha=alpha, diag=0.0)
    gen = torch.Generator(device="cuda").manual_seed(seed + 1)
    c = 1.0 + 0.1 * torch.rand(N, dtype=torch.float64, device="cuda", generator=gen)
    b = torch.randn(N, dtype=torch.float64, device="cuda", generator=gen)
    xs = torch.randn(N, dtype=torch.float64, device="cuda", generator=gen)
    return x, h, alpha, c, b, xs, gen


@pytest.fixture(scope="module")
def c5():
    import torch
    print("[C5] N=262144 leaf1d 4 tol 1e-8: GPU assembly", flush=True)
    out = _problem(262144, 5, 8)
    yield out
    out[1].close()
    torch.cuda.empty_cache()


def _loss_and_grads(h, b, c, xs):
    from hodlr3d_amd import autograd as AG
    bb = b.clone().requires_grad_(True)
    cc = c.clone().requires_grad_(True)
    x = AG.solve(h, bb, cc, **SOLVE)
    L = ((x - xs) ** 2).sum()
    L.backward()
    return x.detach(), bb.grad, cc.grad


def test_c5_solve_gradients_match_finite_differences(c5):
    import torch
    from hodlr3d_amd import autograd as AG
    x, h, alpha, c, b, xs, gen = c5
    x0, gb, gc = _loss_and_grads(h, b, c, xs)
    # the forward solve itself
    h.set_diag(c)
    r = b - h.matvec(x0)
    assert float(torch.linalg.norm(r) / torch.linalg.norm(b)) <= 1e-11

    def dL(xp, xm):
        """L(x+) - L(x-) without cancellation."""
        return float(((xp - xm) * (xp + xm - 2 * xs)).sum())

    def sol(bv, cv):
        with torch.no_grad():
            return AG.solve(h, bv, cv, **SOLVE)

    vb = torch.randn(len(b), dtype=torch.float64, device="cuda", generator=gen)
    # In c the derivative along a random direction is a sum of terms of random
    # sign (|dL/dc . v| << ||dL/dc|| ||v||), which the solves' own error of
    # 1e-12 would swamp; along v = sign(dL/dc) * |w| (w iid N(0,1)) the terms
    # add up, and a wrong gradient (sign, scale or index) still shows.
    vc = torch.sign(gc) * torch.randn(len(b), dtype=torch.float64, device="cuda",
                                      generator=gen).abs()
    eps = 1e-3  # x is linear in b: the quotient is exact up to the solve tolerance
    fd_b = dL(sol(b + eps * vb, c), sol(b - eps * vb, c)) / (2 * eps)
    an_b = float((gb * vb).sum())
    eps = 1e-4  # O(eps^2) truncation in c
    fd_c = dL(sol(b, c + eps * vc), sol(b, c - eps * vc)) / (2 * eps)
    an_c = float((gc * vc).sum())
    print(f"[C5] dL/db.v fd {fd_b:.12e} adjoint {an_b:.12e}; dL/dc.v fd {fd_c:.12e} "
          f"adjoint {an_c:.12e}", flush=True)
    assert abs(fd_b - an_b) <= 1e-8 * abs(fd_b)
    assert abs(fd_c - an_c) <= 1e-6 * abs(fd_c)
    # dL/dc = -lambda * x with lambda = dL/db
    assert torch.allclose(gc, -gb * x0, rtol=0, atol=1e-13 * float(gc.abs().max()))


def test_c5_matvec_backward_is_the_transpose(c5):
    import torch
    from hodlr3d_amd import autograd as AG
    x, h, alpha, c, b, xs, gen = c5
    h.set_diag(None)
    h.set_operator(alpha=1.0, diag=0.0)  # K itself
    try:
        q = torch.randn(len(b), dtype=torch.float64, device="cuda", generator=gen,
                        requires_grad=True)
        g = torch.randn(len(b), dtype=torch.float64, device="cuda", generator=gen)
        (g * AG.matvec(h, q)).sum().backward()
        v = torch.randn(len(b), dtype=torch.float64, device="cuda", generator=gen)
        with torch.no_grad():
            fd = float((g * h.matvec(v)).sum())  # linear: the central difference is exact
        an = float((q.grad * v).sum())
        assert abs(fd - an) <= 1e-12 * float(torch.linalg.norm(g) * torch.linalg.norm(
            h.matvec(v)))
    finally:
        h.set_operator(alpha=alpha, diag=0.0)


def test_solve_gradients_match_a_dense_adjoint_at_4096():
    """N = 4096 (tol 1e-10): autograd's dL/db and dL/dc against numpy's dense
    solves with the operator's own matrix (formed column by column, 1e-9) and
    with the exact kernel matrix (within the ACA error)."""
    import torch
    x, h, alpha, c, b, xs, gen = _problem(4096, 7, 10)
    try:
        from hodlr3d_amd import autograd as AG
        x0, gb, gc = _loss_and_grads(h, b, c, xs)
        N = len(x)
        h.set_diag(None)
        h.set_operator(alpha=alpha, diag=0.0)
        I = torch.eye(N, dtype=torch.float64, device="cuda")
        Ah = h.matvec(I).cpu().numpy() + np.diag(c.cpu().numpy())  # alpha K~ + diag(c)
        d = x[:, None, :] - x[None, :, :]
        with np.errstate(divide="ignore"):
            K = 1.0 / np.sqrt((d[..., 0] ** 2 + d[..., 1] ** 2) + d[..., 2] ** 2)
        np.fill_diagonal(K, 0.0)
        Ae = alpha * K + np.diag(c.cpu().numpy())
        bn, xsn = b.cpu().numpy(), xs.cpu().numpy()
        for A, tol in ((Ah, 1e-9), (Ae, 1e-6)):
            xd = np.linalg.solve(A, bn)
            lam = np.linalg.solve(A.T, 2 * (xd - xsn))
            gbd, gcd = lam, -lam * xd
            eb = np.linalg.norm(gb.cpu().numpy() - gbd) / np.linalg.norm(gbd)
            ec = np.linalg.norm(gc.cpu().numpy() - gcd) / np.linalg.norm(gcd)
            print(f"[C5-4096] vs dense ({tol:g}): dL/db {eb:.2e} dL/dc {ec:.2e}", flush=True)
            assert eb <= tol and ec <= tol
    finally:
        h.close()
