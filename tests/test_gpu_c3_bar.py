"""The 1e-6 bar of the mixed-precision product at N = 2^20 (GPU): the precision
workload's points (2^20 cells of the cell-centred 128^3 lattice of [-1,1]^3,
jitter U(+-h/4), seeded), leaf1d 4 -> 5 levels, K = 1/r, at ACA tolerance
1e-8: at 1e-7 the ACA operator alone exceeds 1e-6 against the direct sum for
some N(0,1) right-hand sides (test_gpu_c3.py), so no recompression of it can
meet the bar on every draw.

The ACA factors are stored in fp32 only (126 GB; the fp64 copy would take
252 GB) and the recompressed copy is built from them:
* the mixed product within 1e-6 of the direct fp64 sum for eight N(0,1)
  right-hand sides (32,768 rows each) and over all 2^20 rows for one;
* A = I + h^3/(4 pi) K solved to 1e-10 by GMRES(30) on the mixed copy inside
  fp64 iterative refinement (outer residual in fp64 arithmetic over the fp32
  factors), the residual checked with an independent product.
"""
import time

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(1200)]

MIXED_TOL = 4e-7
ERR_BAR = 1e-6


@pytest.fixture(scope="module")
def c3_tol8():
    import torch
    import hodlr3d_amd as H3
    from hodlr3d_amd import configs
    x = configs.jittered(1 << 20, seed=0)
    print("[C3 tol 1e-8] assembly (fp32 factors)", flush=True)
    h = H3.HODLR3D(4, 1.0, 8, 7, x, store=("fp32",))
    h.assemble()
    t = time.time()
    info = h.compress(MIXED_TOL)
    print(f"[C3 tol 1e-8] mixed copy from fp32 factors {time.time() - t:.1f} s: {info}",
          flush=True)
    yield h
    h.close()
    torch.cuda.empty_cache()


def test_mixed_product_within_the_bar_on_every_draw(c3_tol8):
    import torch
    h = c3_tol8
    errs = []
    for d in range(8):
        rng = np.random.default_rng(400 + d)
        q = torch.from_numpy(rng.standard_normal(h.N)).cuda()
        rows = np.sort(rng.choice(h.N, 32768, replace=False))
        yd = h.direct(q, rows).cpu().numpy()
        y = h.matvec(q, prec="mixed").cpu().numpy()
        errs.append(float(np.linalg.norm(y[rows] - yd) / np.linalg.norm(yd)))
    print(f"[C3 tol 1e-8] mixed vs direct: {['%.2e' % v for v in errs]}", flush=True)
    assert max(errs) <= ERR_BAR


def test_mixed_product_within_the_bar_on_every_row(c3_tol8):
    import torch
    h = c3_tol8
    q = torch.from_numpy(np.random.default_rng(1).standard_normal(h.N)).cuda()
    rows = np.arange(h.N)
    yd = h.direct(q, rows).cpu().numpy()
    y = h.matvec(q, prec="mixed").cpu().numpy()
    e = float(np.linalg.norm(y - yd) / np.linalg.norm(yd))
    print(f"[C3 tol 1e-8] mixed vs direct over all {h.N} rows: {e:.3e}", flush=True)
    assert e <= ERR_BAR


def test_refined_solve_reaches_1e10(c3_tol8):
    import torch
    from hodlr3d_amd import configs
    h = c3_tol8
    h.set_operator(alpha=configs.second_kind_scale(128), diag=1.0)
    try:
        b = torch.from_numpy(np.random.default_rng(3).standard_normal(h.N)).cuda()
        t = time.time()
        sol, info = h.solve(b, tol=1e-10, restart=30, max_inner=600, inner_tol=1e-5,
                            inner_prec="mixed", max_outer=12)
        torch.cuda.synchronize()
        print(f"[C3 tol 1e-8] refined solve {1e3 * (time.time() - t):.0f} ms: {info}",
              flush=True)
        assert info["rel_residual"] <= 1e-10
        r = b - h.matvec(sol, prec="fp32/acc64")
        assert float(torch.linalg.norm(r) / torch.linalg.norm(b)) <= 2e-10
    finally:
        h.set_operator(alpha=1.0, diag=0.0)
