"""The batching workload (GPU), both runs at their sizes.

* nrhs = 32: N = 262,144 iid uniform points in [-1,1]^3, leaf1d 4, ACA tol
  1e-8, fp32 factors, 32 right-hand sides iid N(0,1): the grouped matrix-core
  product against one product per column and against the direct fp64 sum.
* many instances: instances with N spread log-uniformly over the full range
  [8192, 200000], each iid uniform in its own random axis-aligned box (side in
  [0.5, 2]), leaf1d 4, 1/r, ACA tol 1e-8, fp32 factors, one product of each
  replayed from one captured graph: equal to the eager products bit for bit on
  every replay, and each instance within the fp32 error of its direct sum.
The caller side batched here is HODLR3D::MatVecProduct (src/HODLR3D.cpp:21-27).
"""
import math

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


def test_32_rhs_at_262144_points():
    import torch
    import hodlr3d_amd as H3
    from hodlr3d_amd import configs
    N = 262144
    x = configs.uniform(N, seed=11)
    print("[C4] N=262144 nrhs 32 fp32: assembly", flush=True)
    h = H3.HODLR3D(4, 1.0, 8, 7, x, store=("fp32",))
    h.assemble()
    try:
        gen = torch.Generator(device="cuda").manual_seed(12)
        Q = torch.randn(N, 32, dtype=torch.float64, device="cuda", generator=gen)
        h.autotune_rhs("fp32", nrhs=32, reps=1)
        Y = h.matvec(Q, prec="fp32")
        rows = np.sort(np.random.default_rng(13).choice(N, 2048, replace=False))
        worst_single, worst_direct = 0.0, 0.0
        for j in range(32):
            yj = h.matvec(Q[:, j].contiguous(), prec="fp32")
            worst_single = max(worst_single, float(torch.linalg.norm(Y[:, j] - yj) /
                                                   torch.linalg.norm(yj)))
            if j % 8 == 0:
                d = h.direct(Q[:, j].contiguous(), rows).cpu().numpy()
                e = np.linalg.norm(Y[:, j].cpu().numpy()[rows] - d) / np.linalg.norm(d)
                worst_direct = max(worst_direct, float(e))
        print(f"[C4] 32 RHS: vs single products {worst_single:.2e}, vs direct {worst_direct:.2e}",
              flush=True)
        assert worst_single <= 5e-6  # fp32 sums in two orders
        assert worst_direct <= 5e-6
    finally:
        h.close()
        torch.cuda.empty_cache()


def test_instances_over_the_full_range_in_one_graph():
    import torch
    import hodlr3d_amd as H3
    rng = np.random.default_rng(41)
    Ns = [8192, 19000, 45000, 90000, 150000, 200000]
    hs, b, y, pts = [], [], [], []
    gen = torch.Generator(device="cuda").manual_seed(42)
    for n in Ns:
        half = rng.uniform(0.25, 1.0)
        ctr = rng.uniform(-3.0, 3.0, 3)
        p = ctr + rng.uniform(-half, half, (n, 3))
        h = H3.HODLR3D(4, half, 8, 7, p, center=ctr, store=("fp32",))
        h.assemble()
        print(f"[C4] instance N={n} nLevels {h.nLevels}", flush=True)
        hs.append(h)
        pts.append(p)
        b.append(torch.randn(n, dtype=torch.float64, device="cuda", generator=gen))
        y.append(torch.empty(n, dtype=torch.float64, device="cuda"))
    try:
        eager = [h.matvec(bi, prec="fp32") for h, bi in zip(hs, b)]
        g = H3.MatvecGraph(hs, b, y, prec="fp32")
        for rep in range(3):
            g.launch()
            torch.cuda.synchronize()
            assert all(torch.equal(yi, ei) for yi, ei in zip(y, eager)), rep
            _ = [h.matvec(bi, prec="fp32") for h, bi in zip(hs, b)]  # eager work in between
        for h, bi, yi in zip(hs, b, y):
            rows = np.sort(np.random.default_rng(h.N).choice(h.N, 512, replace=False))
            d = h.direct(bi, rows).cpu().numpy()
            e = np.linalg.norm(yi.cpu().numpy()[rows] - d) / np.linalg.norm(d)
            assert e <= 5e-6, (h.N, e)
        g.close()
    finally:
        for h in hs:
            h.close()
        torch.cuda.empty_cache()
