"""Host-side mirror of the reference's facade, kernel plugin and GMRES.

Names and argument meaning follow SAFRAN-LAB/HODLR3D:
  * ``HODLR3D(nParticlesInLeafAlong1D, L, TOL_POW, Qchoice, particles)``,
    ``assemble()``, ``MatVecProduct(b)``  -- include/HODLR3D.hpp:12-22,
    src/HODLR3D.cpp:3-27;
  * ``userkernel`` (Qchoice catalogue + diagonal) -- include/kernel.hpp:37-115,
    examples/testHODLR3Dsolve.cpp:32-76;
  * ``classGMRES().gmres(H, b, maxIterations, threshold)`` -- include/gmres.hpp:13-27,
    src/gmres.cpp:40-94, returning what the reference writes to its reference
    arguments: ``(x, error, noOfIterations, e)``;
  * ``H.A`` (the tree): ``findMemory()``, ``getMaxRank()``, ``assTime``,
    ``matVecTime`` -- src/HODLR3DTree.cpp:1757-1818, include/HODLR3DTree.hpp:46;
  * ``H.mykernel``: ``getMatrixEntry(i, j)``, ``getRow``, ``getCol``,
    ``getMatrix`` -- include/kernel.hpp:19-33, src/kernel.cpp:7-62;
  * ``userkernel(H3.QCHOICE_PLUGIN)`` selects the kernel plugin built into the
    library (include/h3d_kernel_plugin.hpp).
Every call goes through libhodlr3d.so (the C ABI of include/hodlr3d.h) and runs
on the GPU; invalid input raises ``H3DError`` instead of undefined behaviour.
"""
import ctypes
import math
import os

import numpy as np

from . import _native as nat

PRECISIONS = {"fp64": nat.PREC_FP64, "fp32": nat.PREC_FP32, "bf16": nat.PREC_BF16,
              "mixed": nat.PREC_MIXED}


def _prec(p):
    """'fp64' | 'fp32' | 'bf16', optionally with '/acc64' (fp64 accumulation), or
    'mixed' (the recompressed fp32/fp24/fp16/fp8 copy built by ``compress``)."""
    if isinstance(p, str):
        base, _, acc = p.partition("/")
        v = PRECISIONS[base]
        if acc:
            if acc != "acc64":
                raise ValueError(f"unknown accumulation {acc!r}")
            v |= nat.ACC_FP64
        return v
    return int(p)


def nlevels(N, leaf1d):
    """nLevels = ceil(3 log(N^(1/3)/leaf1d)/log 8) (src/HODLR3D.cpp:6-7)."""
    return nat.lib().h3d_nlevels(int(N), int(leaf1d))


class pts3D:
    """A point (include/definitions.hpp:7-9)."""
    __slots__ = ("x", "y", "z")

    def __init__(self, x=0.0, y=0.0, z=0.0):
        self.x, self.y, self.z = float(x), float(y), float(z)


def as_xyz(particles):
    if isinstance(particles, np.ndarray):
        a = np.ascontiguousarray(particles, dtype=np.float64)
    elif hasattr(particles, "detach"):
        a = np.ascontiguousarray(particles.detach().cpu().numpy(), dtype=np.float64)
    else:
        a = np.array([[p.x, p.y, p.z] for p in particles], dtype=np.float64)
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError("particles must be N x 3")
    return a


class userkernel:
    """Kernel plugin: entry(i,j) = scale * phi_Qchoice(r_ij) for i != j, diag for i == j.

    ``scale=1, diag=0`` is the reference for Qchoice 0..15.  Qchoice 16
    (IE_CUBE_3D) with ``scale=diag=0`` uses h3 = 1/N and Kii = 1 + the
    self-cell integral, as include/kernel.hpp:45-60 does.  ``diag_vec`` sets a
    per-point diagonal (second-kind operators diag(c) + scale*K).
    """

    def __init__(self, Qchoice=7, scale=None, diag=0.0, diag_vec=None):
        self.Qchoice = int(Qchoice)
        if scale is None:
            scale = 0.0 if self.Qchoice == 16 else 1.0
        self.scale = float(scale)
        self.diag = float(diag)
        self.diag_vec = None if diag_vec is None else np.ascontiguousarray(diag_vec, np.float64)


QCHOICE_PLUGIN = nat.QCHOICE_PLUGIN


def plugin_name():
    """Name of the kernel plugin compiled into libhodlr3d.so."""
    return nat.lib().h3d_plugin_name().decode()


class _BoundKernel:
    """Kernel entries of a handle's points, evaluated on the GPU
    (h3d_get_matrix): the reference's kernel::getRow/getCol/getMatrix
    (src/kernel.cpp:7-62) with original indices."""

    def __init__(self, H):
        self._H = H
        self.Qchoice = H.kernel.Qchoice

    def getMatrix(self, row_indices, col_indices):
        r = np.ascontiguousarray(row_indices, dtype=np.int64).ravel()
        c = np.ascontiguousarray(col_indices, dtype=np.int64).ravel()
        out = np.empty((len(c), len(r)))  # column-major nrows x ncols
        nat.check(nat.lib().h3d_get_matrix(self._H._h, r.ctypes.data_as(ctypes.c_void_p), len(r),
                                           c.ctypes.data_as(ctypes.c_void_p), len(c),
                                           out.ctypes.data_as(ctypes.c_void_p), nat.HOST))
        return out.T

    def getMatrixEntry(self, i, j):
        return float(self.getMatrix([i], [j])[0, 0])

    def getRow(self, j, col_indices):
        return self.getMatrix([j], col_indices)[0]

    def getCol(self, k, row_indices=None):
        rows = np.arange(self._H.N) if row_indices is None else row_indices
        return self.getMatrix(rows, [k])[:, 0]


class _Tree:
    """The reference's HODLR3DTree members the drivers read (H->A->...)."""

    def __init__(self, H):
        self._H = H

    @property
    def assTime(self):
        """Seconds of kernel-entry regeneration inside products (the on-the-fly
        near field), accumulated while per-kernel timing is on."""
        return self._H.stats()["ref_ass_time"]

    @property
    def matVecTime(self):
        """Seconds of the low-rank BLAS-2 applies (far field), accumulated the same way."""
        return self._H.stats()["ref_matvec_time"]

    def findMemory(self):
        return self._H.stats()["find_memory"]

    def getMaxRank(self):
        return self._H.stats()["max_rank"]


def _is_torch_cuda(a):
    return hasattr(a, "is_cuda") and a.is_cuda


class HODLR3D:
    """The reference facade (src/HODLR3D.cpp:3-27) on the MI355X engine."""

    def __init__(self, nParticlesInLeafAlong1D, L, TOL_POW, Qchoice=7, particles=None, *,
                 kernel=None, center=None, nlevels=None, workspace_bytes=0, store=("fp64",)):
        if particles is None:
            raise ValueError("particles are required")
        xyz = as_xyz(particles)
        self.N = len(xyz)
        self.kernel = kernel if kernel is not None else userkernel(Qchoice)
        k = nat.Kernel(self.kernel.Qchoice, self.kernel.scale, self.kernel.diag,
                       None if self.kernel.diag_vec is None else
                       self.kernel.diag_vec.ctypes.data_as(ctypes.c_void_p))
        opt = nat.Options()
        c = center if center is not None else (0.0, 0.0, 0.0)
        for i in range(3):
            opt.center[i] = float(c[i])
        opt.nlevels = -1 if nlevels is None else int(nlevels)
        opt.workspace_bytes = int(workspace_bytes)
        opt.store_mask = sum(1 << PRECISIONS[p] for p in store)
        h = ctypes.c_void_p()
        self._h = None
        nat.check(nat.lib().h3d_create(xyz.ctypes.data_as(ctypes.c_void_p), self.N,
                                       int(nParticlesInLeafAlong1D), float(L), int(TOL_POW),
                                       ctypes.byref(k), ctypes.byref(opt), ctypes.byref(h)))
        self._h = h
        self.nParticlesInLeafAlong1D = int(nParticlesInLeafAlong1D)
        self.L = float(L)
        self.TOL_POW = int(TOL_POW)
        self.nLevels = self.stats()["nlevels"]
        self.assembled = False

    # -- saved representation (no reference counterpart: the reference keeps nothing on disk)
    def save(self, path):
        """Write the assembled representation (points, kernel, operator, tree
        permutation, ranks, factors, pivots; checksummed) to ``path``."""
        nat.check(nat.lib().h3d_save(self._h, os.fsencode(path)))

    @classmethod
    def load(cls, path, *, workspace_bytes=0):
        """An assembled handle from a file written by ``save``: its products
        equal the saved handle's bit for bit.  Raises H3DError (status
        H3D_ERR_IO) for a missing, truncated or corrupt file."""
        opt = nat.Options()
        opt.nlevels = -1
        opt.workspace_bytes = int(workspace_bytes)
        h = ctypes.c_void_p()
        nat.check(nat.lib().h3d_load(os.fsencode(path), ctypes.byref(opt), ctypes.byref(h)))
        self = cls.__new__(cls)
        self._h = h
        c = self.config()
        self.N = c["N"]
        self.kernel = userkernel(c["qchoice"], c["kernel_scale"], c["kernel_diag"])
        self.nParticlesInLeafAlong1D = c["leaf1d"]
        self.L = c["L"]
        self.TOL_POW = c["tol_pow"]
        self.nLevels = c["nlevels"]
        self.assembled = True
        return self

    def config(self):
        c = nat.Config()
        nat.check(nat.lib().h3d_get_config(self._h, ctypes.byref(c)))
        return {"N": c.N, "nlevels": c.nlevels, "leaf1d": c.leaf1d, "L": c.L,
                "center": tuple(c.center), "tol_pow": c.tol_pow, "qchoice": c.qchoice,
                "kernel_scale": c.kernel_scale, "kernel_diag": c.kernel_diag, "alpha": c.alpha,
                "has_diag_vec": bool(c.has_diag_vec)}

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None):
            nat.lib().h3d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference API
    def assemble(self):
        nat.check(nat.lib().h3d_assemble(self._h))
        self.assembled = True

    def MatVecProduct(self, b, prec="fp64"):
        """Returns A b (original ordering), like HODLR3D::MatVecProduct."""
        return self.matvec(b, prec=prec)

    def findMemory(self):
        return self.stats()["find_memory"]

    def getMaxRank(self):
        return self.stats()["max_rank"]

    @property
    def A(self):
        return _Tree(self)

    @property
    def mykernel(self):
        return _BoundKernel(self)

    # -- engine API
    def _bind_stream(self):
        import torch
        nat.check(nat.lib().h3d_set_stream(self._h,
                                           ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def matvec(self, b, prec="fp64", transpose=False):
        """y = A b (or A^T b).  b: numpy (host) or torch CUDA tensor, shape (N,) or (N, nrhs)."""
        fn = nat.lib().h3d_matvec_transpose if transpose else nat.lib().h3d_matvec
        if _is_torch_cuda(b):
            import torch
            bt = b.detach().to(torch.float64)
            cols = bt.reshape(self.N, -1).t().contiguous()  # nrhs x N, column-major RHS
            out = torch.empty_like(cols)
            self._bind_stream()
            try:
                nat.check(fn(self._h, ctypes.c_void_p(cols.data_ptr()),
                             ctypes.c_void_p(out.data_ptr()), cols.shape[0], _prec(prec), nat.DEVICE))
            finally:
                nat.lib().h3d_use_own_stream(self._h)
            return out.t().reshape(b.shape)
        a = np.asarray(b, dtype=np.float64)
        if a.shape[0] != self.N:
            raise ValueError("b has the wrong length")
        cols = np.ascontiguousarray(a.reshape(self.N, -1).T)
        out = np.empty_like(cols)
        nat.check(fn(self._h, cols.ctypes.data_as(ctypes.c_void_p),
                     out.ctypes.data_as(ctypes.c_void_p), cols.shape[0], _prec(prec), nat.HOST))
        return out.T.reshape(a.shape)

    def matvec_transpose(self, b, prec="fp64"):
        return self.matvec(b, prec=prec, transpose=True)

    def prepare_precision(self, prec):
        nat.check(nat.lib().h3d_prepare_precision(self._h, _prec(prec)))

    def release_precision(self, prec):
        nat.check(nat.lib().h3d_release_precision(self._h, _prec(prec)))

    def move_precision(self, prec):
        """Convert the factors to `prec` chunk by chunk, freeing the other copies."""
        nat.check(nat.lib().h3d_move_precision(self._h, _prec(prec)))

    def set_operator(self, alpha=1.0, diag=0.0):
        """Apply-time operator y = alpha * K~ b + diag * b (per-point diagonal if set)."""
        nat.check(nat.lib().h3d_set_operator(self._h, float(alpha), float(diag)))

    def autotune(self, prec="fp64", reps=3):
        """Pick the fastest per-lane load width of both far-field kernels for `prec`
        on this problem; returns the measured ms per width (for 'mixed' also the
        fp16 class under 'phase1_fp16' / 'phase2_fp16')."""
        ms = (ctypes.c_double * 16)(*([-1.0] * 16))
        nat.check(nat.lib().h3d_autotune(self._h, _prec(prec), int(reps), ms))
        w = [16, 8, 4, 2]
        out = {"phase1": {w[i]: ms[i] for i in range(4) if ms[i] >= 0},
               "phase2": {w[i]: ms[4 + i] for i in range(4) if ms[4 + i] >= 0}}
        if _prec(prec) == nat.PREC_MIXED:
            out["phase1_fp16"] = {w[i]: ms[8 + i] for i in range(4) if ms[8 + i] >= 0}
            out["phase2_fp16"] = {w[i]: ms[12 + i] for i in range(4) if ms[12 + i] >= 0}
        return out

    def set_mixed_formats(self, formats=("fp16", "fp8", "fp24", "int8")):
        """Formats the next build of the mixed copy may store columns in besides
        fp32: any of 'fp16', 'fp8', 'fp24' and 'int8' (the 8-bit tier in int8
        with a scale per column side instead of e4m3) (h3d_set_mixed_formats)."""
        bits = {"fp16": MIXED_FP16, "fp8": MIXED_FP8, "fp24": MIXED_FP24, "int8": MIXED_INT8}
        mask = 0
        for f in formats:
            if f not in bits:
                raise nat.H3DError(nat.H3D_ERR_INVALID, f"unknown mixed-copy format {f!r}")
            mask |= bits[f]
        nat.check(nat.lib().h3d_set_mixed_formats(self._h, mask))

    def compress(self, tol=None, formats=None):
        """Build the recompressed mixed-precision copy ('mixed') with an expected
        added relative error of `tol` (default 5 * 10^-TOL_POW); returns its stats.
        `formats` (optional) sets the storage formats first (set_mixed_formats)."""
        if tol is None:
            tol = 5.0 * 10.0 ** (-self.TOL_POW)
        if formats is not None:
            self.set_mixed_formats(formats)
        s = nat.MixedStats()
        nat.check(nat.lib().h3d_compress(self._h, float(tol), ctypes.byref(s)))
        return self.mixed_stats()

    def mixed_stats(self):
        s = nat.MixedStats()
        nat.check(nat.lib().h3d_get_mixed_stats(self._h, ctypes.byref(s)))
        d = {k: getattr(s, k) for k, _ in nat.MixedStats._fields_}
        d["u_elems"] = list(s.u_elems)
        d["v_elems"] = list(s.v_elems)
        return d

    def set_tuning(self, prec, p1_bytes, p2_bytes):
        """Pin the per-lane load widths (bytes) of the two far-field kernels for
        `prec`; for 'mixed' phase 1 takes 16 only (H3DError otherwise)."""
        nat.check(nat.lib().h3d_set_tuning(self._h, _prec(prec), int(p1_bytes), int(p2_bytes)))

    def get_tuning(self, prec):
        """The complete tuning state of `prec` (see h3d_get_tuning): six ints."""
        v = (ctypes.c_int * 6)()
        nat.check(nat.lib().h3d_get_tuning(self._h, _prec(prec), v))
        return list(v)

    def set_tuning_state(self, prec, v):
        """Pin a tuning state read back by ``get_tuning`` (h3d_set_tuning_state)."""
        a = (ctypes.c_int * 6)(*[int(t) for t in v])
        nat.check(nat.lib().h3d_set_tuning_state(self._h, _prec(prec), a))

    def set_rhs_block(self, width):
        """Group width (16 or 32; 0: one at a time) of products with several
        right-hand sides on the ACA factor copies (matrix-core path)."""
        nat.check(nat.lib().h3d_set_rhs_block(self._h, int(width)))

    def autotune_rhs(self, prec="fp32", nrhs=32, reps=3):
        """Pick the group width and phase-1 item size of grouped products of
        `nrhs` columns by timing them; returns the choice and every timing."""
        ms = (ctypes.c_double * 4)()
        best = (ctypes.c_int * 2)()
        nat.check(nat.lib().h3d_autotune_rhs(self._h, _prec(prec), int(nrhs), int(reps), ms, best))
        times = {(w, c): ms[i * 2 + j] for i, w in enumerate((16, 32))
                 for j, c in enumerate((1024, 2048)) if ms[i * 2 + j] >= 0}
        return {"width": best[0], "columns": best[1], "ms": times}

    def kernel_timing(self, enable=True):
        """Returns accumulated per-kernel ms since the last call, then (re)arms timing."""
        ms = (ctypes.c_double * 15)()
        cnt = ctypes.c_int()
        nat.check(nat.lib().h3d_kernel_timing(self._h, int(enable), ms, ctypes.byref(cnt)))
        return {"phase1": ms[0], "phase2": ms[1], "near": ms[2], "combine": ms[3],
                "total": ms[4], "phase1_second": ms[5], "phase2_second": ms[6],
                "phase1_kernel": ms[7], "phase1_kernel_second": ms[8],
                "phase1_third": ms[9], "phase2_third": ms[10], "phase1_kernel_third": ms[11],
                "phase1_fourth": ms[12], "phase2_fourth": ms[13], "phase1_kernel_fourth": ms[14],
                "count": cnt.value}

    def set_diag(self, c):
        if _is_torch_cuda(c):
            import torch
            ct = c.detach().to(torch.float64).contiguous()
            self._bind_stream()
            try:
                nat.check(nat.lib().h3d_set_diag(self._h, ctypes.c_void_p(ct.data_ptr()), nat.DEVICE))
            finally:
                nat.lib().h3d_use_own_stream(self._h)
            return
        if c is None:
            nat.check(nat.lib().h3d_set_diag(self._h, None, nat.HOST))
            return
        a = np.ascontiguousarray(c, dtype=np.float64)
        nat.check(nat.lib().h3d_set_diag(self._h, a.ctypes.data_as(ctypes.c_void_p), nat.HOST))

    def solve(self, b, tol=1e-10, restart=30, max_inner=300, inner_tol=1e-5, inner_prec="fp32",
              max_outer=20, transpose=False):
        """GMRES(restart) in `inner_prec` inside fp64 iterative refinement.

        Returns (x, info dict).  Raises on non-convergence."""
        info = nat.SolveInfo()
        if _is_torch_cuda(b):
            import torch
            bt = b.detach().to(torch.float64).contiguous()
            x = torch.empty_like(bt)
            self._bind_stream()
            try:
                rc = nat.lib().h3d_solve_ir(self._h, ctypes.c_void_p(bt.data_ptr()),
                                            ctypes.c_void_p(x.data_ptr()), restart, max_inner,
                                            inner_tol, _prec(inner_prec), tol, max_outer,
                                            int(transpose), ctypes.byref(info), nat.DEVICE)
            finally:
                nat.lib().h3d_use_own_stream(self._h)
        else:
            a = np.ascontiguousarray(b, dtype=np.float64)
            x = np.empty_like(a)
            rc = nat.lib().h3d_solve_ir(self._h, a.ctypes.data_as(ctypes.c_void_p),
                                        x.ctypes.data_as(ctypes.c_void_p), restart, max_inner,
                                        inner_tol, _prec(inner_prec), tol, max_outer,
                                        int(transpose), ctypes.byref(info), nat.HOST)
        d = {"outer_iterations": info.outer_iterations, "inner_iterations": info.inner_iterations,
             "outer_prec": info.outer_prec,
             "rel_residual": info.rel_residual,
             "history": [info.history[i] for i in range(min(info.outer_iterations, 64))]}
        nat.check(rc)
        return x, d

    def direct(self, b, rows):
        """Direct FP64 sum (A b)[rows] from kernel entries (compensated)."""
        r = np.ascontiguousarray(rows, dtype=np.int64)
        if _is_torch_cuda(b):
            import torch
            bt = b.detach().to(torch.float64).contiguous()
            out = torch.empty(len(r), dtype=torch.float64, device=bt.device)
            self._bind_stream()
            try:
                nat.check(nat.lib().h3d_direct(self._h, ctypes.c_void_p(bt.data_ptr()),
                                               r.ctypes.data_as(ctypes.c_void_p), len(r),
                                               ctypes.c_void_p(out.data_ptr()), nat.DEVICE))
            finally:
                nat.lib().h3d_use_own_stream(self._h)
            return out
        a = np.ascontiguousarray(b, dtype=np.float64)
        out = np.empty(len(r))
        nat.check(nat.lib().h3d_direct(self._h, a.ctypes.data_as(ctypes.c_void_p),
                                       r.ctypes.data_as(ctypes.c_void_p), len(r),
                                       out.ctypes.data_as(ctypes.c_void_p), nat.HOST))
        return out

    def stats(self):
        s = nat.Stats()
        nat.check(nat.lib().h3d_get_stats(self._h, ctypes.byref(s)))
        return {"N": s.N, "nlevels": s.nlevels, "max_rank": s.max_rank, "n_blocks": s.n_blocks,
                "rank_sum": s.rank_sum, "factor_elems": s.factor_elems,
                "find_memory": s.find_memory, "near_pairs": s.near_pairs,
                "t_tree_ms": s.t_tree_ms, "t_assemble_ms": s.t_assemble_ms,
                "factor_bytes": list(s.factor_bytes), "u_elems": s.u_elems,
                "v_elems": s.v_elems, "ref_ass_time": s.ref_ass_time,
                "ref_matvec_time": s.ref_matvec_time}

    def geometry_stats(self):
        """Sizes that set the algorithmic bytes of a matvec."""
        return self.stats()

    # -- structure export (bit-exact checks)
    def perm(self):
        p = np.zeros(self.N, dtype=np.int32)
        nat.check(nat.lib().h3d_export_perm(self._h, p.ctypes.data_as(ctypes.c_void_p)))
        return p

    def boxes(self, level):
        nb = 8 ** level
        s = np.zeros(nb, dtype=np.int32)
        c = np.zeros(nb, dtype=np.int32)
        nat.check(nat.lib().h3d_export_boxes(self._h, level, s.ctypes.data_as(ctypes.c_void_p),
                                             c.ctypes.data_as(ctypes.c_void_p)))
        return s, c

    def neighbors(self, level):
        n = np.zeros((8 ** level, 26), dtype=np.int32)
        nat.check(nat.lib().h3d_export_neighbors(self._h, level, n.ctypes.data_as(ctypes.c_void_p)))
        return n

    def il(self, level):
        sz = nat.lib().h3d_il_size(self._h, level)
        if sz < 0:
            raise ValueError("level out of range")
        off = np.zeros(8 ** level + 1, dtype=np.int32)
        idx = np.zeros(max(sz, 1), dtype=np.int32)
        nat.check(nat.lib().h3d_export_il(self._h, level, off.ctypes.data_as(ctypes.c_void_p),
                                          idx.ctypes.data_as(ctypes.c_void_p)))
        return off, idx[:sz]

    def ranks(self, level):
        sz = nat.lib().h3d_il_size(self._h, level)
        r = np.zeros(max(sz, 1), dtype=np.int32)
        s = np.zeros(max(sz, 1), dtype=np.int32)
        nat.check(nat.lib().h3d_export_ranks(self._h, level, r.ctypes.data_as(ctypes.c_void_p),
                                             s.ctypes.data_as(ctypes.c_void_p)))
        return r[:sz], s[:sz]

    def block(self, level, b, rank, n=0, m=0, uv=False):
        I = np.zeros(max(rank, 1), dtype=np.int32)
        J = np.zeros(max(rank, 1), dtype=np.int32)
        U = np.zeros((rank, n)) if uv else None
        V = np.zeros((rank, m)) if uv else None
        vp = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
        nat.check(nat.lib().h3d_export_block(self._h, level, int(b), vp(I), vp(J), vp(U), vp(V)))
        return I[:rank], J[:rank], U, V


class MatvecGraph:
    """One matvec (or transpose) of each handle, b[i] -> y[i], captured into a
    single HIP graph (h3d_graph_create); ``launch()`` replays all of them with
    one launch on torch's current stream.  b and y are CUDA float64 tensors of
    each handle's N, kept alive here; rewrite b in place between launches."""

    def __init__(self, handles, b, y, prec="fp64", transpose=False):
        import torch
        if not (len(handles) == len(b) == len(y)) or not handles:
            raise ValueError("handles, b and y must have the same non-zero length")
        for h, bi, yi in zip(handles, b, y):
            for t in (bi, yi):
                if not (_is_torch_cuda(t) and t.dtype == torch.float64 and t.is_contiguous()
                        and t.numel() == h.N):
                    raise ValueError("b and y must be contiguous CUDA float64 tensors of length N")
        self.handles, self.b, self.y = list(handles), list(b), list(y)
        n = len(handles)
        hs = (ctypes.c_void_p * n)(*[h._h.value for h in handles])
        bp = (ctypes.c_void_p * n)(*[t.data_ptr() for t in b])
        yp = (ctypes.c_void_p * n)(*[t.data_ptr() for t in y])
        g = ctypes.c_void_p()
        nat.check(nat.lib().h3d_graph_create(hs, n, bp, yp, _prec(prec), int(transpose),
                                             ctypes.byref(g)))
        self._g = g

    def launch(self):
        import torch
        nat.check(nat.lib().h3d_graph_launch(
            self._g, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def close(self):
        if getattr(self, "_g", None):
            nat.lib().h3d_graph_destroy(self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


PERSISTENT = 0x100  # h3d_batch_set_tuning flag: the persistent (work-queue) launch
MIXED_FP16 = 1      # h3d_set_mixed_formats bits
MIXED_FP8 = 2
MIXED_FP24 = 4
MIXED_INT8 = 8


class Batch:
    """One product of each of many assembled handles in one launch per step for
    all of them (h3d_batch_create: the instances' work tables merged, factors
    read in place).  Vectors are packed CUDA float64 tensors: instance i owns
    ``b[off[i]:off[i+1]]`` in its own point order (``offsets``; ``pack`` and
    ``unpack`` convert lists of per-instance tensors).  Products run on torch's
    current stream.  Keep the handles and their ``prec`` copy alive while the
    batch is used."""

    def __init__(self, handles, prec="fp32"):
        if not handles:
            raise ValueError("a batch needs at least one handle")
        self.handles = list(handles)
        n = len(self.handles)
        hs = (ctypes.c_void_p * n)(*[h._h.value for h in self.handles])
        bt = ctypes.c_void_p()
        nat.check(nat.lib().h3d_batch_create(hs, n, _prec(prec), ctypes.byref(bt)))
        self._b = bt
        off = np.zeros(n + 1, dtype=np.int64)
        nat.check(nat.lib().h3d_batch_layout(self._b, off.ctypes.data_as(ctypes.c_void_p)))
        self.offsets = off
        self.N = int(off[-1])
        self._graphs = []

    def _bind(self):
        import torch
        nat.check(nat.lib().h3d_batch_set_stream(
            self._b, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def _check(self, t):
        import torch
        if not (_is_torch_cuda(t) and t.dtype == torch.float64 and t.is_contiguous()
                and t.numel() == self.N):
            raise ValueError("packed vectors are contiguous CUDA float64 tensors of the batch's N")

    def pack(self, parts):
        import torch
        if len(parts) != len(self.handles):
            raise ValueError("one vector per instance")
        return torch.cat([p.reshape(-1).to(torch.float64) for p in parts]).contiguous()

    def unpack(self, y):
        o = self.offsets
        return [y[o[i]:o[i + 1]] for i in range(len(self.handles))]

    def matvec(self, b, out=None):
        """y = A_i b_i for every instance i (packed b -> packed y)."""
        import torch
        self._check(b)
        y = torch.empty_like(b) if out is None else out
        self._check(y)
        self._bind()
        nat.check(nat.lib().h3d_batch_matvec(self._b, ctypes.c_void_p(b.data_ptr()),
                                             ctypes.c_void_p(y.data_ptr())))
        return y

    def autotune(self, reps=3):
        """Fastest load width per phase over the merged tables, then the
        persistent launch at that width; ms per width and 'persistent' ms."""
        self._bind()
        ms = (ctypes.c_double * 10)(*([-1.0] * 10))
        nat.check(nat.lib().h3d_batch_autotune(self._b, int(reps), ms))
        w = [16, 8, 4, 2]
        return {"phase1": {w[i]: ms[i] for i in range(4) if ms[i] >= 0},
                "phase2": {w[i]: ms[4 + i] for i in range(4) if ms[4 + i] >= 0},
                "phase1_persistent": ms[8], "phase2_persistent": ms[9]}

    def get_tuning(self):
        """(phase-1, phase-2) load bytes per lane, | PERSISTENT for the persistent launch."""
        a, b = ctypes.c_int(), ctypes.c_int()
        nat.check(nat.lib().h3d_batch_get_tuning(self._b, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_tuning(self, p1_bytes, p2_bytes):
        nat.check(nat.lib().h3d_batch_set_tuning(self._b, int(p1_bytes), int(p2_bytes)))

    def graph(self, b, y):
        """Captures one product b -> y (both kept alive here) into a HIP graph;
        returns a callable that replays it on torch's current stream."""
        import torch
        self._check(b)
        self._check(y)
        self._bind()
        g = ctypes.c_void_p()
        nat.check(nat.lib().h3d_batch_graph_create(self._b, ctypes.c_void_p(b.data_ptr()),
                                                   ctypes.c_void_p(y.data_ptr()), ctypes.byref(g)))
        self._graphs.append((g, b, y))

        def launch():
            nat.check(nat.lib().h3d_graph_launch(
                g, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return launch

    def close(self):
        for g, _, _ in self._graphs:
            nat.lib().h3d_graph_destroy(g)
        self._graphs = []
        if getattr(self, "_b", None):
            nat.lib().h3d_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class classGMRES:
    """src/gmres.cpp:40-94 on the GPU: unrestarted GMRES, x0 = 0, MGS + Givens."""

    def gmres(self, M, b, maxIterations, threshold):
        a = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(a)
        err = ctypes.c_double()
        its = ctypes.c_int()
        hist = np.zeros(maxIterations + 1)
        nat.check(nat.lib().h3d_gmres(M._h, a.ctypes.data_as(ctypes.c_void_p),
                                      x.ctypes.data_as(ctypes.c_void_p), int(maxIterations),
                                      float(threshold), ctypes.byref(err), ctypes.byref(its),
                                      hist.ctypes.data_as(ctypes.c_void_p), nat.HOST))
        n = its.value + 1
        return x, err.value, its.value, hist[:n].tolist()
