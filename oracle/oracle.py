"""ctypes binding of the CPU oracle (liboracle.so, built from h3d_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's ``cpu_baseline`` leg, where it is the checker (or the timed CPU
baseline) beside the HIP path.  The product package never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_d = ctypes.c_double


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        sig = {
            "orc_nlevels": (ctypes.c_int, [_i64, ctypes.c_int]),
            "orc_set_reduction": (None, [ctypes.c_int]),
            "orc_ie_cube_kii": (_d, [_i64]),
            "orc_gauss_legendre": (None, [ctypes.c_int, _p, _p]),
            "orc_create": (_p, [_p, _i64, ctypes.c_int, _d, _p, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, _d, _d, _p, ctypes.c_int]),
            "orc_destroy": (None, [_p]),
            "orc_get_nlevels": (ctypes.c_int, [_p]),
            "orc_get_perm": (None, [_p, _p]),
            "orc_get_boxes": (None, [_p, ctypes.c_int, _p, _p, _p]),
            "orc_get_neighbors": (None, [_p, ctypes.c_int, _p]),
            "orc_il_size": (_i64, [_p, ctypes.c_int]),
            "orc_get_il": (None, [_p, ctypes.c_int, _p, _p]),
            "orc_assemble": (_d, [_p]),
            "orc_assemble_boxes": (_d, [_p, ctypes.c_int, _i64, _p]),
            "orc_get_ranks": (None, [_p, ctypes.c_int, _p, _p]),
            "orc_get_block": (None, [_p, ctypes.c_int, _i64, _p, _p, _p, _p, _p, _p]),
            "orc_aca_positions": (ctypes.c_int, [_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, _p, _p, _p, _p]),
            "orc_matvec": (None, [_p, _p, _p, ctypes.c_int, _p]),
            "orc_timers": (None, [_p, _p, _p]),
            "orc_find_memory": (_d, [_p]),
            "orc_max_rank": (ctypes.c_int, [_p]),
            "orc_gmres": (ctypes.c_int, [_p, _p, ctypes.c_int, _d, _p, _p, _p, _p]),
            "orc_direct_rows": (None, [_p, _i64, ctypes.c_int, _d, _d, _p, _p, _i64, _p, _p]),
            "orc_entry": (_d, [ctypes.c_int, _d, _d, _p, _p, _i64, _i64]),
            "orc_max_threads": (ctypes.c_int, []),
            "orc_time_far_level": (_d, [_p, ctypes.c_int, _p]),
            "orc_time_near": (_d, [_p, _p, _p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def max_threads():
    return lib().orc_max_threads()


def nlevels(N, leaf1d):
    return lib().orc_nlevels(N, leaf1d)


def ie_cube_kii(N):
    return lib().orc_ie_cube_kii(N)


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, _ptr(x), _ptr(w))
    return x, w


def set_reduction(mode):
    """0: the canonical 256-lane order; 1: Eigen's 2-wide packet order."""
    lib().orc_set_reduction(mode)


def entry(q, a, b, i=0, j=1, scale=1.0, diag=0.0):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return lib().orc_entry(q, scale, diag, _ptr(a), _ptr(b), i, j)


def direct_rows(xyz, q, rows, qchoice=7, scale=1.0, diag=0.0, diag_vec=None):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    dv = None if diag_vec is None else np.ascontiguousarray(diag_vec, dtype=np.float64)
    out = np.zeros(len(rows))
    lib().orc_direct_rows(_ptr(xyz), len(xyz), qchoice, scale, diag, _ptr(dv), _ptr(q),
                          len(rows), _ptr(rows), _ptr(out))
    return out


class Oracle:
    """The reference facade restated on the CPU (src/HODLR3D.cpp:3-27)."""

    def __init__(self, xyz, leaf1d, L=1.0, tol_pow=7, qchoice=7, scale=1.0, diag=0.0,
                 diag_vec=None, center=None, nlevels=-1, keep_uv=False):
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.N = len(self.xyz)
        self._dv = None if diag_vec is None else np.ascontiguousarray(diag_vec, dtype=np.float64)
        c = None if center is None else np.ascontiguousarray(center, dtype=np.float64)
        self.h = lib().orc_create(_ptr(self.xyz), self.N, leaf1d, L, _ptr(c), nlevels, tol_pow,
                                  qchoice, scale, diag, _ptr(self._dv), int(keep_uv))
        if not self.h:
            raise ValueError("oracle: invalid input")
        self.nLevels = lib().orc_get_nlevels(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def perm(self):
        p = np.zeros(self.N, dtype=np.int32)
        lib().orc_get_perm(self.h, _ptr(p))
        return p

    def boxes(self, level):
        nb = 8 ** level
        s = np.zeros(nb, dtype=np.int32)
        c = np.zeros(nb, dtype=np.int32)
        ctr = np.zeros((nb, 3))
        lib().orc_get_boxes(self.h, level, _ptr(s), _ptr(c), _ptr(ctr))
        return s, c, ctr

    def neighbors(self, level):
        n = np.zeros((8 ** level, 26), dtype=np.int32)
        lib().orc_get_neighbors(self.h, level, _ptr(n))
        return n

    def il(self, level):
        sz = lib().orc_il_size(self.h, level)
        off = np.zeros(8 ** level + 1, dtype=np.int32)
        idx = np.zeros(max(sz, 1), dtype=np.int32)
        lib().orc_get_il(self.h, level, _ptr(off), _ptr(idx))
        return off, idx[:sz]

    def assemble(self):
        return lib().orc_assemble(self.h)

    def assemble_boxes(self, level, boxes):
        b = np.ascontiguousarray(boxes, dtype=np.int32)
        return lib().orc_assemble_boxes(self.h, level, len(b), _ptr(b))

    def ranks(self, level):
        n = lib().orc_il_size(self.h, level)
        r = np.zeros(n, dtype=np.int32)
        s = np.zeros(n, dtype=np.int32)
        lib().orc_get_ranks(self.h, level, _ptr(r), _ptr(s))
        return r, s

    def block(self, level, b, rank, n=0, m=0, uv=False):
        I = np.zeros(max(rank, 1), dtype=np.int32)
        J = np.zeros(max(rank, 1), dtype=np.int32)
        Lm = np.zeros((rank, rank))
        Rm = np.zeros((rank, rank))
        U = np.zeros((rank, n)) if uv else None
        V = np.zeros((rank, m)) if uv else None
        lib().orc_get_block(self.h, level, b, _ptr(I), _ptr(J), _ptr(Lm), _ptr(Rm), _ptr(U),
                            _ptr(V))
        return I[:rank], J[:rank], Lm, Rm, U, V

    def aca_positions(self, rs, n, cs, m):
        I = np.zeros(max(min(n, m), 1), dtype=np.int32)
        J = np.zeros(max(min(n, m), 1), dtype=np.int32)
        U = np.zeros(max(min(n, m), 1) * n)
        V = np.zeros(max(min(n, m), 1) * m)
        r = lib().orc_aca_positions(self.h, rs, n, cs, m, _ptr(I), _ptr(J), _ptr(U), _ptr(V))
        return r, I[:r], J[:r], U[: r * n].reshape(r, n), V[: r * m].reshape(r, m)

    def matvec(self, q, mode=0, leafmask=None):
        q = np.ascontiguousarray(q, dtype=np.float64)
        y = np.zeros(self.N)
        lm = None if leafmask is None else np.ascontiguousarray(leafmask, dtype=np.uint8)
        lib().orc_matvec(self.h, _ptr(q), _ptr(y), mode, _ptr(lm))
        return y

    def timers(self):
        a = ctypes.c_double()
        m = ctypes.c_double()
        lib().orc_timers(self.h, ctypes.byref(a), ctypes.byref(m))
        return a.value, m.value

    def time_far_level(self, level, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        return lib().orc_time_far_level(self.h, level, _ptr(q))

    def time_near(self, q, leafmask):
        q = np.ascontiguousarray(q, dtype=np.float64)
        lm = np.ascontiguousarray(leafmask, dtype=np.uint8)
        return lib().orc_time_near(self.h, _ptr(q), _ptr(lm))

    def find_memory(self):
        return lib().orc_find_memory(self.h)

    def max_rank(self):
        return lib().orc_max_rank(self.h)

    def gmres(self, b, maxit, thr):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.N)
        err = ctypes.c_double()
        its = ctypes.c_int()
        hist = np.zeros(maxit + 1)
        nh = lib().orc_gmres(self.h, _ptr(b), maxit, thr, _ptr(x), ctypes.byref(err),
                             ctypes.byref(its), _ptr(hist))
        return x, err.value, its.value, hist[:nh]
