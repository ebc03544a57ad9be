"""ctypes binding of libhodlr3d.so (the C ABI declared in include/hodlr3d.h).

The library is built in-tree by ``hodlr3d_amd/csrc/Makefile``; importing this
module fails loudly when it is missing -- there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhodlr3d.so")

H3D_OK = 0
H3D_ERR_INVALID = 1
H3D_ERR_OOM = 2
H3D_ERR_NONFINITE = 3
H3D_ERR_HIP = 4
H3D_ERR_STATE = 5
H3D_ERR_NOT_CONVERGED = 6
H3D_ERR_INTERNAL = 7
H3D_ERR_IO = 8

QCHOICE_PLUGIN = 100
PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16 = 2
PREC_MIXED = 3
ACC_FP64 = 0x100
HOST = 0
DEVICE = 1

_p = ctypes.c_void_p
_i32 = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double


class Kernel(ctypes.Structure):
    _fields_ = [("qchoice", _i32), ("scale", _d), ("diag", _d), ("diag_vec", _p)]


class Options(ctypes.Structure):
    _fields_ = [("center", _d * 3), ("nlevels", _i32), ("workspace_bytes", _i64),
                ("store_mask", _i32)]


class Stats(ctypes.Structure):
    _fields_ = [("N", _i64), ("nlevels", _i32), ("max_rank", _i32), ("n_blocks", _i64),
                ("rank_sum", _i64), ("factor_elems", _i64), ("find_memory", _d),
                ("near_pairs", _i64), ("t_tree_ms", _d), ("t_assemble_ms", _d),
                ("factor_bytes", _i64 * 3), ("u_elems", _i64), ("v_elems", _i64),
                ("ref_ass_time", _d), ("ref_matvec_time", _d)]


class MixedStats(ctypes.Structure):
    _fields_ = [("tol", _d), ("predicted_rel_err", _d), ("err_drop", _d), ("err_fp16", _d),
                ("err_fp32", _d), ("fro_far", _d), ("fro_near", _d),
                ("aca_rank_sum", _i64), ("rank_fp32", _i64), ("rank_fp16", _i64),
                ("dropped", _i64), ("raw_blocks", _i64), ("raw_rank", _i64), ("bytes", _i64),
                ("u_elems", _i64 * 2), ("v_elems", _i64 * 2), ("build_ms", _d),
                ("single_jacobi", _i32), ("formats", _i32), ("err_fp8", _d),
                ("rank_fp8", _i64), ("u_elems_fp8", _i64), ("v_elems_fp8", _i64),
                ("err_fp24", _d), ("rank_fp24", _i64), ("u_elems_fp24", _i64),
                ("v_elems_fp24", _i64), ("err_fp8_model", _d)]


class Config(ctypes.Structure):
    _fields_ = [("N", _i64), ("nlevels", _i32), ("leaf1d", _i32), ("L", _d), ("center", _d * 3),
                ("tol_pow", _i32), ("qchoice", _i32), ("kernel_scale", _d), ("kernel_diag", _d),
                ("alpha", _d), ("has_diag_vec", _i32)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("outer_iterations", _i32), ("inner_iterations", _i32), ("outer_prec", _i32),
                ("rel_residual", _d),
                ("history", _d * 64)]


# name -> (restype, argtypes); the list is also the export contract tested on CPU
SIGNATURES = {
    "h3d_nlevels": (_i32, [_i64, _i32]),
    "h3d_create": (_i32, [_p, _i64, _i32, _d, _i32, ctypes.POINTER(Kernel),
                          ctypes.POINTER(Options), ctypes.POINTER(_p)]),
    "h3d_assemble": (_i32, [_p]),
    "h3d_save": (_i32, [_p, ctypes.c_char_p]),
    "h3d_load": (_i32, [ctypes.c_char_p, ctypes.POINTER(Options), ctypes.POINTER(_p)]),
    "h3d_get_config": (_i32, [_p, ctypes.POINTER(Config)]),
    "h3d_prepare_precision": (_i32, [_p, _i32]),
    "h3d_release_precision": (_i32, [_p, _i32]),
    "h3d_move_precision": (_i32, [_p, _i32]),
    "h3d_set_operator": (_i32, [_p, _d, _d]),
    "h3d_kernel_timing": (_i32, [_p, _i32, _p, ctypes.POINTER(_i32)]),
    "h3d_autotune": (_i32, [_p, _i32, _i32, _p]),
    "h3d_set_tuning": (_i32, [_p, _i32, _i32, _i32]),
    "h3d_get_tuning": (_i32, [_p, _i32, _p]),
    "h3d_matvec_noalloc": (_i32, [_p, _p, _p, _i32]),
    "h3d_debug_fail_alloc": (_i32, [_i64]),
    "h3d_get_matrix": (_i32, [_p, _p, _i64, _p, _i64, _p, _i32]),
    "h3d_plugin_name": (ctypes.c_char_p, []),
    "h3d_set_tuning_state": (_i32, [_p, _i32, _p]),
    "h3d_set_rhs_block": (_i32, [_p, _i32]),
    "h3d_autotune_rhs": (_i32, [_p, _i32, _i32, _i32, _p, _p]),
    "h3d_compress": (_i32, [_p, _d, ctypes.POINTER(MixedStats)]),
    "h3d_get_mixed_stats": (_i32, [_p, ctypes.POINTER(MixedStats)]),
    "h3d_set_mixed_formats": (_i32, [_p, _i32]),
    "h3d_matvec": (_i32, [_p, _p, _p, _i32, _i32, _i32]),
    "h3d_matvec_transpose": (_i32, [_p, _p, _p, _i32, _i32, _i32]),
    "h3d_set_diag": (_i32, [_p, _p, _i32]),
    "h3d_gmres": (_i32, [_p, _p, _p, _i32, _d, ctypes.POINTER(_d), ctypes.POINTER(_i32), _p,
                         _i32]),
    "h3d_solve_ir": (_i32, [_p, _p, _p, _i32, _i32, _d, _i32, _d, _i32, _i32,
                            ctypes.POINTER(SolveInfo), _i32]),
    "h3d_direct": (_i32, [_p, _p, _p, _i64, _p, _i32]),
    "h3d_export_perm": (_i32, [_p, _p]),
    "h3d_export_boxes": (_i32, [_p, _i32, _p, _p]),
    "h3d_export_neighbors": (_i32, [_p, _i32, _p]),
    "h3d_il_size": (_i64, [_p, _i32]),
    "h3d_export_il": (_i32, [_p, _i32, _p, _p]),
    "h3d_export_ranks": (_i32, [_p, _i32, _p, _p]),
    "h3d_export_block": (_i32, [_p, _i32, _i64, _p, _p, _p, _p]),
    "h3d_get_stats": (_i32, [_p, ctypes.POINTER(Stats)]),
    "h3d_graph_create": (_i32, [ctypes.POINTER(_p), _i32, ctypes.POINTER(_p), ctypes.POINTER(_p),
                                _i32, _i32, ctypes.POINTER(_p)]),
    "h3d_graph_launch": (_i32, [_p, _p]),
    "h3d_graph_destroy": (None, [_p]),
    "h3d_batch_create": (_i32, [ctypes.POINTER(_p), _i32, _i32, ctypes.POINTER(_p)]),
    "h3d_batch_layout": (_i32, [_p, _p]),
    "h3d_batch_matvec": (_i32, [_p, _p, _p]),
    "h3d_batch_set_stream": (_i32, [_p, _p]),
    "h3d_batch_synchronize": (_i32, [_p]),
    "h3d_batch_autotune": (_i32, [_p, _i32, _p]),
    "h3d_batch_set_tuning": (_i32, [_p, _i32, _i32]),
    "h3d_batch_get_tuning": (_i32, [_p, ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "h3d_batch_graph_create": (_i32, [_p, _p, _p, ctypes.POINTER(_p)]),
    "h3d_batch_destroy": (None, [_p]),
    "h3d_set_stream": (_i32, [_p, _p]),
    "h3d_use_own_stream": (_i32, [_p]),
    "h3d_synchronize": (_i32, [_p]),
    "h3d_destroy": (None, [_p]),
    "h3d_status_string": (ctypes.c_char_p, [_i32]),
    "h3d_last_error": (ctypes.c_char_p, []),
}

_lib = None


class H3DError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hodlr3d: {msg} (status {code})")
        self.code = code


def _prefer_torch_runtime():
    """PyTorch-ROCm ships its own libamdhip64.so.7.  When torch is present it must
    be loaded first, so that this library binds to the same HIP runtime (one
    runtime per process); otherwise the system runtime from /opt/rocm is used."""
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def lib():
    global _lib
    if _lib is None:
        _prefer_torch_runtime()
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `make -C hodlr3d_amd/csrc` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != H3D_OK:
        L = lib()
        raise H3DError(code, (L.h3d_last_error() or b"").decode() or
                       L.h3d_status_string(code).decode())
