"""MI355X-native HODLR3D engine: octree, ACA assembly, fast matvec and GMRES on the GPU.

The compute path is libhodlr3d.so (hand-written HIP for gfx950) behind the C ABI of
include/hodlr3d.h; this package is its host-side mirror of the reference API.
"""
from .hodlr3d import (HODLR3D, Batch, PERSISTENT, MatvecGraph, classGMRES, userkernel, pts3D, nlevels,  # noqa: F401
                      PRECISIONS, QCHOICE_PLUGIN, plugin_name)
from ._native import H3DError  # noqa: F401

__all__ = ["HODLR3D", "Batch", "PERSISTENT", "MatvecGraph", "classGMRES", "userkernel", "pts3D", "nlevels", "PRECISIONS",
           "QCHOICE_PLUGIN", "plugin_name", "H3DError"]
