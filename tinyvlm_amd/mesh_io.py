"""plot3d meshes through the native reader (tvlm_plot3d_*, MeshIO of vlm/src/vlm_mesh.cpp:34-124)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import capi as H


def read_plot3d(path: str, quarter_chord: bool = True, parse_f64: bool = False) -> np.ndarray:
    """verts (ns+1, nc+1, 4) float64 Fortran order, w = 1: parsed (and quarter-chord shifted) in
    float32 as the VLM/UVLM drivers do, or in float64 as the harmonic-balance driver does."""
    L = H.lib()
    ns, nc = ctypes.c_int64(), ctypes.c_int64()
    p = path.encode()
    H.check(L.tvlm_plot3d_dims(p, ctypes.byref(ns), ctypes.byref(nc)), "tvlm_plot3d_dims")
    v = np.zeros((ns.value + 1, nc.value + 1, 4), dtype=np.float64, order="F")
    H.check(L.tvlm_plot3d_read(p, ctypes.c_int32(1 if quarter_chord else 0), ctypes.c_int32(1 if parse_f64 else 0),
                               ctypes.c_void_p(v.ctypes.data)), "tvlm_plot3d_read")
    return v
