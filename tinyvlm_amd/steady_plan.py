"""Steady solves of a batch of requests on the device: wrapper of tvlm_vlm_* (the C ABI's
batched form of VLM::run, vlm/src/vlm.cpp:97-151).

MODE_PLAIN_FP64 runs the requests one after another in exact FP64 (the drop-in path);
MODE_SERVED assembles all requests' FP64 matrices in the seeded form, factors them in one batched
FP32 LU and refines each solution in FP64 against its FP64 matrix.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import capi as H

MODE_PLAIN_FP64 = 0
MODE_SERVED = 1


class SteadyPlan:
    def __init__(self, verts_list, batch: int, mode: int = MODE_SERVED, stream=None):
        import torch

        self.dims = [(v.shape[0] - 1, v.shape[1] - 1) for v in verts_list]
        self.N = sum(ns * nc for ns, nc in self.dims)
        self.batch = int(batch)
        self.mode = int(mode)
        self.stream = stream if stream is not None else torch.cuda.Stream()
        self._verts = [np.asfortranarray(np.asarray(v, dtype=np.float64)) for v in verts_list]
        n = len(self._verts)
        ns = (ctypes.c_int64 * n)(*[d[0] for d in self.dims])
        nc = (ctypes.c_int64 * n)(*[d[1] for d in self.dims])
        ptrs = (ctypes.c_void_p * n)(*[v.ctypes.data for v in self._verts])
        self._plan = ctypes.c_void_p()
        H.check(H.lib().tvlm_vlm_create(ctypes.byref(self._plan), ctypes.c_int32(n), ns, nc, ptrs,
                                         ctypes.c_int32(self.batch), ctypes.c_int32(self.mode),
                                         ctypes.c_void_p(self.stream.cuda_stream)), "tvlm_vlm_create")

    def solve(self, alpha, beta=None, u_inf=1.0, rho=1.0, cm_ref=(0.25, 0.0, 0.0)):
        """Asynchronous on the plan's stream; alpha (radians) has one entry per request."""
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(alpha, dtype=np.float64), (self.batch,)))
        b = None if beta is None else np.ascontiguousarray(np.broadcast_to(np.asarray(beta, np.float64),
                                                                           (self.batch,)))
        ref = np.ascontiguousarray(cm_ref, dtype=np.float64)
        self._keep = (a, b, ref)
        H.check(H.lib().tvlm_vlm_solve(self._plan, ctypes.c_void_p(a.ctypes.data),
                                        ctypes.c_void_p(None if b is None else b.ctypes.data),
                                        ctypes.c_double(u_inf), ctypes.c_double(rho),
                                        ctypes.c_void_p(ref.ctypes.data)), "tvlm_vlm_solve")

    def results(self, gamma=True):
        """Synchronises; returns dict cl[batch], cm[batch,3], info[batch] and gamma[batch,N]."""
        cl = np.zeros(self.batch)
        cm = np.zeros((self.batch, 3))
        info = np.zeros(self.batch, dtype=np.int32)
        g = np.zeros((self.batch, self.N)) if gamma else None
        H.check(H.lib().tvlm_vlm_results(self._plan, ctypes.c_void_p(cl.ctypes.data),
                                          ctypes.c_void_p(cm.ctypes.data),
                                          ctypes.c_void_p(None if g is None else g.ctypes.data),
                                          ctypes.c_void_p(info.ctypes.data)), "tvlm_vlm_results")
        out = {"cl": cl, "cm": cm, "info": info}
        if g is not None:
            out["gamma"] = g
        return out

    def close(self):
        if self._plan:
            H.lib().tvlm_vlm_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
