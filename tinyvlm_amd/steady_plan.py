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


def _stream_handle(stream) -> int:
    """hipStream_t of a torch stream or of a raw handle (int)."""
    return int(stream) if isinstance(stream, int) else int(stream.cuda_stream)


class SteadyPlan:
    def __init__(self, verts_list, batch: int, mode: int = MODE_SERVED, stream=None):
        """stream: a torch stream or a raw hipStream_t (int, e.g. from tvlm_stream_create); a new
        torch stream when omitted."""
        if stream is None:
            import torch

            stream = torch.cuda.Stream()
        self.dims = [(v.shape[0] - 1, v.shape[1] - 1) for v in verts_list]
        self.N = sum(ns * nc for ns, nc in self.dims)
        self.batch = int(batch)
        self.mode = int(mode)
        self.stream = stream
        self._stream = _stream_handle(stream)
        self._verts = [np.asfortranarray(np.asarray(v, dtype=np.float64)) for v in verts_list]
        n = len(self._verts)
        ns = (ctypes.c_int64 * n)(*[d[0] for d in self.dims])
        nc = (ctypes.c_int64 * n)(*[d[1] for d in self.dims])
        ptrs = (ctypes.c_void_p * n)(*[v.ctypes.data for v in self._verts])
        self._plan = ctypes.c_void_p()
        self.active = self.batch
        H.check(H.lib().tvlm_vlm_create(ctypes.byref(self._plan), ctypes.c_int32(n), ns, nc, ptrs,
                                         ctypes.c_int32(self.batch), ctypes.c_int32(self.mode),
                                         ctypes.c_void_p(self._stream)), "tvlm_vlm_create")

    def solve(self, alpha, beta=None, u_inf=1.0, rho=1.0, cm_ref=(0.25, 0.0, 0.0)):
        """Asynchronous on the plan's stream; alpha (radians) has one entry per request, at most
        batch entries (a scalar fills the whole batch)."""
        al = np.asarray(alpha, dtype=np.float64)
        n = self.batch if al.ndim == 0 else len(al)
        if not 1 <= n <= self.batch:
            raise ValueError(f"{n} requests for a plan of {self.batch}")
        a = np.ascontiguousarray(np.broadcast_to(al, (n,)))
        b = None if beta is None else np.ascontiguousarray(np.broadcast_to(np.asarray(beta, np.float64), (n,)))
        ref = np.ascontiguousarray(cm_ref, dtype=np.float64)
        self._keep = (a, b, ref)
        self.active = n
        H.check(H.lib().tvlm_vlm_solve_n(self._plan, ctypes.c_int32(n), ctypes.c_void_p(a.ctypes.data),
                                          ctypes.c_void_p(None if b is None else b.ctypes.data),
                                          ctypes.c_double(u_inf), ctypes.c_double(rho),
                                          ctypes.c_void_p(ref.ctypes.data)), "tvlm_vlm_solve_n")

    def results(self, gamma=True):
        """Waits for the last solve; returns dict cl[n], cm[n,3], info[n] and gamma[n,N] for the n requests of
        the last solve."""
        n = self.active
        cl = np.zeros(n)
        cm = np.zeros((n, 3))
        info = np.zeros(n, dtype=np.int32)
        g = np.zeros((n, self.N)) if gamma else None
        H.check(H.lib().tvlm_vlm_results(self._plan, ctypes.c_void_p(cl.ctypes.data),
                                          ctypes.c_void_p(cm.ctypes.data),
                                          ctypes.c_void_p(None if g is None else g.ctypes.data),
                                          ctypes.c_void_p(info.ctypes.data)), "tvlm_vlm_results")
        out = {"cl": cl, "cm": cm, "info": info}
        if g is not None:
            out["gamma"] = g
        return out

    def refinement(self):
        """What the served refinement took for each request of the last solve (checks and rescues
        them first if results() has not): dict status[n] (bits: 1 more corrections, 2 FP64
        fallback, 4 re-factored after a stalled cooperative panel, 8 unresolved), corrections[n],
        backward_error[n]."""
        n = self.active
        st = np.zeros(n, dtype=np.int32)
        co = np.zeros(n, dtype=np.int32)
        be = np.zeros(n)
        H.check(H.lib().tvlm_vlm_refinement(self._plan, ctypes.c_void_p(st.ctypes.data), ctypes.c_void_p(co.ctypes.data),
                                            ctypes.c_void_p(be.ctypes.data)), "tvlm_vlm_refinement")
        return {"status": st, "corrections": co, "backward_error": be}

    def done(self) -> bool:
        """True once the plan's latest solve has finished on the device (does not block)."""
        d = ctypes.c_int32()
        H.check(H.lib().tvlm_vlm_query(self._plan, ctypes.byref(d)), "tvlm_vlm_query")
        return bool(d.value)

    def set_stream(self, stream):
        """Later solves queue on `stream` (torch stream or raw handle), after the plan's own
        latest work; the host does not block."""
        h = _stream_handle(stream)
        if h != self._stream:
            H.check(H.lib().tvlm_vlm_set_stream(self._plan, ctypes.c_void_p(h)), "tvlm_vlm_set_stream")
            self.stream, self._stream = stream, h

    def footprint(self) -> int:
        """Device bytes the plan holds."""
        b = ctypes.c_int64()
        H.check(H.lib().tvlm_vlm_footprint(self._plan, ctypes.byref(b)), "tvlm_vlm_footprint")
        return int(b.value)

    def close(self):
        if self._plan:
            H.lib().tvlm_vlm_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
