"""Device-resident UVLM time loop (tvlm_uvlm_* in include/tinyvlm_hip.h) from Python.

The whole of UVLM::run (vlm/src/vlm.cpp:358-509) -- setup, every time step, CL/CM history --
runs inside libtinyvlm_hip.so on one stream; Python only passes the mesh and the motion.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import capi as H

MODE_PLAIN_FP64 = 0
MODE_FAST = 1


def time_step(verts_init, u):
    """dt = last panel chord of surface 0 at span station 0 / |u| (vlm.cpp:380-386)."""
    d = np.asarray(verts_init)[0, -1, :3] - np.asarray(verts_init)[0, -2, :3]
    return math.sqrt(float(np.dot(d, d))) / abs(u)


def motion_pitch_heave(n_surf, t_steps, dt, u, h_amp=0.0, h_phase=0.0, a_amp=0.0, a_phase=0.0, omega=0.0,
                       x_pivot=0.25):
    """Host-side evaluation through tvlm_motion_pitch_heave; returns (T, dT, fs, ref) arrays."""
    L = H.lib()
    T = np.zeros((t_steps, n_surf, 16))
    dT = np.zeros((t_steps, n_surf, 16))
    fs = np.zeros((t_steps, 3))
    ref = np.zeros((t_steps, 3))
    d = ctypes.c_double
    P = ctypes.c_void_p
    H.check(L.tvlm_motion_pitch_heave(d(u), d(h_amp), d(h_phase), d(a_amp), d(a_phase), d(omega), d(x_pivot),
                                      ctypes.c_int32(n_surf), ctypes.c_int64(t_steps), d(dt), P(T.ctypes.data),
                                      P(dT.ctypes.data), P(fs.ctypes.data), P(ref.ctypes.data)),
            "tvlm_motion_pitch_heave")
    return T, dT, fs, ref


class UVLMPlan:
    """One reusable device-resident simulation (buffers, LU, wake, optional captured hipGraph).

    sims > 1 (fast mode): that many simulations of the mesh, each with its own motion, advanced by
    the same launches (tvlm_uvlm_create_batched); ``sim`` arguments pick one, None means all."""

    def __init__(self, verts_init_list, t_steps, dt, mode=MODE_FAST, stream=None, sims=1):
        import torch

        self.sims = int(sims)
        self.n_surf = len(verts_init_list)
        self.t_steps = int(t_steps)
        self.dt = float(dt)
        self.mode = mode
        # a torch stream or a raw hipStream_t (int); the plan's graph is captured on it, and the
        # library's scratch buffers it uses belong to it, so a plan stays on its stream
        self.stream = stream if stream is not None else torch.cuda.Stream()
        self.N = sum((v.shape[0] - 1) * (v.shape[1] - 1) for v in verts_init_list)
        self._verts = [np.asfortranarray(np.asarray(v, dtype=np.float64)) for v in verts_init_list]
        ns = (ctypes.c_int64 * self.n_surf)(*[v.shape[0] - 1 for v in self._verts])
        nc = (ctypes.c_int64 * self.n_surf)(*[v.shape[1] - 1 for v in self._verts])
        ptrs = (ctypes.c_void_p * self.n_surf)(*[v.ctypes.data for v in self._verts])
        self._plan = ctypes.c_void_p()
        H.check(H.lib().tvlm_uvlm_create_batched(ctypes.byref(self._plan), ctypes.c_int32(self.n_surf), ns, nc, ptrs,
                                                 ctypes.c_int64(self.t_steps), ctypes.c_double(self.dt),
                                                 ctypes.c_int32(mode), ctypes.c_int32(self.sims),
                                                 H.stream_ptr(self.stream)), "tvlm_uvlm_create_batched")

    def set_far_wake(self, levels):
        """Far-wake aggregation of the fast mode: levels = [(age, row_group, span_group), ...] (None or
        [] for the exact wake); see tvlm_uvlm_set_far_wake."""
        lv = [] if not levels else [tuple(int(x) for x in t) for t in levels]
        arr = (ctypes.c_int32 * max(1, 3 * len(lv)))(*[x for t in lv for x in t])
        H.check(H.lib().tvlm_uvlm_set_far_wake(self._plan, ctypes.c_int32(len(lv)), arr), "tvlm_uvlm_set_far_wake")
        self.far_wake = lv

    def set_lattice_target(self, groups):
        """Work-group target of the fast mode's wake launches (tvlm_uvlm_set_lattice_target)."""
        H.check(H.lib().tvlm_uvlm_set_lattice_target(self._plan, ctypes.c_int32(int(groups))),
                "tvlm_uvlm_set_lattice_target")
        self.lattice_target = int(groups)

    def set_pitch_heave(self, u, h_amp=0.0, h_phase=0.0, a_amp=0.0, a_phase=0.0, omega=0.0, x_pivot=0.25,
                        sim=None):
        d = ctypes.c_double
        z = -1 if sim is None else int(sim)
        H.check(H.lib().tvlm_uvlm_set_pitch_heave_sim(self._plan, ctypes.c_int32(z), d(u), d(h_amp), d(h_phase),
                                                      d(a_amp), d(a_phase), d(omega), d(x_pivot)),
                "tvlm_uvlm_set_pitch_heave_sim")

    def set_motion(self, T, dT, fs, ref, sim=None):
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (T, dT, fs, ref)]
        ptrs = [ctypes.c_void_p(a.ctypes.data) for a in arrs]
        if sim is None:
            H.check(H.lib().tvlm_uvlm_set_motion(self._plan, *ptrs), "tvlm_uvlm_set_motion")
        else:
            H.check(H.lib().tvlm_uvlm_set_motion_sim(self._plan, ctypes.c_int32(int(sim)), *ptrs),
                    "tvlm_uvlm_set_motion_sim")

    def run(self, use_graph=None):
        """One whole simulation, asynchronous.  use_graph (default: the served mode's hipGraph replay;
        plain mode always runs call by call)."""
        if use_graph is None:
            use_graph = self.mode == MODE_FAST
        H.check(H.lib().tvlm_uvlm_run(self._plan, ctypes.c_int32(int(use_graph))), "tvlm_uvlm_run")

    def results(self, sim=0):
        n = self.t_steps - 1
        cl = np.zeros(n)
        cm = np.zeros((n, 3))
        g = np.zeros(self.N)
        info = ctypes.c_int32(0)
        H.check(H.lib().tvlm_uvlm_results_sim(self._plan, ctypes.c_int32(int(sim)), ctypes.c_void_p(cl.ctypes.data),
                                              ctypes.c_void_p(cm.ctypes.data), ctypes.c_void_p(g.ctypes.data),
                                              ctypes.byref(info)), "tvlm_uvlm_results_sim")
        if info.value != 0:
            raise H.TvlmError(f"singular influence matrix (getrf info {info.value})")
        return {"cl": cl, "cm": cm, "gamma": g}

    def time_wake(self):
        """Instrumented run: (summed wake-kernel ms, launches, receiver x wake-ring pairs)."""
        ms = ctypes.c_double()
        n = ctypes.c_int32()
        pairs = ctypes.c_double()
        H.check(H.lib().tvlm_uvlm_time_wake(self._plan, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(pairs)),
                "tvlm_uvlm_time_wake")
        return ms.value, n.value, pairs.value

    def done(self) -> bool:
        """True once everything queued on the plan's stream has finished (does not block)."""
        d = ctypes.c_int32()
        H.check(H.lib().tvlm_stream_query(H.stream_ptr(self.stream), ctypes.byref(d)), "tvlm_stream_query")
        return bool(d.value)

    def close(self):
        if self._plan:
            H.lib().tvlm_uvlm_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
