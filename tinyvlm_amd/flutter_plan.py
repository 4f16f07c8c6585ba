"""Coupled pitch/heave flutter sweep on the device (tvlm_flutter_* in include/tinyvlm_hip.h): the
reference's UVLM_2DOF driver (tests/uvlm_2dof.cpp:428-865, COUPLED path) for many reduced velocities
of one wing at once, every time step and every fixed-point iteration on the device, the wake kept
as a ring buffer of its newest rows.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import capi as H

PI_F = float(np.float32(3.141593))


class FlutterParams(ctypes.Structure):
    """struct tvlm_flutter_params (UVLM_2DOF_Vars, tests/uvlm_2dof.cpp:23-39, without U_a)."""

    _fields_ = [("b", ctypes.c_double), ("rho", ctypes.c_double), ("a_h", ctypes.c_double),
                ("omega", ctypes.c_double), ("zeta_a", ctypes.c_double), ("zeta_h", ctypes.c_double),
                ("x_a", ctypes.c_double), ("mu", ctypes.c_double), ("r_a", ctypes.c_double),
                ("beta", ctypes.c_double * 4), ("init_pitch", ctypes.c_double)]


def params(b=0.5, rho=1.0, a_h=-0.5, omega=0.2, zeta_a=0.0, zeta_h=0.0, x_a=0.25, mu=100.0, r_a=0.5,
           beta=(0.0, 0.1, 0.0, 40.0), init_pitch_deg=3.0) -> FlutterParams:
    """Defaults are the reference main's values (tests/uvlm_2dof.cpp:884-898); the initial pitch is
    to_radians(init_pitch_deg) (vlm/include/vlm_utils.hpp:68-70) with the float constant PI_f, evaluated
    in FP64 (the plan's precision)."""
    p = FlutterParams()
    p.b, p.rho, p.a_h, p.omega = b, rho, a_h, omega
    p.zeta_a, p.zeta_h, p.x_a, p.mu, p.r_a = zeta_a, zeta_h, x_a, mu, r_a
    for k in range(4):
        p.beta[k] = beta[k]
    p.init_pitch = float(np.float64(init_pitch_deg) * np.float64(np.float32(3.141593)) / np.float64(180.0))
    return p


class FlutterPlan:
    """One wing (verts {ns+1, nc+1, 4}, as read_plot3d returns it), one reduced velocity per airspeed,
    a wake ring of `wake_rows` ring rows, room for `max_steps` steps."""

    def __init__(self, verts, U_a, prm: FlutterParams | None = None, wake_rows: int = 200, max_steps: int = 1000,
                 stream=None):
        import torch

        self.verts = np.asfortranarray(np.asarray(verts, dtype=np.float64))
        self.ns, self.nc = self.verts.shape[0] - 1, self.verts.shape[1] - 1
        self.N = self.ns * self.nc
        self.U_a = np.ascontiguousarray(np.atleast_1d(np.asarray(U_a, dtype=np.float64)))
        self.B = len(self.U_a)
        self.prm = prm if prm is not None else params()
        self.wake_rows, self.max_steps = int(wake_rows), int(max_steps)
        self.stream = stream if stream is not None else torch.cuda.Stream()
        self._plan = ctypes.c_void_p()
        H.check(H.lib().tvlm_flutter_create(ctypes.byref(self._plan), ctypes.c_int64(self.ns), ctypes.c_int64(self.nc),
                                            ctypes.c_void_p(self.verts.ctypes.data), ctypes.byref(self.prm),
                                            ctypes.c_int32(self.B), ctypes.c_void_p(self.U_a.ctypes.data),
                                            ctypes.c_int32(self.wake_rows), ctypes.c_int64(self.max_steps),
                                            H.stream_ptr(self.stream)), "tvlm_flutter_create")

    def run(self, steps: int):
        H.check(H.lib().tvlm_flutter_run(self._plan, ctypes.c_int64(int(steps))), "tvlm_flutter_run")

    def results(self) -> dict:
        """u, v, a: [steps + 1, B, 2] (h, alpha); F: [steps + 1, B, 2] (heave and pitch loads
        -CL/(pi mu), 2 CM_y/(pi mu r_a^2)); iters: [steps, B]; t: [steps + 1]; dt."""
        n = ctypes.c_int64()
        dt = ctypes.c_double()
        H.check(H.lib().tvlm_flutter_results(self._plan, ctypes.byref(n), ctypes.byref(dt), None, None),
                "tvlm_flutter_results")
        steps = int(n.value)
        st = np.zeros((steps + 1, self.B, 8))
        it = np.zeros((max(steps, 1), self.B), dtype=np.int32)
        H.check(H.lib().tvlm_flutter_results(self._plan, None, None, ctypes.c_void_p(st.ctypes.data),
                                             ctypes.c_void_p(it.ctypes.data)), "tvlm_flutter_results")
        return {"u": st[:, :, 0:2], "v": st[:, :, 2:4], "a": st[:, :, 4:6], "F": st[:, :, 6:8], "iters": it[:steps],
                "dt": dt.value, "t": np.arange(steps + 1) * dt.value}

    def cl(self) -> np.ndarray:
        """CL history [steps + 1, B] from the heave load F_h = -CL / (pi mu) (step 0: 0)."""
        return -self.results()["F"][:, :, 0] * (PI_F * self.prm.mu)

    def time_wake(self, launches: int = 5):
        """(ms per wake-induction launch of the next step, (receiver, wake ring) pairs per launch)."""
        ms = ctypes.c_double()
        pairs = ctypes.c_double()
        H.check(H.lib().tvlm_flutter_time_wake(self._plan, ctypes.c_int32(launches), ctypes.byref(ms),
                                               ctypes.byref(pairs)), "tvlm_flutter_time_wake")
        return ms.value / launches, pairs.value

    def close(self):
        if self._plan:
            H.lib().tvlm_flutter_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sweep_airspeeds(n: int = 32, lo: float = 0.6, hi: float = 1.2) -> np.ndarray:
    """Reduced velocities of the flutter sweep: n points evenly over [lo, hi], around the onset of the
    sweep's wing (between 0.7, where a 3-degree pitch decays, and 1.0, where it grows)."""
    return np.linspace(lo, hi, n)


def rectangular_wing(ns: int, nc: int, chord: float = 1.0, half_span: float = 5.0) -> np.ndarray:
    """Flat rectangular half wing, n x n uniform panels, as the plot3d reader would store it (float32
    parse, quarter-chord shift), widened to double."""
    import os
    import tempfile

    from .mesh_io import read_plot3d
    from .serve import write_grid_plot3d

    xs = np.linspace(0.0, chord, nc + 1)
    ys = np.linspace(0.0, half_span, ns + 1)
    x, y = np.meshgrid(xs, ys, indexing="ij")
    fd, path = tempfile.mkstemp(suffix=".x")
    os.close(fd)
    try:
        write_grid_plot3d(path, x, y)
        return read_plot3d(path)
    finally:
        os.remove(path)


__all__ = ["FlutterPlan", "FlutterParams", "params", "sweep_airspeeds", "rectangular_wing"]
