"""Coupled pitch/heave aeroelastic UVLM over the HIP C ABI: the reference's UVLM_2DOF driver
(tests/uvlm_2dof.cpp:48-865, COUPLED path) with its Newmark-beta integrator
(vlm/src/vlm_integrator.cpp:1-68).

A rigid two-surface section (a non-lifting wing and a lifting flap that continues it) pitches about
the elastic axis 1 + a_h and heaves; the structure is M u'' + C u' = F_aero - [(w/U)^2 h,
f(alpha)/U^2] with the cubic pitch spring f.  Each time step runs Newton iterations on the
structural increment: a Newmark step from the current load increment, the pose and its velocity
from the new state, the unsteady VLM solve with the rigid influence matrix factored once, the
loads, the new increment; until |du_k - du| <= EPS_f (at most 50).  The aerodynamics run on the
device in the chosen precision (FP32 as the reference, or FP64); the 2 x 2 structural algebra and
the kinematics are evaluated on the host.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from . import capi as H
from .mesh_io import read_plot3d

PI_F = np.float32(3.141593)
EPS_F = np.float32(1.19209e-07)


@dataclass
class Vars:
    """UVLM_2DOF_Vars (tests/uvlm_2dof.cpp:23-39); defaults are the reference main's values
    (tests/uvlm_2dof.cpp:884-898)."""
    b: float = 0.5          # half chord
    k_a: float = 1000.0     # pitch spring stiffness
    rho: float = 1.0
    a_h: float = -0.5       # elastic axis from mid-chord, half chords
    omega: float = 0.2      # heave / pitch natural frequency ratio
    zeta_a: float = 0.0
    zeta_h: float = 0.0
    x_a: float = 0.25       # elastic axis to centre of mass, half chords
    mu: float = 100.0       # mass ratio
    r_a: float = 0.5        # radius of gyration about the elastic axis
    U_a: float = 7.0        # reduced velocity
    beta: tuple = (0.0, 0.1, 0.0, 40.0)  # torsional_func's coefficients (tests/uvlm_2dof.cpp:285)
    init_pitch_deg: float = 3.0  # initial pitch (tests/uvlm_2dof.cpp:512)


def torsional_func(alpha, beta0=0.0, beta1=0.1, beta2=0.0, beta3=40.0):
    """Cubic pitch spring (tests/uvlm_2dof.cpp:285-287)."""
    return beta0 + beta1 * alpha + beta2 * alpha * alpha + beta3 * alpha * alpha * alpha


class NewmarkBeta:
    """Incremental Newmark-beta (vlm/src/vlm_integrator.cpp:6-68): K_eff = K + M/(beta dt^2) +
    gamma/(beta dt) C; K_eff du = dF + M (v/(beta dt) + a/(2 beta)) + C (gamma/beta v -
    dt (1 - gamma/(2 beta)) a); dv, da from du.  2 x 2 systems, evaluated on the host."""

    def __init__(self, M, Cm, K, dt, beta=0.25, gamma=0.5, dtype=np.float32):
        self.beta, self.gamma, self.dt, self.T = beta, gamma, dt, dtype
        self.x1 = dtype(gamma / (beta * dt))
        self.x0 = dtype(1.0 / (beta * dt * dt))
        self.K_eff = (K + self.x0 * M + self.x1 * Cm).astype(dtype)
        self.xd0, self.xd1 = dtype(1.0 / (beta * dt)), dtype(gamma / beta)
        self.xdd0, self.xdd1 = dtype(1.0 / (2.0 * beta)), dtype(-dt * (1.0 - gamma / (2.0 * beta)))

    def step(self, M, Cm, v_i, a_i, dF):
        T = self.T
        rhs = dF + M @ (self.xd0 * v_i + self.xdd0 * a_i) + Cm @ (self.xd1 * v_i + self.xdd1 * a_i)
        du = np.linalg.solve(self.K_eff.astype(np.float64), rhs.astype(np.float64)).astype(T)
        dv = (self.x1 * du - self.xd1 * v_i - self.xdd1 * a_i).astype(T)
        da = (self.x0 * du - self.xd0 * v_i - self.xdd0 * a_i).astype(T)
        return du, dv, da


def _rotation_y(px, ang, dang):
    """Rotation about the +y axis through (px, 0, 0) by ang, and its time derivative for ang' =
    dang (rotation_matrix, vlm/include/vlm_kinematics.hpp:184-202, with dual angle)."""
    c, s = math.cos(ang), math.sin(ang)
    R = np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])
    dR = np.array([[-s, 0.0, c], [0.0, 0.0, 0.0], [-c, 0.0, -s]]) * dang
    p = np.array([px, 0.0, 0.0])
    M = np.eye(4)
    M[:3, :3] = R
    M[:3, 3] = p - R @ p
    dM = np.zeros((4, 4))
    dM[:3, :3] = dR
    dM[:3, 3] = -dR @ p
    return M, dM


def section_transform(t, h, hd, al, ald, a_h):
    """Pose of the section and its time derivative: freestream T(-t, 0, 0), heave T(0, 0, -h) with
    rate h', pitch about (1 + a_h, 0, 0) by alpha with rate alpha' (tests/uvlm_2dof.cpp:645-680)."""
    F = np.eye(4)
    F[0, 3] = -t
    dF = np.zeros((4, 4))
    dF[0, 3] = -1.0
    Hm = np.eye(4)
    Hm[2, 3] = -h
    dH = np.zeros((4, 4))
    dH[2, 3] = -hd
    P, dP = _rotation_y(1.0 + a_h, al, ald)
    return F @ Hm @ P, dF @ Hm @ P + F @ dH @ P + F @ Hm @ dP


class UVLM2DOF:
    """UVLM_2DOF (tests/uvlm_2dof.cpp:48-128): meshes = [(plot3d path, lifting), ...] read like
    Simulation (FP32 parse, quarter-chord shift, vlm/src/vlm.cpp:17-45)."""

    def __init__(self, backend_name: str, meshes, dtype=np.float32):
        """meshes: [(plot3d path or verts array as read_plot3d returns it, lifting), ...]."""
        if backend_name != "hip":
            raise RuntimeError("Unsupported backend: " + str(backend_name))
        import torch

        H.init(0)
        self.torch = torch
        self.T = np.dtype(dtype).type
        self.tdt = torch.float32 if self.T == np.float32 else torch.float64
        self.lifting = [bool(l) for _, l in meshes]
        self.verts_file = [np.asarray(p, dtype=np.float64) if isinstance(p, np.ndarray)
                           else read_plot3d(p, quarter_chord=True, parse_f64=False) for p, _ in meshes]
        self.dims = [(v.shape[0] - 1, v.shape[1] - 1) for v in self.verts_file]
        self.N = sum(ns * nc for ns, nc in self.dims)

    def run(self, vars: Vars, t_final_nd: float, max_steps: int | None = None, wake_rows: int | None = None,
            step_times: list | None = None):
        """UVLM_2DOF::run (tests/uvlm_2dof.cpp:428-865).  Returns dict t, u (2, n), v, a, F (loads
        [-CL/(pi mu), 2 CM_y/(pi mu r_a^2)]), iters (Newton iterations per step).  wake_rows: only the
        newest wake_rows ring rows induce (truncated wake; None: the whole wake).  step_times: if a list,
        the wall time of every step (seconds) is appended to it."""
        torch, T, tdt = self.torch, self.T, self.tdt
        z = lambda *s: H.colmajor(s, tdt)  # noqa: E731
        b = T(vars.b)
        verts_init_h = []
        for v in self.verts_file:
            w = np.asfortranarray(v.astype(T))
            w[:, :, :3] = (w[:, :, :3] / b).astype(T)  # nondimensionalize_verts (tests/uvlm_2dof.cpp:417-425)
            verts_init_h.append(w)
        verts_init = [H.to_device(w, tdt) for w in verts_init_h]
        dims, N = self.dims, self.N
        verts = [z(ns + 1, nc + 1, 4) for ns, nc in dims]
        colloc = [z(ns, nc, 3) for ns, nc in dims]
        normals = [z(ns, nc, 3) for ns, nc in dims]
        areas = [z(ns, nc) for ns, nc in dims]
        vel = [z(ns, nc, 3) for ns, nc in dims]
        forces = [z(ns, nc, 3) for ns, nc in dims]
        gdelta = [z(ns, nc) for ns, nc in dims]
        gdt = [z(ns, nc) for ns, nc in dims]
        gprev = [z(ns, nc) for ns, nc in dims]
        scal = torch.zeros(6 * len(dims), dtype=tdt, device="cuda")
        a_h = float(T(vars.a_h))
        init_angle = float(T(vars.init_pitch_deg) * PI_F / T(180.0))  # to_radians(3.f)
        T0, _ = _rotation_y(1.0 + a_h, init_angle, 0.0)
        T0d = H.to_device(np.asfortranarray(T0), tdt)
        H.displace_wing([T0d] * len(dims), verts, verts_init)
        H.mesh_metrics(0.0, verts, colloc, normals, areas)
        colloc_h = [c.clone() for c in colloc]
        # time step: the last surface's last chordwise panel (tests/uvlm_2dof.cpp:474-481)
        last = verts_init_h[-1]
        d = last[0, -1, :3] - last[0, -2, :3]
        dt_nd = T(np.sqrt(T(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])))
        t_steps = int(T(t_final_nd) / dt_nd)
        n_loop = t_steps - 2 if max_steps is None else min(t_steps - 2, max_steps)
        wake = [z(ns + 1, t_steps, 4) for ns, nc in dims]
        gamma_wake = [z(ns, t_steps - 1) for ns, nc in dims]
        # structure (tests/uvlm_2dof.cpp:334-412): linear stiffness on the load side
        M = np.array([[1.0, vars.x_a], [vars.x_a / vars.r_a ** 2, 1.0]], dtype=T)
        Cm = np.array([[2.0 * vars.zeta_h * (vars.omega / vars.U_a), 0.0], [0.0, 2.0 * vars.zeta_a / vars.U_a]],
                      dtype=T)
        K = np.zeros((2, 2), dtype=T)
        integ = NewmarkBeta(M, Cm, K, float(dt_nd), dtype=T)
        n = n_loop + 1
        u = np.zeros((2, n + 1), dtype=T)
        v = np.zeros((2, n + 1), dtype=T)
        acc = np.zeros((2, n + 1), dtype=T)
        F = np.zeros((2, n + 1), dtype=T)
        u[1, 0] = T(init_angle)
        kh = T((vars.omega / vars.U_a) ** 2)
        ka = T(1.0 / (vars.U_a * vars.U_a))
        tf = lambda al: torsional_func(al, *vars.beta)  # noqa: E731
        a0 = np.array([-kh * u[0, 0], -ka * T(tf(u[1, 0]))], dtype=T) - Cm @ v[:, 0] - K @ u[:, 0]
        acc[:, 0] = np.linalg.solve(M.astype(np.float64), a0.astype(np.float64)).astype(T)
        # rigid motion of all surfaces together leaves the influence matrix unchanged: factor once
        lhs = z(N, N)
        H.lhs_assemble(lhs, colloc, normals, verts, wake, 0)
        ipiv = torch.zeros(N, dtype=torch.int32, device="cuda")
        info = torch.zeros(1, dtype=torch.int32, device="cuda")
        H.getrf(lhs, ipiv, info)
        if int(info.item()) != 0:
            raise H.TvlmError(f"singular influence matrix (getrf info {int(info.item())})")
        rhs = z(N)
        offs = np.cumsum([0] + [ns * nc for ns, nc in dims])
        gam = [rhs[offs[m]: offs[m + 1]].view(nc, ns).t() for m, (ns, nc) in enumerate(dims)]
        inv_dt = float(T(1.0) / dt_nd)
        # per Newton iteration: the section transform goes up through pinned host memory without
        # blocking, and the surfaces' areas and MACs (host arguments of the coefficients) come back
        # through pinned memory behind an event while the solve runs: one full wait per iteration
        Twd = z(4, 4)
        tw_host = torch.empty((4, 4), dtype=tdt, pin_memory=True).t()  # column-major like Twd
        am_host = torch.empty(2 * len(dims), dtype=tdt, pin_memory=True)
        am_ready = torch.cuda.Event()
        scal_slices: dict = {}

        def S(a, b):
            """scal[a:b], the same tensor at every call (its view is built once)."""
            t = scal_slices.get((a, b))
            if t is None:
                t = scal_slices[(a, b)] = scal[a:b]
            return t

        def solve(dT, it):
            H.kinematic_velocities([dT] * len(dims), colloc_h, vel)
            rhs.zero_()
            H.rhs_assemble_velocities(rhs, normals, vel)
            if wake_rows is None or it <= wake_rows:
                H.rhs_assemble_wake_influence(rhs, gamma_wake, colloc, normals, wake, self.lifting, it)
            else:
                # the newest wake_rows ring rows, as views of R+1 vertex rows and R circulation columns
                r0 = t_steps - 1 - it
                H.rhs_assemble_wake_influence(rhs, [g[:, r0: r0 + wake_rows] for g in gamma_wake], colloc, normals,
                                              [w[:, r0: r0 + wake_rows + 1] for w in wake], self.lifting, wake_rows)
            H.getrs(lhs, ipiv, rhs)

        # the slices gamma_post works on, made once (the same tensors at every call keep their views)
        te = [gam[m][:, nc - 1] for m, (ns, nc) in enumerate(dims)]
        g_head = [gam[m][:, : nc - 1] if nc > 1 else None for m, (ns, nc) in enumerate(dims)]
        gd_tail = [gdelta[m][:, 1:] if nc > 1 else None for m, (ns, nc) in enumerate(dims)]
        gd_first = [gdelta[m][:, 0] for m in range(len(dims))]
        gdt_flat = [gdt[m].t().reshape(-1) for m in range(len(dims))]

        def gamma_post(it, with_dt):
            # update_wake_and_gamma (tests/uvlm_2dof.cpp:289-331, 712-755)
            for m, (ns, nc) in enumerate(dims):
                g = gam[m]
                gamma_wake[m][:, t_steps - 2 - it].copy_(te[m])
                gdelta[m].copy_(g)
                if m == 1:  # the flap's first row continues the wing's trailing edge
                    H.axpy(-1.0, te[0], gd_first[m])
                if nc > 1:
                    H.axpy(-1.0, g_head[m], gd_tail[m])
                if with_dt:
                    gdt[m].copy_(g)
                    H.axpy(-1.0, gprev[m], gdt[m])
                    H.scal(inv_dt, gdt_flat[m])

        # t = 0 (tests/uvlm_2dof.cpp:547-559): freestream only
        H.wake_shed(verts, wake, 0)
        dF0 = np.zeros((4, 4))
        dF0[0, 3] = -1.0
        solve(dF0, 0)
        gamma_post(0, False)
        fs = (1.0, 0.0, 0.0)
        la = (0.0, 0.0, 1.0)
        iters = []
        for i in range(n_loop):
            t_step = time.perf_counter()
            t_nd = T(i) * dt_nd
            for m in range(len(dims)):
                gprev[m].copy_(gam[m])
            dF = np.zeros(2, dtype=T)
            du = np.zeros(2, dtype=T)
            du_k = np.ones(2, dtype=T)
            it = 0
            while float(np.sqrt(np.dot(du_k.astype(np.float64), du_k.astype(np.float64)))) > float(EPS_F):
                du_k = du.copy()
                du, dv, da = integ.step(M, Cm, v[:, i], acc[:, i], dF)
                u[:, i + 1] = u[:, i] + du
                v[:, i + 1] = v[:, i] + dv
                acc[:, i + 1] = acc[:, i] + da
                tt = float(t_nd + dt_nd)
                Tw, dTw = section_transform(tt, float(u[0, i + 1]), float(v[0, i + 1]), float(u[1, i + 1]),
                                            float(v[1, i + 1]), a_h)
                ref = (Tw @ np.array([1.0 + a_h, 0.0, 0.0, 1.0]))[:3]
                tw_host.copy_(torch.from_numpy(np.asarray(Tw)))
                Twd.copy_(tw_host, non_blocking=True)
                H.displace_wing([Twd] * len(dims), verts, verts_init)
                H.mesh_metrics(0.0, verts, colloc, normals, areas)
                for m in range(len(dims)):
                    H.sum_(areas[m], S(2 * m, 2 * m + 1))
                    H.mesh_mac(verts[m], areas[m], S(2 * m + 1, 2 * m + 2))
                am_host.copy_(S(0, 2 * len(dims)), non_blocking=True)
                am_ready.record()
                H.wake_shed(verts, wake, i + 1)
                solve(dTw, i + 1)
                gamma_post(i + 1, True)
                for m in range(len(dims)):
                    H.forces_unsteady(verts[m], gdelta[m], gdt[m], vel[m], areas[m], normals[m], forces[m])
                am_ready.synchronize()
                cl, cmy = self._coefficients(forces, verts, ref, fs, la, vars.rho, S, am_host.tolist())
                F[0, i + 1] = T(-cl / (float(PI_F) * vars.mu))
                F[1, i + 1] = T(2.0 * cmy / (float(PI_F) * vars.mu * vars.r_a ** 2))
                dF = np.array([F[0, i + 1] - F[0, i] - kh * du[0],
                               F[1, i + 1] - F[1, i] - ka * T(tf(u[1, i + 1]) - tf(u[1, i]))],
                              dtype=T)
                du_k = du_k - du
                it += 1
                if it > 50:
                    break
            iters.append(it)
            if step_times is not None:
                step_times.append(time.perf_counter() - t_step)  # each iteration ends with a host read
        return {"t": np.arange(n_loop + 1) * float(dt_nd), "u": u[:, : n_loop + 1], "v": v[:, : n_loop + 1],
                "a": acc[:, : n_loop + 1], "F": F[:, : n_loop + 1], "iters": np.array(iters), "dt": float(dt_nd),
                "t_steps": t_steps}

    @staticmethod
    def _coefficients(forces, verts, ref, fs, la, rho, S, am):
        """coeff_cl_multibody / coeff_cm_multibody (vlm/src/vlm_backend.cpp:71-135) from the surfaces'
        areas and MACs am = [area0, mac0, area1, mac1, ...] (Backend::sum and mesh_mac of the same
        pose): all CL/CM launched together into the device scalars S(a, b) (scal[a:b]), one host
        read."""
        n = len(forces)
        base = 2 * n
        for m in range(n):
            area, mac = am[2 * m], am[2 * m + 1]
            o = base + 4 * m
            H.coeff_cl(forces[m], la, fs, rho, area, S(o, o + 1))
            H.coeff_cm(forces[m], verts[m], ref, fs, rho, area, mac, S(o + 1, o + 4))
        co = S(base, base + 4 * n).tolist()
        cl = cmy = ta = tm = 0.0
        for m in range(n):
            area, mac = am[2 * m], am[2 * m + 1]
            cl += co[4 * m] * area
            cmy += co[4 * m + 2] * area * mac
            ta += area
            tm += mac
        return cl / ta, cmy / (ta * tm)
