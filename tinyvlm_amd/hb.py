"""Harmonic-balance VLM over the HIP C ABI with the Python interface of the reference's libhbvlm3
module (lib/libhbvlm3.cpp:92-642): ``HBVLM(backend_name, [wing_mesh, flap_mesh])``,
``.init(harmonics, scaling)``, ``.run(omega, U, dyn_f, force_t)``.

The periodic circulation is expanded in K = 2H + 1 real Fourier coefficients per panel (an
{N, K} column-major matrix).  One evaluation of the fixed-point map (hb_vlm_iter,
lib/libhbvlm3.cpp:349-410) builds, at each of the K collocation times of the period, the
kinematic rhs and the induction of the flap's wake, whose circulation follows from the flap's
trailing-edge coefficients; the K rhs columns are multiplied by the orthogonal DFT matrix and
solved with the rigid influence matrix, factored once.  Anderson acceleration (history 10,
tolerance 1e-12, at most 100 iterations, vlm/src/vlm_solvers.cpp:6-115) finds the fixed point;
the forces are then evaluated at the K times (lib/libhbvlm3.cpp:424-549).

Differences from the reference's loop, none of which changes a value:
- the pose, panel metrics and kinematic rhs of each collocation time do not depend on the
  circulation, so they are computed once per run instead of once per fixed-point iteration;
- the transforms of all wake-shedding steps are evaluated on the host in one vectorised pass and
  uploaded once.
Every array operation runs on the device through libtinyvlm_hip.so; there is no host fallback.
"""
from __future__ import annotations

import math

import numpy as np

from . import capi as H
from . import kinematics as K
from .mesh_io import read_plot3d

PI_F = float(np.float32(3.141593))  # PI_f (vlm/include/vlm_types.hpp)


def _f32(x):
    return float(np.float32(x))


class Vars:
    """The aerodynamic constants of the binding (lib/libhbvlm3.cpp:558-580).  The reference assigns
    them from float literals, so they carry FP32 rounding; the structural ones play no role in the
    aerodynamic solve."""

    def __init__(self, U):
        self.a = _f32(-0.5)  # elastic axis from mid-chord, semi-chords
        self.b = _f32(0.127)  # semi-chord
        self.c = _f32(0.5)  # flap hinge from mid-chord, semi-chords
        self.omega_alpha = _f32(52.6506)
        self.U = float(U)


def build_kinematics(dyn_f, omega, v: Vars, harmonics: int):
    """Kinematics tree of the binding (lib/libhbvlm3.cpp:587-640): freestream translation, heave
    after it, pitch about (1 + a, 0, 0) after the heave, flap rotation about (1 + c, 0, 0) after
    the pitch.  dyn_f rows are the Fourier coefficients of heave, pitch and flap angle.
    Returns (freestream node, [wing node, flap node])."""
    r = np.array(dyn_f, dtype=np.float64)
    Hh = int(harmonics)
    tree = K.KinematicsTree()
    u = v.U / (v.b * v.omega_alpha)
    fs = tree.add(lambda t: K.translation(-u * t, 0.0, 0.0))

    def series(row, t):
        s = K.Dual(r[row, 0])
        for h in range(Hh):
            k = float(h + 1)
            s = s + K.dcos(omega * t * k) * r[row, 2 * h + 1]
            s = s + K.dsin(omega * t * k) * r[row, 2 * h + 2]
        return s

    heave = tree.add(lambda t: K.translation(0.0, 0.0, -series(0, t))).after(fs)
    pitch = tree.add(lambda t: K.rotation((1.0 + v.a, 0.0, 0.0), (0.0, 1.0, 0.0), series(1, t))).after(heave)
    control = tree.add(lambda t: K.rotation((1.0 + v.c, 0.0, 0.0), (0.0, 1.0, 0.0), series(2, t))).after(pitch)
    return fs, [pitch, control]


class AndersonSolver:
    """anderson_acceleration (vlm/src/vlm_solvers.cpp:6-115) on device vectors: the same sequence
    of BLAS copies/axpy/gemv and the LSQ solve (gels) of the C ABI; the residual norm is read back
    once per iteration, the by-value contract of BLAS::norm.  The least-squares info (which the
    reference does not check) shares the norm's device buffer: the norm read of the next iteration
    also returns it, so a rank-deficient history still raises, one map evaluation later, without a
    second host read per iteration."""

    def __init__(self, n, m):
        import torch

        z = lambda *s: H.colmajor(s, torch.float64)  # noqa: E731
        self.torch = torch
        self.n, self.m = n, m
        self.X, self.G, self.Gk = z(n, m), z(n, m), z(n, m)
        self.x_curr, self.x_new, self.g_curr, self.g_new, self.gamma = z(n), z(n), z(n), z(n), z(n)
        # the history's column slices, made once (the same tensors at every iteration keep their views)
        self.lead = {name: [B[:, :j] for j in range(m + 1)] for name, B in (("X", self.X), ("G", self.G),
                                                                           ("Gk", self.Gk))}
        self.lead["gamma"] = [self.gamma[:j] for j in range(m + 1)]
        self.col = {name: [B[:, j] for j in range(m)] for name, B in (("X", self.X), ("G", self.G))}
        # [0]: the residual norm (nrm2), [1]: the last gels info (int32 in its first four bytes)
        self.flags = torch.zeros(2, dtype=torch.float64, device="cuda")
        self.scal = self.flags[0:1]
        self.info = self.flags[1:2].view(torch.int32)[0:1]

    def _read(self):
        """(norm, info) from one host read of the flags buffer."""
        v = self.flags.cpu()
        return float(v[0]), int(v[1:2].view(self.torch.int32)[0])

    @staticmethod
    def _check(info):
        if info != 0:
            raise H.TvlmError(f"Anderson least squares: rank-deficient history (gels info {info})")

    def run(self, x0, f, max_iter=100, tol_res=1e-6):
        """f(x, y) writes the map's value at x into y.  x0 is overwritten with the fixed point;
        returns the iteration counter k of the reference's loop."""
        X, G, Gk = self.X, self.G, self.Gk
        x_curr, x_new, g_curr, g_new, gamma = self.x_curr, self.x_new, self.g_curr, self.g_new, self.gamma
        m = self.m
        x_curr.copy_(x0)
        f(x_curr, x_new)
        g_curr.copy_(x_new)
        H.axpy(-1.0, x_curr, g_curr)
        X[:, 0].copy_(g_curr)
        x_curr.copy_(x_new)
        f(x_curr, x_new)
        g_new.copy_(x_new)
        H.axpy(-1.0, x_curr, g_new)
        G[:, 0].copy_(g_new)
        H.axpy(-1.0, g_curr, G[:, 0])
        g_curr.copy_(g_new)
        k = 1
        pending = False  # a gels whose info has not been read yet
        while k < max_iter:
            H.nrm2(g_curr, self.scal)
            res, info = self._read()
            if pending:
                self._check(info)
                pending = False
            if not res > tol_res:
                break
            m_k = min(m, k)
            # X[:, :m_k], G[:, :m_k], Gk[:, :m_k], gamma[:m_k]
            Xm, Gm, Gkm, gm = (self.lead[name][m_k] for name in ("X", "G", "Gk", "gamma"))
            gamma.copy_(g_curr)
            Gkm.copy_(Gm)
            H.gels(Gkm, gamma, self.info)
            pending = True
            x_new.copy_(x_curr)
            H.axpy(1.0, g_curr, x_new)
            H.gemv(-1.0, Xm, gm, 1.0, x_new)
            H.gemv(-1.0, Gm, gm, 1.0, x_new)
            Xb, Gb = self.col["X"][k % m], self.col["G"][k % m]
            Xb.copy_(x_new)
            H.axpy(-1.0, x_curr, Xb)
            x_curr.copy_(x_new)
            f(x_curr, x_new)
            g_new.copy_(x_new)
            H.axpy(-1.0, x_curr, g_new)
            Gb.copy_(g_new)
            H.axpy(-1.0, g_curr, Gb)
            g_curr.copy_(g_new)
            k += 1
        if pending:
            self._check(self._read()[1])
        x0.copy_(x_curr)
        return k


class HBVLM:
    """HBVLM (lib/libhbvlm3.cpp:92-552) for two surfaces: a non-lifting wing and a lifting flap
    whose wake alone is shed.  The Fourier coefficients persist between run() calls, so a
    continuation warm-starts from the previous point (lib/libhbvlm3.cpp:213-214, 413-414)."""

    def __init__(self, backend_name: str, meshes):
        if backend_name != "hip":
            raise RuntimeError("Unsupported backend: " + str(backend_name))
        if len(meshes) != 2:
            raise ValueError("HBVLM takes two meshes: [wing, flap]")
        import torch

        H.init(0)
        self.torch = torch
        # plot3d parsed in FP64 without quarter-chord shift (lib/libhbvlm3.cpp:158-167)
        self.verts_file = [read_plot3d(m, quarter_chord=False, parse_f64=True) for m in meshes]
        self.dims = [(v.shape[0] - 1, v.shape[1] - 1) for v in self.verts_file]
        self.N = sum(ns * nc for ns, nc in self.dims)
        self.iterations = 0
        self.harmonics = None
        # the solve runs on a stream of its own; the map evaluated in every Anderson iteration is
        # replayed from a hipGraph (use_graph = False: launched call by call, the same kernels)
        self.stream = torch.cuda.Stream()
        self.use_graph = True
        # wake induction of the map: "exact" (Backend::rhs_assemble_wake_influence ring by ring, the
        # reference's evaluation) or "lattice" (the same sum over the wake's vortex lattice, every
        # segment once, FP64 seeded; agrees with "exact" to rounding)
        self.wake_form = "exact"

    # ---- HBVLM::init (lib/libhbvlm3.cpp:168-233) ------------------------------------------------
    def init(self, harmonics: int, scaling: float):
        torch = self.torch
        self.harmonics = int(harmonics)
        Kc = 2 * self.harmonics + 1
        self.K = Kc
        z = lambda *s: H.colmajor(s, torch.float64)  # noqa: E731
        self.verts_init = []
        for v in self.verts_file:
            d = H.to_device(v, torch.float64)
            # x, y, z planes are the first 3 (ns+1)(nc+1) values of the column-major buffer
            H.scal(float(scaling), d.permute(2, 1, 0)[0:3].reshape(-1))
            self.verts_init.append(d)
        N = self.N
        self.verts = [z(ns + 1, nc + 1, 4) for ns, nc in self.dims]
        self.colloc0 = [z(ns, nc, 3) for ns, nc in self.dims]
        self.normals0 = [z(ns, nc, 3) for ns, nc in self.dims]
        self.areas0 = [z(ns, nc) for ns, nc in self.dims]
        self.lhs = z(N, N)
        self.rhs = z(N, Kc)
        self.gamma_coeffs = z(N, Kc)  # zero once; warm start of every later run
        self.gamma_wing = z(N, Kc)
        self.dgamma_wing = z(N, Kc)
        self.gd = [z(ns, nc) for ns, nc in self.dims]
        self.dg = [z(ns, nc) for ns, nc in self.dims]
        self.vel = [z(ns, nc, 3) for ns, nc in self.dims]
        self.forces = [z(ns, nc, 3) for ns, nc in self.dims]
        self.ipiv = torch.zeros(N, dtype=torch.int32, device="cuda")
        self.info = torch.zeros(1, dtype=torch.int32, device="cuda")
        self.anderson = AndersonSolver(N * Kc, 10)
        # per collocation time: pose and panel metrics
        self.inst = [{
            "verts": [z(ns + 1, nc + 1, 4) for ns, nc in self.dims],
            "colloc": [z(ns, nc, 3) for ns, nc in self.dims],
            "normals": [z(ns, nc, 3) for ns, nc in self.dims],
            "areas": [z(ns, nc) for ns, nc in self.dims],
        } for _ in range(Kc)]
        self.rhs_vel = z(N, Kc)
        # the buffers were filled on the caller's current stream; the solve runs on self.stream, a
        # stream that does not wait for the null stream or any other: finish the fills here
        torch.cuda.current_stream().synchronize()

    def reset(self):
        """Forget the warm start: the next run starts from zero Fourier coefficients, as the first
        run after init does (a served request is then independent of the ones before it)."""
        if self.harmonics is None:
            raise RuntimeError("HBVLM.init must be called before reset")
        self.gamma_coeffs.zero_()
        # on the caller's current stream: done before run() starts the solve on self.stream (a served
        # worker resets before it makes self.stream current)
        self.torch.cuda.current_stream().synchronize()
        self.iterations = 0

    # ---- the binding's run (lib/libhbvlm3.cpp:552-641) -----------------------------------------
    def run(self, omega: float, U: float, dyn_f, force_t):
        """force_t (numpy, 3 x K) receives CL, CM_y about the elastic axis and the flap's CM_y
        about its hinge at the K collocation times."""
        if self.harmonics is None:
            raise RuntimeError("HBVLM.init must be called before run")
        v = Vars(U)
        cycles = 2.0
        t_final = cycles * 2.0 * PI_F / omega
        fs_node, nodes = build_kinematics(dyn_f, omega, v, self.harmonics)
        self.torch.cuda.current_stream().synchronize()
        with self.torch.cuda.stream(self.stream):
            out = self.solve(t_final, omega, v, fs_node, nodes, force_t)
        self.stream.synchronize()
        return out

    @staticmethod
    def _poses(nodes, times):
        """The dual transform of every node at every time of `times`, evaluated over the array of
        times in one pass: [[(T, dT/dt) per node] per time], 4 x 4 numpy arrays."""
        tt = np.asarray(times, dtype=np.float64)
        vg = [(d.value(), d.grad()) for d in (n.transform_dual(tt) for n in nodes)]
        return [[(v[:, :, i], g[:, :, i]) for v, g in vg] for i in range(len(tt))]

    def _pose(self, pose, verts, colloc, normals, areas):
        """displace_wing to `pose` (one (T, dT/dt) per surface) and the panel metrics there."""
        torch = self.torch
        tr = [H.to_device(np.asfortranarray(T), torch.float64) for T, _ in pose]
        H.displace_wing(tr, verts, self.verts_init)
        H.mesh_metrics(0.0, verts, colloc, normals, areas)

    def solve(self, t_start, omega, v: Vars, fs_node, nodes, force_t):
        """HBVLM::run (lib/libhbvlm3.cpp:262-552)."""
        torch = self.torch
        if self.wake_form not in ("exact", "lattice"):
            raise ValueError(f"wake_form {self.wake_form!r}: expected 'exact' or 'lattice'")
        Kc, Hh, N = self.K, self.harmonics, self.N
        period = 2.0 * PI_F / omega
        period_interval = period / Kc
        rho = 1.0
        for vv, vi in zip(self.verts, self.verts_init):
            vv.copy_(vi)
        H.mesh_metrics(0.0, self.verts, self.colloc0, self.normals0, self.areas0)
        # time step: the flap's last chordwise panel at the root, an integer number of steps per
        # collocation interval (lib/libhbvlm3.cpp:277-286)
        vf = H.to_host(self.verts[1])
        dx, dy, dz = (vf[0, -1, k] - vf[0, -2, k] for k in range(3))
        chord = math.sqrt(dx * dx + dy * dy + dz * dz)
        if not period_interval > chord:
            raise RuntimeError(f"period interval {period_interval} must exceed the last panel chord {chord}")
        dt = period_interval / math.ceil(period_interval / chord)
        nsteps = int((t_start + period + dt) / dt)
        wake = [H.colmajor((ns + 1, nsteps, 4), torch.float64) for ns, nc in self.dims]
        gamma_wake = [[H.colmajor((ns, nsteps - 1), torch.float64) for ns, nc in self.dims] for _ in range(Kc)]
        # rigid influence matrix without wake, factored once (lib/libhbvlm3.cpp:301-304)
        self.lhs.zero_()
        H.lhs_assemble(self.lhs, self.colloc0, self.normals0, self.verts, wake, 0)
        H.getrf(self.lhs, self.ipiv, self.info)
        info = int(self.info.item())
        if info != 0:
            raise H.TvlmError(f"singular influence matrix (getrf info {info})")
        # wake geometry of the whole run (lib/libhbvlm3.cpp:306-319): transforms of every step
        # in one host pass, one upload
        times = np.arange(nsteps, dtype=np.float64) * dt
        T_all = np.stack([n.transform(times) for n in nodes], axis=2)  # (4, 4, 2, nsteps)
        T_dev = H.to_device(np.asfortranarray(T_all.reshape(4, 4, 2 * nsteps, order="F")), torch.float64)
        # (the per-step displace_wing + wake_shed pairs in one launch per surface; the wing is then
        # left at the last step's pose, as after the loop)
        H.wake_geometry(T_dev, self.verts_init, wake, nsteps)
        H.displace_wing([T_dev[:, :, 2 * nsteps - 2], T_dev[:, :, 2 * nsteps - 1]], self.verts, self.verts_init)
        # orthogonal DFT matrix of the K times and its time derivative (lib/libhbvlm3.cpp:321-346)
        s0 = 1.0 / math.sqrt(float(Kc))
        s2 = math.sqrt(2.0 / float(Kc))
        dft = np.zeros((Kc, Kc))
        ddft = np.zeros((Kc, Kc))
        dft[:, 0] = s0
        for j in range(1, Kc, 2):
            k = (float(j) + 1.0) / 2.0
            for i in range(Kc):
                tn = (float(i) / float(Kc)) * period
                dft[i, j] = math.cos(omega * tn * k) * s2
                dft[i, j + 1] = math.sin(omega * tn * k) * s2
                ddft[i, j] = -omega * k * math.sin(omega * tn * k) * s2
                ddft[i, j + 1] = omega * k * math.cos(omega * tn * k) * s2
        dft_d = H.to_device(np.asfortranarray(dft), torch.float64)
        ddft_d = H.to_device(np.asfortranarray(ddft), torch.float64)
        # circulation-independent part of each collocation time: pose, metrics, kinematic rhs
        steps = []
        for s in range(Kc):
            tf = t_start + period * float(s) / float(Kc)
            ts = int(round(tf / dt))
            steps.append((ts, float(ts) * dt))
        poses = self._poses(nodes, [t for _, t in steps])
        self.rhs_vel.zero_()
        for s in range(Kc):
            I = self.inst[s]
            self._pose(poses[s], I["verts"], I["colloc"], I["normals"], I["areas"])
            H.kinematic_velocities([g for _, g in poses[s]], self.colloc0, self.vel)
            H.rhs_assemble_velocities(self.rhs_vel[:, s], I["normals"], self.vel)
        n0 = self.dims[0][0] * self.dims[0][1]
        ns1, nc1 = self.dims[1]
        lifting = [False, True]
        wake_rhs = H.rhs_assemble_wake_influence if self.wake_form == "exact" else \
            H.rhs_assemble_wake_influence_lattice

        def hb_vlm_iter(x_in, x_out):
            g_in = x_in.view(Kc, N).t()  # {N, K} column-major
            g_out = x_out.view(Kc, N).t()
            te = g_in[n0 + (nc1 - 1) * ns1: n0 + nc1 * ns1, :]  # flap trailing-edge row, {ns1, K}
            self.rhs.copy_(self.rhs_vel)
            for s in range(Kc):
                ts, t = steps[s]
                I = self.inst[s]
                H.gamma_wake_from_coeffs(gamma_wake[s][1], te, Hh, t, omega, dt, ts)
                wake_rhs(self.rhs[:, s], gamma_wake[s], I["colloc"], I["normals"], wake, lifting, ts)
            # A G D^T = R  <=>  A G = R D  (D orthogonal)
            H.gemm(1.0, self.rhs, dft_d, 0.0, g_out)
            H.getrs(self.lhs, self.ipiv, g_out)

        # The map's launches are the same at every call of this solve (same buffers, poses and wake
        # steps): the first call runs them one by one (which also sizes the library's scratch
        # buffers on this stream), the second captures them into a hipGraph, which every later
        # call replays.  Calls on other buffers run one by one.
        state = {"calls": 0, "graph": None, "ptrs": None}

        def hb_map(x_in, x_out):
            state["calls"] += 1
            ptrs = (x_in.data_ptr(), x_out.data_ptr())
            if not self.use_graph or state["calls"] == 1 or (state["ptrs"] not in (None, ptrs)):
                hb_vlm_iter(x_in, x_out)
                return
            if state["graph"] is None:
                g = torch.cuda.CUDAGraph()
                # thread-local capture: other threads keep calling the runtime while this one
                # captures (the capture itself first synchronizes the device, so concurrent solvers in
                # other threads run with use_graph = False, as the serving loop's do)
                with torch.cuda.graph(g, stream=self.stream, capture_error_mode="thread_local"):
                    hb_vlm_iter(x_in, x_out)
                state["graph"], state["ptrs"] = g, ptrs
            state["graph"].replay()

        x = self.gamma_coeffs.t().reshape(-1)  # the {N, K} buffer as one vector (a view)
        assert x.data_ptr() == self.gamma_coeffs.data_ptr()
        self.iterations = self.anderson.run(x, hb_map, 100, 1e-12)
        state["graph"] = None

        # forces at the K times of the first period (lib/libhbvlm3.cpp:424-549)
        H.gemm(1.0, self.gamma_coeffs, dft_d, 0.0, self.gamma_wing, trans_b=True)
        H.gemm(1.0, self.gamma_coeffs, ddft_d, 0.0, self.dgamma_wing, trans_b=True)
        (ns0, nc0) = self.dims[0]
        out = force_t
        times = [float(int(round(period * float(s) / float(Kc) / dt))) * dt for s in range(Kc)]
        poses = self._poses(nodes, times)
        fs_all = self._poses([fs_node], times)
        # Two passes over the K times, the poses kept in the instance buffers (free once the fixed
        # point is found) and every area, MAC and coefficient in a device row per time (area0, mac0,
        # area1, mac1, CL0, CL1, CM0, CM1, flap CM): two host reads in all, the same kernels and host
        # arithmetic as coeff_cl_multibody / coeff_cm_multibody one time after another.
        coef = torch.zeros(Kc, 16, dtype=torch.float64, device="cuda")
        for s in range(Kc):
            I = self.inst[s]
            self._pose(poses[s], I["verts"], I["colloc"], I["normals"], I["areas"])
            for m in range(2):
                H.sum_(I["areas"][m], coef[s, 2 * m: 2 * m + 1])
                H.mesh_mac(I["verts"][m], I["areas"][m], coef[s, 2 * m + 1: 2 * m + 2])
        am = coef[:, :4].tolist()
        for s in range(Kc):
            pose, I = poses[s], self.inst[s]
            verts, normals, areas = I["verts"], I["normals"], I["areas"]
            H.kinematic_velocities([g for _, g in pose], self.colloc0, self.vel)
            g_fs = fs_all[s][0][1]  # KinematicNode.linear_velocity at the origin
            fs = -(g_fs[:3, :3] @ np.zeros(3) + g_fs[:3, 3])
            g0 = self.gamma_wing[:n0, s].view(nc0, ns0).t()
            g1 = self.gamma_wing[n0:, s].view(nc1, ns1).t()
            self.gd[0].copy_(g0)
            self.gd[1].copy_(g1)
            self.dg[0].copy_(self.dgamma_wing[:n0, s].view(nc0, ns0).t())
            self.dg[1].copy_(self.dgamma_wing[n0:, s].view(nc1, ns1).t())
            H.axpy(-1.0, g0[:, nc0 - 1], self.gd[1][:, 0])
            if nc0 > 1:
                H.axpy(-1.0, g0[:, : nc0 - 1], self.gd[0][:, 1:])
            if nc1 > 1:
                H.axpy(-1.0, g1[:, : nc1 - 1], self.gd[1][:, 1:])
            for m in range(2):
                H.forces_unsteady(verts[m], self.gd[m], self.dg[m], self.vel[m], areas[m], normals[m], self.forces[m])
            T0, T1 = pose[0][0], pose[1][0]
            ref_a = (T0 @ np.array([1.0 + v.a, 0.0, 0.0, 1.0]))[:3]
            ref_b = (T1 @ np.array([1.0 + v.c, 0.0, 0.0, 1.0]))[:3]
            la = self._lift_axis(fs)
            for m in range(2):
                area, mac = am[s][2 * m], am[s][2 * m + 1]
                H.coeff_cl(self.forces[m], la, fs, rho, area, coef[s, 4 + m: 5 + m])
                H.coeff_cm(self.forces[m], verts[m], ref_a, fs, rho, area, mac, coef[s, 6 + 3 * m: 9 + 3 * m])
            H.coeff_cm(self.forces[1], verts[1], ref_b, fs, 1.0, am[s][2], am[s][3], coef[s, 12:15])
        co = coef.tolist()
        for s in range(Kc):
            out[0, s] = self._cl_multibody(am[s], co[s])
            out[1, s] = self._cm_multibody(am[s], co[s])[1]
            out[2, s] = co[s][13]
        return out

    @staticmethod
    def _lift_axis(fs):
        la = np.cross(np.asarray(fs, dtype=np.float64), np.array([0.0, 1.0, 0.0]))
        return la / math.sqrt(float(la[0] * la[0] + la[1] * la[1] + la[2] * la[2]))

    @staticmethod
    def _cl_multibody(am, co):
        """coeff_cl_multibody (vlm/src/vlm_backend.cpp:71-98) from the surfaces' areas (am[0], am[2])
        and CL (co[4], co[5])."""
        cl = total = 0.0
        for m in range(2):
            area = am[2 * m]
            cl += co[4 + m] * area
            total += area
        return cl / total

    @staticmethod
    def _cm_multibody(am, co):
        """coeff_cm_multibody (vlm/src/vlm_backend.cpp:100-134) from the surfaces' areas and MACs and
        CM vectors (co[6:9], co[9:12])."""
        cm = np.zeros(3)
        ta = tm = 0.0
        for m in range(2):
            area, mac = am[2 * m], am[2 * m + 1]
            cm += np.array(co[6 + 3 * m: 9 + 3 * m]) * area * mac
            ta += area
            tm += mac
        return cm / (ta * tm)

