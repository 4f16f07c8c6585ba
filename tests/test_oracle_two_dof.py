"""CPU pins of the coupled pitch/heave oracle (oracle/two_dof.py) and of the host-side logic of
tinyvlm_amd.two_dof (Newmark integrator, section kinematics).

- Newmark-beta (average acceleration) on an undamped oscillator x'' + w^2 x = 0 reproduces its
  exact discrete solution x_n = cos(w_N n dt), tan(w_N dt / 2) = w dt / 2 (amplitude preserved,
  period elongated), for the oracle's and the product's integrator;
- the closed-form section pose and its time derivative equal the dual-number kinematics chain
  freestream * heave * pitch (the reference's construction, tests/uvlm_2dof.cpp:645-680);
- physics of the coupled oracle on the reference's wing/flap section: with the pitch spring's
  linear coefficient 0.1 the linear flutter speed scales to about 6.285 * sqrt(0.1) = 2.0, so a
  3-degree initial pitch decays at U_a = 1 and grows into a limit cycle at U_a = 3.
"""
import math
import os

import numpy as np
import pytest

MESHES = [("mesh/infinite_rectangular_2x2_c0.x", False), ("mesh/infinite_rectangular_2x2_c1.x", True)]


@pytest.fixture(scope="module")
def o2(oracle, root):
    os.chdir(root)
    from oracle import two_dof

    return two_dof


def _oscillator(integ_cls, dtype, tol):
    w, dt, nsteps = 0.7, 0.3, 400
    M = np.eye(1, dtype=dtype)
    Cm = np.zeros((1, 1), dtype=dtype)
    K = np.full((1, 1), w * w, dtype=dtype)
    integ = integ_cls(M, Cm, K, dt, dtype=dtype)
    u, v, a = np.ones(1, dtype=dtype), np.zeros(1, dtype=dtype), np.full(1, -w * w, dtype=dtype)
    wn = 2.0 / dt * math.atan(w * dt / 2.0)
    err = 0.0
    for n in range(1, nsteps + 1):
        du, dv, da = integ.step(M, Cm, v, a, np.zeros(1, dtype=dtype))
        u, v, a = u + du, v + dv, a + da
        err = max(err, abs(float(u[0]) - math.cos(wn * n * dt)))
    assert err < tol, err


def test_newmark_oracle_exact_discrete_solution(o2):
    _oscillator(o2.NewmarkBeta, np.float64, 1e-12)
    _oscillator(o2.NewmarkBeta, np.float32, 2e-4)


def test_newmark_product_exact_discrete_solution():
    from tinyvlm_amd.two_dof import NewmarkBeta

    _oscillator(NewmarkBeta, np.float64, 1e-12)
    _oscillator(NewmarkBeta, np.float32, 2e-4)


def test_section_pose_matches_dual_number_chain(o2):
    from tinyvlm_amd import kinematics as K
    from tinyvlm_amd.two_dof import section_transform

    a_h = -0.5
    for (t, h, hd, al, ald) in ((0.0, 0.0, 0.0, 0.05, 0.0), (3.5, 0.12, -0.03, -0.2, 0.07), (40.0, -1.1, 0.4, 0.3, -0.5)):
        fs = K.KinematicNode(lambda tt: K.translation(-tt, 0.0, 0.0))
        heave = K.KinematicNode(lambda tau: K.translation(0.0, 0.0, -(h + hd * tau)))
        pitch = K.KinematicNode(lambda tau: K.rotation((1.0 + a_h, 0.0, 0.0), (0.0, 1.0, 0.0), al + ald * tau))
        d = fs.transform_dual(t) @ heave.transform_dual(0.0) @ pitch.transform_dual(0.0)
        for fn in (section_transform, o2.wing_transform):
            T, dT = fn(t, h, hd, al, ald, a_h)
            assert np.max(np.abs(T - d.value())) < 1e-14
            assert np.max(np.abs(dT - d.grad())) < 1e-14


def test_coupled_section_flutter_boundary(o2):
    below = o2.run(MESHES, o2.Vars(U_a=1.0), 150.0, np.float64)
    above = o2.run(MESHES, o2.Vars(U_a=3.0), 150.0, np.float64)
    a0 = math.radians(3.0)
    n = below["u"].shape[1]
    assert np.max(np.abs(below["u"][1, 3 * n // 4:])) < 0.6 * a0
    assert np.max(np.abs(above["u"][1, 3 * n // 4:])) > 1.3 * a0
    assert below["iters"].max() <= 10 and above["iters"].max() <= 10


def test_product_rejects_unknown_backend():
    from tinyvlm_amd.two_dof import UVLM2DOF

    with pytest.raises(RuntimeError, match="Unsupported backend"):
        UVLM2DOF("cpu", MESHES)


def test_wake_truncation_is_zeroed_old_rings(oracle):
    """Truncating the wake to its newest R ring rows (a wake view of those rows: vertex rows
    [nwv-1-it, nwv-1-it+R], circulation columns [nwv-1-it, nwv-1-it+R), all R rings inducing) is the
    full wake with the older rings' circulation set to zero, bit for bit: a zero-strength ring adds
    exactly 0 to each receiver's sum.  (Row nwv-1-s holds the vertices shed at step s: the newest
    rings have the smallest row indices.)"""
    rng = np.random.default_rng(4)
    ns, nc, rows = 7, 3, 30
    v = oracle.rectangular_wing(ns, nc, half_span=2.0)
    wake = np.zeros((ns + 1, rows + 1, 4), order="F")
    for r in range(rows + 1):
        wake[:, r] = v[:, -1]
        wake[:, r, 0] += 0.13 * (rows - r) + 0.01 * rng.standard_normal(ns + 1)
        wake[:, r, 2] += 0.04 * math.sin(0.4 * r)
    gw = np.asfortranarray(rng.standard_normal((ns, rows)))
    s = oracle.Surface(verts=v, verts_wake=np.asfortranarray(wake), gamma_wake=gw)
    oracle.mesh_metrics(0.0, [s])
    nwv = rows + 1
    for it, R in ((25, 10), (30, 1), (12, 11)):
        r0 = nwv - 1 - it
        sub = oracle.Surface(verts=v, verts_wake=np.asfortranarray(wake[:, r0: r0 + R + 1]),
                             gamma_wake=np.asfortranarray(gw[:, r0: r0 + R]), colloc=s.colloc, normals=s.normals,
                             areas=s.areas)
        trunc = np.zeros(ns * nc)
        oracle.rhs_assemble_wake_influence(trunc, [sub], R)
        zeroed = oracle.Surface(verts=v, verts_wake=s.verts_wake, gamma_wake=gw.copy(order="F"), colloc=s.colloc,
                                normals=s.normals, areas=s.areas)
        # ring rows used by iteration it: [nwv - 1 - it, nwv - 2]; the newest R of them (the first R) stay
        zeroed.gamma_wake[:, r0 + R: nwv - 1] = 0.0
        full = np.zeros(ns * nc)
        oracle.rhs_assemble_wake_influence(full, [zeroed], it)
        assert np.array_equal(trunc, full), (it, R)


def test_oracle_two_dof_truncation_limits(o2, oracle):
    """A ring of R rows changes nothing while the wake is at most R rows long: step i induces with i + 1
    ring rows, so the first R states are bitwise the full wake's and the first load that differs is
    that of step R (state R + 1); a ring as long as the run changes nothing at all."""
    v = oracle.rectangular_wing(6, 4, chord=1.0, half_span=3.0)
    vars = o2.Vars(mu=20.0, omega=0.5, a_h=-0.3, U_a=0.9)
    full = o2.run([(v, True)], vars, t_final_nd=40 * 0.5 + 0.2, dtype=np.float64)
    n = full["F"].shape[1] - 1
    same = o2.run([(v, True)], vars, t_final_nd=40 * 0.5 + 0.2, dtype=np.float64, wake_rows=n + 5)
    assert np.array_equal(full["F"], same["F"]) and np.array_equal(full["u"], same["u"])
    for R in (5, 12):
        tr = o2.run([(v, True)], vars, t_final_nd=40 * 0.5 + 0.2, dtype=np.float64, wake_rows=R)
        differ = np.flatnonzero(np.any(tr["F"] != full["F"], axis=0))
        assert differ.size and differ[0] == R + 1, (R, differ[:3])
