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
