"""GPU checks of the batched flutter sweep (tvlm_flutter_*: the reference's coupled pitch/heave
driver UVLM_2DOF for many reduced velocities at once, resident on the device, wake ring buffer).

- against the oracle (oracle/two_dof.py, FP64, the same wake truncation) on a 12 x 6 wing with a
  15-row wake ring over 40 steps (the ring wraps): states and loads within 1e-9 of their scale,
  equal fixed-point iteration counts at every step;
- against the plain FP64 drop-in driver (tinyvlm_amd.two_dof over the Backend entry points: exact
  Biot-Savart, LU substitution) on the same case, to the same bar;
- one airspeed alone equals the same airspeed in the batch, bit for bit;
- full size (80 chordwise x 100 spanwise, 8000 panels, 200-row ring, past the first wrap): finite,
  converged, and within |dCL| <= 1e-6 of the plain FP64 driver.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C5 = dict(mu=20.0, omega=0.5, a_h=-0.3, x_a=0.25, r_a=0.5)


@pytest.fixture(scope="module")
def F():
    from tinyvlm_amd import capi as H
    from tinyvlm_amd import flutter_plan

    H.init(0)
    return flutter_plan


def _oracle_runs(oracle_two_dof, v, airspeeds, steps, R, dt):
    out = []
    for Ua in airspeeds:
        vars = oracle_two_dof.Vars(U_a=Ua, **C5)
        out.append(oracle_two_dof.run([(v, True)], vars, t_final_nd=(steps + 2) * dt + 0.5 * dt, dtype=np.float64,
                                      wake_rows=R))
    return out


def _rel_close(a, b, tol):
    scale = max(float(np.max(np.abs(b))), 1e-30)
    return float(np.max(np.abs(a - b))) <= tol * scale


def test_flutter_matches_oracle_and_plain(F, oracle):
    from oracle import two_dof as O

    from tinyvlm_amd.two_dof import UVLM2DOF, Vars

    v = oracle.rectangular_wing(12, 6, chord=1.0, half_span=3.0)
    airspeeds = [0.7, 1.0, 2.0]
    steps, R = 40, 15
    print(f"[case] flutter 12x6 R={R} {steps} steps, U_a {airspeeds}", flush=True)
    plan = F.FlutterPlan(v, airspeeds, F.params(**C5), wake_rows=R, max_steps=steps)
    plan.run(steps)
    res = plan.results()
    plan.close()
    dt = res["dt"]
    assert res["iters"].shape == (steps, 3)
    orc = _oracle_runs(O, v, airspeeds, steps, R, dt)
    drop = UVLM2DOF("hip", [(v, True)], dtype=np.float64)
    for b, Ua in enumerate(airspeeds):
        o = orc[b]
        assert o["u"].shape[1] == steps + 1, o["u"].shape
        assert abs(o["dt"] - dt) <= 1e-15
        assert np.array_equal(res["iters"][:, b], o["iters"]), (Ua, res["iters"][:, b], o["iters"])
        for key in ("u", "v", "F"):
            assert _rel_close(res[key][:, b, :].T, o[key], 1e-9), (Ua, key)
        p = drop.run(Vars(U_a=Ua, **C5), t_final_nd=(steps + 2) * dt + 0.5 * dt, wake_rows=R)
        assert np.array_equal(res["iters"][:, b], p["iters"]), Ua
        for key in ("u", "F"):
            assert _rel_close(res[key][:, b, :].T, p[key], 1e-9), (Ua, key)


def test_flutter_alone_equals_batched(F, oracle):
    v = oracle.rectangular_wing(10, 5, chord=1.0, half_span=2.5)
    batch = [0.6, 0.85, 1.1, 1.3]
    print("[case] flutter alone vs batch", flush=True)
    pb = F.FlutterPlan(v, batch, F.params(**C5), wake_rows=12, max_steps=30)
    pb.run(30)
    rb = pb.results()
    pb.close()
    pa = F.FlutterPlan(v, [batch[2]], F.params(**C5), wake_rows=12, max_steps=30)
    pa.run(30)
    ra = pa.results()
    pa.close()
    for key in ("u", "v", "a", "F"):
        assert np.array_equal(ra[key][:, 0], rb[key][:, 2]), key
    assert np.array_equal(ra["iters"][:, 0], rb["iters"][:, 2])


def test_flutter_full_size_vs_plain(F):
    """80 chordwise x 100 spanwise panels (half span 5 chords), 200-row wake ring, 210 steps (the ring
    wraps after 200): two airspeeds in the batch; the first also through the plain FP64 driver."""
    from tinyvlm_amd.two_dof import UVLM2DOF, Vars

    v = F.rectangular_wing(100, 80, chord=1.0, half_span=5.0)
    airspeeds = [0.8, 1.1]
    steps, R = 210, 200
    print(f"[case] flutter full size 80x100 R={R} {steps} steps", flush=True)
    plan = F.FlutterPlan(v, airspeeds, F.params(**C5), wake_rows=R, max_steps=steps)
    plan.run(steps)
    res = plan.results()
    plan.close()
    assert np.all(np.isfinite(res["u"])) and np.all(np.isfinite(res["F"]))
    assert np.all(res["iters"] >= 1) and np.all(res["iters"] <= 51)
    dt = res["dt"]
    assert abs(dt - 2.0 / 80) < 1e-6  # the last chordwise panel of the dimensionless wing
    print(f"[case]   mean iterations {res['iters'].mean(axis=0)}", flush=True)
    p = UVLM2DOF("hip", [(v, True)], dtype=np.float64).run(Vars(U_a=airspeeds[0], **C5),
                                                           t_final_nd=(steps + 2) * dt + 0.5 * dt, wake_rows=R)
    cl_plan = -res["F"][:, 0, 0] * (F.PI_F * C5["mu"])
    cl_plain = -p["F"][0] * (F.PI_F * C5["mu"])
    dcl = float(np.max(np.abs(cl_plan - cl_plain)))
    print(f"[case]   |dCL| vs plain FP64 {dcl:.2e}", flush=True)
    assert dcl <= 1e-6
    assert np.array_equal(res["iters"][:, 0], p["iters"])
