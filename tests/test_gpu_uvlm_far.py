"""GPU checks of the served UVLM loop with far-wake aggregation (tvlm_uvlm_set_far_wake) on the
pitch + heave workload (AR-4 rectangular wing, 10 x 40 panels, dt = c/10).

- against the oracle's aggregated wake (orc_uvlm_run_far, the same levels): CL(t), CM(t) and the
  last gamma within 1e-11 relative -- the same bar as the exact served loop against the exact oracle;
- the headline schedule {(80, 2, 4), (160, 4, 8), (320, 8, 8)} over 500 steps against the plain
  FP64 exact wake: max |dCL| <= 1e-6 at every step (the acceptance bar; the oracle measures 2.65e-7);
- graph replays reproduce bits; switching the aggregation off restores the exact served loop bit
  for bit; bad level lists are refused.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32 = np.float32
A_AMP = float(F32(2) * F32(3.141593) / F32(180))
MOTION = dict(u=1.0, h_amp=0.1, h_phase=0.0, a_amp=A_AMP, a_phase=math.pi / 2, omega=1.0, x_pivot=0.25)
MOT = [MOTION[k] for k in ("u", "h_amp", "h_phase", "a_amp", "a_phase", "omega", "x_pivot")]
HEADLINE = [(80, 2, 4), (160, 4, 8), (320, 8, 8)]


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def wing(oracle):
    from tinyvlm_amd import capi

    capi.init(0)
    return oracle.rectangular_wing(40, 10, chord=1.0, half_span=2.0)


def _plan(wing, t_steps, mode, far=None):
    from tinyvlm_amd.uvlm_plan import UVLMPlan, time_step

    p = UVLMPlan([wing], t_steps, time_step(wing, 1.0), mode=mode)
    p.set_pitch_heave(**MOTION)
    if far is not None:
        p.set_far_wake(far)
    return p


def test_far_wake_matches_oracle_aggregation(oracle, wing):
    ts = 161
    far = [(20, 2, 4), (40, 4, 8), (80, 8, 8)]
    print(f"[case] uvlm far wake vs oracle {far} steps {ts - 1}", flush=True)
    p = _plan(wing, ts, 1, far)
    p.run(use_graph=True)
    r = p.results()
    cl_o, cm_o, g_o, _ = oracle.uvlm_run([wing], MOT, t_steps=ts, far_wake=far)
    cl_x, _, _, _ = oracle.uvlm_run([wing], MOT, t_steps=ts)
    print(f"[case]   rel dCL vs oracle far {rel(r['cl'][1:], cl_o[1:]):.2e}, far vs exact oracle "
          f"{rel(cl_o[1:], cl_x[1:]):.2e}", flush=True)
    assert rel(r["cl"][1:], cl_o[1:]) < 1e-11
    assert rel(r["cm"][1:, 1], cm_o[1:]) < 1e-10
    assert rel(r["gamma"], g_o) < 1e-11
    # the aggregation is really on: it moves CL by far more than the tolerance above
    assert rel(cl_o[1:], cl_x[1:]) > 1e-8
    p.close()


def test_far_wake_headline_schedule_accepted(oracle, wing):
    ts = 501
    print(f"[case] uvlm far wake headline {HEADLINE} vs plain exact, 500 steps", flush=True)
    plain = _plan(wing, ts, 0)
    plain.run(use_graph=False)
    rp = plain.results()
    fast = _plan(wing, ts, 1, HEADLINE)
    fast.run(use_graph=True)
    rf = fast.results()
    dcl = float(np.max(np.abs(rf["cl"] - rp["cl"])))
    dcm = float(np.max(np.abs(rf["cm"] - rp["cm"])))
    print(f"[case]   max |dCL| {dcl:.3e} max |dCM| {dcm:.3e}", flush=True)
    assert dcl <= 1e-6 and dcm <= 1e-6
    cl_o, _, _, _ = oracle.uvlm_run([wing], MOT, t_steps=ts, far_wake=HEADLINE)
    assert rel(rf["cl"][1:], cl_o[1:]) < 1e-11
    # replay and the uncaptured launch sequence give the same bits
    fast.run(use_graph=True)
    assert np.array_equal(fast.results()["cl"], rf["cl"])
    fast.run(use_graph=False)
    assert np.array_equal(fast.results()["cl"], rf["cl"])
    plain.close()
    fast.close()


def test_far_wake_off_restores_the_exact_served_loop(wing):
    ts = 121
    print("[case] uvlm far wake on then off", flush=True)
    ref = _plan(wing, ts, 1)
    ref.run(use_graph=True)
    r0 = ref.results()
    p = _plan(wing, ts, 1, [(10, 2, 2)])
    p.run(use_graph=True)
    r1 = p.results()
    assert not np.array_equal(r1["cl"], r0["cl"])
    p.set_far_wake(None)
    p.run(use_graph=True)
    assert np.array_equal(p.results()["cl"], r0["cl"])
    ref.close()
    p.close()


def test_far_wake_time_wake_counts_coarse_rings(wing):
    """tvlm_uvlm_time_wake's pair count follows the aggregation: exact rows x 40 rings plus each
    level's blocks x ceil(40 / S) coarse rings, summed over steps 1..t_steps-2."""
    ts = 121
    far = [(10, 2, 4), (30, 4, 8)]
    p = _plan(wing, ts, 1, far)
    _, _, pairs = p.time_wake()
    expect = 0
    for it in range(1, ts - 1):
        hi, rings = it, 0
        for A, G, S in reversed(far):
            C = (it - A) // G if it >= A else 0
            nb = max(0, C - (it - hi) // G)
            rings += nb * -(-40 // S)
            hi -= nb * G
        expect += (rings + hi * 40) * 400
    assert pairs == expect
    p.close()


def test_far_wake_rejects_bad_levels(wing):
    import ctypes

    from tinyvlm_amd import capi as H

    p = _plan(wing, 31, 1)
    lib = H.lib()
    for lv in ([(10, 2, 2), (5, 4, 4)], [(10, 3, 2), (20, 4, 2)], [(10, 0, 2)], [(10, 2, 0)], [(-1, 2, 2)]):
        arr = (ctypes.c_int32 * (3 * len(lv)))(*[x for t in lv for x in t])
        assert lib.tvlm_uvlm_set_far_wake(p._plan, ctypes.c_int32(len(lv)), arr) != 0, lv
    plain = _plan(wing, 31, 0)
    arr = (ctypes.c_int32 * 3)(10, 2, 2)
    assert lib.tvlm_uvlm_set_far_wake(plain._plan, ctypes.c_int32(1), arr) != 0  # the plain mode stays exact
    p.close()
    plain.close()
