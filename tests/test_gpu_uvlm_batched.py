"""GPU checks of the UVLM plan that holds several simulations of one mesh
(tvlm_uvlm_create_batched) on the pitch + heave workload (AR-4 rectangular wing, 10 x 40 panels,
dt = c/10).

- every simulation of a batched plan against a one-simulation plan with the same motion: CL(t),
  CM(t) and the last gamma within 1e-12 relative (the batched LU may block its trailing updates
  differently, so bits are not promised), exact wake and far-wake aggregation;
- one simulation of the batch against the oracle's UVLM::run restatement within 1e-11 relative;
- graph replays of a batched plan reproduce bits; motions set per simulation; a batch whose
  motions are not all set, or a plain-mode batch, is refused.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32 = np.float32
A_AMP = float(F32(2) * F32(3.141593) / F32(180))
BASE = dict(u=1.0, h_amp=0.1, h_phase=0.0, a_amp=A_AMP, a_phase=math.pi / 2, omega=1.0, x_pivot=0.25)
# four motions: the base case, a stronger pitch, another phase, a faster and weaker heave
MOTIONS = [
    BASE,
    dict(BASE, a_amp=2.0 * A_AMP),
    dict(BASE, a_phase=0.0, h_phase=0.3),
    dict(BASE, omega=1.7, h_amp=0.05),
]


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def bits_equal(a, b):
    return bool(np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64)))


@pytest.fixture(scope="module")
def wing(oracle):
    from tinyvlm_amd import capi

    capi.init(0)
    return oracle.rectangular_wing(40, 10, chord=1.0, half_span=2.0)


def _single(wing, t_steps, motion, far=None):
    from tinyvlm_amd.uvlm_plan import UVLMPlan, time_step

    p = UVLMPlan([wing], t_steps, time_step(wing, 1.0), mode=1)
    p.set_pitch_heave(**motion)
    if far:
        p.set_far_wake(far)
    p.run(use_graph=True)
    r = p.results()
    p.close()
    return r


def _batched(wing, t_steps, motions, far=None):
    from tinyvlm_amd.uvlm_plan import UVLMPlan, time_step

    p = UVLMPlan([wing], t_steps, time_step(wing, 1.0), mode=1, sims=len(motions))
    for z, m in enumerate(motions):
        p.set_pitch_heave(**m, sim=z)
    if far:
        p.set_far_wake(far)
    return p


@pytest.mark.parametrize("far", [None, [(20, 2, 4), (40, 4, 8), (80, 8, 8)]], ids=["exact", "far"])
def test_batched_matches_single_plans(wing, far):
    ts = 161
    print(f"[case] uvlm batched {len(MOTIONS)} sims vs single plans, far {far}, steps {ts - 1}", flush=True)
    p = _batched(wing, ts, MOTIONS, far)
    p.run(use_graph=True)
    got = [p.results(sim=z) for z in range(len(MOTIONS))]
    for z, m in enumerate(MOTIONS):
        ref = _single(wing, ts, m, far)
        e_cl, e_cm, e_g = rel(got[z]["cl"][1:], ref["cl"][1:]), rel(got[z]["cm"][1:], ref["cm"][1:]), \
            rel(got[z]["gamma"], ref["gamma"])
        print(f"[case]   sim {z}: rel dCL {e_cl:.2e} dCM {e_cm:.2e} dgamma {e_g:.2e} "
              f"bits {bits_equal(got[z]['cl'], ref['cl'])}", flush=True)
        assert e_cl < 1e-12 and e_cm < 1e-12 and e_g < 1e-12
    # the motions really differ (a mix-up of simulations would not pass the checks above)
    assert rel(got[1]["cl"][1:], got[0]["cl"][1:]) > 1e-3
    p.close()


def test_batched_simulation_matches_oracle(oracle, wing):
    ts = 81
    print(f"[case] uvlm batched sim 2 of {len(MOTIONS)} vs oracle, steps {ts - 1}", flush=True)
    p = _batched(wing, ts, MOTIONS)
    p.run(use_graph=False)
    r = p.results(sim=2)
    m = MOTIONS[2]
    mot = [m[k] for k in ("u", "h_amp", "h_phase", "a_amp", "a_phase", "omega", "x_pivot")]
    cl_o, cm_o, g_o, _ = oracle.uvlm_run([wing], mot, t_steps=ts)
    print(f"[case]   rel dCL {rel(r['cl'][1:], cl_o[1:]):.2e} dgamma {rel(r['gamma'], g_o):.2e}", flush=True)
    assert rel(r["cl"][1:], cl_o[1:]) < 1e-11
    assert rel(r["cm"][1:, 1], cm_o[1:]) < 1e-10
    assert rel(r["gamma"], g_o) < 1e-11
    p.close()


def test_batched_graph_replays_and_motion_updates(wing):
    ts = 121
    print(f"[case] uvlm batched graph replays, steps {ts - 1}", flush=True)
    p = _batched(wing, ts, MOTIONS[:3], far=[(20, 2, 4)])
    p.run(use_graph=True)
    first = [p.results(sim=z) for z in range(3)]
    p.run(use_graph=True)
    again = [p.results(sim=z) for z in range(3)]
    for a, b in zip(first, again):
        assert bits_equal(a["cl"], b["cl"]) and bits_equal(a["gamma"], b["gamma"])
    # a new motion for simulation 1 only: the graph reads the motion buffers, so its replay follows
    p.set_pitch_heave(**MOTIONS[3], sim=1)
    p.run(use_graph=True)
    upd = [p.results(sim=z) for z in range(3)]
    ref = _single(wing, ts, MOTIONS[3], far=[(20, 2, 4)])
    assert rel(upd[1]["cl"][1:], ref["cl"][1:]) < 1e-12
    assert bits_equal(upd[0]["cl"], first[0]["cl"]) and bits_equal(upd[2]["cl"], first[2]["cl"])
    p.close()


def test_batched_plan_refusals(wing):
    from tinyvlm_amd import capi as H
    from tinyvlm_amd.uvlm_plan import UVLMPlan, time_step

    dt = time_step(wing, 1.0)
    with pytest.raises(H.TvlmError):
        UVLMPlan([wing], 11, dt, mode=0, sims=2)
    with pytest.raises(H.TvlmError):
        UVLMPlan([wing], 11, dt, mode=1, sims=0)
    p = UVLMPlan([wing], 11, dt, mode=1, sims=2)
    p.set_pitch_heave(**BASE, sim=0)
    with pytest.raises(H.TvlmError):
        p.run(use_graph=False)  # simulation 1 has no motion
    with pytest.raises(H.TvlmError):
        p.set_pitch_heave(**BASE, sim=2)
    with pytest.raises(H.TvlmError):
        p.results(sim=-1)
    p.close()


def test_batched_two_surfaces_match_single_plans_and_oracle(oracle):
    """A wing split spanwise into two lifting surfaces (per-surface offsets inside every simulation's
    block), three simulations, far-wake aggregation on both surfaces: each simulation against a
    one-simulation plan (1e-12) and, without aggregation, simulation 1 against the oracle (1e-11)."""
    from tinyvlm_amd import capi
    from tinyvlm_amd.uvlm_plan import UVLMPlan, time_step

    capi.init(0)
    v = oracle.rectangular_wing(12, 4, chord=1.0, half_span=2.0)
    a, b = np.asfortranarray(v[:7]), np.asfortranarray(v[6:])
    ts, dt = 61, time_step(a, 1.0)
    far = [(10, 2, 2), (20, 4, 4)]
    mots = MOTIONS[:3]
    print(f"[case] uvlm batched two surfaces, {len(mots)} sims, steps {ts - 1}", flush=True)
    for agg in (far, None):
        p = UVLMPlan([a, b], ts, dt, mode=1, sims=len(mots))
        for z, m in enumerate(mots):
            p.set_pitch_heave(**m, sim=z)
        if agg:
            p.set_far_wake(agg)
        p.run(use_graph=True)
        got = [p.results(sim=z) for z in range(len(mots))]
        p.close()
        for z, m in enumerate(mots):
            q = UVLMPlan([a, b], ts, dt, mode=1)
            q.set_pitch_heave(**m)
            if agg:
                q.set_far_wake(agg)
            q.run(use_graph=True)
            ref = q.results()
            q.close()
            assert rel(got[z]["cl"][1:], ref["cl"][1:]) < 1e-12 and rel(got[z]["gamma"], ref["gamma"]) < 1e-12
        if agg is None:
            m = mots[1]
            mot = [m[k] for k in ("u", "h_amp", "h_phase", "a_amp", "a_phase", "omega", "x_pivot")]
            cl_o, cm_o, g_o, _ = oracle.uvlm_run([a, b], mot, t_steps=ts)
            assert rel(got[1]["cl"][1:], cl_o[1:]) < 1e-11
            assert rel(got[1]["gamma"], g_o) < 1e-11
