"""GPU checks of the tile-size autotuner (tinyvlm_amd.autotune): it picks one of its candidates,
leaves the plan on that target (the same CL(t) bit for bit as a plan set to it by hand, and the
default target's within 1e-12 relative: the candidates differ only in tile order), reuses a cached
choice without timing, and the ring-by-ring wake tuner restores rhs and keeps the result."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32 = np.float32
A_AMP = float(F32(2) * F32(3.141593) / F32(180))
MOTION = dict(u=1.0, h_amp=0.1, h_phase=0.0, a_amp=A_AMP, a_phase=math.pi / 2, omega=1.0, x_pivot=0.25)


@pytest.fixture(scope="module")
def torch():
    import torch as t

    return t


@pytest.fixture(scope="module")
def H():
    from tinyvlm_amd import capi

    capi.init(0)
    return capi


def _plan(oracle, ts=121, far=None):
    from tinyvlm_amd import capi
    from tinyvlm_amd.uvlm_plan import UVLMPlan, time_step

    capi.init(0)
    v = oracle.rectangular_wing(40, 10, chord=1.0, half_span=2.0)
    p = UVLMPlan([v], ts, time_step(v, 1.0), mode=1)
    p.set_pitch_heave(**MOTION)
    if far:
        p.set_far_wake(far)
    return p


def test_tune_uvlm_picks_a_candidate_and_caches_it(oracle):
    from tinyvlm_amd import autotune as A

    print("[case] autotune uvlm lattice target", flush=True)
    cache = A.Cache(None)
    p = _plan(oracle, far=[(20, 2, 4)])
    e = A.tune_uvlm(p, candidates=(512, 1024, 2048), cache=cache)
    assert not e["cached"] and e["best"] in (512, 1024, 2048) and set(e["sim_ms"]) == {"512", "1024", "2048"}
    p.run(use_graph=True)
    cl = p.results()["cl"]
    q = _plan(oracle, far=[(20, 2, 4)])
    q.set_lattice_target(e["best"])
    q.run(use_graph=True)
    assert np.array_equal(q.results()["cl"], cl)
    d = _plan(oracle, far=[(20, 2, 4)])
    d.run(use_graph=True)
    cl_d = d.results()["cl"]
    assert np.max(np.abs(cl - cl_d)) <= 1e-12 * np.max(np.abs(cl_d))
    # a second plan of the same workload takes the cached choice without running anything
    r = _plan(oracle, far=[(20, 2, 4)])
    runs = []
    orig = r.run
    r.run = lambda *a, **k: (runs.append(1), orig(*a, **k))
    e2 = A.tune_uvlm(r, candidates=(512, 1024, 2048), cache=cache)
    assert e2["cached"] and e2["best"] == e["best"] and not runs
    for x in (p, q, d, r):
        x.close()


def test_tune_wake_influence_restores_rhs(H, torch, oracle):
    from tinyvlm_amd import autotune as A

    print("[case] autotune ring-by-ring wake target", flush=True)
    dt = torch.float64
    ns, nc, rows = 24, 6, 60
    v = oracle.rectangular_wing(ns, nc, chord=1.0, half_span=3.0)
    verts = [H.to_device(v, dt)]
    colloc, normals, areas = [H.colmajor((ns, nc, 3), dt)], [H.colmajor((ns, nc, 3), dt)], [H.colmajor((ns, nc), dt)]
    H.mesh_metrics(0.05, verts, colloc, normals, areas)
    rng = np.random.default_rng(0)
    wv = np.zeros((ns + 1, rows + 1, 4), order="F")
    for r in range(rows + 1):
        wv[:, r, 0] = 1.2 + 0.1 * r
        wv[:, r, 1] = np.linspace(0, 3.0, ns + 1)
        wv[:, r, 3] = 1
    gw = np.asfortranarray(rng.standard_normal((ns, rows)))
    wake, gwake = [H.to_device(wv, dt)], [H.to_device(gw, dt)]
    rhs = torch.zeros(ns * nc, dtype=dt, device="cuda")
    H.rhs_assemble_wake_influence(rhs, gwake, colloc, normals, wake, [True], rows)
    ref = rhs.clone()
    try:
        e = A.tune_wake_influence(rhs, gwake, colloc, normals, wake, [True], rows, cache=A.Cache(None))
        torch.cuda.synchronize()
        assert torch.equal(rhs, ref)  # restored
        assert A.get_targets()[0] == e["best"]
        rhs.zero_()
        H.rhs_assemble_wake_influence(rhs, gwake, colloc, normals, wake, [True], rows)
        assert float((rhs - ref).abs().max() / ref.abs().max()) <= 1e-13
    finally:
        A.set_targets(0, 0)
