"""GPU checks of continuous batching (tinyvlm_amd.serve.Server) on the steady-solve mix, and on
steady and unsteady requests served together.

- alone vs in the mix: every request served in a mix of meshes, arriving while others run,
  equals the same request solved alone (a one-request served plan): circulation within 1e-10
  relative to its largest value, |dCL| <= 1e-10, CM within 1e-10;
- against the plain FP64 path and the oracle (VLM::run restatement) for requests on the smallest
  mesh: circulation within 1e-10 relative, |dCL| <= 1e-10;
- overlap: requests were admitted while others were still on the device (more than one lane busy
  at once, batches of more than one request);
- memory admission: with a budget that holds only two plans, plans are evicted and re-created and
  every result is still the alone result;
- unsteady requests (whole UVLM simulations, 10 x 8 wing, 60 steps, different motions) served
  between steady ones on the same lanes: every CL(t) history equals the same simulation run alone
  (bitwise: the same kernels on the same operands) and the oracle's UVLM::run within 1e-9, and the
  steady requests still equal their alone results.
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def mix(tmp_path_factory):
    from tinyvlm_amd import capi as H
    from tinyvlm_amd import serve as S

    H.init(0)
    d = str(tmp_path_factory.mktemp("mix"))
    t = S.read_trace()
    meshes, reqs = {}, []
    for i in range(len(t["gap"])):
        n = int(t["n"][i])
        if n > 45:
            continue
        key = f"{S.MIX_FAMILIES[t['family'][i]]}_{n}"
        if key not in meshes:
            meshes[key] = [S.mix_mesh(S.MIX_FAMILIES[t["family"][i]], n, d)]
        reqs.append((i, key, float(np.deg2rad(np.float64(t["alpha_deg"][i]))), float(t["gap"][i])))
        if len(reqs) == 72:
            break
    return meshes, reqs


def _requests(reqs, gap_scale):
    from tinyvlm_amd.serve import Request

    out, t = [], 0.0
    for rid, key, a, g in reqs:
        t += g * gap_scale
        out.append(Request(rid, key, a, arrival=t))
    return out


def _alone(meshes, r):
    from tinyvlm_amd.steady_plan import MODE_SERVED, SteadyPlan

    p = SteadyPlan(meshes[r.mesh], 1, mode=MODE_SERVED)
    p.solve([r.alpha])
    res = p.results()
    p.close()
    return res


def _check_same(r, alone):
    g = alone["gamma"][0]
    assert r.info == 0
    assert np.max(np.abs(r.gamma - g)) <= TOL * np.max(np.abs(g)), r.rid
    assert abs(r.cl - alone["cl"][0]) <= TOL, r.rid
    assert np.max(np.abs(r.cm - alone["cm"][0])) <= TOL, r.rid


def test_mix_equals_alone_and_overlaps(mix):
    from tinyvlm_amd.serve import Server

    meshes, reqs = mix
    print(f"[case] serve mix {len(reqs)} requests over {len(meshes)} meshes", flush=True)
    requests = _requests(reqs, 2e-3)
    with Server(meshes, lanes=4, max_batch=8, keep_gamma=True) as srv:
        st = srv.run(requests)
    assert all(np.isfinite(r.finished) for r in requests)
    assert st.max_running >= 2 and max(st.batch_sizes) >= 2, (st.max_running, st.batch_sizes)
    # some request was admitted while another one, admitted earlier, was still running
    spans = sorted((r.admitted, r.finished, r.lane) for r in requests)
    assert any(b[0] < a[1] and b[2] != a[2] for a, b in zip(spans, spans[1:]))
    for r in requests:
        assert r.admitted >= r.arrival - 1e-9 and r.finished > r.admitted
        _check_same(r, _alone(meshes, r))


def test_mix_smallest_mesh_matches_plain_and_oracle(mix, oracle):
    from tinyvlm_amd.serve import Server
    from tinyvlm_amd.steady_plan import MODE_PLAIN_FP64, SteadyPlan

    meshes, reqs = mix
    requests = _requests(reqs, 0.0)
    with Server(meshes, lanes=3, max_batch=16, keep_gamma=True) as srv:
        srv.run(requests, realtime=False)
    small = [r for r in requests if r.mesh == "elliptic_16"][:4]
    assert small
    print(f"[case] serve vs plain/oracle elliptic_16 x {len(small)}", flush=True)
    p = SteadyPlan(meshes["elliptic_16"], len(small), mode=MODE_PLAIN_FP64)
    p.solve([r.alpha for r in small])
    plain = p.results()
    p.close()
    for j, r in enumerate(small):
        g = plain["gamma"][j]
        assert np.max(np.abs(r.gamma - g)) <= TOL * np.max(np.abs(g))
        assert abs(r.cl - plain["cl"][j]) <= TOL
        go, clo, _, _ = oracle.vlm_run(meshes["elliptic_16"], r.alpha)
        assert np.max(np.abs(r.gamma - go)) <= TOL * np.max(np.abs(go))
        assert abs(r.cl - clo) <= TOL


def test_mix_under_small_memory_budget(mix):
    from tinyvlm_amd.serve import Server

    meshes, reqs = mix
    reqs = reqs[:36]
    sizes = {}
    with Server(meshes, lanes=2, max_batch=4, budget_bytes=1 << 40) as probe:
        for k in meshes:
            sizes[k] = probe._estimate(k)
    budget = 2 * max(sizes.values()) + 1
    print(f"[case] serve mix with budget {budget} bytes", flush=True)
    requests = _requests(reqs, 1e-3)
    with Server(meshes, lanes=2, max_batch=4, budget_bytes=budget, keep_gamma=True) as srv:
        st = srv.run(requests)
        assert srv.used_bytes <= budget
    assert st.plans_evicted > 0 and st.peak_bytes <= budget
    for r in requests:
        _check_same(r, _alone(meshes, r))


def test_unsteady_requests_in_the_mix(mix, oracle):
    from tinyvlm_amd.serve import Request, Server
    from tinyvlm_amd.uvlm_plan import MODE_FAST, UVLMPlan, time_step

    meshes, reqs = mix
    wing = oracle.rectangular_wing(8, 10, chord=1.0, half_span=2.0)
    ts, dt = 61, time_step(wing, 1.0)
    meshes = dict(meshes, wing10x8=[wing])
    a_amp = 2.0 * math.pi / 180.0
    motions = [dict(u=1.0, h_amp=0.1 * (1 + k % 3), h_phase=0.0, a_amp=a_amp * (k % 2 + 1), a_phase=math.pi / 2,
                    omega=1.0 + 0.25 * k, x_pivot=0.25) for k in range(6)]
    requests = []
    t = 0.0
    for j, (rid, key, a, g) in enumerate(reqs[:30]):
        t += g * 2e-3
        requests.append(Request(rid, key, a, arrival=t))
        if j % 5 == 0:
            requests.append(Request(10000 + j, "wing10x8", arrival=t, kind="unsteady", motion=motions[j // 5]))
    print(f"[case] serve {len(requests)} steady + unsteady requests", flush=True)
    with Server(meshes, lanes=4, max_batch=8, keep_gamma=True, unsteady={"wing10x8": (ts, dt)}) as srv:
        st = srv.run(requests)
    assert st.max_running >= 2
    unsteady = [r for r in requests if r.kind == "unsteady"]
    assert len(unsteady) == 6
    for r in unsteady:
        p = UVLMPlan([wing], ts, dt, mode=MODE_FAST)
        p.set_pitch_heave(**r.motion)
        p.run(use_graph=True)
        alone = p.results()
        p.close()
        assert np.array_equal(r.cl_t, alone["cl"]), r.rid
        assert np.array_equal(r.cm_t, alone["cm"]), r.rid
        mot = [r.motion[k] for k in ("u", "h_amp", "h_phase", "a_amp", "a_phase", "omega", "x_pivot")]
        cl_o, _, _, _ = oracle.uvlm_run([wing], mot, t_steps=ts)
        assert np.max(np.abs(r.cl_t[1:] - cl_o[1:])) <= 1e-9, r.rid
    for r in requests:
        if r.kind == "steady":
            _check_same(r, _alone(meshes, r))


def test_unsteady_requests_batched_per_lane(oracle):
    """unsteady_batch = 4, 2 lanes, 11 simulations of different motions waiting from the start: they
    run as batches of several simulations per plan; each CL(t), CM(t) agrees with its alone
    one-simulation plan to 1e-12 relative (the batched LU rounds differently) and with the oracle."""
    from tinyvlm_amd.serve import Request, Server
    from tinyvlm_amd.uvlm_plan import MODE_FAST, UVLMPlan, time_step

    wing = oracle.rectangular_wing(8, 10, chord=1.0, half_span=2.0)
    ts, dt = 61, time_step(wing, 1.0)
    a_amp = 2.0 * math.pi / 180.0
    motions = [dict(u=1.0, h_amp=0.05 * (1 + k % 4), h_phase=0.1 * k, a_amp=a_amp * (k % 3 + 1), a_phase=math.pi / 2,
                    omega=1.0 + 0.125 * k, x_pivot=0.25) for k in range(11)]
    requests = [Request(k, "w", kind="unsteady", motion=m) for k, m in enumerate(motions)]
    print(f"[case] serve {len(requests)} unsteady requests, batches of up to 4 on 2 lanes", flush=True)
    with Server({"w": [wing]}, lanes=2, unsteady={"w": (ts, dt)}, unsteady_batch=4) as srv:
        st = srv.run(requests, realtime=False)
        assert all(k[2] in (1, 2, 4) for k in srv.uslots)
    assert max(st.batch_sizes) == 4 and sum(st.batch_sizes) == 11
    for r in requests:
        p = UVLMPlan([wing], ts, dt, mode=MODE_FAST)
        p.set_pitch_heave(**r.motion)
        p.run(use_graph=True)
        alone = p.results()
        p.close()
        scale = np.max(np.abs(alone["cl"]))
        assert np.max(np.abs(r.cl_t - alone["cl"])) <= 1e-12 * scale, r.rid
        assert np.max(np.abs(r.cm_t - alone["cm"])) <= 1e-12 * max(np.max(np.abs(alone["cm"])), 1e-300), r.rid
        mot = [r.motion[k] for k in ("u", "h_amp", "h_phase", "a_amp", "a_phase", "omega", "x_pivot")]
        cl_o, _, _, _ = oracle.uvlm_run([wing], mot, t_steps=ts)
        assert np.max(np.abs(r.cl_t[1:] - cl_o[1:])) <= 1e-9, r.rid


def test_large_mix_meshes_through_the_server(tmp_path):
    """The mix's large meshes (N = 4096, 8100, 16384: rectangular 64 x 64, swept 90 x 90, the
    pivot-heavy elliptic 128 x 128) served beside small ones through Server: every request equals
    the same request solved alone, and the plain FP64 solve within 1e-10 (circulation, relative;
    CL absolute); the refinement test passed on the fixed corrections for all of them."""
    from tinyvlm_amd import capi as H
    from tinyvlm_amd import serve as S
    from tinyvlm_amd.steady_plan import MODE_PLAIN_FP64, SteadyPlan

    H.init(0)
    d = str(tmp_path)
    keys = [("rectangular", 64), ("swept25", 90), ("elliptic", 128), ("elliptic", 16), ("swept25", 32)]
    meshes = {f"{f}_{n}": [S.mix_mesh(f, n, d)] for f, n in keys}
    spec = [("elliptic_128", 4.0), ("rectangular_64", -3.0), ("swept25_90", 7.5), ("elliptic_16", 1.0),
            ("elliptic_128", -1.5), ("swept25_32", 9.0), ("rectangular_64", 2.5), ("swept25_90", 0.5)]
    requests = [S.Request(i, k, float(np.deg2rad(a)), arrival=1e-3 * i) for i, (k, a) in enumerate(spec)]
    print(f"[case] serve large mix meshes: {len(requests)} requests", flush=True)
    with S.Server(meshes, lanes=3, keep_gamma=True) as srv:
        st = srv.run(requests)
    assert st.max_running >= 2
    for r in requests:
        assert r.refine == 0, (r.rid, r.mesh, r.refine)
        alone = _alone(meshes, r)
        _check_same(r, alone)
        p = SteadyPlan(meshes[r.mesh], 1, mode=MODE_PLAIN_FP64)
        p.solve([r.alpha])
        plain = p.results()
        p.close()
        g = plain["gamma"][0]
        print(f"[case]   {r.mesh} rid {r.rid}: served vs plain |dCL| {abs(r.cl - plain['cl'][0]):.2e}", flush=True)
        assert np.max(np.abs(r.gamma - g)) <= TOL * np.max(np.abs(g)), r.rid
        assert abs(r.cl - plain["cl"][0]) <= TOL, r.rid


@pytest.mark.parametrize("processes,wake", [(False, "exact"), (True, "exact"), (True, "lattice")],
                         ids=["threads", "processes", "processes-lattice"])
def test_hb_requests_served_beside_steady(root, oracle, processes, wake):
    """Harmonic-balance requests (the reference script's wing 9x5 + flap 3x5, H = 2, several
    airspeeds) served on worker threads, or on two worker processes, beside steady requests: every
    request starts cold, so its loads equal (bit for bit) the same solve on a fresh HBVLM alone; the
    steady results match their alone solves as in the other mix tests.  With wake = "lattice" the
    configuration's solvers evaluate the wake induction over the vortex lattice (HBVLM.wake_form)."""
    import os

    from tinyvlm_amd import capi
    from tinyvlm_amd import serve as S
    from tinyvlm_amd.hb import HBVLM

    capi.init(0)
    paths = [os.path.join(root, "mesh", m) for m in ("3dof_wing_9x5.x", "3dof_flap_3x5.x")]
    H, scale = 2, 1.0 / 0.127
    K = 2 * H + 1
    X = np.zeros((3, K))
    X[0, 1] = 0.01 / 0.127
    X[1, 2] = 0.02
    X[2, 1] = 0.05
    speeds = [5.0, 5.5, 6.0, 6.5, 7.0, 7.5]
    print(f"[case] serve: hb requests beside steady, worker processes {processes}, wake {wake}", flush=True)
    alone = {}
    for U in speeds:
        h = HBVLM("hip", paths)
        h.wake_form = wake
        h.init(H, scale)
        h.use_graph = False  # as the serving loop's solvers (the replay gives the same bits anyway)
        F = np.zeros((3, K))
        h.run(0.8, U, X, F)
        alone[U] = (F, h.iterations)
    v = oracle.rectangular_wing(16, 8, chord=1.0, half_span=4.0)
    reqs = [S.Request(i, "r16", float(np.deg2rad(a)), arrival=0.0) for i, a in enumerate((1.0, 2.0, 3.0, 4.0))]
    reqs += [S.Request(100 + k, "wf", kind="hb", hb={"omega": 0.8, "U": U, "dyn": X}) for k, U in enumerate(speeds)]
    with S.Server({"r16": [v]}, lanes=2, hb={"wf": {"meshes": paths, "harmonics": H, "scaling": scale, "wake": wake}},
                  hb_workers=2 if processes else 4, hb_processes=processes) as srv:
        srv.run(reqs, realtime=False)
    for r in reqs[4:]:
        F, it = alone[r.hb["U"]]
        assert np.array_equal(r.forces, F), (r.hb["U"], np.max(np.abs(r.forces - F)))
        assert r.iterations == it
    for r in reqs[:4]:
        g, cl, cm, _ = oracle.vlm_run([v], r.alpha)
        assert abs(r.cl - cl) <= 1e-10 * abs(cl)
