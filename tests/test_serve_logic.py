"""CPU checks of the serving loop's host logic (tinyvlm_amd.serve.Server) with stand-in plans in
place of the device plans: every request is answered from its own batch slot, batches respect the
mesh's cap and the lane count, no plan takes a second batch while one is running, the memory budget
holds (least recently used idle plans are evicted), admission is oldest-first across meshes, and
unsteady requests run on plans bound to one lane."""
import itertools

import numpy as np
import pytest


class FakePlan:
    live = []

    def __init__(self, verts_list, batch, mode=1, stream=None):
        self.N = sum((v.shape[0] - 1) * (v.shape[1] - 1) for v in verts_list)
        self.batch, self.stream = batch, stream
        self.polls = 0
        self.busy = False
        self.closed = False
        self.alphas = None
        FakePlan.live.append(self)

    def solve(self, alpha, u_inf=1.0, rho=1.0):
        assert not self.busy and not self.closed
        assert 1 <= len(alpha) <= self.batch
        self.alphas = np.asarray(alpha, dtype=np.float64)
        self.busy = True
        self.polls = 3

    def done(self):
        self.polls -= 1
        return self.polls <= 0

    def results(self, gamma=True):
        assert self.busy
        self.busy = False
        a = self.alphas
        out = {"cl": 2 * np.pi * a, "cm": np.stack([a, -a, 0 * a], axis=1), "info": np.zeros(len(a), np.int32)}
        if gamma:
            out["gamma"] = np.repeat(a[:, None], self.N, axis=1)
        return out

    def refinement(self):
        # a request with alpha < -0.05 pretends its refinement needed the FP64 fallback
        a = self.alphas
        return {"status": np.where(a < -0.05, 2, 0).astype(np.int32), "corrections": np.full(len(a), 4, np.int32),
                "backward_error": np.zeros(len(a))}

    def set_stream(self, s):
        assert not self.busy
        self.stream = s

    def footprint(self):
        return 12 * self.N * self.N * self.batch + 400 * self.N * self.batch

    def close(self):
        assert not self.busy
        self.closed = True


class FakeUVLM:
    live = []

    def __init__(self, verts_list, t_steps, dt, mode=1, stream=None, sims=1):
        self.stream, self.t_steps, self.sims = stream, t_steps, sims
        self.motion = [None] * sims
        self.runs = 0
        self.polls = 0
        self.closed = False
        FakeUVLM.live.append(self)

    def set_pitch_heave(self, sim=None, **m):
        for z in range(self.sims) if sim is None else [sim]:
            self.motion[z] = m

    def run(self, use_graph=True):
        assert all(m is not None for m in self.motion)
        self.runs += 1
        self.polls = 2

    def done(self):
        self.polls -= 1
        return self.polls <= 0

    def results(self, sim=0):
        n = self.t_steps - 1
        cl = np.full(n, self.motion[sim]["h_amp"])
        return {"cl": cl, "cm": np.zeros((n, 3)), "gamma": np.zeros(4)}

    def close(self):
        self.closed = True


@pytest.fixture
def S(monkeypatch):
    from tinyvlm_amd import serve

    counter = itertools.count(1000)
    monkeypatch.setattr(serve, "stream_create", lambda: next(counter))
    monkeypatch.setattr(serve, "stream_destroy", lambda s: None)
    monkeypatch.setattr(serve, "SteadyPlan", FakePlan)
    monkeypatch.setattr(serve, "UVLMPlan", FakeUVLM)
    FakePlan.live = []
    FakeUVLM.live = []
    return serve


def _mesh(n):
    v = np.zeros((n + 1, n + 1, 4))
    return [v]


def _requests(S, spec):
    return [S.Request(i, key, alpha, arrival=0.0) for i, (key, alpha) in enumerate(spec)]


def test_results_batches_and_lanes(S):
    meshes = {"a": _mesh(16), "b": _mesh(32), "c": _mesh(64)}
    rng = np.random.default_rng(0)
    spec = [(str(rng.choice(list(meshes))), float(rng.uniform(-0.1, 0.2))) for _ in range(200)]
    reqs = _requests(S, spec)
    with S.Server(meshes, lanes=3, max_batch=5, budget_bytes=1 << 40, keep_gamma=True) as srv:
        st = srv.run(reqs, realtime=False)
        assert st.max_running <= 3
        assert max(st.batch_sizes) <= 5 and sum(st.batch_sizes) == 200
    for r in reqs:
        assert r.cl == pytest.approx(2 * np.pi * r.alpha)
        assert r.cm[0] == pytest.approx(r.alpha) and r.cm[1] == pytest.approx(-r.alpha)
        assert np.all(r.gamma == r.alpha)
        assert r.finished >= r.admitted >= r.arrival


def test_batch_cap_follows_mesh_size(S):
    assert S.batch_cap(256) == 32 and S.batch_cap(4096, plan_bytes=4 << 30) == 21
    assert S.batch_cap(8100, plan_bytes=4 << 30) == 4 and S.batch_cap(16384, plan_bytes=4 << 30) == 1
    assert S.batch_cap(4096) == 32 and S.batch_cap(8100) == 21 and S.batch_cap(16384) == 2
    for n in (1500, 2025, 8100, 16384):  # batches under 8 keep the LU's panels cooperative
        for pb in (1 << 30, 4 << 30, 16 << 30):
            cap = S.batch_cap(n, plan_bytes=pb)
            assert cap >= 8 or cap * ((n + 511) // 512) <= 64


def test_memory_budget_and_eviction(S):
    meshes = {"a": _mesh(32), "b": _mesh(32), "c": _mesh(32)}  # any two plans fit, three do not
    with S.Server(meshes, lanes=2, max_batch=4, budget_bytes=1 << 40) as probe:
        est = {k: probe._estimate(k) for k in meshes}
    budget = 2 * max(est.values())
    spec = [(k, 0.01 * i) for i, k in enumerate(["a", "b", "c"] * 20)]
    reqs = _requests(S, spec)
    with S.Server(meshes, lanes=2, max_batch=4, budget_bytes=budget) as srv:
        st = srv.run(reqs, realtime=False)
        assert srv.used_bytes <= budget
    assert st.plans_evicted > 0 and st.peak_bytes <= budget
    assert all(np.isfinite(r.cl) for r in reqs)
    with pytest.raises(RuntimeError, match="exceeds the memory budget"):
        with S.Server(meshes, lanes=1, budget_bytes=10) as srv:
            srv.run(_requests(S, [("a", 0.0)]), realtime=False)


def test_oldest_request_first(S):
    meshes = {"a": _mesh(8), "b": _mesh(8)}
    reqs = [S.Request(0, "b", 0.1, arrival=0.0), S.Request(1, "a", 0.2, arrival=0.001),
            S.Request(2, "b", 0.3, arrival=0.002)]
    with S.Server(meshes, lanes=1, max_batch=8, budget_bytes=1 << 40) as srv:
        srv.run(reqs, realtime=False)
    # one lane: the b batch (requests 0 and 2, both waiting) goes first, then a
    assert reqs[0].admitted == reqs[2].admitted < reqs[1].admitted
    assert reqs[0].batch == 2 and reqs[1].batch == 1


def test_unsteady_requests_stay_on_their_lane(S):
    meshes = {"w": _mesh(8), "s": _mesh(16)}
    reqs = []
    for i in range(12):
        reqs.append(S.Request(i, "w", kind="unsteady", motion=dict(u=1.0, h_amp=0.01 * i)))
        reqs.append(S.Request(100 + i, "s", 0.05 * i))
    with S.Server(meshes, lanes=3, budget_bytes=1 << 40, unsteady={"w": (11, 0.1)}) as srv:
        srv.run(reqs, realtime=False)
        for (key, lane, sims), slot in srv.uslots.items():
            assert slot.plan.stream == srv.lanes[lane] and sims == 1
    for r in reqs:
        if r.kind == "unsteady":
            assert len(r.cl_t) == 10 and np.all(r.cl_t == r.motion["h_amp"])
        else:
            assert r.cl == pytest.approx(2 * np.pi * r.alpha)


def test_unsteady_batches_take_waiting_simulations_of_one_mesh(S):
    """unsteady_batch = 4: 11 waiting simulations run as 4 + 4 + 3 on one lane's plan of 4
    simulations (3 waiting take the power of two above), each result from its own simulation."""
    meshes = {"w": _mesh(8)}
    reqs = [S.Request(i, "w", kind="unsteady", motion=dict(u=1.0, h_amp=0.01 * (i + 1))) for i in range(11)]
    with S.Server(meshes, lanes=1, budget_bytes=1 << 40, unsteady={"w": (11, 0.1)}, unsteady_batch=4) as srv:
        st = srv.run(reqs, realtime=False)
        assert sorted(k[2] for k in srv.uslots) == [4] and FakeUVLM.live[0].runs == 3
    assert st.batch_sizes == [4, 4, 3]
    for r in reqs:
        assert len(r.cl_t) == 10 and np.all(r.cl_t == r.motion["h_amp"])
    assert [r.batch for r in reqs] == [4] * 8 + [3] * 3


def test_unsteady_batch_sizes_are_powers_of_two(S):
    meshes = {"w": _mesh(8)}
    with S.Server(meshes, lanes=1, budget_bytes=1 << 40, unsteady={"w": (11, 0.1)}, unsteady_batch=6) as srv:
        assert [srv.usims(n) for n in (1, 2, 3, 4, 5, 6, 7, 100)] == [1, 2, 4, 4, 6, 6, 6, 6]
        e1, e2, e4 = (srv._uestimate("w", k) for k in (1, 2, 4))
        assert e4 - e1 == 3 * (e2 - e1) > 0  # linear in the simulations held
    with pytest.raises(ValueError):
        S.Server(meshes, lanes=1, budget_bytes=1 << 40, unsteady_batch=0)


def test_partial_unsteady_batch_repeats_the_first_motion(S):
    meshes = {"w": _mesh(8)}
    reqs = [S.Request(i, "w", kind="unsteady", motion=dict(u=1.0, h_amp=0.1 * (i + 1))) for i in range(3)]
    with S.Server(meshes, lanes=1, budget_bytes=1 << 40, unsteady={"w": (11, 0.1)}, unsteady_batch=8) as srv:
        srv.run(reqs, realtime=False)
        (key, slot), = srv.uslots.items()
        assert key[2] == 4 and slot.plan.runs == 1
        assert [m["h_amp"] for m in slot.plan.motion] == [0.1, 0.2, 0.30000000000000004, 0.1]
    assert [float(r.cl_t[0]) for r in reqs] == [0.1, 0.2, 0.30000000000000004]


def test_request_errors(S):
    meshes = {"w": _mesh(8)}
    with S.Server(meshes, lanes=1, budget_bytes=1 << 40) as srv:
        with pytest.raises(KeyError):
            srv.run([S.Request(0, "nope", 0.0)], realtime=False)
        with pytest.raises(KeyError):
            srv.run([S.Request(0, "w", kind="unsteady", motion={"u": 1.0})], realtime=False)
        with pytest.raises(ValueError):
            srv.run([S.Request(0, "w", kind="other")], realtime=False)
    with pytest.raises(KeyError):
        S.Server(meshes, lanes=1, budget_bytes=1 << 40, unsteady={"x": (11, 0.1)})


def test_refinement_status_is_recorded_and_counted(S):
    meshes = {"a": _mesh(16)}
    alphas = [-0.08, 0.02, -0.06, 0.1, 0.0]
    reqs = _requests(S, [("a", x) for x in alphas])
    with S.Server(meshes, lanes=1, max_batch=2, budget_bytes=1 << 40) as srv:
        st = srv.run(reqs, realtime=False)
    assert [r.refine for r in reqs] == [2 if x < -0.05 else 0 for x in alphas]
    assert st.fallbacks == 2 and st.more_corrections == 0 and st.refactored == 0 and st.unresolved == 0


def test_unsteady_plan_larger_than_budget_raises(S):
    meshes = {"w": _mesh(8)}
    with S.Server(meshes, lanes=1, budget_bytes=1000, unsteady={"w": (11, 0.1)}) as srv:
        with pytest.raises(RuntimeError, match="exceeds the memory budget"):
            srv.run([S.Request(0, "w", kind="unsteady", motion=dict(u=1.0, h_amp=0.0))], realtime=False)


def test_idle_unsteady_plans_are_evicted_for_steady_plans(S):
    """An idle unsteady plan holding most of the budget is evicted when a steady plan needs room
    (before the fix, steady requests waited for ever behind it)."""
    meshes = {"w": _mesh(40), "s": _mesh(40)}
    with S.Server(meshes, lanes=1, max_batch=1, budget_bytes=1 << 40, unsteady={"w": (11, 0.1)}) as probe:
        u_need, s_need = probe._uestimate("w"), probe._estimate("s")
    budget = max(u_need, s_need) + min(u_need, s_need) // 2  # one plan of either kind fits, not both
    reqs = [S.Request(0, "w", kind="unsteady", motion=dict(u=1.0, h_amp=0.5)),
            S.Request(1, "s", 0.05, arrival=0.001)]
    with S.Server(meshes, lanes=1, max_batch=1, budget_bytes=budget, unsteady={"w": (11, 0.1)}) as srv:
        st = srv.run(reqs, realtime=False)
        assert srv.used_bytes <= budget
    assert st.plans_evicted == 1 and FakeUVLM.live[0].closed
    assert np.all(reqs[0].cl_t == 0.5) and reqs[1].cl == pytest.approx(2 * np.pi * 0.05)


def test_stalled_admission_raises_instead_of_spinning(S, monkeypatch):
    """If nothing runs and no waiting queue can be placed, nothing will ever free a lane or memory:
    the loop must raise, not sleep for ever."""
    meshes = {"a": _mesh(8)}
    with S.Server(meshes, lanes=2, budget_bytes=1 << 40) as srv:
        monkeypatch.setattr(srv, "_slot_for", lambda key, lane, now: None)
        with pytest.raises(RuntimeError, match="serving stalled"):
            srv.run(_requests(S, [("a", 0.0), ("a", 0.1)]), realtime=False)


class FakeHB:
    instances = []
    lock = None

    def __init__(self, backend, meshes):
        import threading

        assert backend == "hip" and len(meshes) == 2
        self.thread = threading.get_ident()
        self.stream = None  # torch.cuda.stream(None) is a no-op
        self.use_graph = True
        self.K = None
        self.iterations = 0
        self.warm = 0.0
        FakeHB.instances.append(self)

    def init(self, harmonics, scaling):
        self.K = 2 * harmonics + 1

    def reset(self):
        self.warm = 0.0

    def run(self, omega, U, dyn, force_t):
        import time as _t

        assert threading_ident() == self.thread  # a worker's solver stays in its thread
        assert not self.use_graph  # served solvers never capture
        _t.sleep(0.002)
        force_t[0, :] = U + self.warm  # a result that would show a missing reset
        force_t[1, :] = omega
        force_t[2, :] = dyn[0, 1]
        self.warm += 1.0  # the next run would be warm-started
        self.iterations = 7
        if U < 0:
            raise RuntimeError("HB solve diverged")
        return force_t


def threading_ident():
    import threading

    return threading.get_ident()


def test_hb_requests_run_on_workers_cold_and_beside_steady(S, monkeypatch):
    import tinyvlm_amd.hb as hbmod

    monkeypatch.setattr(hbmod, "HBVLM", FakeHB)
    FakeHB.instances = []
    meshes = {"a": _mesh(16)}
    hbcfg = {"wf": {"meshes": ["w.x", "f.x"], "harmonics": 2, "scaling": 1.0}}
    dyn = np.zeros((3, 5))
    dyn[0, 1] = 0.1
    reqs = [S.Request(i, "a", 0.01 * i, arrival=0.0) for i in range(10)]
    reqs += [S.Request(100 + k, "wf", kind="hb", hb={"omega": 0.8, "U": 1.0 + k, "dyn": dyn}, arrival=0.0)
             for k in range(9)]
    with S.Server(meshes, lanes=2, max_batch=4, budget_bytes=1 << 40, hb=hbcfg, hb_workers=3) as srv:
        srv.run(reqs, realtime=False)
    for r in reqs[10:]:
        # each request started cold (a stale warm start would add to the first row)
        assert np.all(r.forces[0] == r.hb["U"]) and np.all(r.forces[1] == 0.8) and np.all(r.forces[2] == 0.1)
        assert r.iterations == 7 and r.finished >= r.admitted >= 0
    for r in reqs[:10]:
        assert r.cl == pytest.approx(2 * np.pi * r.alpha)
    assert 1 <= len(FakeHB.instances) <= 3  # one solver per worker thread and configuration


def test_hb_request_errors(S, monkeypatch):
    import tinyvlm_amd.hb as hbmod

    monkeypatch.setattr(hbmod, "HBVLM", FakeHB)
    hbcfg = {"wf": {"meshes": ["w.x", "f.x"], "harmonics": 1, "scaling": 1.0}}
    with S.Server({"a": _mesh(8)}, lanes=1, budget_bytes=1 << 40, hb=hbcfg) as srv:
        with pytest.raises(KeyError):
            srv.run([S.Request(0, "nope", kind="hb", hb={"omega": 1, "U": 1, "dyn": np.zeros((3, 3))})])
        with pytest.raises(KeyError):
            srv.run([S.Request(0, "wf", kind="hb")])
        with pytest.raises(RuntimeError, match="diverged"):
            srv.run([S.Request(0, "wf", kind="hb", hb={"omega": 1, "U": -1, "dyn": np.zeros((3, 3))})],
                    realtime=False)
