"""CPU checks of the tile-size autotuner's bookkeeping (tinyvlm_amd.autotune) and of the library's
tile-target setters: the JSON cache round trip, a corrupt cache starting empty, workload keys that
separate shapes, step counts and far-wake schedules, and the setters' range checks (no GPU work)."""
import json

import numpy as np


class _FakePlan:
    def __init__(self, ns, nc, t_steps, far=None, sims=1):
        self._verts = [np.zeros((ns + 1, nc + 1, 4))]
        self.t_steps = t_steps
        self.far_wake = far
        self.sims = sims


def test_cache_round_trip_and_corrupt_file(tmp_path):
    from tinyvlm_amd import autotune as A

    p = tmp_path / "sub" / "tune.json"
    c = A.Cache(str(p))
    assert c.get("k") is None
    c.put("k", {"best": 1024, "sim_ms": {"512": 2.0, "1024": 1.5}})
    c2 = A.Cache(str(p))
    assert c2.get("k")["best"] == 1024
    assert json.loads(p.read_text())["k"]["sim_ms"]["1024"] == 1.5
    p.write_text("{not json")
    assert A.Cache(str(p)).data == {}
    m = A.Cache(None)  # memory only
    m.put("x", {"best": 2})
    assert m.get("x") == {"best": 2}


def test_uvlm_keys_separate_workloads():
    from tinyvlm_amd import autotune as A

    dev = "MI355X|gfx950|cu256"
    k = A.uvlm_key(_FakePlan(40, 10, 501), dev)
    assert k == A.uvlm_key(_FakePlan(40, 10, 501), dev)
    assert k != A.uvlm_key(_FakePlan(40, 10, 201), dev)
    assert k != A.uvlm_key(_FakePlan(40, 12, 501), dev)
    assert k != A.uvlm_key(_FakePlan(40, 10, 501, [(80, 2, 4)]), dev)
    assert k != A.uvlm_key(_FakePlan(40, 10, 501), "other")
    assert k != A.uvlm_key(_FakePlan(40, 10, 501, sims=64), dev)


def test_tile_target_setters_round_trip_and_range():
    from tinyvlm_amd import autotune as A
    from tinyvlm_amd import capi

    lib = capi.lib()
    w0, l0 = A.get_targets()
    assert (w0, l0) == (2048, 1024)
    try:
        A.set_targets(4096, 512)
        assert A.get_targets() == (4096, 512)
        A.set_targets(0, 0)
        assert A.get_targets() == (2048, 1024)
        import ctypes

        assert lib.tvlm_set_tile_targets(ctypes.c_int32(10), ctypes.c_int32(0)) != 0
        assert lib.tvlm_set_tile_targets(ctypes.c_int32(0), ctypes.c_int32(1 << 20)) != 0
        assert A.get_targets() == (2048, 1024)
    finally:
        A.set_targets(0, 0)
