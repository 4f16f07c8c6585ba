"""CPU checks of the LU factorisation plan (tvlm_getrf_plan, host only): the path a getrf call takes,
the form of every panel and the exchange buffers of the cooperative panels.

A batch that started with one-group panels and switched to cooperative ones part-way once ran
with exchange buffers sized for the first panel only.  The plan is now computed before anything is
queued and getrf executes exactly it; these checks sweep (n, batch, precision) and assert that
every cooperative panel has a counter set of its own inside the allocation, that a factorisation
never switches to cooperative panels part-way, and that the group counts fit one launch.
"""
import ctypes

import numpy as np
import pytest

NB = 64
MAX_GROUPS = 64  # work-groups of one cooperative launch
MIN_ROWS = 1024  # shorter panels stay in one work-group
ROWS_PER_GROUP = {4: 512, 8: 256}
PATH_LOOKAHEAD, PATH_COOP, PATH_PANELS = 0, 1, 2


def plan(n, batch, elem_bytes):
    from tinyvlm_amd import capi

    lib = capi.lib()
    path, steps = ctypes.c_int32(), ctypes.c_int32()
    cap = (n + NB - 1) // NB + 1
    groups = (ctypes.c_int32 * cap)()
    ctr = (ctypes.c_int32 * cap)()
    ws = ctypes.c_int64()
    capi.check(lib.tvlm_getrf_plan(ctypes.c_int64(n), ctypes.c_int32(batch), ctypes.c_int32(elem_bytes),
                                   ctypes.byref(path), ctypes.byref(steps), groups, ctr, ctypes.c_int32(cap),
                                   ctypes.byref(ws)), "tvlm_getrf_plan")
    k = steps.value
    return path.value, np.array(groups[:k]), np.array(ctr[:k]), ws.value


SIZES = [1, 63, 64, 65, 1000, 1024, 1025, 1100, 2048, 2049, 4096, 5000, 6144, 8100, 8192, 8193, 9000, 16384]
BATCHES = [1, 2, 3, 4, 5, 7, 8, 9, 16, 21, 32, 64]


@pytest.mark.parametrize("elem_bytes", [4, 8])
def test_plan_sweep_invariants(elem_bytes):
    rows = ROWS_PER_GROUP[elem_bytes]
    for n in SIZES:
        for batch in BATCHES:
            path, groups, ctr, ws = plan(n, batch, elem_bytes)
            case = (n, batch, elem_bytes)
            if batch >= 8 and n <= 8192:
                assert path == PATH_LOOKAHEAD and len(groups) == 0, case
                continue
            if path == PATH_PANELS:
                assert len(groups) == 0 and ws == 0, case
                # one-group panels throughout only when the first panel cannot be cooperative
                g0 = -(-n // rows)
                assert n <= MIN_ROWS or g0 < 2 or g0 * batch > MAX_GROUPS, case
                continue
            assert path == PATH_COOP, case
            assert len(groups) == -(-n // NB), case
            assert groups[0] > 0, case  # cooperative from the first panel on
            coop = np.flatnonzero(groups > 0)
            # a contiguous prefix: once a panel is one-group, every later (shorter) one is too
            assert np.array_equal(coop, np.arange(len(coop))), case
            # every cooperative panel: its own counter set, in order, inside the ones allocated
            assert np.array_equal(ctr[coop], np.arange(len(coop))), case
            assert np.all(ctr[groups == 0] == -1), case
            for k in coop:
                m = n - k * NB
                g = groups[k]
                assert m > MIN_ROWS and g >= 2 and g * batch <= MAX_GROUPS and g * rows >= m, (case, k)
            # the exchange buffers getrf allocates cover the counters of every set plus the gather lists
            ints_min = (len(coop) * batch * (NB + 4) + 2 * batch + batch * NB * MAX_GROUPS
                        + batch * len(groups) * n)
            assert ws >= 4 * ints_min, case


@pytest.mark.parametrize("n,batch,elem_bytes,path", [
    (5000, 7, 8, PATH_PANELS),   # the sequence that faulted: too tall to start cooperative
    (5000, 7, 4, PATH_PANELS),
    (16384, 1, 4, PATH_COOP), (16384, 2, 4, PATH_COOP), (16384, 1, 8, PATH_COOP), (16384, 2, 8, PATH_PANELS),
    (4096, 64, 4, PATH_LOOKAHEAD), (8100, 21, 4, PATH_LOOKAHEAD), (2048, 1, 4, PATH_COOP), (1024, 1, 4, PATH_PANELS),
])
def test_plan_known_cases(n, batch, elem_bytes, path):
    assert plan(n, batch, elem_bytes)[0] == path


def test_plan_rejects_bad_arguments():
    from tinyvlm_amd import capi

    lib = capi.lib()
    p, s = ctypes.c_int32(), ctypes.c_int32()
    assert lib.tvlm_getrf_plan(ctypes.c_int64(100), ctypes.c_int32(1), ctypes.c_int32(2), ctypes.byref(p),
                               ctypes.byref(s), None, None, ctypes.c_int32(0), None) != 0
    assert lib.tvlm_getrf_plan(ctypes.c_int64(100), ctypes.c_int32(0), ctypes.c_int32(4), ctypes.byref(p),
                               ctypes.byref(s), None, None, ctypes.c_int32(0), None) != 0
    # too little room for the panel list
    g = (ctypes.c_int32 * 2)()
    assert lib.tvlm_getrf_plan(ctypes.c_int64(16384), ctypes.c_int32(1), ctypes.c_int32(4), ctypes.byref(p),
                               ctypes.byref(s), g, None, ctypes.c_int32(2), None) != 0
