"""GPU parity of the batched steady plan (tvlm_vlm_*): VLM::run for many angles of attack.

Plain FP64 mode against the oracle's VLM::run restatement (same mesh, same angle): circulation
within 1e-12 relative, CL and CM within 1e-12.  Served mode (seeded FP64 matrices, batched FP32 LU,
FP64 refinement) against the plain mode: circulation within 1e-10 relative, |dCL| <= 1e-10 (the
acceptance bar is 1e-6).  Includes the full-size elliptic 64x64 wing (N = 4096) at 5 degrees and a
two-surface split mesh.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _alpha(deg):
    return float(np.float32(deg) * np.float32(3.141593) / np.float32(180))


def _run(meshes, alphas, mode):
    from tinyvlm_amd import capi as H
    from tinyvlm_amd.steady_plan import SteadyPlan

    H.init(0)
    p = SteadyPlan(meshes, len(alphas), mode=mode)
    p.solve(alphas)
    r = p.results()
    p.close()
    return r


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("mesh", ["elliptic_45x45.x", "swept_4x4.x"])
def test_plain_plan_matches_oracle(root, oracle, mesh):
    from tinyvlm_amd.steady_plan import MODE_PLAIN_FP64

    v = oracle.read_mesh(os.path.join(root, "mesh", mesh))
    degs = [0.0, 1.0, 5.0, 10.0]
    alphas = [_alpha(d) for d in degs]
    print(f"[case] steady plain {mesh}", flush=True)
    r = _run([v], alphas, MODE_PLAIN_FP64)
    assert np.all(r["info"] == 0)
    for b, a in enumerate(alphas):
        g, cl, cm, info = oracle.vlm_run([v], a)
        if degs[b] == 0.0:
            assert np.max(np.abs(r["gamma"][b])) == 0.0 and r["cl"][b] == 0.0
            continue
        assert _rel(r["gamma"][b], g) < 1e-12
        assert abs(r["cl"][b] - cl) <= 1e-12 * abs(cl)
        assert np.max(np.abs(r["cm"][b] - cm)) <= 1e-12 * np.max(np.abs(cm))


@pytest.mark.parametrize("mesh", ["elliptic_45x45.x", "rectangular_4x4.x"])
def test_served_plan_matches_plain(root, oracle, mesh):
    from tinyvlm_amd.steady_plan import MODE_PLAIN_FP64, MODE_SERVED

    v = oracle.read_mesh(os.path.join(root, "mesh", mesh))
    # 9 requests: the batched FP32 LU takes its look-ahead form (batches of 8 or more)
    alphas = [_alpha(d) for d in (-5.0, -1.0, 0.5, 2.0, 3.0, 5.0, 6.0, 7.5, 10.0)]
    print(f"[case] steady served vs plain {mesh}", flush=True)
    rp = _run([v], alphas, MODE_PLAIN_FP64)
    rs = _run([v], alphas, MODE_SERVED)
    assert np.all(rs["info"] == 0)
    for b in range(len(alphas)):
        assert _rel(rs["gamma"][b], rp["gamma"][b]) < 1e-10
        assert abs(rs["cl"][b] - rp["cl"][b]) <= 1e-10
        assert np.max(np.abs(rs["cm"][b] - rp["cm"][b])) <= 1e-10


def test_served_plan_two_surfaces(root, oracle):
    from tinyvlm_amd.steady_plan import MODE_SERVED

    parts = [oracle.read_mesh(os.path.join(root, "mesh", f"infinite_rectangular_2x2_part{k}.x")) for k in (0, 1)]
    alphas = [_alpha(d) for d in (1.0, 3.0, 15.0)]
    print("[case] steady served two surfaces", flush=True)
    r = _run(parts, alphas, MODE_SERVED)
    for b, a in enumerate(alphas):
        g, cl, cm, _ = oracle.vlm_run(parts, a)
        assert _rel(r["gamma"][b], g) < 1e-10
        assert abs(r["cl"][b] - cl) <= 1e-10 * abs(cl)
        assert np.max(np.abs(r["cm"][b] - cm)) <= 1e-10 * np.max(np.abs(cm))


def test_served_plan_elliptic_64_full_size(root, oracle):
    """The steady configuration at full size: elliptic 64 x 64 (N = 4096), alpha = 5 degrees, as a
    batch of two requests (5 degrees and 3 degrees), against the oracle."""
    from tinyvlm_amd.steady_plan import MODE_SERVED

    v = oracle.read_mesh(os.path.join(root, "mesh", "elliptic_64x64.x"))
    alphas = [_alpha(5.0), _alpha(3.0)]
    print("[case] steady served elliptic 64x64", flush=True)
    r = _run([v], alphas, MODE_SERVED)
    assert np.all(r["info"] == 0)
    for b, a in enumerate(alphas):
        g, cl, cm, _ = oracle.vlm_run([v], a)
        assert _rel(r["gamma"][b], g) < 1e-10
        assert abs(r["cl"][b] - cl) <= 1e-10 * abs(cl)


def _stacked_wings(oracle, dz, ns=8, nc=6):
    """Two copies of one rectangular wing, the second lifted by dz: as dz shrinks, every receiver
    of one copy nearly coincides with its twin, and the influence matrix's condition number grows
    like 1/dz^2 (dz = 3e-4: 2e6; 5e-5: 7e7; 1e-5: 2e9)."""
    v = oracle.rectangular_wing(ns, nc, chord=1.0, half_span=5.0)
    v2 = np.asfortranarray(v.copy())
    v2[:, :, 2] += dz
    return [v, v2]


@pytest.mark.parametrize("dz,expect", [(1e-2, "fast"), (3e-4, "rescued"), (5e-5, "rescued"), (1e-5, "rescued")])
def test_served_refinement_closed_loop(oracle, dz, expect):
    """The served path tests every returned x (||b - A x|| <= sqrt(N) eps ||x|| ||A||, LAPACK
    dsgesv's test) and rescues the requests that fail it: more corrections, then the FP64 LU.
    On well-conditioned stacked wings the fixed corrections pass (status 0); on ill-conditioned ones
    (kappa 2e6 .. 2e9, kappa eps32 from 0.1 to 100) the fixed corrections do not, and the rescued
    result equals the plain FP64 solve: |dCL| <= 1e-6 (the acceptance bar), gamma within 1e-6
    relative (the FP64 solve itself is only good to kappa eps64 here)."""
    from tinyvlm_amd.steady_plan import MODE_PLAIN_FP64, MODE_SERVED, SteadyPlan

    meshes = _stacked_wings(oracle, dz)
    alphas = [_alpha(2.0), _alpha(6.0)]
    print(f"[case] served refinement stacked wings dz={dz}", flush=True)
    rp = _run(meshes, alphas, MODE_PLAIN_FP64)
    p = SteadyPlan(meshes, len(alphas), mode=MODE_SERVED)
    p.solve(alphas)
    rs = p.results()
    ref = p.refinement()
    p.close()
    print(f"[case]   status {ref['status']} corrections {ref['corrections']} berr {ref['backward_error']}", flush=True)
    assert np.all(rp["info"] == 0) and np.all(rs["info"] == 0)
    for b in range(len(alphas)):
        assert abs(rs["cl"][b] - rp["cl"][b]) <= 1e-6
        assert _rel(rs["gamma"][b], rp["gamma"][b]) <= 1e-6
        if expect == "fast":
            assert ref["status"][b] == 0 and ref["corrections"][b] == 4
            assert ref["backward_error"][b] <= np.sqrt(96) * 2.3e-16
        else:
            assert ref["status"][b] & 3, ref  # more corrections or the FP64 fallback
            if ref["status"][b] & 2 == 0:
                assert 4 < ref["corrections"][b] <= 8 and ref["backward_error"][b] <= np.sqrt(96) * 2.3e-16


def test_served_recovers_from_stalled_cooperative_panel(root, oracle):
    """A cooperative LU panel whose bounded wait gives up leaves invalid factors and info -999.
    Forced from the host (the wait bound set to one spin: every wait that does not find all groups
    already there gives up at once, and the kernel still runs to its end), the served plan
    re-factors that matrix with one-group panels and returns the result of the normal run."""
    import ctypes

    from tinyvlm_amd import capi as H
    from tinyvlm_amd.steady_plan import MODE_SERVED, SteadyPlan

    v = oracle.rectangular_wing(64, 32, chord=1.0, half_span=5.0)  # N = 2048: cooperative FP32 panels
    alphas = [_alpha(3.0), _alpha(-2.0)]
    print("[case] served stall recovery N=2048", flush=True)
    normal = _run([v], alphas, MODE_SERVED)
    lib = H.lib()
    p = SteadyPlan([v], len(alphas), mode=MODE_SERVED)
    try:
        H.check(lib.tvlm_lu_set_coop_spin_limit(ctypes.c_uint32(1)), "spin limit")
        p.solve(alphas)
        r = p.results()
        ref = p.refinement()
    finally:
        H.check(lib.tvlm_lu_set_coop_spin_limit(ctypes.c_uint32(0)), "spin limit")
        p.close()
    print(f"[case]   status {ref['status']} corrections {ref['corrections']}", flush=True)
    assert np.any(ref["status"] & 4), ref  # the forced stall happened and was recovered
    assert np.all(r["info"] == 0)
    for b in range(len(alphas)):
        assert _rel(r["gamma"][b], normal["gamma"][b]) < 1e-10
        assert abs(r["cl"][b] - normal["cl"][b]) <= 1e-10
