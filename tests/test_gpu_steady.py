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
