"""GPU parity of tvlm_rhs_assemble_wake_influence_lattice_f64: Backend::rhs_assemble_wake_influence
(include/vlm_backend.hpp:46) evaluated over the wake's vortex lattice (every segment once, with the
difference of the circulations of the rings beside it) instead of ring by ring.

Bar (FP64): the golden vectors, the oracle's ring-by-ring restatement and the ring-by-ring GPU call
within LATTICE_TOL = 2e-12 of the largest entry, over the tile shapes the launch picks at several
work-group targets (one to four rows per tile, spans split), several surfaces (one without a wake),
iteration 0 (no work) and a rejected FP32 call.  The two sums are equal in exact arithmetic; the
bar is set by rounding.  The lattice form takes 1/|r| and 1/(4 pi |r1 x r2|^2) from the hardware
estimates refined by one Newton step (a few ulp, against 0.5 ulp for a correctly rounded sqrt and
division), and the segment formula's (r0.r1/|r1| - r0.r2/|r2|) cancels for a receiver near the
extension of a segment (here the receivers of the surfaces beside a wake, in line with its spanwise
segments).  On the three-surface case at 8 rows, against a long-double evaluation of the same sum:
the oracle is 4.4e-14 off, the lattice sum with correctly rounded operations 8.5e-14, and with its
reciprocals perturbed by 1, 4 and 16 ulp 2.5e-13, 6.3e-13 and 1.5e-12 (the GPU call measured
6.3e-13 from the oracle).  The ring-by-ring call keeps its 1e-13 bar.

The harmonic-balance solver with wake_form = "lattice" against the oracle's HBVLM::run at the bar
of the exact form (1e-9 of the largest load).
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LATTICE_TOL = 2e-12


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def H():
    import torch

    from tinyvlm_amd import capi

    assert torch.cuda.is_available()
    capi.init(0)
    return capi


@pytest.fixture
def lattice_target():
    """Sets the library-wide lattice work-group target for one test and restores it after."""
    from tinyvlm_amd import autotune

    w, l = autotune.get_targets()
    yield lambda t: autotune.set_targets(w, t)
    autotune.set_targets(w, l)


@pytest.mark.parametrize("it", [1, 3, 5])
def test_lattice_vs_golden(H, golden, it):
    import torch

    print(f"[case] wake lattice golden it={it}", flush=True)
    dt = torch.float64
    c = H.to_device(golden["wake4_colloc"], dt)
    n = H.to_device(golden["wake4_normals"], dt)
    w = H.to_device(golden["wake4_verts_wake"], dt)
    g = H.to_device(golden["wake4_gamma_wake"], dt)
    rhs = torch.zeros(c.shape[0] * c.shape[1], dtype=dt, device="cuda")
    H.rhs_assemble_wake_influence_lattice(rhs, [g], [c], [n], [w], [True], it)
    assert rel(H.to_host(rhs), golden[f"wake4_rhs_it{it}"]) < LATTICE_TOL


def _surface(rng, ns, nc, rows, y0):
    v = np.zeros((ns + 1, nc + 1, 4), order="F")
    for i in range(ns + 1):
        v[i, :, 0] = np.linspace(0, 1, nc + 1) + 0.1 * i / ns
        v[i, :, 1] = y0 + 2.0 * i / ns
        v[i, :, 2] = 0.02 * np.sin(3 * i / ns) * np.linspace(0, 1, nc + 1)
    v[:, :, 3] = 1
    wake = np.zeros((ns + 1, rows + 1, 4), order="F")
    for r in range(rows + 1):
        wake[:, r] = v[:, -1]
        wake[:, r, 0] += 0.1 * (rows - r) + 0.01 * rng.standard_normal(ns + 1)
        wake[:, r, 2] += 0.05 * math.sin(0.3 * r)
    return v, wake, rng.standard_normal((ns, rows))


@pytest.mark.parametrize("target", [64, 1024, 65536])
def test_lattice_multi_surface_vs_oracle_and_exact(H, oracle, lattice_target, target):
    """Three surfaces (the middle one without a wake), spans wider than one lattice tile (70 > 64),
    iterations below the wake rows; targets 64 (four rows per tile), 1024 (the default) and 65536
    (spans split): the oracle and the ring-by-ring GPU call within LATTICE_TOL, bitwise repeatable."""
    import torch

    lattice_target(target)
    rng = np.random.default_rng(11)
    specs = [(70, 3, 40, 0.0), (5, 2, 9, 2.5), (12, 4, 17, 5.0)]
    lifting = [True, False, True]
    osurf, dev = [], []
    for (ns, nc, rows, y0), lift in zip(specs, lifting):
        v, wake, gw = _surface(rng, ns, nc, rows, y0)
        osurf.append(oracle.Surface(verts=v, verts_wake=wake, gamma_wake=np.asfortranarray(gw), lifting=lift))
    oracle.mesh_metrics(0.0, osurf)
    dt = torch.float64
    for s in osurf:
        dev.append([H.to_device(x, dt) for x in (s.gamma_wake, s.colloc, s.normals, s.verts_wake)])
    args = [[d[k] for d in dev] for k in range(4)]
    N = sum(s.ns * s.nc for s in osurf)
    for it in (1, 8, 17):
        print(f"[case] wake lattice 3 surfaces target={target} it={it}", flush=True)
        rhs_o = np.zeros(N)
        oracle.rhs_assemble_wake_influence(rhs_o, osurf, it)
        lat = torch.zeros(N, dtype=dt, device="cuda")
        H.rhs_assemble_wake_influence_lattice(lat, *args, lifting, it)
        ring = torch.zeros(N, dtype=dt, device="cuda")
        H.rhs_assemble_wake_influence(ring, *args, lifting, it)
        e_o, e_r = rel(H.to_host(lat), rhs_o), rel(H.to_host(lat), H.to_host(ring))
        print(f"[case]   rel vs oracle {e_o:.2e} vs ring-by-ring {e_r:.2e}", flush=True)
        assert e_o < LATTICE_TOL and e_r < LATTICE_TOL
        again = torch.zeros(N, dtype=dt, device="cuda")
        H.rhs_assemble_wake_influence_lattice(again, *args, lifting, it)
        assert torch.equal(lat, again)
    # the receivers of the surface without a wake still see the others' wakes
    assert np.max(np.abs(rhs_o[70 * 3: 70 * 3 + 10])) > 0


def test_lattice_accumulates_and_edge_cases(H, golden):
    import torch

    dt = torch.float64
    c = H.to_device(golden["wake4_colloc"], dt)
    n = H.to_device(golden["wake4_normals"], dt)
    w = H.to_device(golden["wake4_verts_wake"], dt)
    g = H.to_device(golden["wake4_gamma_wake"], dt)
    N = c.shape[0] * c.shape[1]
    print("[case] wake lattice edge cases", flush=True)
    # iteration 0: no wake rows, rhs untouched
    rhs = torch.full((N,), 2.5, dtype=dt, device="cuda")
    H.rhs_assemble_wake_influence_lattice(rhs, [g], [c], [n], [w], [True], 0)
    assert torch.all(rhs == 2.5)
    # a surface without a wake: untouched
    H.rhs_assemble_wake_influence_lattice(rhs, [g], [c], [n], [w], [False], 3)
    assert torch.all(rhs == 2.5)
    # the induction is subtracted from what rhs holds, as the ring-by-ring call does
    ref = torch.full((N,), 2.5, dtype=dt, device="cuda")
    H.rhs_assemble_wake_influence(ref, [g], [c], [n], [w], [True], 3)
    H.rhs_assemble_wake_influence_lattice(rhs, [g], [c], [n], [w], [True], 3)
    assert rel(H.to_host(rhs), H.to_host(ref)) < LATTICE_TOL
    # more wake rows than the wake holds, and FP32 tensors, are refused
    with pytest.raises(H.TvlmError):
        H.rhs_assemble_wake_influence_lattice(rhs, [g], [c], [n], [w], [True], w.shape[1])
    with pytest.raises(H.TvlmError):
        H.rhs_assemble_wake_influence_lattice(rhs.float(), [g.float()], [c.float()], [n.float()], [w.float()],
                                              [True], 1)


def test_hb_lattice_wake_matches_oracle(root, oracle):
    """HBVLM with wake_form = "lattice" on the 9x5 wing + 3x5 flap, H = 2, against the oracle's
    HBVLM::run (ring-by-ring wake) and against the same solver with the exact form."""
    from oracle import hb as ohb
    from tinyvlm_amd.hb import HBVLM

    paths = [os.path.join(root, "mesh", m) for m in ("3dof_wing_9x5.x", "3dof_flap_3x5.x")]
    K = 5
    X = np.zeros((3, K))
    X[0, 1] = 0.01 / 0.127
    X[1, 2] = 0.02
    X[2, 1] = 0.05
    X[0, 3] = 0.002
    X[1, 3] = 0.005
    X[2, 4] = -0.01
    loads = {}
    for form in ("exact", "lattice"):
        print(f"[case] hb 9x5+3x5 H=2 wake {form}", flush=True)
        p = HBVLM("hip", paths)
        p.wake_form = form
        p.init(2, 1.0 / 0.127)
        F = np.zeros((3, K))
        p.run(0.8, 6.0, X, F)
        loads[form] = (F, p.iterations)
    o = ohb.HBVLMOracle(paths)
    o.init(2, 1.0 / 0.127)
    Fo = np.zeros((3, K))
    o.run(0.8, 6.0, X, Fo)
    scale = np.max(np.abs(Fo))
    F, iters = loads["lattice"]
    print(f"[case]   |dF| lattice vs oracle {np.max(np.abs(F - Fo)):.2e}, vs exact "
          f"{np.max(np.abs(F - loads['exact'][0])):.2e} (scale {scale:.3e})", flush=True)
    assert np.max(np.abs(F - Fo)) <= 1e-9 * scale
    assert np.max(np.abs(F - loads["exact"][0])) <= 1e-9 * scale
    assert abs(iters - o.iterations) <= 1
    with pytest.raises(ValueError):
        p = HBVLM("hip", paths)
        p.wake_form = "ring"
        p.init(2, 1.0 / 0.127)
        p.run(0.8, 6.0, X, np.zeros((3, K)))
