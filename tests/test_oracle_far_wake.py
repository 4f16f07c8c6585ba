"""CPU pins of the oracle's far-wake aggregation (orc_rhs_assemble_wake_influence_far /
orc_uvlm_run_far): the restatement the served UVLM loop's aggregated wake is checked against.

- groups of one ring (G = S = 1), or age thresholds beyond the wake, are the exact wake bit for bit;
- uniform circulation on a flat, regular lattice: the interior segments of every merged block
  cancel and its edges are collinear runs of the fine edges, so the coarse rings ARE the fine rings
  (agreement to rounding); a wrong corner vertex, row block or mean breaks this;
- far from the wake the field is the dipole sheet's, which the mean circulation preserves: the
  aggregated field agrees to O((block size / distance)^2) while a circulation sum (instead of the
  mean) is off by the block's ring count;
- in a UVLM run the CL error falls as the exact (near) region grows.
"""
import numpy as np
import pytest

from oracle import cpuref as O


def _flat_wake(ns, rows, dx=0.1, dy=0.25, x0=1.0, wavy=False, seed=0):
    rng = np.random.default_rng(seed)
    nwv = rows + 1
    vw = np.zeros((ns + 1, nwv, 4), order="F")
    for r in range(nwv):
        vw[:, r, 0] = x0 + r * dx
        vw[:, r, 1] = np.arange(ns + 1) * dy
        vw[:, r, 3] = 1.0
        if wavy:
            vw[:, r, 2] = 0.05 * np.sin(0.3 * r) + 0.01 * rng.standard_normal(ns + 1)
    return vw


def _receivers(ns, nc, seed=1, scale=1.0, offset=(0.0, 0.0, 0.0)):
    rng = np.random.default_rng(seed)
    verts = O.rectangular_wing(ns, nc, chord=1.0, half_span=ns * 0.25)
    colloc = np.asfortranarray(rng.uniform(-0.5, 1.0, (ns, nc, 3)) * scale + np.asarray(offset))
    nrm = rng.standard_normal((ns, nc, 3))
    nrm /= np.linalg.norm(nrm, axis=2, keepdims=True)
    return verts, colloc, np.asfortranarray(nrm)


def _wake_rhs(vw, gamma, colloc, normals, verts, it, far=None):
    s = O.Surface(verts=verts, verts_wake=vw, gamma_wake=np.asfortranarray(gamma), colloc=colloc, normals=normals)
    rhs = np.zeros(colloc.shape[0] * colloc.shape[1])
    return O.rhs_assemble_wake_influence(rhs, [s], it, far_wake=far)


NS, ROWS = 12, 40


def test_far_wake_groups_of_one_are_the_exact_wake():
    vw = _flat_wake(NS, ROWS, wavy=True)
    gamma = np.random.default_rng(3).standard_normal((NS, ROWS))
    verts, c, n = _receivers(NS, 4)
    for it in (1, 17, ROWS):
        exact = _wake_rhs(vw, gamma, c, n, verts, it)
        for far in ([(3, 1, 1), (7, 1, 1)], [(0, 1, 1)], [(ROWS + 5, 2, 3)]):
            assert np.array_equal(_wake_rhs(vw, gamma, c, n, verts, it, far), exact), (it, far)


@pytest.mark.parametrize("far", [[(5, 2, 3), (12, 4, 6)], [(0, 4, 5)], [(2, 2, 1), (6, 4, 2), (10, 8, 12)]])
def test_far_wake_uniform_circulation_equals_fine_lattice(far):
    vw = _flat_wake(NS, ROWS)
    gamma = np.full((NS, ROWS), 1.7)
    verts, c, n = _receivers(NS, 4)
    for it in (9, 25, ROWS):
        exact = _wake_rhs(vw, gamma, c, n, verts, it)
        agg = _wake_rhs(vw, gamma, c, n, verts, it, far)
        assert np.max(np.abs(agg - exact)) <= 1e-12 * np.max(np.abs(exact)), (far, it)


def test_far_wake_uniform_circulation_detects_a_wrong_mean():
    """The uniform-circulation pin is not vacuous: the same blocks carrying the SUM of their rings'
    circulations (a plausible slip) are far off."""
    vw = _flat_wake(NS, ROWS)
    gamma = np.full((NS, ROWS), 1.7)
    verts, c, n = _receivers(NS, 4)
    far = [(5, 2, 3)]
    exact = _wake_rhs(vw, gamma, c, n, verts, ROWS)
    # 6-ring blocks with the sum of their circulations == blocks of 6x the circulation: add the
    # coarse part five more times
    near_only = _wake_rhs(vw, np.where(np.arange(ROWS)[None, :] < 6, gamma, 0.0), c, n, verts, ROWS)
    coarse = _wake_rhs(vw, gamma, c, n, verts, ROWS, far) - near_only
    wrong = near_only + 6 * coarse
    assert np.max(np.abs(wrong - exact)) > 0.1 * np.max(np.abs(exact))


def test_far_wake_far_field_keeps_the_dipole_sheet():
    """Receivers ~40 wake lengths away: the aggregated wake's field matches the exact one to
    (block size / distance)^2 ~ 1e-6 relative for smoothly varying circulation."""
    vw = _flat_wake(NS, ROWS, wavy=False)
    rows = np.arange(ROWS)[None, :]
    span = np.arange(NS)[:, None]
    gamma = 1.0 + 0.3 * np.sin(0.2 * rows) + 0.2 * np.cos(0.4 * span)
    verts, c, n = _receivers(NS, 4, scale=2.0, offset=(0.0, 0.0, 150.0))
    far = [(0, 4, 4)]  # every row aggregated
    exact = _wake_rhs(vw, gamma, c, n, verts, ROWS)
    agg = _wake_rhs(vw, gamma, c, n, verts, ROWS, far)
    rel = np.max(np.abs(agg - exact)) / np.max(np.abs(exact))
    assert rel < 1e-4, rel
    summed = _wake_rhs(vw, gamma * 16.0, c, n, verts, ROWS, far)  # what a sum instead of a mean gives
    assert np.max(np.abs(summed - exact)) / np.max(np.abs(exact)) > 1.0


def test_far_wake_uvlm_error_falls_with_the_exact_region():
    """UVLM pitch/heave of a 4 x 12 wing, 160 steps: max |dCL| against the exact wake for row pairs
    aggregated beyond 10, 30, 60 steps decreases (and vanishes when nothing is aggregated)."""
    v = O.rectangular_wing(12, 4, chord=1.0, half_span=2.0)
    a = float(np.float32(2) * np.float32(3.141593) / np.float32(180))
    mot = np.array([1.0, 0.1, 0.0, a, np.pi / 2, 1.0, 0.25])
    cl0, _, _, _ = O.uvlm_run([v], mot, t_steps=161)
    errs = []
    for A in (10, 30, 60):
        cl, _, _, _ = O.uvlm_run([v], mot, t_steps=161, far_wake=[(A, 2, 2)])
        errs.append(np.max(np.abs(cl - cl0)))
    assert errs[0] > errs[1] > errs[2] > 0, errs
    cl, _, _, _ = O.uvlm_run([v], mot, t_steps=161, far_wake=[(500, 2, 2)])
    assert np.array_equal(cl, cl0)
