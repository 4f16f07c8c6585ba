"""GPU parity of the coupled pitch/heave driver (tinyvlm_amd.two_dof.UVLM2DOF) against the oracle's
restatement of tests/uvlm_2dof.cpp on the reference's wing/flap section (2 x 2 + 2 x 2 panels).

FP64: structural states u, v and loads F within 1e-9 of their largest value over 150 time steps,
Newton iteration counts equal.  FP32 (the reference's precision): the same within 2e-3 over 60
steps (GPU and CPU reductions round differently; the coupled loop feeds the difference back).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MESHES = [("mesh/infinite_rectangular_2x2_c0.x", False), ("mesh/infinite_rectangular_2x2_c1.x", True)]


def _runs(root, dtype, U_a, steps):
    from oracle import two_dof as o2
    from tinyvlm_amd.two_dof import UVLM2DOF, Vars

    meshes = [(os.path.join(root, p), l) for p, l in MESHES]
    p = UVLM2DOF("hip", meshes, dtype=dtype).run(Vars(U_a=U_a), 1000.0, max_steps=steps)
    os.chdir(root)
    o = o2.run(MESHES, o2.Vars(U_a=U_a), 1000.0, dtype, max_steps=steps)
    return p, o


def _close(p, o, key, tol):
    scale = float(np.max(np.abs(o[key])))
    err = float(np.max(np.abs(p[key].astype(np.float64) - o[key].astype(np.float64))))
    assert err <= tol * scale, (key, err, scale)


@pytest.mark.parametrize("U_a", [1.0, 7.0])
def test_two_dof_fp64_matches_oracle(root, oracle, U_a):
    print(f"[case] 2dof fp64 U_a={U_a}", flush=True)
    p, o = _runs(root, np.float64, U_a, 150)
    assert p["dt"] == o["dt"] and p["t_steps"] == o["t_steps"]
    for key in ("u", "v", "F"):
        _close(p, o, key, 1e-9)
    assert np.array_equal(p["iters"], o["iters"])


def test_two_dof_fp32_matches_oracle(root, oracle):
    print("[case] 2dof fp32 U_a=7", flush=True)
    p, o = _runs(root, np.float32, 7.0, 60)
    for key in ("u", "v", "F"):
        _close(p, o, key, 2e-3)
    assert np.max(np.abs(p["iters"] - o["iters"])) <= 1
