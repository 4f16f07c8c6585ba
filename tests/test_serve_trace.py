"""CPU checks of the steady-solve request mix (tinyvlm_amd.serve): the committed trace is what its
generator produces (PCG64 seed 0), its marginals are the stated ones, and the generated meshes of
the mix have the stated planforms (chord 1, half-span 5, 25-degree leading-edge sweep)."""
import math

import numpy as np


def test_committed_trace_regenerates():
    from tinyvlm_amd import serve as S

    t, c = S.steady_mix_trace(4096, 0), S.read_trace()
    for k in ("gap", "family", "n", "alpha_deg"):
        assert np.array_equal(np.asarray(t[k]), c[k]), k
    assert c["alpha_deg"].dtype == np.float32


def test_trace_marginals():
    from tinyvlm_amd import serve as S

    c = S.read_trace()
    assert len(c["gap"]) == 4096
    assert abs(float(np.mean(c["gap"])) - 1.0) < 0.05  # unit-rate exponential: mean 1, sd 1/64
    assert set(np.unique(c["n"])) == set(S.MIX_SIZES)
    assert set(np.unique(c["family"])) == {0, 1, 2}
    assert c["alpha_deg"].min() >= -5.0 and c["alpha_deg"].max() <= 10.0
    for counts in (np.bincount(c["family"]), np.unique(c["n"], return_counts=True)[1]):
        assert counts.min() > 0.8 * counts.mean()


def test_mix_meshes_planform(tmp_path):
    from tinyvlm_amd import serve as S
    from tinyvlm_amd.mesh_io import read_plot3d

    for fam in ("elliptic", "rectangular", "swept25"):
        v = S.mix_mesh(fam, 16, str(tmp_path))
        raw = read_plot3d(str(tmp_path / f"tvlm_mix_{fam}_16x16.x"), quarter_chord=False) if fam != "elliptic" else None
        assert v.shape == (17, 17, 4)
        assert abs(v[:, :, 1].max() - 5.0) < 1e-6 and abs(v[0, :, 1]).max() < 1e-6
        if raw is not None:
            sweep = math.tan(math.radians(25.0)) if fam == "swept25" else 0.0
            le = raw[:, 0, :2]
            assert np.allclose(le[:, 0], le[:, 1] * sweep, atol=1e-6)
            assert np.allclose(raw[:, -1, 0] - raw[:, 0, 0], 1.0, atol=1e-6)
