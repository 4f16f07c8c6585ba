"""GPU checks of tvlm_wake_geometry_*: the wake rows of a prescribed motion written in one launch
per surface against the per-step loop it replaces (Backend::displace_wing then Backend::wake_shed
for every step, lib/libhbvlm3.cpp:306-319), bit for bit, FP64 and FP32; two surfaces of different
sizes, fewer steps than wake rows (the other rows untouched), and refused shapes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _surface(ns, nc, y0):
    v = np.zeros((ns + 1, nc + 1, 4), order="F")
    for i in range(ns + 1):
        v[i, :, 0] = np.linspace(0.0, 1.0, nc + 1) + 0.05 * i / ns
        v[i, :, 1] = y0 + 2.0 * i / ns
        v[i, :, 2] = 0.01 * np.sin(2.0 * i / ns)
    v[:, :, 3] = 1.0
    return v


def _transforms(rng, steps, n_surf):
    """A rotation about y and a translation per (step, surface), step-major: {4, 4, steps * n_surf}."""
    T = np.zeros((4, 4, steps * n_surf), order="F")
    for q in range(steps * n_surf):
        a = 0.3 * rng.standard_normal()
        c, s = np.cos(a), np.sin(a)
        T[:, :, q] = [[c, 0, s, -0.1 * q], [0, 1, 0, 0.01 * q], [-s, 0, c, 0.2 * rng.standard_normal()], [0, 0, 0, 1]]
    return T


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_wake_geometry_equals_displace_and_shed_loop(prec):
    import torch

    from tinyvlm_amd import capi as H

    H.init(0)
    dt = torch.float64 if prec == "f64" else torch.float32
    rng = np.random.default_rng(5)
    dims = [(20, 6, 0.0), (9, 3, 2.5)]
    nwv, steps = 40, 31
    print(f"[case] wake geometry {prec}, surfaces {dims}, {steps} of {nwv} rows", flush=True)
    init = [H.to_device(_surface(ns, nc, y0), dt) for ns, nc, y0 in dims]
    verts = [H.colmajor(tuple(v.shape), dt) for v in init]
    T = H.to_device(_transforms(rng, steps, len(dims)), dt)
    fill = 7.25
    loop = [H.colmajor((ns + 1, nwv, 4), dt).fill_(fill) for ns, _, _ in dims]
    one = [H.colmajor((ns + 1, nwv, 4), dt).fill_(fill) for ns, _, _ in dims]
    for i in range(steps):
        H.displace_wing([T[:, :, 2 * i], T[:, :, 2 * i + 1]], verts, init)
        H.wake_shed(verts, loop, i)
    H.wake_geometry(T, init, one, steps)
    for a, b in zip(loop, one):
        assert torch.equal(a, b)
        assert torch.all(b[:, : nwv - steps] == fill)  # rows of steps not taken are left alone
    assert not torch.equal(one[0][:, nwv - 1], one[0][:, nwv - steps])  # the poses differ


def test_wake_geometry_refusals():
    import torch

    from tinyvlm_amd import capi as H

    H.init(0)
    init = [H.to_device(_surface(4, 2, 0.0), torch.float64)]
    T = H.to_device(_transforms(np.random.default_rng(0), 6, 1), torch.float64)
    wake = [H.colmajor((5, 5, 4), torch.float64)]
    with pytest.raises(H.TvlmError):  # more steps than wake rows
        H.wake_geometry(T, init, wake, 6)
    with pytest.raises(H.TvlmError):  # fewer transforms than steps x surfaces
        H.wake_geometry(T[:, :, :3], init, [H.colmajor((5, 6, 4), torch.float64)], 4)
    H.wake_geometry(T, init, wake, 0)  # nothing to do
