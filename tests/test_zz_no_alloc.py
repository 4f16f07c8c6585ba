"""GPU check of the allocation-free wake induction (tvlm_rhs_assemble_wake_influence_noalloc_f64)
against the allocating entry point it mirrors (tvlm_rhs_assemble_wake_influence_f64), through the C ABI.

Both run the same kernels on the same operands, so the results must be equal bit for bit (integer-
exact bar: the tile partials only live in a different buffer).  Two calls in a row, the second on a
smaller input, show that the shared fixed scratch carries nothing from one call to the next.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    import torch

    from tinyvlm_amd import capi

    assert torch.cuda.is_available()
    capi.init(0)
    return capi


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
    gw = rng.standard_normal((ns, rows))
    colloc = np.zeros((ns, nc, 3), order="F")
    normals = np.zeros((ns, nc, 3), order="F")
    colloc[:, :, :] = 0.25 * (v[:-1, :-1, :3] + v[1:, :-1, :3] + v[:-1, 1:, :3] + v[1:, 1:, :3])
    normals[:, :, 2] = 1.0
    return colloc, normals, wake, gw


def _case(H, torch, rng, specs, lifting):
    dev = []
    for ns, nc, rows, y0 in specs:
        c, n, w, g = _surface(rng, ns, nc, rows, y0)
        dev.append(dict(c=H.to_device(c, torch.float64), n=H.to_device(n, torch.float64),
                        w=H.to_device(w, torch.float64), g=H.to_device(g, torch.float64)))
    N = sum(ns * nc for ns, nc, _, _ in specs)
    rhs0 = torch.from_numpy(rng.standard_normal(N)).to("cuda")  # accumulate semantics: nonzero start
    return dev, rhs0, lifting


def _both(H, torch, dev, rhs0, lifting, it):
    args = ([d["g"] for d in dev], [d["c"] for d in dev], [d["n"] for d in dev], [d["w"] for d in dev], lifting, it)
    ref = rhs0.clone()
    H.rhs_assemble_wake_influence(ref, *args)
    mine = rhs0.clone()
    H.rhs_assemble_wake_influence_noalloc(mine, *args)
    torch.cuda.synchronize()
    return ref, mine


def test_noalloc_wake_equals_allocating_entry_point_twice(H):
    import torch

    rng = np.random.default_rng(11)
    # first call: 3 surfaces (one non-lifting), spans wider than a tile, 40-row wake
    big = _case(H, torch, rng, [(70, 3, 40, 0.0), (5, 2, 9, 2.5), (12, 4, 17, 5.0)], [True, False, True])
    # second call: one small surface, fewer tiles and receivers than the first
    small = _case(H, torch, rng, [(9, 2, 6, 0.0)], [True])
    print("[case] noalloc wake: big it=17, then small it=5", flush=True)
    ref, mine = _both(H, torch, *big, 17)
    assert torch.equal(ref, mine)
    assert float((ref - big[1]).abs().max()) > 0.0  # the wake did contribute
    ref, mine = _both(H, torch, *small, 5)
    assert torch.equal(ref, mine)
    # and back to the large case: still equal (nothing left over from the small call matters)
    ref, mine = _both(H, torch, *big, 8)
    assert torch.equal(ref, mine)


def test_noalloc_wake_rejects_oversize_before_any_work(H):
    """A call whose tile partials exceed the fixed scratch is refused on the host: the error names
    the limit and rhs is untouched (nothing was queued)."""
    import torch

    rng = np.random.default_rng(3)
    ns, nc, rows = 64, 300, 4000  # 19200 receivers x >= 1000 tiles x 8 B >> 32 MiB
    c, n, w, g = _surface(rng, ns, nc, rows, 0.0)
    dc, dn = H.to_device(c, torch.float64), H.to_device(n, torch.float64)
    dw, dg = H.to_device(w, torch.float64), H.to_device(g, torch.float64)
    rhs = torch.zeros(ns * nc, dtype=torch.float64, device="cuda")
    print("[case] noalloc wake oversize rejection", flush=True)
    with pytest.raises(H.TvlmError, match="fixed scratch"):
        H.rhs_assemble_wake_influence_noalloc(rhs, [dg], [dc], [dn], [dw], [True], rows)
    torch.cuda.synchronize()
    assert float(rhs.abs().max()) == 0.0
    with pytest.raises(H.TvlmError, match="float64"):
        H.rhs_assemble_wake_influence_noalloc(rhs.float(), [dg], [dc], [dn], [dw], [True], 1)
