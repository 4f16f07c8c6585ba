"""GPU checks of the grouped influence-matrix assembly (tvlm_lhs_group_*): many matrices of uneven
size -- different meshes, one and two surfaces, different wake iterations -- in one launch.

Every matrix of a group must be bitwise the per-section launches' result (tvlm_lhs_assemble_f64 for
the exact form, tvlm_lhs_assemble_seeded_f64 for the seeded form): the grouped kernel runs the same
tile code on the same tiles.  Also: a prefix of the items, accumulation, large and small matrices in
one work list, and the argument checks.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32 = np.float32
CASES = [("rect2", 1), ("rect4", 1), ("swept4", 1), ("split2", 2)]


def to_radians(deg):
    return float(F32(deg) * F32(3.141593) / F32(180.0))


@pytest.fixture(scope="module")
def H():
    from tinyvlm_amd import capi

    capi.init(0)
    return capi


def _surfaces(H, golden, name, nsurf):
    import torch

    out = []
    for k in range(nsurf):
        key = f"{name}_a5_s{k}"
        v, w = golden[key + "_verts"], golden[key + "_wake"]
        ns, nc = v.shape[0] - 1, v.shape[1] - 1
        out.append(dict(verts=H.to_device(v, torch.float64), wake=H.to_device(w, torch.float64),
                        colloc=H.colmajor((ns, nc, 3), torch.float64), normals=H.colmajor((ns, nc, 3), torch.float64),
                        areas=H.colmajor((ns, nc), torch.float64), ns=ns, nc=nc))
    H.mesh_metrics(to_radians(5), [s["verts"] for s in out], [s["colloc"] for s in out],
                   [s["normals"] for s in out], [s["areas"] for s in out])
    return out


def _wing(H, oracle, ns, nc):
    """A rectangular wing with a one-row wake 100 chords behind its trailing edge."""
    import torch

    v = oracle.rectangular_wing(ns, nc, chord=1.0, half_span=5.0)
    w = np.zeros((ns + 1, 2, 4))
    w[:, 0, :] = v[:, -1, :]
    w[:, 1, :] = v[:, -1, :]
    w[:, 1, 0] += 100.0
    s = dict(verts=H.to_device(v, torch.float64), wake=H.to_device(w, torch.float64),
             colloc=H.colmajor((ns, nc, 3), torch.float64), normals=H.colmajor((ns, nc, 3), torch.float64),
             areas=H.colmajor((ns, nc), torch.float64), ns=ns, nc=nc)
    H.mesh_metrics(0.05, [s["verts"]], [s["colloc"]], [s["normals"]], [s["areas"]])
    return [s]


def _item(H, S, iteration):
    import torch

    N = sum(s["ns"] * s["nc"] for s in S)
    lhs = H.colmajor((N, N), torch.float64)
    return (lhs, [s["colloc"] for s in S], [s["normals"] for s in S], [s["verts"] for s in S],
            [s["wake"] for s in S], iteration)


def _alone(H, item, seeded):
    import torch

    lhs, c, n, v, w, it = item
    ref = H.colmajor(tuple(lhs.shape), torch.float64)
    if seeded:
        H.lhs_assemble_seeded(ref, c, n, v, w, it)
    else:
        H.lhs_assemble(ref, c, n, v, w, it)
    return H.to_host(ref)


def _bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.parametrize("seeded", [False, True], ids=["exact", "seeded"])
def test_group_of_uneven_meshes_is_bitwise_the_per_matrix_assembly(H, golden, seeded):
    import torch

    print(f"[case] lhs group of uneven meshes, seeded={seeded}", flush=True)
    surf = {name: _surfaces(H, golden, name, ns) for name, ns in CASES}
    items = [_item(H, surf[name], 1) for name, _ in CASES] + [_item(H, surf["rect4"], 0),
                                                               _item(H, surf["split2"], 0)]
    sizes = [it[0].shape[0] for it in items]
    assert len(set(sizes)) >= 3  # uneven N
    g = H.LhsGroup(items)
    g.assemble(seeded=seeded)
    torch.cuda.synchronize()
    for it in items:
        assert _bits(H.to_host(it[0]), _alone(H, it, seeded))
    # a prefix of the items: the rest stays untouched
    for it in items:
        it[0].zero_()
    g.assemble(n_items=2, seeded=seeded)
    torch.cuda.synchronize()
    assert _bits(H.to_host(items[0][0]), _alone(H, items[0], seeded))
    assert _bits(H.to_host(items[1][0]), _alone(H, items[1], seeded))
    assert all(float(it[0].abs().max()) == 0.0 for it in items[2:])
    # accumulation, as lhs_assemble: twice the same assembly
    g.assemble(n_items=1, seeded=seeded)
    lhs, c, n, v, w, k = items[0]
    twice = H.colmajor(tuple(lhs.shape), torch.float64)
    for _ in range(2):
        (H.lhs_assemble_seeded if seeded else H.lhs_assemble)(twice, c, n, v, w, k)
    assert _bits(H.to_host(lhs), H.to_host(twice))
    g.close()


def test_group_of_large_and_small_matrices(H, golden, oracle):
    """Two wings of 1600 and 2400 panels and the golden meshes: 25 x 50 + 38 x 75 + ... tiles, so the
    work list's section lookup crosses many 64-tile chunks and section boundaries inside chunks."""
    import torch

    print("[case] lhs group of large and small matrices", flush=True)
    items = [_item(H, _wing(H, oracle, 40, 40), 1), _item(H, _wing(H, oracle, 60, 40), 1),
             _item(H, _surfaces(H, golden, "split2", 2), 1)]
    g = H.LhsGroup(items)
    g.assemble(seeded=True)
    torch.cuda.synchronize()
    for it in items:
        assert _bits(H.to_host(it[0]), _alone(H, it, True))
    g.close()


def test_group_argument_checks(H, golden):
    print("[case] lhs group argument checks", flush=True)
    S = _surfaces(H, golden, "rect4", 1)
    g = H.LhsGroup([_item(H, S, 1)])
    with pytest.raises(H.TvlmError):
        g.assemble(n_items=2)
    with pytest.raises(H.TvlmError):
        g.assemble(seeded=True, sigma=0.1)
    g.close()
    with pytest.raises(H.TvlmError):  # wake with fewer rows than iteration + 1
        H.LhsGroup([_item(H, S, 5)])
