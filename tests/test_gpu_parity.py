"""GPU parity of the HIP backend (through the C ABI) against the oracle and the golden vectors.

Tolerances (written per test):
  FP64 kernels      : max |gpu - ref| <= 1e-13 * max |ref| for assembled quantities; the exact
                      FP64 forms differ from the oracle only by FMA contraction.
  FP64 solve (gamma): <= 1e-12 relative (the plain FP64 path's acceptance bar).
  FP32 kernels      : <= 2e-5 relative to the block max (the fast form uses rsqrt/rcp).
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32 = np.float32


def to_radians(deg):
    return float(F32(deg) * F32(3.141593) / F32(180.0))


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


@pytest.fixture(scope="module")
def torch():
    import torch

    return torch


def DT(torch, name):
    return torch.float64 if name == "f64" else torch.float32


CASES = [("rect2", 1), ("rect4", 1), ("swept4", 1), ("split2", 2)]


def _golden_surfaces(H, golden, torch, name, tag, nsurf, dt):
    out = []
    for k in range(nsurf):
        key = f"{name}_{tag}_s{k}"
        v = golden[key + "_verts"]
        w = golden[key + "_wake"]
        ns, nc = v.shape[0] - 1, v.shape[1] - 1
        out.append(dict(key=key, verts=H.to_device(v, dt), wake=H.to_device(w, dt),
                        colloc=H.colmajor((ns, nc, 3), dt), normals=H.colmajor((ns, nc, 3), dt),
                        areas=H.colmajor((ns, nc), dt), ns=ns, nc=nc))
    return out


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("name,nsurf", CASES)
@pytest.mark.parametrize("tag,alpha", [("a0", 0.0), ("a5", to_radians(5))])
def test_metrics_lhs_solve_vs_golden(H, torch, golden, oracle, prec, name, nsurf, tag, alpha):
    print(f"[case] metrics/lhs/solve {name} {tag} {prec}", flush=True)
    dt = DT(torch, prec)
    S = _golden_surfaces(H, golden, torch, name, tag, nsurf, dt)
    H.mesh_metrics(alpha, [s["verts"] for s in S], [s["colloc"] for s in S], [s["normals"] for s in S],
                   [s["areas"] for s in S])
    tol_m = 1e-14 if prec == "f64" else 2e-6
    for s in S:
        assert rel(H.to_host(s["colloc"]), golden[s["key"] + "_colloc"]) < tol_m
        assert rel(H.to_host(s["normals"]), golden[s["key"] + "_normals"]) < tol_m
        assert rel(H.to_host(s["areas"]), golden[s["key"] + "_areas"]) < tol_m
    N = sum(s["ns"] * s["nc"] for s in S)
    lhs = H.colmajor((N, N), dt)
    H.lhs_assemble(lhs, [s["colloc"] for s in S], [s["normals"] for s in S], [s["verts"] for s in S],
                   [s["wake"] for s in S], 1)
    L = H.to_host(lhs)
    gl = golden[f"{name}_{tag}_lhs"]
    assert rel(L, gl) < (1e-13 if prec == "f64" else 2e-5)
    # the same inputs through the oracle (FP32 restatement for the f32 path)
    npdt = np.float64 if prec == "f64" else np.float32
    osurf = [oracle.Surface(verts=np.asfortranarray(golden[s["key"] + "_verts"].astype(npdt)),
                            verts_wake=np.asfortranarray(golden[s["key"] + "_wake"].astype(npdt))) for s in S]
    oracle.mesh_metrics(alpha, osurf)
    Lo = oracle.lhs_assemble(osurf, 1)
    assert rel(L, Lo) < (1e-13 if prec == "f64" else 2e-5)
    if prec == "f64":
        # the served steady path's seeded form (hardware 1/sqrt and 1/x estimates + one Newton step)
        # assembles the same matrix to FP64 accuracy
        lhs_s = H.colmajor((N, N), dt)
        H.lhs_assemble_seeded(lhs_s, [s["colloc"] for s in S], [s["normals"] for s in S], [s["verts"] for s in S],
                              [s["wake"] for s in S], 1)
        assert rel(H.to_host(lhs_s), gl) < 1e-13
    # solve with the freestream rhs
    fs = [math.cos(alpha), 0.0, math.sin(alpha)]
    vel = []
    for s in S:
        v = H.colmajor((s["ns"], s["nc"], 3), dt)
        for k in range(3):
            v[:, :, k] = fs[k]
        vel.append(v)
    rhs = torch.zeros(N, dtype=dt, device="cuda")
    H.rhs_assemble_velocities(rhs, [s["normals"] for s in S], vel)
    ipiv = torch.zeros(N, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    H.getrf(lhs, ipiv, info)
    H.getrs(lhs, ipiv, rhs)
    assert int(info.item()) == 0
    assert rel(H.to_host(rhs), golden[f"{name}_{tag}_gamma"]) < (1e-12 if prec == "f64" else 1e-4)


def test_lhs_accumulates_and_iteration_zero(H, torch, golden):
    dt = torch.float64
    S = _golden_surfaces(H, golden, torch, "rect4", "a5", 1, dt)
    s = S[0]
    H.mesh_metrics(to_radians(5), [s["verts"]], [s["colloc"]], [s["normals"]], [s["areas"]])
    N = s["ns"] * s["nc"]
    lhs1 = H.colmajor((N, N), dt)
    H.lhs_assemble(lhs1, [s["colloc"]], [s["normals"]], [s["verts"]], [s["wake"]], 1)
    H.lhs_assemble(lhs1, [s["colloc"]], [s["normals"]], [s["verts"]], [s["wake"]], 1)
    lhs0 = H.colmajor((N, N), dt)
    H.lhs_assemble(lhs0, [s["colloc"]], [s["normals"]], [s["verts"]], [s["wake"]], 0)
    L1, L0 = H.to_host(lhs1), H.to_host(lhs0)
    gl = golden["rect4_a5_lhs"]
    assert rel(L1, 2 * gl) < 1e-13
    te = (s["nc"] - 1) * s["ns"]
    assert rel(L0[:, :te], gl[:, :te]) < 1e-13  # no wake term outside the trailing edge
    assert np.max(np.abs(L0[:, te:] - gl[:, te:])) > 1e-6  # the wake ring contributes at the TE


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("it", [1, 3, 5])
def test_wake_influence_vs_golden(H, torch, golden, oracle, prec, it):
    print(f"[case] wake influence it={it} {prec}", flush=True)
    dt = DT(torch, prec)
    c = H.to_device(golden["wake4_colloc"], dt)
    n = H.to_device(golden["wake4_normals"], dt)
    w = H.to_device(golden["wake4_verts_wake"], dt)
    g = H.to_device(golden["wake4_gamma_wake"], dt)
    rhs = torch.zeros(c.shape[0] * c.shape[1], dtype=dt, device="cuda")
    H.rhs_assemble_wake_influence(rhs, [g], [c], [n], [w], [True], it)
    assert rel(H.to_host(rhs), golden[f"wake4_rhs_it{it}"]) < (1e-13 if prec == "f64" else 2e-5)


def _random_wake_surface(rng, ns, nc, rows, y0=0.0):
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
    return v, wake, gw


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_wake_multi_surface_and_tiles(H, torch, oracle, prec):
    """3 surfaces (one non-lifting), spans wider than a 64-span tile, iteration < rows."""
    dt = DT(torch, prec)
    npdt = np.float64 if prec == "f64" else np.float32
    rng = np.random.default_rng(7)
    specs = [(70, 3, 40, 0.0), (5, 2, 9, 2.5), (12, 4, 17, 5.0)]
    lifting = [True, False, True]
    osurf, osurf64, dev = [], [], []
    for (ns, nc, rows, y0), lift in zip(specs, lifting):
        v, wake, gw = _random_wake_surface(rng, ns, nc, rows, y0)
        arrs = [np.asfortranarray(x.astype(npdt)) for x in (v, wake, gw)]
        osurf.append(oracle.Surface(verts=arrs[0], verts_wake=arrs[1], gamma_wake=arrs[2], lifting=lift))
        # FP64 evaluation of exactly the same (rounded) inputs: the accuracy yardstick for FP32
        a64 = [np.asfortranarray(x.astype(np.float64)) for x in arrs]
        osurf64.append(oracle.Surface(verts=a64[0], verts_wake=a64[1], gamma_wake=a64[2], lifting=lift))
    oracle.mesh_metrics(0.0, osurf)
    oracle.mesh_metrics(0.0, osurf64)
    for s in osurf:
        dev.append(dict(c=H.to_device(s.colloc, dt), n=H.to_device(s.normals, dt), w=H.to_device(s.verts_wake, dt),
                        g=H.to_device(s.gamma_wake, dt)))
    N = sum(s.ns * s.nc for s in osurf)
    for it in (1, 8, 17):
        rhs_o = np.zeros(N, dtype=npdt)
        oracle.rhs_assemble_wake_influence(rhs_o, osurf, it)
        rhs = torch.zeros(N, dtype=dt, device="cuda")
        H.rhs_assemble_wake_influence(rhs, [d["g"] for d in dev], [d["c"] for d in dev], [d["n"] for d in dev],
                                      [d["w"] for d in dev], lifting, it)
        if prec == "f64":
            assert rel(H.to_host(rhs), rhs_o) < 1e-13
        else:
            # FP32: no worse than 3x the FP32 restatement's own error against FP64 on the same inputs
            rhs64 = np.zeros(N)
            oracle.rhs_assemble_wake_influence(rhs64, osurf64, it)
            e_gpu, e_orc = rel(H.to_host(rhs), rhs64), rel(rhs_o, rhs64)
            print(f"[case] f32 wake it={it}: gpu err {e_gpu:.2e} fp32-oracle err {e_orc:.2e}", flush=True)
            assert e_gpu < max(3 * e_orc, 1e-6)
        # deterministic: the same call gives bitwise the same result
        rhs2 = torch.zeros(N, dtype=dt, device="cuda")
        H.rhs_assemble_wake_influence(rhs2, [d["g"] for d in dev], [d["c"] for d in dev], [d["n"] for d in dev],
                                      [d["w"] for d in dev], lifting, it)
        assert torch.equal(rhs, rhs2)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("n", [1, 5, 64, 65, 130, 700, 1025, 2100, 4100, 9000])
def test_lu_random(H, torch, prec, n):
    import scipy.linalg as sl

    print(f"[case] lu n={n} {prec}", flush=True)
    dt = DT(torch, prec)
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + (0.0 if n < 100 else 0.0)
    B = rng.standard_normal((n, 3))
    a = H.to_device(A, dt)
    x = H.to_device(B, dt)
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    H.getrf(a, ipiv, info)
    H.getrs(a, ipiv, x)
    assert int(info.item()) == 0
    X = H.to_host(x)
    if prec == "f64":
        lu_ref, piv_ref = sl.lu_factor(A)
        assert np.array_equal(ipiv.cpu().numpy() - 1, piv_ref)
        assert rel(H.to_host(a), lu_ref) < 1e-11
        res = np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X))
        assert res < 4 * n * np.finfo(np.float64).eps
    else:
        res = np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X))
        assert res < 4 * n * np.finfo(np.float32).eps


def test_lu_singular_info(H, torch):
    A = np.array([[1.0, 2.0, 0.0], [2.0, 4.0, 0.0], [0.0, 0.0, 1.0]])
    a = H.to_device(A, torch.float64)
    ipiv = torch.zeros(3, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    H.getrf(a, ipiv, info)
    assert int(info.item()) == 2


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("n,batch", [(300, 1), (2048, 1), (200, 8), (1500, 3)])
def test_lu_nan_columns_keep_pivots_in_range(H, torch, prec, n, batch):
    """A column of NaNs wins no pivot comparison.  Every pivot must still be a row of the matrix
    (1 <= ipiv <= n; a NaN column takes its diagonal row, as LAPACK's idamax stays on the first
    entry), so that the interchanges and the solve stay inside the matrix; the other matrices of a
    batch are factored exactly as without the NaN member.  Paths: one-group panels (300), cooperative
    panels (2048), the batched look-ahead (200 x 8), cooperative panels over a batch (1500 x 3)."""
    import scipy.linalg as sl

    print(f"[case] lu NaN column n={n} batch={batch} {prec}", flush=True)
    dt = DT(torch, prec)
    rng = np.random.default_rng(77 + n)
    As = [rng.standard_normal((n, n)) for _ in range(batch)]
    As[0][:, n // 3] = np.nan  # member 0: NaN from column n/3 on, through the elimination
    a = H.colmajor((n, n, batch), dt)
    x = H.colmajor((n, 1, batch), dt)
    for b in range(batch):
        a[:, :, b] = H.to_device(As[b], dt)
        x[:, :, b] = H.to_device(np.ones((n, 1)), dt)
    ipiv = torch.zeros(n * batch, dtype=torch.int32, device="cuda")
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    H.getrf(a, ipiv, info)
    H.getrs(a, ipiv, x)
    torch.cuda.synchronize()
    piv = ipiv.cpu().numpy().reshape(batch, n)
    assert piv.min() >= 1 and piv.max() <= n
    # pivots never point above their own row: row k is exchanged with a row k' >= k
    assert np.all(piv - 1 >= np.arange(n)[None, :])
    assert not np.all(np.isfinite(H.to_host(x[:, :, 0])))
    for b in range(1, batch):
        X = H.to_host(x[:, :, b])
        res = np.linalg.norm(As[b] @ X - 1.0) / (np.linalg.norm(As[b]) * np.linalg.norm(X))
        eps = np.finfo(np.float64 if prec == "f64" else np.float32).eps
        assert res < 4 * n * eps
        if prec == "f64":
            assert np.array_equal(piv[b] - 1, sl.lu_factor(As[b])[1])


def test_getrs_many_rhs(H, torch):
    n, k = 600, 11  # harmonic-balance shape: one factor, 2H+1 right-hand sides
    rng = np.random.default_rng(11)
    A = rng.standard_normal((n, n)) + n * np.eye(n) * 0.05
    B = rng.standard_normal((n, k))
    a = H.to_device(A, torch.float64)
    x = H.to_device(B, torch.float64)
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    H.getrf(a, ipiv, info)
    H.getrs(a, ipiv, x)
    assert rel(H.to_host(x), np.linalg.solve(A, B)) < 1e-11


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_forces_coeffs_mac_sum(H, torch, oracle, prec):
    dt = DT(torch, prec)
    npdt = np.float64 if prec == "f64" else np.float32
    tol = 1e-13 if prec == "f64" else 1e-5
    rng = np.random.default_rng(3)
    ns, nc = 37, 9
    v = oracle.rectangular_wing(ns, nc, chord=1.0, half_span=3.0).astype(npdt)
    v[:, :, 2] += (0.01 * rng.standard_normal((ns + 1, nc + 1))).astype(npdt)
    v = np.asfortranarray(v)
    s = oracle.Surface(verts=v)
    oracle.mesh_metrics(0.0, [s])
    gd = np.asfortranarray(rng.standard_normal((ns, nc)).astype(npdt))
    gdt = np.asfortranarray(rng.standard_normal((ns, nc)).astype(npdt))
    vel = np.asfortranarray((rng.standard_normal((ns, nc, 3)) + [1, 0, 0]).astype(npdt))
    Fs = oracle.forces_steady(v, gd, vel)
    Fu = oracle.forces_unsteady(v, gd, gdt, vel, s.areas, s.normals)
    dv = {k: H.to_device(x, dt) for k, x in dict(v=v, gd=gd, gdt=gdt, vel=vel, a=s.areas, n=s.normals).items()}
    f = H.colmajor((ns, nc, 3), dt)
    H.forces_steady(dv["v"], dv["gd"], dv["vel"], f)
    assert rel(H.to_host(f), Fs) < tol
    H.forces_unsteady(dv["v"], dv["gd"], dv["gdt"], dv["vel"], dv["a"], dv["n"], f)
    assert rel(H.to_host(f), Fu) < tol
    out = torch.zeros(8, dtype=dt, device="cuda")
    fs = [1.0, 0.0, 0.05]
    la = [-0.05, 0.0, 1.0]
    area = oracle.sum_(s.areas)
    mac = oracle.mesh_mac(v, s.areas)
    H.sum_(dv["a"], out[0:1])
    H.mesh_mac(dv["v"], dv["a"], out[1:2])
    H.coeff_cl(f, la, fs, 1.0, float(area), out[2:3])
    H.coeff_cm(f, dv["v"], [0.25, 0.0, 0.0], fs, 1.0, float(area), float(mac), out[3:6])
    got = out.cpu().numpy().astype(np.float64)
    assert abs(got[0] - area) <= tol * abs(area)
    assert abs(got[1] - mac) <= tol * abs(mac)
    cl = oracle.coeff_cl(Fu, la, fs, 1.0, area)
    cm = oracle.coeff_cm(Fu, v, [0.25, 0, 0], fs, 1.0, area, mac)
    assert abs(got[2] - cl) <= 10 * tol * abs(cl)
    assert rel(got[3:6], cm) < 10 * tol


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_shed_displace_coeff_wake_kinematics(H, torch, oracle, prec):
    dt = DT(torch, prec)
    npdt = np.float64 if prec == "f64" else np.float32
    v = np.asfortranarray(oracle.rectangular_wing(6, 3).astype(npdt))
    wk = np.zeros((7, 5, 4), dtype=npdt, order="F")
    dv, dw = H.to_device(v, dt), H.to_device(wk, dt)
    H.wake_shed([dv], [dw], 2)
    oracle.wake_shed(v, wk, 2)
    assert np.array_equal(H.to_host(dw), wk)
    T, dT = oracle.motion_transform([1.0, 0.1, 0.3, 0.05, 0.2, 1.1, 0.25], 0.77)
    out = H.colmajor(v.shape, dt)
    H.displace_wing([H.to_device(T, dt)], [out], [dv])
    assert rel(H.to_host(out), oracle.displace_wing(np.asfortranarray(T.astype(npdt)), v)) < (1e-15 if prec == "f64" else 1e-6)
    # gamma_wake_from_coeffs
    rng = np.random.default_rng(9)
    gc = np.asfortranarray(rng.standard_normal((6, 11)).astype(npdt))
    gw = np.zeros((6, 9), dtype=npdt, order="F")
    dgw = H.to_device(gw, dt)
    H.gamma_wake_from_coeffs(dgw, H.to_device(gc, dt), 5, 0.7, 1.3, 0.05, 6)
    oracle.gamma_wake_from_coeffs(gw, gc, 5, 0.7, 1.3, 0.05, 6)
    assert rel(H.to_host(dgw), gw) < (1e-13 if prec == "f64" else 1e-5)
    # kinematic velocities = -(dT [x;1]) at the collocation points
    s = oracle.Surface(verts=v)
    oracle.mesh_metrics(0.0, [s])
    c = H.to_device(s.colloc, dt)
    vel = H.colmajor(s.colloc.shape, dt)
    H.kinematic_velocities([dT], [c], [vel])
    pts = s.colloc.reshape(-1, 3, order="F").astype(np.float64)
    ref = -(pts @ dT[:3, :3].T + dT[:3, 3])
    assert rel(H.to_host(vel).reshape(-1, 3, order="F"), ref) < (1e-15 if prec == "f64" else 1e-6)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_blas(H, torch, prec):
    dt = DT(torch, prec)
    tol = 1e-13 if prec == "f64" else 1e-5
    rng = np.random.default_rng(2)
    for ta in (0, 1):
        for tb in (0, 1):
            A = rng.standard_normal((23, 17) if not ta else (17, 23))
            B = rng.standard_normal((17, 31) if not tb else (31, 17))
            C = rng.standard_normal((23, 31))
            c = H.to_device(C, dt)
            H.gemm(0.7, H.to_device(A, dt), H.to_device(B, dt), 0.3, c, ta, tb)
            ref = 0.7 * (A.T if ta else A) @ (B.T if tb else B) + 0.3 * C
            assert rel(H.to_host(c), ref) < tol
    A = rng.standard_normal((40, 25))
    for tr in (0, 1):
        x = rng.standard_normal(40 if tr else 25)
        y = rng.standard_normal(25 if tr else 40)
        dy = H.to_device(y, dt)
        H.gemv(1.5, H.to_device(A, dt), H.to_device(x, dt), -0.5, dy, tr)
        assert rel(H.to_host(dy), 1.5 * ((A.T if tr else A) @ x) - 0.5 * y) < tol
    X = rng.standard_normal((9, 4))
    Y = rng.standard_normal((9, 4))
    dY = H.to_device(Y, dt)
    H.axpy(-2.0, H.to_device(X, dt)[:, 1:3], dY[:, 0:2])
    Yr = Y.copy()
    Yr[:, 0:2] += -2.0 * X[:, 1:3]
    assert rel(H.to_host(dY), Yr) < tol
    x = rng.standard_normal(100)
    dx = H.to_device(x, dt)
    H.scal(3.0, dx)
    assert rel(H.to_host(dx), 3 * x) < tol
    out = torch.zeros(2, dtype=dt, device="cuda")
    H.nrm2(dx, out[0:1])
    H.sum_(dx, out[1:2])
    o = out.cpu().numpy()
    assert abs(o[0] - np.linalg.norm(3 * x)) < 10 * tol * np.linalg.norm(3 * x)
    assert abs(o[1] - np.sum(3 * x)) < 100 * tol * np.sum(np.abs(3 * x))


def test_steady_end_to_end_vs_oracle(H, torch, oracle, root):
    """VLM::run call sequence on the reference's own fixtures, FP64, against the oracle."""
    from tinyvlm_amd.steady import SteadyVLM

    for meshes in (["rectangular_4x4"], ["infinite_rectangular_2x2_part0", "infinite_rectangular_2x2_part1"],
                   ["elliptic_16x16"], ["swept_4x4"]):
        vs = [oracle.read_mesh(os.path.join(root, "mesh", m + ".x")) for m in meshes]
        s = SteadyVLM(vs, dtype=torch.float64)
        for deg in (0, 5, 15):
            a = to_radians(deg)
            g, cl, cm = s.run(a)
            g_o, cl_o, cm_o, _ = oracle.vlm_run(vs, a)
            print(f"[case] steady {meshes} {deg}: cl {cl:.15f} vs {cl_o:.15f}", flush=True)
            assert rel(g, g_o) < 1e-12
            assert abs(cl - cl_o) <= 1e-12 * max(abs(cl_o), 1e-3)
            assert rel(cm, cm_o) < 1e-11


def test_steady_elliptic_64_full_size(H, torch, oracle, root):
    """The steady workload at full size: elliptic 64x64 (N=4096), alpha = 5 deg."""
    from tinyvlm_amd.steady import SteadyVLM

    v = oracle.read_mesh(os.path.join(root, "mesh", "elliptic_64x64.x"))
    a = to_radians(5)
    s = SteadyVLM([v], dtype=torch.float64)
    g, cl, cm = s.run(a)
    g_o, cl_o, cm_o, _ = oracle.vlm_run([v], a)
    print(f"[case] elliptic64: cl gpu {cl:.15f} oracle {cl_o:.15f} gamma rel {rel(g, g_o):.3e}", flush=True)
    assert rel(g, g_o) < 1e-12
    assert abs(cl - cl_o) <= 1e-12 * abs(cl_o)
    # analytic known answer of tests/vlm_elliptic_coeffs.cpp (3%)
    ar = 100 / (0.5 * float(F32(3.141593)) * 5.0)
    an = 2 * float(F32(3.141593)) * a / (1 + 2 / ar)
    assert abs(cl - an) / an < 0.03


def test_uvlm_fp64_vs_oracle(H, torch, oracle):
    """UVLM::run call sequence, pitch + heave, FP64, 40 steps on the 10 x 40 AR-4 wing."""
    from tinyvlm_amd.kinematics import pitch_heave
    from tinyvlm_amd.unsteady import UVLM

    v = oracle.rectangular_wing(40, 10, chord=1.0, half_span=2.0)
    a_amp = float(F32(2) * F32(3.141593) / F32(180))
    mot = [1.0, 0.1, 0.0, a_amp, math.pi / 2, 1.0, 0.25]
    tree, fs, pt = pitch_heave(u=1.0, h_amp=0.1, a_amp=a_amp, a_phase=math.pi / 2, omega=1.0, x_pivot=0.25)
    sim = UVLM([v], [pt], fs, dtype=torch.float64)
    out = sim.run(t_steps=41)
    cl_o, cm_o, g_o, n = oracle.uvlm_run([v], mot, t_steps=41)
    print(f"[case] uvlm f64: cl[-1] {out['cl'][-1]:.15f} oracle {cl_o[-1]:.15f}", flush=True)
    assert rel(out["cl"][1:], cl_o[1:]) < 1e-11
    assert rel(out["cm_y"][1:], cm_o[1:]) < 1e-10
    assert rel(out["gamma"], g_o) < 1e-11


def test_error_status_no_launch(H, torch):
    """Bad strides are rejected on the host with TVLM_ERR_INVALID (no kernel runs)."""
    from tinyvlm_amd import capi

    c = torch.zeros(3, 4, 5, dtype=torch.float64, device="cuda")  # row-major: stride[0] != 1
    lhs = capi.colmajor((60, 60), torch.float64)
    with pytest.raises(capi.TvlmError, match="status 1"):
        capi.lhs_assemble(lhs, [c], [c], [c], [c], 0)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("n,batch", [(130, 5), (1500, 3), (200, 9), (1100, 8), (2200, 8), (3000, 8), (6400, 2), (5000, 7)])
def test_lu_batched(H, torch, prec, n, batch):
    """A batch of independent matrices in one getrf/getrs pass equals factoring each alone
    (LAPACK pivots for FP64, per-matrix info, solutions to the single-matrix bar).  Batches of 8 or
    more take the look-ahead factorisation (64-column panels, split trailing updates; while the
    batch's trailing matrices exceed 256 MB, two panels per rank-128 update: 2200 x 8 FP64 and
    3000 x 8 start that way).  Smaller
    batches of tall matrices take cooperative panels (several work-groups per matrix) when the
    first panel's groups fit (1500 x 3); while the trailing matrix is taller than 6144 rows they
    are taken two at a time with one rank-128 trailing update per pair (6400 x 2: two pairs, then
    single panels); n = 5000, batch 7 has too many groups for that and keeps single-group panels."""
    import scipy.linalg as sl

    print(f"[case] lu batched n={n} batch={batch} {prec}", flush=True)
    dt = DT(torch, prec)
    rng = np.random.default_rng(1000 + n)
    As = [rng.standard_normal((n, n)) for _ in range(batch)]
    As[1][:, 7] = 0.0  # exactly singular member: info = 8 for it only
    Bs = [rng.standard_normal((n, 2)) for _ in range(batch)]
    a = H.colmajor((n, n, batch), dt)
    x = H.colmajor((n, 2, batch), dt)
    for b in range(batch):
        a[:, :, b] = H.to_device(As[b], dt)
        x[:, :, b] = H.to_device(Bs[b], dt)
    ipiv = torch.zeros(n * batch, dtype=torch.int32, device="cuda")
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    H.getrf(a, ipiv, info)
    H.getrs(a, ipiv, x)
    inf = info.cpu().numpy()
    assert inf[1] == 8 and all(inf[b] == 0 for b in range(batch) if b != 1)
    piv = ipiv.cpu().numpy().reshape(batch, n)
    for b in range(batch):
        if b == 1:
            continue
        X = H.to_host(x[:, :, b])
        res = np.linalg.norm(As[b] @ X - Bs[b]) / (np.linalg.norm(As[b]) * np.linalg.norm(X))
        if prec == "f64":
            lu_ref, piv_ref = sl.lu_factor(As[b])
            assert np.array_equal(piv[b] - 1, piv_ref)
            assert rel(H.to_host(a[:, :, b]), lu_ref) < 1e-11
            assert res < 4 * n * np.finfo(np.float64).eps
        else:
            assert res < 4 * n * np.finfo(np.float32).eps


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("n,batch", [(2200, 1), (3000, 2), (6400, 2)])
def test_lu_coop_single_exchange_attempt(H, torch, prec, n, batch):
    """Cooperative panels first try one exchange per panel (the first work-group pivots among its
    own rows, the others eliminate against its U11 and check that no row of theirs would have been
    the pivot) and fall back to the column-by-column exchange when the check fails.  Matrix 0 is
    column diagonally dominant (LAPACK never interchanges: every attempt stands); matrix 1 has rows
    interchanged so that LAPACK pivots once inside the first group's rows (300 <-> 400: the attempt
    stands with a row interchange) and twice far below them in consecutive panels (1000 <-> 1900,
    1070 <-> 1870: two failed attempts in a row, after which the rest of that matrix takes the
    exchange loop).  Pivots equal LAPACK's, factors within 1e-11 (FP64), solutions to the
    single-matrix bar."""
    import scipy.linalg as sl

    print(f"[case] lu coop attempt n={n} batch={batch} {prec}", flush=True)
    dt = DT(torch, prec)
    rng = np.random.default_rng(77 + n)
    As = []
    for b in range(batch):
        A = rng.standard_normal((n, n)) + n * np.eye(n)
        if b == 1:
            for r, s in ((300, 400), (1000, 1900), (1070, 1870)):
                A[[r, s]] = A[[s, r]]
        As.append(A)
    Bs = [rng.standard_normal((n, 1)) for _ in range(batch)]
    a = H.colmajor((n, n, batch), dt)
    x = H.colmajor((n, 1, batch), dt)
    for b in range(batch):
        a[:, :, b] = H.to_device(As[b], dt)
        x[:, :, b] = H.to_device(Bs[b], dt)
    ipiv = torch.zeros(n * batch, dtype=torch.int32, device="cuda")
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    H.getrf(a, ipiv, info)
    H.getrs(a, ipiv, x)
    assert np.all(info.cpu().numpy() == 0)
    piv = ipiv.cpu().numpy().reshape(batch, n)
    for b in range(batch):
        lu_ref, piv_ref = sl.lu_factor(As[b])
        assert np.array_equal(piv[b] - 1, piv_ref)
        if b == 1:
            assert piv_ref[300] == 400 and piv_ref[1000] == 1900 and piv_ref[1070] == 1870
        X = H.to_host(x[:, :, b])
        res = np.linalg.norm(As[b] @ X - Bs[b]) / (np.linalg.norm(As[b]) * np.linalg.norm(X))
        if prec == "f64":
            assert rel(H.to_host(a[:, :, b]), lu_ref) < 1e-11
            assert res < 4 * n * np.finfo(np.float64).eps
        else:
            assert rel(H.to_host(a[:, :, b]), lu_ref) < 1e-4
            assert res < 4 * n * np.finfo(np.float32).eps


def test_current_stream_handle_matches_torch(H, torch):
    """capi.current_stream_handle (the launches' default stream) is torch's current stream, outside
    and inside a stream context."""
    assert H.current_stream_handle() == torch.cuda.current_stream().cuda_stream
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        assert H.current_stream_handle() == s.cuda_stream == torch.cuda.current_stream().cuda_stream
    assert H.current_stream_handle() == torch.cuda.current_stream().cuda_stream
    assert H.current_stream_handle() != s.cuda_stream
