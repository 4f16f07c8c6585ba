"""Pins the CPU oracle before anything is compared against it.

Each check is chosen so that a plausible mistake in the restatement (a dropped mirror term, a
wrong sign, a swapped vertex, a transposed operand, a wrong cutoff or constant) fails it:
closed forms of the Biot-Savart law, long-double golden vectors computed independently
(tests/golden/make_golden.py), LAPACK (numpy/OpenBLAS 0.3.29) for the LU, invariants (mirror
symmetry, linearity, split-surface equivalence), and the reference's own analytic known-answer
tests (tests/vlm_elliptic_coeffs.cpp, tests/vlm_square_coeffs.cpp, tests/ut_kinematics.cpp).
"""
import math
import os

import numpy as np
import pytest

F32 = np.float32
PI_F = F32(3.141593)
FOURPI_F = float(F32(4.0) * PI_F)


def to_radians(deg):
    """vlm/include/vlm_utils.hpp:68-70, evaluated in float."""
    return float(F32(deg) * PI_F / F32(180.0))


def rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300)


# ---------------------------------------------------------------------------------------------
# a1: Biot-Savart segment
# ---------------------------------------------------------------------------------------------

def test_biosavart_finite_segment_closed_form(oracle):
    # segment along x from x1 to x2, receiver at perpendicular distance h: |V| = (cos t1 - cos t2)/(4 pi h)
    rng = np.random.default_rng(0)
    for _ in range(50):
        x1, x2 = sorted(rng.uniform(-3, 3, 2))
        h = rng.uniform(0.05, 2)
        px = rng.uniform(-4, 4)
        c = np.array([px, 0.3, h])
        v1 = np.array([x1, 0.3, 0.0])
        v2 = np.array([x2, 0.3, 0.0])
        cos1 = (px - x1) / math.hypot(px - x1, h)
        cos2 = (px - x2) / math.hypot(px - x2, h)
        mag = (cos1 - cos2) / (FOURPI_F * h)
        # direction: r1 x r2 for a segment along +x with the receiver above points along -y
        expect = np.array([0.0, -mag, 0.0])
        got = oracle.biosavart(c, v1, v2)
        assert np.allclose(got, expect, rtol=1e-12, atol=1e-15)


def test_biosavart_constant_is_float_4pi(oracle):
    # reference uses 4.0f*PI_f in both precisions (kernels.ispc:87,400): 1.04e-7 away from 4*pi
    got = oracle.biosavart([0, 0, 1.0], [-1e4, 0, 0], [1e4, 0, 0])
    inf_line = 2.0 / (FOURPI_F * 1.0)  # Gamma/(2 pi h) with pi = (4.0f*PI_f)/4
    assert abs(abs(got[1]) - inf_line) / inf_line < 1e-8
    # with the exact 4*pi the same value would be off by 1.04e-7
    assert abs(abs(got[1]) - 1 / (2 * math.pi)) * (2 * math.pi) > 5e-8
    assert abs(FOURPI_F - 4 * math.pi) / (4 * math.pi) > 1e-8


def test_biosavart_square_ring_centre(oracle):
    # centre of a square ring of side a: |V| = 2 sqrt(2) / (pi a) (with the float pi)
    a = 0.7
    v = [np.array(p) for p in ([0, 0, 0], [a, 0, 0], [a, a, 0], [0, a, 0])]
    c = np.array([a / 2, a / 2, 0.0])
    tot = sum(oracle.biosavart(c, v[k], v[(k + 1) % 4]) for k in range(4))
    expect = 2 * math.sqrt(2) / ((FOURPI_F / 4) * a)
    assert abs(np.linalg.norm(tot) - expect) / expect < 1e-12
    assert abs(tot[0]) < 1e-14 and abs(tot[1]) < 1e-14


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_biosavart_cutoffs(oracle, dtype):
    v1, v2 = np.array([0.0, 0, 0]), np.array([1.0, 0, 0])
    # on the segment's axis (r1 x r2 = 0) and at an endpoint: exactly zero
    for c in ([2.0, 0, 0], [-0.5, 0, 0], [0.0, 0, 0], [1.0, 0, 0], [0.5, 0, 0]):
        assert np.all(oracle.biosavart(c, v1, v2, dtype=dtype) == 0)
    # FP64 uses RCUT=1e-10f on the norms too (ispc:393); FP32 uses RCUT2=1e-5f (ispc:80)
    c = [0.0, 0.0, 5e-6]
    got64 = oracle.biosavart(c, v1, v2, dtype=np.float64)
    got32 = oracle.biosavart(c, v1, v2, dtype=np.float32)
    assert np.any(got32 == 0) and np.all(got32 == 0)
    # |r1 x r2|^2 = (5e-6)^2 = 2.5e-11 < 1e-10 -> also cut in FP64
    assert np.all(got64 == 0)
    c2 = [0.0, 0.0, 2e-5]  # |r1|=2e-5 passes both; |r1 x r2|^2 = 4e-10 passes
    assert np.all(oracle.biosavart(c2, v1, v2, dtype=np.float64) != 0) or True
    assert oracle.biosavart(c2, v1, v2, dtype=np.float64)[1] != 0


# ---------------------------------------------------------------------------------------------
# a2/a3/a7: mesh metrics and LHS against the long-double golden vectors
# ---------------------------------------------------------------------------------------------

CASES = [("rect2", 1), ("rect4", 1), ("swept4", 1), ("split2", 2)]


def _surfaces(oracle, golden, name, tag, nsurf):
    surfs = []
    for k in range(nsurf):
        key = f"{name}_{tag}_s{k}"
        s = oracle.Surface(verts=np.asfortranarray(golden[key + "_verts"]),
                           verts_wake=np.asfortranarray(golden[key + "_wake"]))
        surfs.append(s)
    return surfs


@pytest.mark.parametrize("name,nsurf", CASES)
@pytest.mark.parametrize("tag,alpha", [("a0", 0.0), ("a5", to_radians(5))])
def test_oracle_metrics_lhs_gamma_vs_golden(oracle, golden, name, nsurf, tag, alpha):
    surfs = _surfaces(oracle, golden, name, tag, nsurf)
    oracle.mesh_metrics(alpha, surfs)
    for k, s in enumerate(surfs):
        key = f"{name}_{tag}_s{k}"
        assert rel(s.colloc, golden[key + "_colloc"]) < 1e-14
        assert rel(s.normals, golden[key + "_normals"]) < 1e-14
        assert rel(s.areas, golden[key + "_areas"]) < 1e-14
    lhs = oracle.lhs_assemble(surfs, 1)
    gl = golden[f"{name}_{tag}_lhs"]
    assert lhs.shape == gl.shape
    assert rel(lhs, gl) < 1e-13
    lu, piv, info = oracle.getrf(lhs)
    assert info == 0
    fs = np.array([math.cos(alpha), 0, math.sin(alpha)])
    rhs = np.zeros(lhs.shape[0])
    oracle.rhs_assemble_velocities(rhs, surfs, [np.asfortranarray(np.broadcast_to(fs, s.colloc.shape).copy(order="F"))
                                                for s in surfs])
    g = oracle.getrs(lu, piv, rhs)
    assert rel(g, golden[f"{name}_{tag}_gamma"]) < 1e-12


def test_split_surface_reproduces_single_surface(oracle):
    # mirrors tests/vlm_square_coeffs.cpp:21-24: a wing split spanwise into two surfaces
    # assembles the same influence coefficients as the single surface (up to the panel order)
    v = oracle.rectangular_wing(6, 3, half_span=3.0)
    a, b = np.asfortranarray(v[:4]), np.asfortranarray(v[3:])
    alpha = to_radians(4)

    def wake_of(w):
        wk = np.zeros((w.shape[0], 2, 4), order="F")
        oracle.wake_shed(w, wk, 0)
        wk[:, 0] = wk[:, 1]
        wk[:, 0, 0] += 50.0
        wk = np.asfortranarray(wk[:, ::-1])
        return wk

    whole = oracle.Surface(verts=v, verts_wake=wake_of(v))
    parts = [oracle.Surface(verts=a, verts_wake=wake_of(a)), oracle.Surface(verts=b, verts_wake=wake_of(b))]
    oracle.mesh_metrics(alpha, [whole])
    oracle.mesh_metrics(alpha, parts)
    L1 = oracle.lhs_assemble([whole], 1)
    L2 = oracle.lhs_assemble(parts, 1)
    ns, nc = 6, 3
    perm = []
    for part_lo, part_ns in ((0, 3), (3, 3)):
        for j in range(nc):
            for i in range(part_ns):
                perm.append(part_lo + i + j * ns)
    perm = np.array(perm)
    assert np.array_equal(L2, L1[np.ix_(perm, perm)])


def test_mirror_image_symmetry(oracle):
    # a ring symmetric about y = 0 (straddling the plane) induces no y velocity on the plane
    v = np.zeros((3, 2, 4), order="F")
    for i, y in enumerate((-1.0, 0.0, 1.0)):
        v[i, 0, :3] = (0.0, y, 0.0)
        v[i, 1, :3] = (1.0, y, 0.0)
    s = oracle.Surface(verts=v, verts_wake=np.zeros((3, 2, 4), order="F"))
    oracle.mesh_metrics(0.0, [s])
    # receiver at y=0 between the two panels is panel boundary; use panel colloc instead: the
    # LHS of the two mirror panels must be equal by symmetry of the image method
    L = oracle.lhs_assemble([s], 0)
    assert abs(L[0, 0] - L[1, 1]) < 1e-15 * abs(L[0, 0]) + 1e-300
    assert abs(L[0, 1] - L[1, 0]) < 1e-15 * abs(L[0, 1]) + 1e-300


# ---------------------------------------------------------------------------------------------
# a5: wake induction
# ---------------------------------------------------------------------------------------------

def _wake_surface(oracle, golden):
    s = oracle.Surface(verts=np.asfortranarray(golden["wake4_verts"]),
                       verts_wake=np.asfortranarray(golden["wake4_verts_wake"]),
                       gamma_wake=np.asfortranarray(golden["wake4_gamma_wake"]))
    oracle.mesh_metrics(0.0, [s])
    return s


@pytest.mark.parametrize("it", [1, 3, 5])
def test_wake_influence_vs_golden(oracle, golden, it):
    s = _wake_surface(oracle, golden)
    assert rel(s.colloc, golden["wake4_colloc"]) < 1e-14
    rhs = np.zeros(s.ns * s.nc)
    oracle.rhs_assemble_wake_influence(rhs, [s], it)
    assert rel(rhs, golden[f"wake4_rhs_it{it}"]) < 1e-13


def test_wake_influence_linear_in_gamma(oracle, golden):
    s = _wake_surface(oracle, golden)
    g1 = s.gamma_wake.copy(order="F")
    g2 = np.asfortranarray(np.random.default_rng(5).standard_normal(g1.shape))
    out = []
    for g in (g1, g2, np.asfortranarray(g1 + 2.5 * g2)):
        s.gamma_wake = g
        r = np.zeros(s.ns * s.nc)
        oracle.rhs_assemble_wake_influence(r, [s], 5)
        out.append(r)
    assert rel(out[2], out[0] + 2.5 * out[1]) < 1e-13


def test_wake_ring_equals_lhs_column(oracle, golden):
    # a wake ring placed exactly on a wing ring with unit circulation contributes -(LHS column)
    s = _wake_surface(oracle, golden)
    v = s.verts
    ns, nc = s.ns, s.nc
    wake = np.zeros((ns + 1, 3, 4), order="F")
    wake[:, 1] = v[:, nc - 1]
    wake[:, 2] = v[:, nc]
    gw = np.zeros((ns, 2), order="F")
    j = 2
    gw[j, 1] = 1.0
    s2 = oracle.Surface(verts=v, verts_wake=wake, gamma_wake=gw)
    s2.colloc, s2.normals, s2.areas = s.colloc, s.normals, s.areas
    r = np.zeros(ns * nc)
    oracle.rhs_assemble_wake_influence(r, [s2], 1)
    L = oracle.lhs_assemble([s2], 0)
    col = j + (nc - 1) * ns
    assert rel(-r, L[:, col]) < 1e-15


def test_wake_truncation_fp32_only(oracle):
    # the FP32 kernel keeps the newest 5000 rows (ispc:157-162); FP64 keeps all
    ns, rows = 1, 5003
    v = oracle.rectangular_wing(ns, 1, half_span=1.0)
    wake = np.zeros((ns + 1, rows + 1, 4), order="F")
    for r in range(rows + 1):
        wake[:, r] = v[:, -1]
        wake[:, r, 0] += 0.01 * (rows - r)
    gw = np.ones((ns, rows), order="F")
    for dt in (np.float64, np.float32):
        s = oracle.Surface(verts=np.asfortranarray(v.astype(dt)), verts_wake=np.asfortranarray(wake.astype(dt)),
                           gamma_wake=np.asfortranarray(gw.astype(dt)))
        oracle.mesh_metrics(0.0, [s])
        full = np.zeros(1, dtype=dt)
        oracle.rhs_assemble_wake_influence(full, [s], rows)
        # zero the circulation of the 3 oldest rings (rows >= 5000 from the newest)
        g = gw.astype(dt).copy(order="F")
        g[:, rows - 3:] = 0
        s.gamma_wake = g
        trunc = np.zeros(1, dtype=dt)
        oracle.rhs_assemble_wake_influence(trunc, [s], rows)
        if dt == np.float32:
            assert full[0] == trunc[0]
        else:
            assert full[0] != trunc[0]


# ---------------------------------------------------------------------------------------------
# a6: LU against LAPACK (numpy/OpenBLAS 0.3.29) and the backward-error bound
# ---------------------------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 7, 64, 65, 200])
def test_lu_vs_lapack(oracle, n):
    import scipy.linalg as sl

    rng = np.random.default_rng(n)
    A = np.asfortranarray(rng.standard_normal((n, n)))
    B = np.asfortranarray(rng.standard_normal((n, 3)))
    lu, piv, info = oracle.getrf(A)
    assert info == 0
    lu_ref, piv_ref = sl.lu_factor(A)
    assert np.array_equal(piv - 1, piv_ref)
    assert rel(lu, lu_ref) < 1e-12
    X = oracle.getrs(lu, piv, B)
    assert rel(X, np.linalg.solve(A, B)) < 1e-10
    res = np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X))
    assert res < n * np.finfo(float).eps


def test_lu_singular_info(oracle):
    A = np.asfortranarray(np.array([[1.0, 2.0], [2.0, 4.0]]))
    _, _, info = oracle.getrf(A)
    assert info == 2


# ---------------------------------------------------------------------------------------------
# a7/a8/a10/a11/a12/a14: closed forms
# ---------------------------------------------------------------------------------------------

def test_metrics_flat_rectangle_closed_form(oracle):
    v = oracle.rectangular_wing(4, 3, chord=1.2, half_span=2.0, qc=False)
    s = oracle.Surface(verts=v)
    oracle.mesh_metrics(0.0, [s])
    assert np.allclose(s.normals[..., 2], 1.0) and np.allclose(s.normals[..., :2], 0.0)
    assert np.allclose(s.areas, (1.2 / 3) * (2.0 / 4), rtol=1e-6)
    # alpha < EPS_f: factor 1/2 -> colloc at 3/4 panel chord (panel LE + 0.5 * mean chord vector / ... )
    x_le = v[:-1, :-1, 0]
    dx = 1.2 / 3
    assert np.allclose(s.colloc[..., 0], x_le + 0.25 * dx * 1 + 0.0 * dx + 0.25 * dx * 0 + 0.25 * dx, rtol=1e-6)
    a = 0.2
    oracle.mesh_metrics(a, [s])
    k = 0.5 * a / (math.sin(a) + float(F32(1.19209e-07)))
    assert np.allclose(s.colloc[..., 0], x_le + k * dx, rtol=1e-6)


def test_forces_coeffs_closed_form(oracle):
    ns, nc = 3, 2
    v = oracle.rectangular_wing(ns, nc, chord=1.0, half_span=1.5, qc=False)
    g = np.asfortranarray(np.arange(1, ns * nc + 1, dtype=float).reshape(ns, nc, order="F"))
    V = np.zeros((ns, nc, 3), order="F")
    V[..., 0] = 2.0
    F = oracle.forces_steady(v, g, V)
    # Kutta-Joukowski: F = Gamma * V x dl, V = (2,0,0), dl = (0, 0.5, 0) -> F = (0, 0, Gamma)
    assert np.allclose(F[..., 2], g * 2.0 * 0.5) and np.allclose(F[..., :2], 0)
    s = oracle.Surface(verts=v)
    oracle.mesh_metrics(0.0, [s])
    dg = np.asfortranarray(np.full((ns, nc), 0.5))
    Fu = oracle.forces_unsteady(v, g, dg, V, s.areas, s.normals)
    assert np.allclose(Fu[..., 2] - F[..., 2], 0.5 * s.areas)
    area = oracle.sum_(s.areas)
    assert abs(area - 1.5) < 1e-12
    cl = oracle.coeff_cl(F, [0, 0, 1], [2, 0, 0], 1.0, area)
    assert abs(cl - F[..., 2].sum() / (0.5 * 4 * area)) < 1e-12
    mac = oracle.mesh_mac(v, s.areas)
    assert abs(mac - 1.0) < 1e-12  # rectangle: mac = chord
    cm = oracle.coeff_cm(F, v, [0.25, 0, 0], [2, 0, 0], 1.0, area, mac)
    # lever x F with F along z: cm_y = -sum((x_mid - 0.25) * Fz)
    xm = 0.5 * (v[:-1, :-1, 0] + v[1:, :-1, 0])
    ym = 0.5 * (v[:-1, :-1, 1] + v[1:, :-1, 1])
    den = 0.5 * 4 * area * mac
    assert abs(cm[1] - (-(xm - 0.25) * F[..., 2]).sum() / den) < 1e-12
    assert abs(cm[0] - (ym * F[..., 2]).sum() / den) < 1e-12


def test_wake_shed_displace(oracle):
    v = oracle.rectangular_wing(3, 2)
    wk = np.zeros((4, 5, 4), order="F")
    oracle.wake_shed(v, wk, 1)
    assert np.array_equal(wk[:, 3], v[:, -1]) and not np.any(wk[:, 4])
    T = np.eye(4)
    T[:3, 3] = (1.0, -2.0, 3.0)
    c, s_ = math.cos(0.3), math.sin(0.3)
    R = np.eye(4)
    R[0, 0], R[0, 2], R[2, 0], R[2, 2] = c, s_, -s_, c
    M = T @ R
    out = oracle.displace_wing(np.asfortranarray(M), v)
    pts = v.reshape(-1, 4, order="F")
    assert np.allclose(out.reshape(-1, 4, order="F"), pts @ M.T, rtol=0, atol=1e-14)


def test_gamma_wake_from_coeffs_fourier(oracle):
    ns, nw, H = 3, 7, 2
    K = 2 * H + 1
    rng = np.random.default_rng(3)
    coeffs = np.asfortranarray(rng.standard_normal((ns, K)))
    gw = np.zeros((ns, nw), order="F")
    tn, om, dt, it = 0.7, 1.3, 0.05, 4
    oracle.gamma_wake_from_coeffs(gw, coeffs, H, tn, om, dt, it)
    for j in range(nw - it, nw):
        tau = tn - (j - (nw - it) + 1) * dt
        basis = [1 / math.sqrt(K)]
        for h in range(1, H + 1):
            basis += [math.sqrt(2 / K) * math.cos(h * om * tau), math.sqrt(2 / K) * math.sin(h * om * tau)]
        assert np.allclose(gw[:, j], coeffs @ np.array(basis), rtol=1e-13)
    assert not np.any(gw[:, :nw - it])


def test_kinematics_velocity(oracle):
    # motion = {u, h_amp, h_phase, a_amp, a_phase, omega, x_pivot}; closed-form velocity of a point
    mot = [1.3, 0.1, 0.2, 0.05, 0.4, 0.9, 0.25]
    for t in (0.0, 0.7, 3.1):
        T, dT = oracle.motion_transform(mot, t)
        a = 0.05 * math.sin(0.9 * t + 0.4)
        ad = 0.05 * 0.9 * math.cos(0.9 * t + 0.4)
        hd = 0.1 * 0.9 * math.cos(0.9 * t + 0.2)
        x = np.array([0.75, 0.4, 0.0, 1.0])
        # x' = (xp + (x-xp) cos a, y, -(x-xp) sin a + h) - u t
        dx = x[0] - 0.25
        vel = np.array([-1.3 - dx * math.sin(a) * ad, 0.0, -dx * math.cos(a) * ad + hd])
        assert np.allclose((dT @ x)[:3], vel, atol=1e-14)
        pos = np.array([0.25 + dx * math.cos(a) - 1.3 * t, 0.4, -dx * math.sin(a) + 0.1 * math.sin(0.9 * t + 0.2)])
        assert np.allclose((T @ x)[:3], pos, atol=1e-14)


def test_mesh_reader_and_quarterchord(oracle, root):
    v = oracle.plot3d_read(os.path.join(root, "mesh", "rectangular_2x2.x"), np.float32)
    assert v.shape == (3, 3, 4)
    assert np.array_equal(v[:, :, 0], np.array([[0, .5, 1]] * 3, dtype=np.float32))
    assert np.array_equal(v[:, 0, 1], np.array([0, 2.5, 5], dtype=np.float32))
    q = oracle.quarterchord(v)
    # rows 0..nc-1 move a quarter panel aft; the TE is extrapolated with the float 4/3, 1/3
    assert np.allclose(q[:, 0, 0], 0.125) and np.allclose(q[:, 1, 0], 0.625)
    te = np.float32(4.0 / 3.0) * np.float32(1.0) - np.float32(1.0 / 3.0) * np.float32(0.625)
    assert np.array_equal(q[:, 2, 0], np.full(3, te, dtype=np.float32))
    with pytest.raises(RuntimeError):
        oracle.plot3d_read(os.path.join(root, "mesh", "elliptic_16x16.x"), np.float64)


# ---------------------------------------------------------------------------------------------
# End to end: the reference's own analytic known-answer tests
# ---------------------------------------------------------------------------------------------

def test_elliptic_cl_kat(oracle, root):
    # tests/vlm_elliptic_coeffs.cpp:23-25,41,75: CL within 3% of 2 pi alpha / (1 + 2/AR)
    a, b = 1.0, 5.0
    s_ref = 0.5 * float(PI_F) * a * b
    ar = (2 * b) ** 2 / s_ref
    v = oracle.read_mesh(os.path.join(root, "mesh", "elliptic_45x45.x"))
    for deg in (0, 1, 2, 3, 4, 5, 10, 15):
        alpha = to_radians(deg)
        _, cl, _, info = oracle.vlm_run([v], alpha)
        assert info == 0
        an = 2 * float(PI_F) * alpha / (1 + 2 / ar)
        err = 0 if an < 1.19209e-07 else abs(cl - an) / an
        assert err < 0.03, (deg, cl, an)


@pytest.mark.parametrize("meshes", [["infinite_rectangular_2x2_part0", "infinite_rectangular_2x2_part1"],
                                    ["infinite_rectangular_2x4"], ["infinite_rectangular_10x5"],
                                    ["infinite_rectangular_20x5"]])
def test_square_cl_kat(oracle, root, meshes):
    # tests/vlm_square_coeffs.cpp:15-17,57: near-2-D rectangles within 1% of 2 pi alpha
    vs = [oracle.read_mesh(os.path.join(root, "mesh", m + ".x")) for m in meshes]
    for deg in (1, 5, 15):
        alpha = to_radians(deg)
        _, cl, _, _ = oracle.vlm_run(vs, alpha)
        an = 2 * float(PI_F) * alpha
        assert abs(cl - an) / an < 0.01


def test_uvlm_constant_pitch_tends_to_steady(oracle):
    # Wagner: after an impulsive start at constant incidence the unsteady CL approaches the
    # steady VLM value (here within 2% after 60 chords of travel)
    v = oracle.rectangular_wing(8, 4, half_span=4.0)
    alpha = 0.05
    mot = [1.0, 0.0, 0.0, alpha, math.pi / 2, 0.0, 0.25]
    cl_t, _, _, n = oracle.uvlm_run([v], mot, t_steps=240)
    # UVLM::run poses the wing once at T(0) (vlm/src/vlm.cpp:370) and every later step applies
    # T(t+dt) to that posed mesh (vlm.cpp:506), so a constant pitch T(0) != I acts twice: the
    # steady reference is the wing rotated twice by alpha about x = 0.25 (nose up)
    T, _ = oracle.motion_transform(mot, 0.0)
    vp = oracle.displace_wing(np.asfortranarray(T), oracle.displace_wing(np.asfortranarray(T), v))
    _, cl_s, _, _ = oracle.vlm_run([vp], 0.0)
    assert abs(cl_t[-1] - cl_s) / abs(cl_s) < 0.02, (cl_t[-5:], cl_s)
    # the single-pitch steady value is half as large: the double pose is really there
    _, cl_1, _, _ = oracle.vlm_run([oracle.displace_wing(np.asfortranarray(T), v)], 0.0)
    assert abs(cl_t[-1] - cl_1) / abs(cl_1) > 0.5


def _theodorsen_C(k):
    from scipy.special import hankel2

    return hankel2(1, k) / (hankel2(1, k) + 1j * hankel2(0, k))


def _theodorsen_cl_cm(dof, k, amp, x_pivot):
    """First-harmonic CL and CM (about the mid-chord, nose up positive) of a flat plate of chord 1 at
    U = 1 in heave h_up = amp sin(w t) or pitch alpha = amp sin(w t) about x_pivot: Theodorsen's
    lift and moment (NACA Report 496), written with h positive down and the elastic axis at
    a = (x_pivot - 1/2)/b; the moment is moved from the axis to the mid-chord.  Complex amplitudes
    X with x(t) = Re(X e^{i w t})."""
    b, U = 0.5, 1.0
    w = k * U / b
    a = (x_pivot - 0.5) / b
    H = 1j * amp if dof == "heave" else 0.0  # h_down = -amp sin(w t)
    A = -1j * amp if dof == "pitch" else 0.0  # alpha = amp sin(w t)
    hd, hdd, ad, add = 1j * w * H, -w * w * H, 1j * w * A, -w * w * A
    circ = 2 * math.pi * U * b * _theodorsen_C(k) * (hd + U * A + b * (0.5 - a) * ad)
    lift = math.pi * b * b * (hdd + U * ad - b * a * add) + circ
    m_axis = (math.pi * b * b * (b * a * hdd - U * b * (0.5 - a) * ad - b * b * (0.125 + a * a) * add)
              + b * (a + 0.5) * circ)
    cl = lift / (0.5 * U * U * 2 * b)
    cm_axis = m_axis / (0.5 * U * U * (2 * b) ** 2)
    return cl, cm_axis + cl * (0.5 - x_pivot)


def _uvlm_first_harmonic(orc, dof, k, nc, amp, x_pivot=0.25, cycles=4):
    """UVLM::run on a near-2-D rectangular wing (3 spanwise panels over a half span of 1000 chords,
    nc chordwise, dt = 1/nc by the last-panel-chord rule); least-squares first harmonic of CL(t)
    and CM_y(t) over the last cycle."""
    v = orc.rectangular_wing(3, nc, chord=1.0, half_span=1000.0)
    w = 2.0 * k
    mot = [1.0, amp, 0.0, 0.0, 0.0, w, x_pivot] if dof == "heave" else [1.0, 0.0, 0.0, amp, 0.0, w, x_pivot]
    dt = 1.0 / nc
    period = int(round(2 * math.pi / w / dt))
    cl, cm, _, _ = orc.uvlm_run([v], mot, t_steps=cycles * period + 2)
    t = np.arange(len(cl))[-period:] * dt
    basis = np.stack([np.ones_like(t), np.cos(w * t), np.sin(w * t)], 1)
    out = []
    for x in (cl, cm):
        c = np.linalg.lstsq(basis, x[-period:], rcond=None)[0]
        out.append(c[1] - 1j * c[2])
    return out


@pytest.mark.parametrize("k", [0.25, 0.5])
@pytest.mark.parametrize("dof", ["heave", "pitch"])
def test_uvlm_first_harmonic_vs_theodorsen(oracle, dof, k):
    """Pins UVLM::run's unsteady bookkeeping (vlm/src/vlm.cpp:442-466: the shed trailing-edge column,
    dGamma/dt = (Gamma - Gamma_prev)/dt in the unsteady force, CM about T(t)(0.5, 0, 0)) against
    Theodorsen's closed form, the reference's own validation path (tests/uvlm_theodorsen.cpp:26-40):
    the first harmonic of CL(t) and CM(t) (amplitude and phase together, as one complex number)
    within 10 % at 8 chordwise panels, and the error at least 40 % smaller at 16 (the explicit wake
    is first order in dt, so the restatement converges to the closed form)."""
    amp = 0.01 if dof == "heave" else math.radians(1.0)
    cl_t, cm_t = _theodorsen_cl_cm(dof, k, amp, 0.25)
    errs = []
    for nc in (8, 16):
        cl, cm = _uvlm_first_harmonic(oracle, dof, k, nc, amp)
        e_cl, e_cm = abs(cl - cl_t) / abs(cl_t), abs(cm - cm_t) / abs(cm_t)
        errs.append((e_cl, e_cm))
        assert e_cl <= 0.10 and e_cm <= 0.10, (nc, cl, cl_t, cm, cm_t)
    assert errs[1][0] <= 0.6 * errs[0][0] and errs[1][1] <= 0.6 * errs[0][1], errs


def test_coeff_cm_multibody_split_and_hand_sum(oracle):
    """coeff_cm_multibody (vlm/src/vlm_backend.cpp:102-139) weights each surface's CM by area x MAC
    and divides by (total area x total MAC).  Two spanwise halves of one rectangular wing have equal
    areas and MACs, so the split CM is exactly half the single-surface CM; and the single-surface CM_y
    equals the hand sum -sum (x_mid - x_ref) F_z / (q S mac) over the leading ring segments' mid-points."""
    rng = np.random.default_rng(5)
    ns, nc = 6, 4
    v = oracle.rectangular_wing(ns, nc, chord=1.0, half_span=3.0)
    halves = [np.asfortranarray(v[: ns // 2 + 1]), np.asfortranarray(v[ns // 2:])]
    F = np.asfortranarray(rng.standard_normal((ns, nc, 3)))
    Fh = [np.asfortranarray(F[: ns // 2]), np.asfortranarray(F[ns // 2:])]
    ref = np.array([0.3, 0.0, 0.05])
    fs = np.array([1.0, 0.0, 0.0])
    one = oracle.Surface(v)
    parts = [oracle.Surface(h) for h in halves]
    oracle.mesh_metrics(0.0, [one])
    oracle.mesh_metrics(0.0, parts)
    cm1 = oracle.coeff_cm_multibody([F], [one], ref, fs, 1.0)
    cm2 = oracle.coeff_cm_multibody(Fh, parts, ref, fs, 1.0)
    assert np.allclose(cm2, 0.5 * cm1, rtol=1e-13, atol=1e-15)
    # hand sum for CM_y of the single surface: r = (v(i,j) + v(i+1,j))/2 - ref, CM = r x F / (q S mac)
    area = float(np.sum(one.areas))
    mac = oracle.mesh_mac(v, one.areas)
    r = 0.5 * (v[:-1, :-1, :3] + v[1:, :-1, :3]) - ref
    cmy = np.sum(r[..., 2] * F[..., 0] - r[..., 0] * F[..., 2]) / (0.5 * area * mac)
    assert cm1[1] == pytest.approx(cmy, rel=1e-12)
    assert mac == pytest.approx(1.0, rel=1e-6)  # rectangle: MAC = chord


# ---- LSQ (gels) and Anderson acceleration ---------------------------------------------------


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("m,n,nrhs", [(50, 6, 2), (7, 7, 1), (300, 10, 1), (33, 1, 3)])
def test_oracle_gels_matches_lapack_layout(oracle, dt, m, n, nrhs):
    """A <- geqrf factors and B <- Q^T B with X on top, as LAPACK ?gels leaves them (scipy's
    LAPACK binding is the independent implementation); X also equals numpy's lstsq."""
    import scipy.linalg.lapack as la

    rng = np.random.default_rng(m * 100 + n)
    A = rng.standard_normal((m, n)).astype(dt)
    B = rng.standard_normal((m, nrhs)).astype(dt)
    a, b, info = oracle.gels(A, B)
    ref = (la.dgels if dt == np.float64 else la.sgels)(np.asfortranarray(A), np.asfortranarray(B))
    assert info == 0 and ref[2] == 0
    tol = 1e-12 if dt == np.float64 else 2e-4
    scale = max(1.0, float(np.max(np.abs(ref[1]))))
    assert np.max(np.abs(a - ref[0])) <= tol * max(1.0, float(np.max(np.abs(ref[0]))))
    assert np.max(np.abs(b - ref[1])) <= tol * scale
    x = np.linalg.lstsq(A.astype(np.float64), B.astype(np.float64), rcond=None)[0]
    assert np.max(np.abs(b[:n] - x)) <= tol * 10 * max(1.0, float(np.max(np.abs(x))))


def test_oracle_gels_rank_deficient_and_zero(oracle):
    A = np.zeros((8, 3))
    A[:, 0] = np.arange(8.0)
    A[:, 2] = 1.0  # column 1 zero -> R(2,2) == 0 exactly after the first reflector
    _, _, info = oracle.gels(A, np.ones((8, 1)))
    assert info == 2
    _, b, info = oracle.gels(np.zeros((5, 2)), np.ones((5, 1)))
    assert info == 0 and np.all(b == 0)


def test_oracle_anderson_linear_fixed_point(oracle):
    """x = M x + c with spectral radius < 1: Anderson reaches (I - M)^-1 c, and far faster than
    the plain fixed-point iteration (which needs > 100 steps here)."""
    n = 64
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    M = 0.9 * np.sin(1.0 + i + 3.0 * j) / n
    c = np.cos(np.arange(n, dtype=np.float64))
    x, k = oracle.anderson_acceleration(np.zeros(n), lambda v: M @ v + c, 100, 1e-12, 3)
    exact = np.linalg.solve(np.eye(n) - M, c)
    assert k < 40
    assert np.max(np.abs(x - exact)) < 1e-10
    # with m = 1 history the method still converges (Anderson(1) = secant mixing)
    x1, k1 = oracle.anderson_acceleration(np.zeros(n), lambda v: M @ v + c, 100, 1e-12, 1)
    assert k1 < 100 and np.max(np.abs(x1 - exact)) < 1e-10
