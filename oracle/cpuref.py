"""ctypes front end of the CPU restatement of the reference backend (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may import
this module: it is the checker, never the thing measured or shipped.

Arrays follow the reference's column-major TensorView layout (vlm/include/vlm_memory.hpp:306-319):
numpy arrays in Fortran order with shapes ``verts (ns+1, nc+1, 4)``, ``colloc/normals (ns, nc, 3)``,
``areas (ns, nc)``, ``verts_wake (ns+1, nwv, 4)``, ``gamma_wake (ns, nwv-1)``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libvlm_oracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile the restatement with its Makefile (gcc, OpenMP, -ffp-contract=off)."""
    if force and os.path.exists(_LIB_PATH):
        os.remove(_LIB_PATH)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run oracle/Makefile)")
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32


def _declare(L):
    for s, R in (("d", ctypes.c_double), ("f", ctypes.c_float)):
        getattr(L, f"orc_coeff_cl_{s}").restype = R
        getattr(L, f"orc_sum_{s}").restype = R
        getattr(L, f"orc_mesh_mac_{s}").restype = R
        getattr(L, f"orc_coeff_cl_multibody_{s}").restype = R
        getattr(L, f"orc_getrf_{s}").restype = ctypes.c_int
        getattr(L, f"orc_gels_{s}").restype = ctypes.c_int
        getattr(L, f"orc_vlm_run_{s}").restype = ctypes.c_int
        getattr(L, f"orc_uvlm_run_{s}").restype = ctypes.c_int


def _sfx(dtype) -> str:
    return "d" if np.dtype(dtype) == np.float64 else "f"


def _real(dtype):
    return ctypes.c_double if np.dtype(dtype) == np.float64 else ctypes.c_float


def _ptr(a: np.ndarray) -> int:
    assert a.flags["F_CONTIGUOUS"] or a.ndim <= 1, "arrays must be column-major"
    return a.ctypes.data


def _ptrs(arrs) -> ctypes.Array:
    return (ctypes.c_void_p * len(arrs))(*[_ptr(a) for a in arrs])


def _i64s(vals) -> ctypes.Array:
    return (ctypes.c_int64 * len(vals))(*vals)


# ---------------------------------------------------------------------------------------------
# Mesh input: plot3d reader and quarter-chord shift (vlm/src/vlm_mesh.cpp:15-124)
# ---------------------------------------------------------------------------------------------

def plot3d_read(path: str, dtype=np.float64) -> np.ndarray:
    """Single-block 2-D plot3d -> verts (ns+1, nc+1, 4), w = 1 (vlm_mesh.cpp:50-92, vlm.cpp:41-43).

    The file stores ni (chordwise points) x nj (spanwise points); x, y, z planes, i fastest.
    Values are parsed directly in ``dtype`` as the reference's ``f >> x`` does for T.
    """
    with open(path) as f:
        tok = f.read().split()
    blocks = int(tok[0])
    if blocks != 1:
        raise RuntimeError("Only single block plot3d mesh is supported")
    ni, nj, nk = int(tok[1]), int(tok[2]), int(tok[3])
    if nk != 1:
        raise RuntimeError("Only 2D plot3d mesh is supported")
    n = ni * nj
    vals = np.array([dtype(t) if np.dtype(dtype) == np.float32 else float(t) for t in tok[4:4 + 3 * n]], dtype=dtype)
    verts = np.zeros((nj, ni, 4), dtype=dtype, order="F")
    for k in range(3):
        plane = vals[k * n:(k + 1) * n].reshape(nj, ni)  # j major, i fastest
        verts[:, :, k] = plane
    verts[:, :, 3] = 1
    first_y = verts[0, 0, 1]
    if np.any(verts[0, 1:, 1] != first_y):
        raise RuntimeError("Mesh vertices should be ordered in chordwise direction")
    return verts


def quarterchord(verts: np.ndarray) -> np.ndarray:
    """mesh_quarterchord through the C restatement (float constants 0.75f/0.25f, 4.f/3.f, 1.f/3.f)."""
    v = np.asfortranarray(verts.copy())
    ns, nc = v.shape[0] - 1, v.shape[1] - 1
    getattr(lib(), f"orc_quarterchord_{_sfx(v.dtype)}")(_I64(ns), _I64(nc), _P(_ptr(v)))
    return v


def read_mesh(path: str, dtype=np.float64, qc: bool = True) -> np.ndarray:
    """Mesh as the reference drivers see it: parsed and quarter-chord shifted in float32 (the
    Simulation buffers are f32, vlm/include/vlm.hpp:44-49), then widened to ``dtype``."""
    v = plot3d_read(path, np.float32)
    v = quarterchord(v) if qc else v
    return np.asfortranarray(v.astype(dtype))


# ---------------------------------------------------------------------------------------------
# Backend method restatements
# ---------------------------------------------------------------------------------------------

@dataclass
class Surface:
    """Per-surface buffers of one lifting surface (the MultiTensorView entries of one index)."""

    verts: np.ndarray  # (ns+1, nc+1, 4)
    verts_wake: np.ndarray | None = None  # (ns+1, nwv, 4)
    gamma_wake: np.ndarray | None = None  # (ns, nwv-1)
    colloc: np.ndarray = field(default=None)
    normals: np.ndarray = field(default=None)
    areas: np.ndarray = field(default=None)
    lifting: bool = True

    @property
    def ns(self):
        return self.verts.shape[0] - 1

    @property
    def nc(self):
        return self.verts.shape[1] - 1

    @property
    def nwv(self):
        return 0 if self.verts_wake is None else self.verts_wake.shape[1]


def biosavart(c, v1, v2, sigma=0.0, dtype=np.float64):
    c, v1, v2 = (np.ascontiguousarray(x, dtype=dtype) for x in (c, v1, v2))
    out = np.zeros(3, dtype=dtype)
    R = _real(dtype)
    getattr(lib(), f"orc_biosavart_{_sfx(dtype)}")(_P(_ptr(c)), _P(_ptr(v1)), _P(_ptr(v2)), R(sigma), _P(_ptr(out)))
    return out


def mesh_metrics(alpha: float, surfs: list[Surface]):
    dt = surfs[0].verts.dtype
    for s in surfs:
        s.colloc = np.zeros((s.ns, s.nc, 3), dtype=dt, order="F")
        s.normals = np.zeros((s.ns, s.nc, 3), dtype=dt, order="F")
        s.areas = np.zeros((s.ns, s.nc), dtype=dt, order="F")
    getattr(lib(), f"orc_mesh_metrics_{_sfx(dt)}")(
        _real(dt)(alpha), _I32(len(surfs)), _i64s([s.ns for s in surfs]), _i64s([s.nc for s in surfs]),
        _ptrs([s.verts for s in surfs]), _ptrs([s.colloc for s in surfs]), _ptrs([s.normals for s in surfs]),
        _ptrs([s.areas for s in surfs]))


def lhs_assemble(surfs: list[Surface], iteration: int, lhs: np.ndarray | None = None, sigma=0.0):
    dt = surfs[0].verts.dtype
    N = sum(s.ns * s.nc for s in surfs)
    if lhs is None:
        lhs = np.zeros((N, N), dtype=dt, order="F")
    wakes = [s.verts_wake if s.verts_wake is not None else np.zeros((s.ns + 1, 2, 4), dtype=dt, order="F") for s in surfs]
    getattr(lib(), f"orc_lhs_assemble_{_sfx(dt)}")(
        _P(_ptr(lhs)), _I64(lhs.shape[0]), _I32(len(surfs)), _i64s([s.ns for s in surfs]),
        _i64s([s.nc for s in surfs]), _i64s([w.shape[1] for w in wakes]), _ptrs([s.colloc for s in surfs]),
        _ptrs([s.normals for s in surfs]), _ptrs([s.verts for s in surfs]), _ptrs(wakes), _I32(iteration),
        _real(dt)(sigma))
    return lhs


def rhs_assemble_velocities(rhs, surfs: list[Surface], velocities):
    dt = rhs.dtype
    getattr(lib(), f"orc_rhs_assemble_velocities_{_sfx(dt)}")(
        _P(_ptr(rhs)), _I32(len(surfs)), _i64s([s.ns for s in surfs]), _i64s([s.nc for s in surfs]),
        _ptrs([s.normals for s in surfs]), _ptrs(velocities))
    return rhs


def rhs_assemble_wake_influence(rhs, surfs: list[Surface], iteration: int, sigma=0.0, far_wake=None):
    """far_wake: far-wake aggregation levels as in uvlm_run (None: the reference's exact wake)."""
    dt = rhs.dtype
    args = (_P(_ptr(rhs)), _I32(len(surfs)), _i64s([s.ns for s in surfs]), _i64s([s.nc for s in surfs]),
            _i64s([s.nwv for s in surfs]), _ptrs([s.gamma_wake for s in surfs]), _ptrs([s.colloc for s in surfs]),
            _ptrs([s.normals for s in surfs]), _ptrs([s.verts_wake for s in surfs]),
            (ctypes.c_int32 * len(surfs))(*[int(s.lifting) for s in surfs]), _I32(iteration), _real(dt)(sigma))
    if far_wake is None:
        getattr(lib(), f"orc_rhs_assemble_wake_influence_{_sfx(dt)}")(*args)
    else:
        far = far_levels_array(far_wake)
        getattr(lib(), f"orc_rhs_assemble_wake_influence_far_{_sfx(dt)}")(*args, _P(_ptr(far)))
    return rhs


def forces_steady(verts, gamma_delta, velocities):
    ns, nc = gamma_delta.shape
    f = np.zeros((ns, nc, 3), dtype=verts.dtype, order="F")
    getattr(lib(), f"orc_forces_steady_{_sfx(verts.dtype)}")(
        _I64(ns), _I64(nc), _P(_ptr(verts)), _P(_ptr(gamma_delta)), _P(_ptr(velocities)), _P(_ptr(f)))
    return f


def forces_unsteady(verts, gamma_delta, dgamma_dt, velocities, areas, normals):
    ns, nc = gamma_delta.shape
    f = np.zeros((ns, nc, 3), dtype=verts.dtype, order="F")
    getattr(lib(), f"orc_forces_unsteady_{_sfx(verts.dtype)}")(
        _I64(ns), _I64(nc), _P(_ptr(verts)), _P(_ptr(gamma_delta)), _P(_ptr(dgamma_dt)), _P(_ptr(velocities)),
        _P(_ptr(areas)), _P(_ptr(normals)), _P(_ptr(f)))
    return f


def coeff_cl(forces, lift_axis, freestream, rho, area):
    dt = forces.dtype
    ns, nc = forces.shape[:2]
    la = np.ascontiguousarray(lift_axis, dtype=dt)
    fs = np.ascontiguousarray(freestream, dtype=dt)
    R = _real(dt)
    return getattr(lib(), f"orc_coeff_cl_{_sfx(dt)}")(_I64(ns), _I64(nc), _P(_ptr(forces)), _P(_ptr(la)),
                                                      _P(_ptr(fs)), R(rho), R(area))


def coeff_cm(forces, verts, ref_pt, freestream, rho, area, mac):
    dt = forces.dtype
    ns, nc = forces.shape[:2]
    rp = np.ascontiguousarray(ref_pt, dtype=dt)
    fs = np.ascontiguousarray(freestream, dtype=dt)
    out = np.zeros(3, dtype=dt)
    R = _real(dt)
    getattr(lib(), f"orc_coeff_cm_{_sfx(dt)}")(_I64(ns), _I64(nc), _P(_ptr(forces)), _P(_ptr(verts)), _P(_ptr(rp)),
                                               _P(_ptr(fs)), R(rho), R(area), R(mac), _P(_ptr(out)))
    return out


def coeff_cm_multibody(forces_list, surfs: list[Surface], ref_pt, freestream, rho):
    """coeff_cm_multibody (vlm/src/vlm_backend.cpp:102-139) over surfaces whose metrics are set."""
    dt = forces_list[0].dtype
    rp = np.ascontiguousarray(ref_pt, dtype=dt)
    fs = np.ascontiguousarray(freestream, dtype=dt)
    out = np.zeros(3, dtype=dt)
    forces_list = [np.asfortranarray(f) for f in forces_list]
    getattr(lib(), f"orc_coeff_cm_multibody_{_sfx(dt)}")(
        _I32(len(surfs)), _i64s([s.ns for s in surfs]), _i64s([s.nc for s in surfs]), _ptrs(forces_list),
        _ptrs([s.verts for s in surfs]), _ptrs([s.areas for s in surfs]), _P(_ptr(rp)), _P(_ptr(fs)), _real(dt)(rho),
        _P(_ptr(out)))
    return out


def sum_(x):
    x = np.asfortranarray(x)
    return getattr(lib(), f"orc_sum_{_sfx(x.dtype)}")(_I64(x.size), _P(_ptr(x)))


def mesh_mac(verts, areas):
    ns, nc = areas.shape
    return getattr(lib(), f"orc_mesh_mac_{_sfx(verts.dtype)}")(_I64(ns), _I64(nc), _P(_ptr(verts)), _P(_ptr(areas)))


def wake_shed(verts, verts_wake, iteration):
    ns, nc = verts.shape[0] - 1, verts.shape[1] - 1
    getattr(lib(), f"orc_wake_shed_{_sfx(verts.dtype)}")(_I64(ns), _I64(nc), _I64(verts_wake.shape[1]),
                                                         _P(_ptr(verts)), _P(_ptr(verts_wake)), _I32(iteration))


def displace_wing(transform, verts_init):
    out = np.zeros_like(verts_init, order="F")
    nv = verts_init.shape[0] * verts_init.shape[1]
    T = np.asfortranarray(transform, dtype=verts_init.dtype)
    getattr(lib(), f"orc_displace_wing_{_sfx(verts_init.dtype)}")(_I64(nv), _P(_ptr(T)), _P(_ptr(verts_init)),
                                                                   _P(_ptr(out)))
    return out


def gamma_wake_from_coeffs(gamma_wake, gamma_coeffs, harmonics, tn, omega, dt_, iteration):
    dt = gamma_wake.dtype
    R = _real(dt)
    getattr(lib(), f"orc_gamma_wake_from_coeffs_{_sfx(dt)}")(
        _I64(gamma_wake.shape[0]), _I64(gamma_wake.shape[1]), _P(_ptr(gamma_wake)), _P(_ptr(gamma_coeffs)),
        _I32(harmonics), R(tn), R(omega), R(dt_), _I64(iteration))
    return gamma_wake


def getrf(a):
    a = np.asfortranarray(a.copy())
    n = a.shape[0]
    ipiv = np.zeros(n, dtype=np.int32)
    info = getattr(lib(), f"orc_getrf_{_sfx(a.dtype)}")(_I64(n), _P(_ptr(a)), _I64(n), _P(_ptr(ipiv)))
    return a, ipiv, info


def getrs(lu, ipiv, b):
    b = np.asfortranarray(b.copy())
    b2 = b.reshape(b.shape[0], -1, order="F")
    getattr(lib(), f"orc_getrs_{_sfx(lu.dtype)}")(_I64(lu.shape[0]), _I64(b2.shape[1]), _P(_ptr(lu)),
                                                  _I64(lu.shape[0]), _P(_ptr(ipiv)), _P(_ptr(b2)),
                                                  _I64(b2.shape[0]))
    return b2.reshape(b.shape, order="F")


def gels(a, b):
    """LSQ::solve restatement (LAPACKE_?gels 'N'); returns (A overwritten by its QR factors,
    B overwritten by Q^T B with X in rows 0..N-1, info)."""
    a = np.asfortranarray(a.copy())
    b = np.asfortranarray(b.copy())
    b2 = b.reshape(b.shape[0], -1, order="F")
    info = getattr(lib(), f"orc_gels_{_sfx(a.dtype)}")(_I64(a.shape[0]), _I64(a.shape[1]), _I64(b2.shape[1]),
                                                       _P(_ptr(a)), _I64(a.shape[0]), _P(_ptr(b2)),
                                                       _I64(b2.shape[0]))
    return a, b2.reshape(b.shape, order="F"), info


def anderson_acceleration(x0, f, max_iter=100, tol_res=1e-6, m=3):
    """anderson_acceleration restatement (vlm/src/vlm_solvers.cpp:6-115): type-II Anderson mixing
    with a circular history of m columns, least squares by gels.  f(x) returns the fixed-point map's
    value.  Returns (x, iterations k)."""
    dt = x0.dtype
    n = x0.shape[0]
    X = np.zeros((n, m), dtype=dt, order="F")
    G = np.zeros((n, m), dtype=dt, order="F")
    x_curr = x0.copy()
    x_new = f(x_curr)
    g_curr = x_new - x_curr
    X[:, 0] = g_curr
    x_curr = x_new.copy()
    x_new = f(x_curr)
    g_new = x_new - x_curr
    G[:, 0] = g_new - g_curr
    g_curr = g_new.copy()
    k = 1
    while k < max_iter and np.sqrt(np.dot(g_curr, g_curr)) > tol_res:
        mk = min(m, k)
        _, sol, info = gels(G[:, :mk], g_curr.copy())
        gam = sol[:mk]
        x_new = x_curr + g_curr - X[:, :mk] @ gam - G[:, :mk] @ gam
        X[:, k % m] = x_new - x_curr
        x_curr = x_new.copy()
        x_new = f(x_curr)
        g_new = x_new - x_curr
        G[:, k % m] = g_new - g_curr
        g_curr = g_new.copy()
        k += 1
    return x_curr, k


def vlm_run(verts_list, alpha, beta=0.0, u_inf=1.0, rho=1.0, cm_ref=(0.25, 0.0, 0.0)):
    """VLM::run restatement; returns (gamma[N], CL, CM[3], info)."""
    dt = verts_list[0].dtype
    N = sum((v.shape[0] - 1) * (v.shape[1] - 1) for v in verts_list)
    gamma = np.zeros(N, dtype=dt)
    coeffs = np.zeros(4, dtype=dt)
    ref = np.ascontiguousarray(cm_ref, dtype=dt)
    R = _real(dt)
    info = getattr(lib(), f"orc_vlm_run_{_sfx(dt)}")(
        _I32(len(verts_list)), _i64s([v.shape[0] - 1 for v in verts_list]), _i64s([v.shape[1] - 1 for v in verts_list]),
        _ptrs(verts_list), R(alpha), R(beta), R(u_inf), R(rho), _P(_ptr(ref)), _P(_ptr(gamma)), _P(_ptr(coeffs)))
    return gamma, float(coeffs[0]), coeffs[1:].copy(), info


def motion_transform(motion, t):
    m = np.ascontiguousarray(motion, dtype=np.float64)
    T = np.zeros(16)
    dT = np.zeros(16)
    lib().orc_motion_transform_d(_P(_ptr(m)), ctypes.c_double(t), _P(_ptr(T)), _P(_ptr(dT)))
    return T.reshape(4, 4, order="F"), dT.reshape(4, 4, order="F")


def far_levels_array(far_wake):
    """[(A, G, S), ...] (or one (A, G, S)) -> int32 {L, A_1, G_1, S_1, ...}, checked: A increasing,
    G_l >= 1 a multiple of G_{l-1}, S_l >= 1, at most 16 levels."""
    lv = [tuple(int(x) for x in far_wake)] if np.ndim(far_wake) == 1 else [tuple(int(x) for x in t) for t in far_wake]
    assert 1 <= len(lv) <= 16
    for k, (A, G, S) in enumerate(lv):
        assert A >= 0 and G >= 1 and S >= 1
        if k:
            assert A > lv[k - 1][0] and G % lv[k - 1][1] == 0
    return np.ascontiguousarray([len(lv)] + [x for t in lv for x in t], dtype=np.int32)


def uvlm_run(verts_list, motion, t_steps=0, t_final=0.0, far_wake=None):
    """UVLM::run restatement for the prescribed motion; returns (cl[t], cm_y[t], gamma_last, t_steps).

    far_wake = [(A_1, G_1, S_1), ...] evaluates the wake with far-wake aggregation: levels of age
    threshold A_l, row group G_l (each a multiple of the one before) and span group S_l; ring rows older
    than A_l steps are merged into G_l x S_l rings of mean circulation (one tuple: one level).  None is
    the exact wake of the reference."""
    dt = verts_list[0].dtype
    N = sum((v.shape[0] - 1) * (v.shape[1] - 1) for v in verts_list)
    cap = t_steps if t_steps > 0 else 1_000_000
    cl = np.zeros(cap, dtype=dt)
    cm = np.zeros(cap, dtype=dt)
    gamma = np.zeros(N, dtype=dt)
    mot = np.ascontiguousarray(motion, dtype=dt)
    ts = ctypes.c_int64(0)
    R = _real(dt)
    args = (_I32(len(verts_list)), _i64s([v.shape[0] - 1 for v in verts_list]),
            _i64s([v.shape[1] - 1 for v in verts_list]), _ptrs(verts_list), _P(_ptr(mot)), _I64(t_steps), R(t_final),
            _P(_ptr(cl)), _P(_ptr(cm)), _P(_ptr(gamma)), ctypes.byref(ts))
    if far_wake is None:
        info = getattr(lib(), f"orc_uvlm_run_{_sfx(dt)}")(*args)
    else:
        far = far_levels_array(far_wake)
        info = getattr(lib(), f"orc_uvlm_run_far_{_sfx(dt)}")(*args, _P(_ptr(far)))
    if info != 0:
        raise RuntimeError(f"singular LHS (getrf info {info})")
    n = ts.value - 1
    return cl[:n].copy(), cm[:n].copy(), gamma, ts.value


# ---------------------------------------------------------------------------------------------
# Synthetic meshes of the unsteady/served workloads (flat wings, z = 0)
# ---------------------------------------------------------------------------------------------

def rectangular_wing(ns: int, nc: int, chord=1.0, half_span=2.0, sweep_deg=0.0, dtype=np.float64, qc=True):
    """Flat rectangular (optionally LE-swept) half wing, uniform panels, y in [0, half_span].

    Built as the plot3d reader would store it (float32), quarter-chord shifted in float32 like every
    mesh file, then widened to ``dtype``.
    """
    v = np.zeros((ns + 1, nc + 1, 4), dtype=np.float64, order="F")
    ys = np.linspace(0.0, half_span, ns + 1)
    xs = np.linspace(0.0, chord, nc + 1)
    tan_s = np.tan(np.radians(sweep_deg))
    for i in range(ns + 1):
        v[i, :, 0] = xs + ys[i] * tan_s
        v[i, :, 1] = ys[i]
    v[:, :, 3] = 1
    v = np.asfortranarray(v.astype(np.float32))
    v = quarterchord(v) if qc else v
    return np.asfortranarray(v.astype(dtype))
