"""ctypes binding of the C ABI in include/tinyvlm_hip.h (libtinyvlm_hip.so).

This is the Python entry to the HIP backend used by the tests, the smoke check and the
benchmark.  Device memory and streams come from PyTorch; every computation happens in the HIP
library.  If the library is missing the import of any compute function raises: there is no
CPU fallback.

Tensors follow the reference TensorView layout (column-major, vlm/include/vlm_memory.hpp:306-319):
a reference shape ``(ns, nc, 3)`` is a torch tensor with strides ``(1, ns, ns*nc)``; use
``colmajor()`` / ``to_device()`` to create them.
"""
from __future__ import annotations

import ctypes
import weakref
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtinyvlm_hip.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "tinyvlm_hip.h")

MAX_SURFACES = 16


class View(ctypes.Structure):
    """struct tvlm_view: device pointer, shape[3], stride[3] (elements)."""

    _fields_ = [("ptr", ctypes.c_void_p), ("shape", ctypes.c_int64 * 3), ("stride", ctypes.c_int64 * 3)]


class TvlmError(RuntimeError):
    pass


_lib = None


def header_symbols() -> list[str]:
    """Every entry point declared in include/tinyvlm_hip.h (macro-expanded for both precisions)."""
    text = open(HEADER).read()
    names = set(re.findall(r"^\s*(?:int|const char\*)\s+(tvlm_\w+)\s*\(", text, re.M))
    for base in re.findall(r"int\s+(tvlm_\w+)_##S\s*\(", text):
        names.add(base + "_f64")
        names.add(base + "_f32")
    return sorted(names)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TvlmError(f"HIP backend library not built: {LIB_PATH} (run __graft_entry__.build())")
        # PyTorch ships its own libamdhip64.so.7 / libhsa-runtime64.so.1 (same sonames as /opt/rocm).
        # Load it first so that this library binds to the runtime PyTorch uses: one HIP runtime
        # per process, and device pointers from torch tensors are valid in our kernels.
        import torch  # noqa: F401

        _lib = ctypes.CDLL(LIB_PATH)
        _lib.tvlm_last_error.restype = ctypes.c_char_p
        _lib.tvlm_backend_name.restype = ctypes.c_char_p
        for name in header_symbols():
            fn = getattr(_lib, name)
            if name not in ("tvlm_last_error", "tvlm_backend_name"):
                fn.restype = ctypes.c_int
    return _lib


def check(status: int, what: str = ""):
    if status != 0:
        msg = lib().tvlm_last_error().decode(errors="replace")
        raise TvlmError(f"{what}: status {status}: {msg}")


# ---------------------------------------------------------------------------------------------
# torch <-> view helpers
# ---------------------------------------------------------------------------------------------

def _torch():
    import torch

    return torch


def colmajor(shape, dtype, device="cuda", zero=True):
    """Column-major tensor of the given reference shape (stride[0] == 1)."""
    torch = _torch()
    rev = tuple(reversed(shape))
    t = torch.zeros(rev, dtype=dtype, device=device) if zero else torch.empty(rev, dtype=dtype, device=device)
    return t.permute(*reversed(range(len(shape))))


def to_device(a: np.ndarray, dtype=None, device="cuda"):
    """numpy array (any order) -> column-major torch device tensor with the same shape."""
    torch = _torch()
    a = np.asarray(a)
    if dtype is None:
        dtype = torch.float64 if a.dtype == np.float64 else torch.float32 if a.dtype == np.float32 else None
    t = colmajor(a.shape, dtype, device, zero=False)
    host = torch.from_numpy(np.ascontiguousarray(a.transpose(*reversed(range(a.ndim))))).to(dtype)
    t.permute(*reversed(range(a.ndim))).copy_(host)
    return t


def to_host(t) -> np.ndarray:
    """Device tensor -> numpy array in Fortran order with the same shape."""
    a = t.detach().cpu().numpy()
    return np.asfortranarray(a)


# View of a tensor, kept per tensor object (weakly) while its address, shape and strides stay the
# same: the host-driven loops pass the same tensors at every launch, and building the ctypes
# structure costs several microseconds.  Callers only read a View (by reference during a call, or
# copied into an array or a struct), never change it.
# Keyed by id(t) with a weak reference to t (tensors compare elementwise, so they cannot be
# dictionary keys themselves); the reference's callback drops the entry when t is freed.
_VIEW_CACHE: dict = {}


def view(t) -> View:
    key = (t.data_ptr(), t.shape, t.stride())
    tid = id(t)
    hit = _VIEW_CACHE.get(tid)
    if hit is not None and hit[0]() is t and hit[1] == key:
        return hit[2]
    v = _make_view(t)
    try:
        ref = weakref.ref(t, lambda _r, k=tid: _VIEW_CACHE.pop(k, None))
    except TypeError:  # not weakly referenceable: no cache
        return v
    _VIEW_CACHE[tid] = (ref, key, v)
    return v


def _make_view(t) -> View:
    nd = t.dim()
    s, st = t.shape, t.stride()
    # unused trailing dimensions: shape 1, stride = extent product (harmless for shape-1 dims)
    if nd == 3:
        return View(t.data_ptr(), (s[0], s[1], s[2]), (st[0], st[1], st[2]))
    if nd == 2:
        return View(t.data_ptr(), (s[0], s[1], 1), (st[0], st[1], st[1] * s[1]))
    if nd == 1:
        e = st[0] * s[0]
        return View(t.data_ptr(), (s[0], 1, 1), (st[0], e, e))
    v = View()
    v.ptr = t.data_ptr()
    shp = list(t.shape) + [1] * (3 - t.dim())
    st = list(t.stride()) + [0] * (3 - t.dim())
    if t.dim() >= 1:
        for d in range(t.dim(), 3):
            st[d] = st[d - 1] * shp[d - 1] if d > 0 else 1
    for k in range(3):
        v.shape[k] = int(shp[k])
        v.stride[k] = int(st[k])
    return v


def views(ts) -> ctypes.Array:
    arr = (View * len(ts))()
    for k, t in enumerate(ts):
        arr[k] = view(t)
    return arr


_RAW_STREAM = None  # (getCurrentRawStream, getDevice) of torch, or False where torch lacks them


def current_stream_handle() -> int:
    """hipStream_t of torch's current stream on the current device, as torch.cuda.current_stream()
    .cuda_stream gives it, through torch's raw-stream lookup where it exists (the public call costs
    ~10 us of Python per launch of a host-driven loop)."""
    global _RAW_STREAM
    if _RAW_STREAM is None:
        C = _torch()._C
        ok = hasattr(C, "_cuda_getCurrentRawStream") and hasattr(C, "_cuda_getDevice")
        _RAW_STREAM = (C._cuda_getCurrentRawStream, C._cuda_getDevice) if ok else False
    if _RAW_STREAM:
        raw, dev = _RAW_STREAM
        return int(raw(dev()))
    return _torch().cuda.current_stream().cuda_stream


def stream_ptr(stream=None):
    """hipStream_t of a torch stream, of a raw handle (int, e.g. from tvlm_stream_create) or of
    torch's current stream."""
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    if stream is None:
        return ctypes.c_void_p(current_stream_handle())
    return ctypes.c_void_p(stream.cuda_stream)


def _real(t):
    torch = _torch()
    return ctypes.c_double if t.dtype == torch.float64 else ctypes.c_float


def _sfx(t):
    torch = _torch()
    if t.dtype == torch.float64:
        return "f64"
    if t.dtype == torch.float32:
        return "f32"
    raise TvlmError(f"unsupported dtype {t.dtype}")


_FNS: dict = {}  # (entry point, dtype) -> the library function


def _fn(name, t):
    f = _FNS.get((name, t.dtype))
    if f is None:
        f = _FNS[(name, t.dtype)] = getattr(lib(), f"tvlm_{name}_{_sfx(t)}")
    return f


def _vec3(t, vals):
    R = _real(t)
    return (R * 3)(*[float(x) for x in vals])


# ---------------------------------------------------------------------------------------------
# Backend methods (vlm/include/vlm_backend.hpp:31-55), one Python call per C entry point
# ---------------------------------------------------------------------------------------------

def init(device: int = 0):
    check(lib().tvlm_init(ctypes.c_int32(device)), "tvlm_init")


def lhs_assemble(lhs, colloc, normals, verts_wing, verts_wake, iteration, sigma=0.0, stream=None):
    n = len(colloc)
    check(_fn("lhs_assemble", lhs)(ctypes.byref(view(lhs)), ctypes.c_int32(n), views(colloc), views(normals),
                                  views(verts_wing), views(verts_wake), ctypes.c_int32(iteration),
                                  _real(lhs)(sigma), stream_ptr(stream)), "lhs_assemble")


def lhs_assemble_seeded(lhs, colloc, normals, verts_wing, verts_wake, iteration, stream=None):
    """FP64 seeded-form assembly (tvlm_lhs_assemble_seeded_f64), the served steady path's kernel."""
    import torch

    if lhs.dtype != torch.float64:
        raise TypeError("the seeded form is FP64")
    check(lib().tvlm_lhs_assemble_seeded_f64(ctypes.byref(view(lhs)), ctypes.c_int32(len(colloc)), views(colloc),
                                             views(normals), views(verts_wing), views(verts_wake),
                                             ctypes.c_int32(iteration), stream_ptr(stream)), "lhs_assemble_seeded")


class LhsItem(ctypes.Structure):
    """struct tvlm_lhs_item: one influence matrix of a grouped assembly."""

    _fields_ = [("lhs", View), ("n_surf", ctypes.c_int32), ("colloc", ctypes.POINTER(View)),
                ("normals", ctypes.POINTER(View)), ("verts_wing", ctypes.POINTER(View)),
                ("verts_wake", ctypes.POINTER(View)), ("iteration", ctypes.c_int32)]


class LhsGroup:
    """Grouped FP64 assembly (tvlm_lhs_group_*): items = [(lhs, colloc, normals, verts_wing, verts_wake,
    iteration), ...], each as lhs_assemble takes them, meshes of any sizes; assemble() fills every
    matrix (the first n_items) in one launch, accumulating like lhs_assemble."""

    def __init__(self, items, stream=None):
        self._keep = []
        arr = (LhsItem * len(items))()
        for k, (lhs, colloc, normals, verts_wing, verts_wake, iteration) in enumerate(items):
            vs = [views(colloc), views(normals), views(verts_wing), views(verts_wake)]
            self._keep.extend(vs)
            arr[k] = LhsItem(view(lhs), len(colloc), vs[0], vs[1], vs[2], vs[3], int(iteration))
        self._g = ctypes.c_void_p()
        check(lib().tvlm_lhs_group_create(ctypes.byref(self._g), ctypes.c_int32(len(items)), arr,
                                          stream_ptr(stream)), "tvlm_lhs_group_create")
        self.n_items = len(items)

    def assemble(self, n_items=None, seeded=False, sigma=0.0, stream=None):
        k = self.n_items if n_items is None else int(n_items)
        check(lib().tvlm_lhs_group_assemble_f64(self._g, ctypes.c_int32(k), ctypes.c_int32(int(seeded)),
                                                ctypes.c_double(sigma), stream_ptr(stream)),
              "tvlm_lhs_group_assemble_f64")

    def close(self):
        if self._g:
            lib().tvlm_lhs_group_destroy(self._g)
            self._g = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rhs_assemble_velocities(rhs, normals, velocities, stream=None):
    check(_fn("rhs_assemble_velocities", rhs)(ctypes.byref(view(rhs)), ctypes.c_int32(len(normals)), views(normals),
                                             views(velocities), stream_ptr(stream)), "rhs_assemble_velocities")


def rhs_assemble_wake_influence(rhs, gamma_wake, colloc, normals, verts_wake, lifting, iteration, sigma=0.0,
                                stream=None):
    n = len(colloc)
    lift = (ctypes.c_uint8 * n)(*[1 if x else 0 for x in lifting])
    check(_fn("rhs_assemble_wake_influence", rhs)(ctypes.byref(view(rhs)), ctypes.c_int32(n), views(gamma_wake),
                                                 views(colloc), views(normals), views(verts_wake), lift,
                                                 ctypes.c_int32(iteration), _real(rhs)(sigma), stream_ptr(stream)),
          "rhs_assemble_wake_influence")


def rhs_assemble_wake_influence_lattice(rhs, gamma_wake, colloc, normals, verts_wake, lifting, iteration,
                                        stream=None):
    """tvlm_rhs_assemble_wake_influence_lattice_f64: the same wake induction evaluated on the wake's vortex
    lattice in the FP64 seeded form (no core model); float64 tensors only."""
    torch = _torch()
    if rhs.dtype != torch.float64:
        raise TvlmError("rhs_assemble_wake_influence_lattice: float64 only")
    n = len(colloc)
    lift = (ctypes.c_uint8 * n)(*[1 if x else 0 for x in lifting])
    check(lib().tvlm_rhs_assemble_wake_influence_lattice_f64(ctypes.byref(view(rhs)), ctypes.c_int32(n),
                                                            views(gamma_wake), views(colloc), views(normals),
                                                            views(verts_wake), lift, ctypes.c_int32(iteration),
                                                            stream_ptr(stream)),
          "rhs_assemble_wake_influence_lattice")


def rhs_assemble_wake_influence_noalloc(rhs, gamma_wake, colloc, normals, verts_wake, lifting, iteration, sigma=0.0,
                                        stream=None):
    """tvlm_rhs_assemble_wake_influence_noalloc_f64: the FP64 wake induction with its tile partials in
    the library's fixed device scratch (no device allocation); float64 tensors only."""
    torch = _torch()
    if rhs.dtype != torch.float64:
        raise TvlmError("rhs_assemble_wake_influence_noalloc: float64 only")
    n = len(colloc)
    lift = (ctypes.c_uint8 * n)(*[1 if x else 0 for x in lifting])
    check(lib().tvlm_rhs_assemble_wake_influence_noalloc_f64(
        ctypes.byref(view(rhs)), ctypes.c_int32(n), views(gamma_wake), views(colloc), views(normals),
        views(verts_wake), lift, ctypes.c_int32(iteration), ctypes.c_double(sigma), stream_ptr(stream)),
        "rhs_assemble_wake_influence_noalloc")


def mesh_metrics(alpha, verts_wing, colloc, normals, areas, stream=None):
    check(_fn("mesh_metrics", colloc[0])(_real(colloc[0])(alpha), ctypes.c_int32(len(colloc)), views(verts_wing),
                                        views(colloc), views(normals), views(areas), stream_ptr(stream)),
          "mesh_metrics")


def forces_steady(verts_wing, gamma_delta, velocities, forces, stream=None):
    check(_fn("forces_steady", forces)(ctypes.byref(view(verts_wing)), ctypes.byref(view(gamma_delta)),
                                      ctypes.byref(view(velocities)), ctypes.byref(view(forces)), stream_ptr(stream)),
          "forces_steady")


def forces_unsteady(verts_wing, gamma_delta, dgamma_dt, velocities, areas, normals, forces, stream=None):
    check(_fn("forces_unsteady", forces)(ctypes.byref(view(verts_wing)), ctypes.byref(view(gamma_delta)),
                                        ctypes.byref(view(dgamma_dt)), ctypes.byref(view(velocities)),
                                        ctypes.byref(view(areas)), ctypes.byref(view(normals)),
                                        ctypes.byref(view(forces)), stream_ptr(stream)), "forces_unsteady")


def coeff_cl(forces, lift_axis, freestream, rho, area, out, stream=None):
    R = _real(forces)
    check(_fn("coeff_cl", forces)(ctypes.byref(view(forces)), _vec3(forces, lift_axis), _vec3(forces, freestream),
                                 R(rho), R(area), ctypes.c_void_p(out.data_ptr()), stream_ptr(stream)), "coeff_cl")


def coeff_cm(forces, verts_wing, ref_pt, freestream, rho, area, mac, out, stream=None):
    R = _real(forces)
    check(_fn("coeff_cm", forces)(ctypes.byref(view(forces)), ctypes.byref(view(verts_wing)), _vec3(forces, ref_pt),
                                 _vec3(forces, freestream), R(rho), R(area), R(mac), ctypes.c_void_p(out.data_ptr()),
                                 stream_ptr(stream)), "coeff_cm")


def mesh_mac(verts_wing, areas, out, stream=None):
    check(_fn("mesh_mac", areas)(ctypes.byref(view(verts_wing)), ctypes.byref(view(areas)),
                                ctypes.c_void_p(out.data_ptr()), stream_ptr(stream)), "mesh_mac")


def sum_(x, out, stream=None):
    check(_fn("sum", x)(ctypes.byref(view(x)), ctypes.c_int32(x.dim()), ctypes.c_void_p(out.data_ptr()),
                        stream_ptr(stream)), "sum")


def nrm2(x, out, stream=None):
    check(_fn("nrm2", x)(ctypes.byref(view(x)), ctypes.c_void_p(out.data_ptr()), stream_ptr(stream)), "nrm2")


def wake_shed(verts_wing, verts_wake, iteration, stream=None):
    check(_fn("wake_shed", verts_wing[0])(ctypes.c_int32(len(verts_wing)), views(verts_wing), views(verts_wake),
                                         ctypes.c_int32(iteration), stream_ptr(stream)), "wake_shed")


def displace_wing(transforms, verts_wing, verts_wing_init, stream=None):
    check(_fn("displace_wing", verts_wing[0])(ctypes.c_int32(len(verts_wing)), views(transforms), views(verts_wing),
                                             views(verts_wing_init), stream_ptr(stream)), "displace_wing")


def wake_geometry(transforms, verts_wing_init, verts_wake, steps, stream=None):
    """tvlm_wake_geometry_*: wake rows of `steps` prescribed steps in one launch per surface
    (transforms {4, 4, steps * n_surf}, step-major)."""
    check(_fn("wake_geometry", transforms)(ctypes.c_int32(len(verts_wing_init)), ctypes.byref(view(transforms)),
                                           views(verts_wing_init), views(verts_wake), ctypes.c_int32(steps),
                                           stream_ptr(stream)), "wake_geometry")


def gamma_wake_from_coeffs(gamma_wake, gamma_coeffs, harmonics, tn, omega, dt, iteration, stream=None):
    R = _real(gamma_wake)
    check(_fn("gamma_wake_from_coeffs", gamma_wake)(ctypes.byref(view(gamma_wake)), ctypes.byref(view(gamma_coeffs)),
                                                    ctypes.c_int32(harmonics), R(tn), R(omega), R(dt),
                                                    ctypes.c_int64(iteration), stream_ptr(stream)),
          "gamma_wake_from_coeffs")


def getrf(a, ipiv, info, stream=None):
    check(_fn("getrf", a)(ctypes.byref(view(a)), ctypes.c_void_p(ipiv.data_ptr()), ctypes.c_void_p(info.data_ptr()),
                          stream_ptr(stream)), "getrf")


def getrs(a, ipiv, x, stream=None):
    """x {N}, {N, nrhs} or, for a batched a {N, N, B}, {N, nrhs, B}."""
    x2 = x.unsqueeze(1) if x.dim() == 1 else x
    check(_fn("getrs", a)(ctypes.byref(view(a)), ctypes.c_void_p(ipiv.data_ptr()), ctypes.byref(view(x2)),
                          stream_ptr(stream)), "getrs")


def gels(a, b, info, stream=None):
    """LSQ::solve: a {M,N} <- QR factors, b {M,NRHS} (or {M}) <- Q^T b with X in rows 0..N-1."""
    b2 = b if b.dim() == 2 else b.unsqueeze(1)
    check(_fn("gels", a)(ctypes.byref(view(a)), ctypes.byref(view(b2)), ctypes.c_void_p(info.data_ptr()),
                         stream_ptr(stream)), "gels")


def kinematic_velocities(dT_list, colloc, velocities, stream=None):
    """dT_list: one 4x4 numpy time derivative per surface."""
    R = _real(colloc[0])
    flat = np.concatenate([np.asarray(d, dtype=np.float64).reshape(16, order="F") for d in dT_list])
    arr = (R * len(flat))(*flat.tolist())
    check(_fn("kinematic_velocities", colloc[0])(ctypes.c_int32(len(colloc)), arr, views(colloc), views(velocities),
                                                 stream_ptr(stream)), "kinematic_velocities")


def gemm(alpha, a, b, beta, c, trans_a=False, trans_b=False, stream=None):
    R = _real(c)
    check(_fn("gemm", c)(R(alpha), ctypes.byref(view(a)), ctypes.byref(view(b)), R(beta), ctypes.byref(view(c)),
                         ctypes.c_int32(int(trans_a)), ctypes.c_int32(int(trans_b)), stream_ptr(stream)), "gemm")


def gemv(alpha, a, x, beta, y, trans=False, stream=None):
    R = _real(y)
    check(_fn("gemv", y)(R(alpha), ctypes.byref(view(a)), ctypes.byref(view(x)), R(beta), ctypes.byref(view(y)),
                         ctypes.c_int32(int(trans)), stream_ptr(stream)), "gemv")


def axpy(alpha, x, y, stream=None):
    R = _real(y)
    check(_fn("axpy", y)(R(alpha), ctypes.byref(view(x)), ctypes.byref(view(y)), ctypes.c_int32(x.dim()),
                         stream_ptr(stream)), "axpy")


def scal(alpha, x, stream=None):
    check(_fn("scal", x)(_real(x)(alpha), ctypes.byref(view(x)), stream_ptr(stream)), "scal")
