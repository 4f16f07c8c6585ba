"""Steady VLM solve over the HIP C ABI: the call sequence of VLM::run (vlm/src/vlm.cpp:97-151).

Every arithmetic step runs in libtinyvlm_hip.so; this module only allocates device tensors and
issues the Backend calls in the reference's order.
"""
from __future__ import annotations

import math

import numpy as np

from . import capi as H


def _torch():
    import torch

    return torch


class SteadyVLM:
    """Buffers of VLM::alloc_buffers (vlm/src/vlm.cpp:60-95) for a list of surfaces."""

    def __init__(self, verts_init_list, dtype=None, device="cuda"):
        torch = _torch()
        self.dtype = dtype or torch.float64
        self.device = device
        self.dims = [(v.shape[0] - 1, v.shape[1] - 1) for v in verts_init_list]
        self.N = sum(ns * nc for ns, nc in self.dims)
        D = self.dtype
        self.verts_init = [H.to_device(np.asarray(v, dtype=np.float64), D) for v in verts_init_list]
        self.verts = [H.colmajor((ns + 1, nc + 1, 4), D) for ns, nc in self.dims]
        self.wake = [H.colmajor((ns + 1, 2, 4), D) for ns, nc in self.dims]
        for w in self.wake:
            w[:, :, 3] = 1
        self.colloc = [H.colmajor((ns, nc, 3), D) for ns, nc in self.dims]
        self.normals = [H.colmajor((ns, nc, 3), D) for ns, nc in self.dims]
        self.areas = [H.colmajor((ns, nc), D) for ns, nc in self.dims]
        self.vel = [H.colmajor((ns, nc, 3), D) for ns, nc in self.dims]
        self.forces = [H.colmajor((ns, nc, 3), D) for ns, nc in self.dims]
        self.gdelta = [H.colmajor((ns, nc), D) for ns, nc in self.dims]
        self.transforms = [H.colmajor((4, 4), D) for _ in self.dims]
        self.lhs = H.colmajor((self.N, self.N), D)
        self.rhs = torch.zeros(self.N, dtype=D, device=device)
        self.ipiv = torch.zeros(self.N, dtype=torch.int32, device=device)
        self.info = torch.zeros(1, dtype=torch.int32, device=device)
        self.scal = torch.zeros(8, dtype=D, device=device)

    def run(self, alpha, beta=0.0, u_inf=1.0, rho=1.0, cm_ref=(0.25, 0.0, 0.0)):
        """Returns (gamma[N] numpy, CL, CM[3]).  CM is taken about cm_ref moved with the wing."""
        torch = _torch()
        self.lhs.zero_()
        self.rhs.zero_()
        for vi, v in zip(self.verts_init, self.verts):
            v.copy_(vi)
        ca, sa = math.cos(alpha), math.sin(alpha)
        tx, tz = -100.0 * u_inf * ca, -100.0 * u_inf * sa
        T = np.eye(4)
        T[0, 3], T[2, 3] = tx, tz
        for tr in self.transforms:
            tr.copy_(H.to_device(T, self.dtype))
        H.wake_shed(self.verts, self.wake, 0)
        H.displace_wing(self.transforms, self.verts, self.verts_init)
        H.wake_shed(self.verts, self.wake, 1)
        H.mesh_metrics(alpha, self.verts, self.colloc, self.normals, self.areas)
        H.lhs_assemble(self.lhs, self.colloc, self.normals, self.verts, self.wake, 1)
        fs = (u_inf * ca * math.cos(beta), -u_inf * ca * math.sin(beta), u_inf * sa)
        for v in self.vel:
            for k in range(3):
                v[:, :, k] = fs[k]
        H.rhs_assemble_velocities(self.rhs, self.normals, self.vel)
        H.getrf(self.lhs, self.ipiv, self.info)
        H.getrs(self.lhs, self.ipiv, self.rhs)
        off = 0
        for (ns, nc), gd, v, vel, f in zip(self.dims, self.gdelta, self.verts, self.vel, self.forces):
            g = self.rhs[off:off + ns * nc].view(nc, ns).t()  # column-major {ns, nc} view of the slice
            gd.copy_(g)
            if nc > 1:
                H.axpy(-1.0, g[:, 0:nc - 1], gd[:, 1:nc])
            H.forces_steady(v, gd, vel, f)
            off += ns * nc
        cl, cm = self._coefficients(fs, rho, T, cm_ref)
        info = int(self.info.item())
        if info != 0:
            raise H.TvlmError(f"singular influence matrix (getrf info {info})")
        return H.to_host(self.rhs).copy(), cl, cm

    def _coefficients(self, fs, rho, T, cm_ref):
        """coeff_cl_multibody / coeff_cm_multibody (vlm/src/vlm_backend.cpp:71-139)."""
        la = np.cross(np.array(fs), np.array([0.0, 1.0, 0.0]))
        la = la / np.sqrt(np.dot(la, la))
        ref = (T @ np.array([*cm_ref, 1.0]))[:3]
        cl = 0.0
        tot_a = 0.0
        tot_mac = 0.0
        cm = np.zeros(3)
        s = self.scal
        for a, f, v in zip(self.areas, self.forces, self.verts):
            H.sum_(a, s[0:1])
            H.mesh_mac(v, a, s[1:2])
            area, mac = (float(x) for x in s[0:2].cpu())
            H.coeff_cl(f, la, fs, rho, area, s[2:3])
            H.coeff_cm(f, v, ref, fs, rho, area, mac, s[3:6])
            vals = s[2:6].cpu().numpy()
            cl += float(vals[0]) * area
            cm += vals[1:4] * area * mac
            tot_a += area
            tot_mac += mac
        return cl / tot_a, cm / (tot_a * tot_mac)


def vlm_solve_f64(verts_init_list, alpha, **kw):
    torch = _torch()
    s = SteadyVLM(verts_init_list, dtype=torch.float64)
    g, cl, _ = s.run(alpha, **kw)
    return g, cl
