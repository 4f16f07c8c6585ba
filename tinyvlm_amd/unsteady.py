"""Unsteady VLM over the HIP C ABI: the call sequence of UVLM::run (vlm/src/vlm.cpp:358-509).

Each Backend call is one C-ABI entry point; the per-panel kinematic velocities are computed on the
device from the host-evaluated dual 4x4 (replacing the host loop at vlm.cpp:419-436).  This is
the drop-in call pattern, one launch per reference call and host reads of the per-surface areas
and MAC each step, like the reference's by-value returns.
"""
from __future__ import annotations

import math

import numpy as np

from . import capi as H
from .kinematics import KinematicNode


def _torch():
    import torch

    return torch


class UVLM:
    def __init__(self, verts_init_list, surface_nodes: list[KinematicNode], assembly_node: KinematicNode,
                 dtype=None, device="cuda"):
        torch = _torch()
        self.dtype = dtype or torch.float64
        self.device = device
        self.nodes = surface_nodes
        self.assembly = assembly_node
        self.verts_init_h = [np.asarray(v, dtype=np.float64) for v in verts_init_list]
        self.dims = [(v.shape[0] - 1, v.shape[1] - 1) for v in verts_init_list]
        self.N = sum(ns * nc for ns, nc in self.dims)
        D = self.dtype
        self.verts_init = [H.to_device(v, D) for v in self.verts_init_h]
        self.verts_pos = [H.colmajor((ns + 1, nc + 1, 4), D) for ns, nc in self.dims]
        self.verts = [H.colmajor((ns + 1, nc + 1, 4), D) for ns, nc in self.dims]
        self.colloc = [H.colmajor((ns, nc, 3), D) for ns, nc in self.dims]
        self.colloc_h = [H.colmajor((ns, nc, 3), D) for ns, nc in self.dims]
        self.normals = [H.colmajor((ns, nc, 3), D) for ns, nc in self.dims]
        self.areas = [H.colmajor((ns, nc), D) for ns, nc in self.dims]
        self.vel = [H.colmajor((ns, nc, 3), D) for ns, nc in self.dims]
        self.forces = [H.colmajor((ns, nc, 3), D) for ns, nc in self.dims]
        self.gdelta = [H.colmajor((ns, nc), D) for ns, nc in self.dims]
        self.gprev = [H.colmajor((ns, nc), D) for ns, nc in self.dims]
        self.gdt = [H.colmajor((ns, nc), D) for ns, nc in self.dims]
        self.transforms = [H.colmajor((4, 4), D) for _ in self.dims]
        self.lhs = H.colmajor((self.N, self.N), D)
        self.rhs = torch.zeros(self.N, dtype=D, device=device)
        self.ipiv = torch.zeros(self.N, dtype=torch.int32, device=device)
        self.info = torch.zeros(1, dtype=torch.int32, device=device)
        self.scal = torch.zeros(8, dtype=D, device=device)

    def _set_transforms(self, t):
        for node, tr in zip(self.nodes, self.transforms):
            tr.copy_(H.to_device(node.transform(t), self.dtype))

    def time_step(self):
        """dt = last panel chord of surface 0 at span station 0 / body speed (vlm.cpp:380-386)."""
        v = self.verts_init_h[0]
        d = v[0, -1, :3] - v[0, -2, :3]
        chord = math.sqrt(float(np.dot(d, d)))
        speed = float(np.linalg.norm(self.assembly.linear_velocity(0.0, (0.0, 0.0, 0.0))))
        return chord / speed

    def run(self, t_final=None, t_steps=None, rho=1.0):
        """Returns dict(cl, cm_y, t) arrays (entries of step 0 are 0, as the reference skips it) and the
        last gamma."""
        torch = _torch()
        D = self.dtype
        self._set_transforms(0.0)
        H.displace_wing(self.transforms, self.verts_pos, self.verts_init)
        for a, b in zip(self.verts, self.verts_pos):
            a.copy_(b)
        H.mesh_metrics(0.0, self.verts, self.colloc, self.normals, self.areas)
        for ch, c in zip(self.colloc_h, self.colloc):
            ch.copy_(c)
        self.lhs.zero_()
        dt = self.time_step()
        if t_steps is None:
            t_steps = int(t_final / dt)
        self.wake = [H.colmajor((ns + 1, t_steps, 4), D) for ns, nc in self.dims]
        self.gamma_wake = [H.colmajor((ns, t_steps - 1), D) for ns, nc in self.dims]
        H.lhs_assemble(self.lhs, self.colloc, self.normals, self.verts, self.wake, 0)
        H.getrf(self.lhs, self.ipiv, self.info)
        lifting = [True] * len(self.dims)
        cl = np.zeros(t_steps - 1)
        cm = np.zeros(t_steps - 1)
        ts = np.zeros(t_steps - 1)
        for i in range(t_steps - 1):
            t = i * dt
            ts[i] = t
            fs = -self.assembly.linear_velocity(t, (0.0, 0.0, 0.0))
            H.wake_shed(self.verts, self.wake, i)
            H.kinematic_velocities([n.transform_dual(t).grad() for n in self.nodes], self.colloc_h, self.vel)
            self.rhs.zero_()
            H.rhs_assemble_velocities(self.rhs, self.normals, self.vel)
            H.rhs_assemble_wake_influence(self.rhs, self.gamma_wake, self.colloc, self.normals, self.wake, lifting, i)
            H.getrs(self.lhs, self.ipiv, self.rhs)
            off = 0
            for (ns, nc), gd, gp, gt, gw in zip(self.dims, self.gdelta, self.gprev, self.gdt, self.gamma_wake):
                g = self.rhs[off:off + ns * nc].view(nc, ns).t()
                gd.copy_(g)
                gw[:, t_steps - 2 - i].copy_(g[:, nc - 1])
                if nc > 1:
                    H.axpy(-1.0, g[:, 0:nc - 1], gd[:, 1:nc])
                gt.copy_(g)
                H.axpy(-1.0, gp, gt)
                H.scal(1.0 / dt, gt.t().reshape(-1))
                off += ns * nc
            if i > 0:
                for v, gd, gt, vel, a, n, f in zip(self.verts, self.gdelta, self.gdt, self.vel, self.areas,
                                                   self.normals, self.forces):
                    H.forces_unsteady(v, gd, gt, vel, a, n, f)
                T0 = self.nodes[0].transform(t)
                ref = (T0 @ np.array([0.5, 0.0, 0.0, 1.0]))[:3]
                cl[i], cmv = self._coefficients(fs, rho, ref)
                cm[i] = cmv[1]
            off = 0
            for (ns, nc), gp in zip(self.dims, self.gprev):
                gp.copy_(self.rhs[off:off + ns * nc].view(nc, ns).t())
                off += ns * nc
            self._set_transforms(t + dt)
            H.displace_wing(self.transforms, self.verts, self.verts_pos)
            H.mesh_metrics(0.0, self.verts, self.colloc, self.normals, self.areas)
        if int(self.info.item()) != 0:
            raise H.TvlmError("singular influence matrix")
        return {"cl": cl, "cm_y": cm, "t": ts, "gamma": H.to_host(self.rhs).copy(), "t_steps": t_steps}

    def _coefficients(self, fs, rho, ref):
        la = np.cross(fs, np.array([0.0, 1.0, 0.0]))
        la = la / np.sqrt(np.dot(la, la))
        s = self.scal
        cl = tot_a = tot_mac = 0.0
        cm = np.zeros(3)
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
