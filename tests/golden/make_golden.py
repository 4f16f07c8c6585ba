"""Writes tests/golden/golden_small.npz: independent long-double evaluations on tiny meshes.

The values are computed here from the mathematics (Katz-Plotkin eq. 10.115 segments with the
y = 0 image, vortex rings, panel metrics, dense solve by Gaussian elimination with partial
pivoting), vectorised in numpy ``longdouble`` (80-bit on x86).  Nothing in this script calls the
oracle or the HIP library, so the fixtures pin both.  The reference's float-typed constants are
reproduced where the reference uses them (4.0f*PI_f, RCUT = 1e-10f, EPS_f), because the FP64
reference backend uses them too (vlm_backend_cpu_kernels.ispc:393-400, vlm_backend_cpud.cpp:290).

Inputs are the mesh fixtures the reference's own tests use (mesh/*.x), parsed in float32 and
quarter-chord shifted in float32 as the reference drivers do (vlm/src/vlm_mesh.cpp:15-124).

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os

import numpy as np

LD = np.longdouble
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
FOURPI = LD(np.float32(4.0) * np.float32(3.141593))  # the float product 4.0f*PI_f
RCUT = LD(np.float32(1e-10))
EPS_F = LD(np.float32(1.19209e-07))


def read_plot3d_f32(path):
    tok = open(path).read().split()
    ni, nj = int(tok[1]), int(tok[2])
    n = ni * nj
    vals = np.array([np.float32(t) for t in tok[4:4 + 3 * n]], dtype=np.float32)
    v = np.zeros((nj, ni, 4), dtype=np.float32)
    for k in range(3):
        v[:, :, k] = vals[k * n:(k + 1) * n].reshape(nj, ni)
    v[:, :, 3] = 1
    return v


def qc_f32(v):
    v = v.copy()
    nc1 = v.shape[1]
    for j in range(nc1 - 1):
        v[:, j, :3] = np.float32(0.75) * v[:, j, :3] + np.float32(0.25) * v[:, j + 1, :3]
    j = nc1 - 2
    v[:, j + 1, :3] = np.float32(4.0 / 3.0) * v[:, j + 1, :3] - np.float32(1.0 / 3.0) * v[:, j, :3]
    return v


def segment(c, a, b):
    """Velocity at points c (..., 3) of unit segments a->b (..., 3), long double, FP64 cutoffs."""
    r0 = b - a
    r1 = c - a
    r2 = c - b
    cr = np.cross(r1, r2)
    r1n = np.sqrt(np.sum(r1 * r1, axis=-1))
    r2n = np.sqrt(np.sum(r2 * r2, axis=-1))
    sq = np.sum(cr * cr, axis=-1)
    num = np.sum(r0 * r1, axis=-1) * r2n - np.sum(r0 * r2, axis=-1) * r1n
    den = FOURPI * sq * r1n * r2n  # sigma = 0: sqrt(sq^2) = sq
    ok = (sq >= RCUT) & (r1n >= RCUT) & (r2n >= RCUT)
    coeff = np.where(ok, num / np.where(ok, den, LD(1)), LD(0))
    return cr * coeff[..., None]


def ring_velocity(c, v0, v1, v2, v3):
    """Ring (v0,v1,v2,v3) plus its y-mirror image response at receivers c.  Broadcasts."""
    cm = c.copy()
    cm[..., 1] = -cm[..., 1]
    flip = np.array([1, -1, 1], dtype=LD)
    tot = np.zeros(np.broadcast_shapes(c.shape, v0.shape), dtype=LD)
    for a, b in ((v0, v1), (v1, v2), (v2, v3), (v3, v0)):
        tot = tot + segment(c, a, b) + flip * segment(cm, a, b)
    return tot


def metrics(v, alpha):
    v = v.astype(LD)
    v0 = v[:-1, :-1, :3]
    v1 = v[1:, :-1, :3]
    v2 = v[1:, 1:, :3]
    v3 = v[:-1, 1:, :3]
    nrm = np.cross(v3 - v1, v2 - v0)
    nrm = nrm / np.sqrt(np.sum(nrm * nrm, axis=-1))[..., None]
    f, b, e = v3 - v0, v2 - v0, v1 - v0
    area = LD(0.5) * (np.sqrt(np.sum(np.cross(f, b) ** 2, axis=-1)) + np.sqrt(np.sum(np.cross(b, e) ** 2, axis=-1)))
    alpha = LD(alpha)
    factor = LD(0.5) if alpha < EPS_F else LD(0.5) * (alpha / (np.sin(alpha) + EPS_F))
    colloc = LD(0.5) * (v0 + v1) + factor * (LD(0.5) * (v2 + v3 - v0 - v1))
    return colloc, nrm, area


def rings_of(v):
    """All rings of a vertex grid (ns+1, nr+1, .) in panel order p = i + j*ns."""
    v = v.astype(LD)[:, :, :3]
    ns = v.shape[0] - 1
    r = [v[:-1, :-1], v[1:, :-1], v[1:, 1:], v[:-1, 1:]]
    return [x.reshape(-1, 3, order="F") for x in r], ns


def lhs_ld(surfs, wakes, iteration):
    """surfs: list of (verts, colloc, normals); wakes: verts_wake per surface."""
    recv_c = np.concatenate([s[1].reshape(-1, 3, order="F") for s in surfs])
    recv_n = np.concatenate([s[2].reshape(-1, 3, order="F") for s in surfs])
    cols = []
    for (verts, _, _), wake in zip(surfs, wakes):
        (r0, r1, r2, r3), ns = rings_of(verts)
        nc = verts.shape[1] - 1
        vel = ring_velocity(recv_c[:, None, :], r0[None], r1[None], r2[None], r3[None])  # (N, np, 3)
        block = np.sum(vel * recv_n[:, None, :], axis=-1)
        w = wake.astype(LD)[:, :, :3]
        nwv = w.shape[1]
        for k in range(iteration):
            row = nwv - 2 - k
            a0, a1, a2, a3 = w[:-1, row], w[1:, row], w[1:, row + 1], w[:-1, row + 1]
            velw = ring_velocity(recv_c[:, None, :], a0[None], a1[None], a2[None], a3[None])
            block[:, (nc - 1) * ns:] += np.sum(velw * recv_n[:, None, :], axis=-1)
        cols.append(block)
    return np.concatenate(cols, axis=1)


def wake_rhs_ld(colloc, normals, verts_wake, gamma_wake, iteration):
    c = colloc.reshape(-1, 3, order="F").astype(LD)
    n = normals.reshape(-1, 3, order="F").astype(LD)
    w = verts_wake.astype(LD)[:, :, :3]
    nwv = w.shape[1]
    tot = np.zeros_like(c)
    for row in range(nwv - 1 - iteration, nwv - 1):
        a0, a1, a2, a3 = w[:-1, row], w[1:, row], w[1:, row + 1], w[:-1, row + 1]
        vel = ring_velocity(c[:, None, :], a0[None], a1[None], a2[None], a3[None])
        tot += np.sum(vel * gamma_wake[:, row].astype(LD)[None, :, None], axis=1)
    return -np.sum(tot * n, axis=-1)


def solve_ld(A, b):
    A = A.copy()
    b = b.copy()
    n = A.shape[0]
    for k in range(n):
        p = k + int(np.argmax(np.abs(A[k:, k])))
        if p != k:
            A[[k, p]] = A[[p, k]]
            b[[k, p]] = b[[p, k]]
        for i in range(k + 1, n):
            f = A[i, k] / A[k, k]
            A[i, k:] -= f * A[k, k:]
            b[i] -= f * b[k]
    x = np.zeros(n, dtype=LD)
    for k in range(n - 1, -1, -1):
        x[k] = (b[k] - np.dot(A[k, k + 1:], x[k + 1:])) / A[k, k]
    return x


def vlm_wake(v, alpha, u_inf=1.0):
    """Wake of VLM::run (vlm/src/vlm.cpp:103-114): row 1 = original TE, row 0 = TE moved by
    -100*u_inf*(cos a, 0, sin a); the wing itself is translated by the same vector."""
    a32 = np.float32(alpha)
    tx = np.float64(np.float32(-100.0) * np.float32(u_inf) * np.cos(a32))
    tz = np.float64(np.float32(-100.0) * np.float32(u_inf) * np.sin(a32))
    wing = v.astype(np.float64).copy()
    wake = np.zeros((v.shape[0], 2, 4))
    wake[:, 1] = wing[:, -1]
    wing[:, :, 0] += tx
    wing[:, :, 2] += tz
    wake[:, 0] = wing[:, -1]
    return wing, wake, (tx, tz)


def main():
    out = {}
    rng = np.random.default_rng(1234)
    alpha5 = float(np.float32(5) * np.float32(3.141593) / np.float32(180))
    cases = [("rect2", ["rectangular_2x2"]), ("rect4", ["rectangular_4x4"]), ("swept4", ["swept_4x4"]),
             ("split2", ["infinite_rectangular_2x2_part0", "infinite_rectangular_2x2_part1"])]
    for name, meshes in cases:
        for alpha, tag in ((0.0, "a0"), (alpha5, "a5")):
            surfs, wakes = [], []
            for k, m in enumerate(meshes):
                v = qc_f32(read_plot3d_f32(os.path.join(ROOT, "mesh", m + ".x")))
                wing, wake, _ = vlm_wake(v, alpha)
                c, n, a = metrics(wing, alpha)
                surfs.append((wing, c, n))
                wakes.append(wake)
                key = f"{name}_{tag}_s{k}"
                out[key + "_verts"] = wing
                out[key + "_wake"] = wake
                out[key + "_colloc"] = c.astype(np.float64)
                out[key + "_normals"] = n.astype(np.float64)
                out[key + "_areas"] = a.astype(np.float64)
            A = lhs_ld(surfs, wakes, 1)
            out[f"{name}_{tag}_lhs"] = A.astype(np.float64)
            fs = np.array([np.cos(LD(alpha)), 0, np.sin(LD(alpha))], dtype=LD)
            rhs = -np.concatenate([s[2].reshape(-1, 3, order="F") @ fs for s in surfs])
            out[f"{name}_{tag}_gamma"] = solve_ld(A, rhs).astype(np.float64)
    # a wake of 4 shed rows with random circulation acting on rectangular_4x4
    v = qc_f32(read_plot3d_f32(os.path.join(ROOT, "mesh", "rectangular_4x4.x"))).astype(np.float64)
    c, n, _ = metrics(v, 0.0)
    nwv = 6
    wake = np.zeros((v.shape[0], nwv, 4))
    for r in range(nwv):
        wake[:, r] = v[:, -1]
        wake[:, r, 0] += 0.25 * (nwv - 1 - r)
        wake[:, r, 2] += 0.01 * np.sin(r)
    gw = rng.standard_normal((v.shape[0] - 1, nwv - 1))
    out["wake4_verts"] = v
    out["wake4_verts_wake"] = wake
    out["wake4_gamma_wake"] = gw
    out["wake4_colloc"] = c.astype(np.float64)
    out["wake4_normals"] = n.astype(np.float64)
    for it in (1, 3, 5):
        out[f"wake4_rhs_it{it}"] = wake_rhs_ld(c.astype(np.float64), n.astype(np.float64), wake, gw, it).astype(np.float64)
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
