"""Rigid-body kinematics with forward-mode dual numbers (caller side of the hot path).

Mirrors vlm/include/vlm_kinematics.hpp:54-282 and headeronly/tinyad.hpp: each node maps a dual
time t + 1*eps to a 4x4 homogeneous transform whose dual part is dT/dt; nodes compose
parent * child (KinematicNode::transform_dual, vlm_kinematics.hpp:214-222).  The host evaluates
one 4x4 (value + derivative) per node per time step; the per-panel velocities
V = -(dT/dt) [x; 1] are then computed on the device (tvlm_kinematic_velocities_*).
"""
from __future__ import annotations

import math
from typing import Callable

import numpy as np


def _num(x):
    return x if isinstance(x, np.ndarray) else float(x)


class Dual:
    """value + derivative; only the operations the transforms need.  v and d are floats or, to
    evaluate many times at once, numpy arrays of times (elementwise the same arithmetic)."""

    __slots__ = ("v", "d")

    def __init__(self, v, d=0.0):
        self.v = _num(v)
        self.d = _num(d)

    def __add__(self, o):
        o = o if isinstance(o, Dual) else Dual(o)
        return Dual(self.v + o.v, self.d + o.d)

    __radd__ = __add__

    def __sub__(self, o):
        o = o if isinstance(o, Dual) else Dual(o)
        return Dual(self.v - o.v, self.d - o.d)

    def __rsub__(self, o):
        return Dual(o) - self

    def __mul__(self, o):
        o = o if isinstance(o, Dual) else Dual(o)
        return Dual(self.v * o.v, o.v * self.d + self.v * o.d)

    __rmul__ = __mul__

    def __neg__(self):
        return Dual(-self.v, -self.d)


def _sin(x):
    return np.sin(x) if isinstance(x, np.ndarray) else math.sin(x)


def _cos(x):
    return np.cos(x) if isinstance(x, np.ndarray) else math.cos(x)


def dsin(x: Dual) -> Dual:
    return Dual(_sin(x.v), _cos(x.v) * x.d)


def dcos(x: Dual) -> Dual:
    return Dual(_cos(x.v), -_sin(x.v) * x.d)


class DMat4:
    """4x4 dual matrix, m[row][col]."""

    def __init__(self, m=None):
        self.m = m if m is not None else [[Dual(1.0 if r == c else 0.0) for c in range(4)] for r in range(4)]

    def __matmul__(self, o: "DMat4") -> "DMat4":
        # column j of the product = sum_k a.col(k) * b(k, j), k = 0..3 in order (linalg::mul)
        r = [[None] * 4 for _ in range(4)]
        for j in range(4):
            for i in range(4):
                s = self.m[i][0] * o.m[0][j]
                for k in range(1, 4):
                    s = s + self.m[i][k] * o.m[k][j]
                r[i][j] = s
        return DMat4(r)

    def _stack(self, part) -> np.ndarray:
        ent = [[getattr(x, part) for x in row] for row in self.m]
        n = [np.shape(e) for row in ent for e in row if isinstance(e, np.ndarray)]
        if not n:
            return np.array(ent)
        out = np.empty((4, 4) + n[0])
        for i in range(4):
            for j in range(4):
                out[i, j] = ent[i][j]
        return out

    def value(self) -> np.ndarray:
        """4x4 values, or (4, 4, len(t)) when evaluated at an array of times."""
        return self._stack("v")

    def grad(self) -> np.ndarray:
        return self._stack("d")


def translation(x, y, z) -> DMat4:
    """translation_matrix (vlm_kinematics.hpp:149-157)."""
    M = DMat4()
    for r, c in ((0, x), (1, y), (2, z)):
        M.m[r][3] = c if isinstance(c, Dual) else Dual(c)
    return M


def rotation(point, axis, angle: Dual) -> DMat4:
    """rotation_matrix, Rodrigues (vlm_kinematics.hpp:184-202): R = I + sin K + (1-cos) K^2,
    translation part (I - R) p, K the cross-product matrix of axis."""
    ax, ay, az = axis
    K = [[0.0, -az, ay], [az, 0.0, -ax], [-ay, ax, 0.0]]
    K2 = [[sum(K[i][k] * K[k][j] for k in range(3)) for j in range(3)] for i in range(3)]
    s, c = dsin(angle), dcos(angle)
    R = [[Dual(1.0 if i == j else 0.0) + s * K[i][j] + (1.0 - c) * K2[i][j] for j in range(3)] for i in range(3)]
    M = DMat4()
    for i in range(3):
        for j in range(3):
            M.m[i][j] = R[i][j]
    for i in range(3):
        t = Dual(0.0)
        for j in range(3):
            t = t + (Dual(1.0 if i == j else 0.0) - R[i][j]) * point[j]
        M.m[i][3] = t
    return M


class KinematicNode:
    """KinematicNode (vlm_kinematics.hpp:54-112): transform(t_dual) composed after its parent."""

    def __init__(self, fn: Callable[[Dual], DMat4] | None = None):
        self.fn = fn or (lambda t: DMat4())
        self.parent: KinematicNode | None = None

    def after(self, parent: "KinematicNode") -> "KinematicNode":
        self.parent = parent
        return self

    def transform_dual(self, t) -> DMat4:
        r = self.fn(Dual(t, np.ones_like(t) if isinstance(t, np.ndarray) else 1.0))
        return self.parent.transform_dual(t) @ r if self.parent else r

    def transform(self, t) -> np.ndarray:
        return self.transform_dual(t).value()

    def linear_velocity(self, t: float, vertex) -> np.ndarray:
        d = self.transform_dual(t).grad()
        return d[:3, :3] @ np.asarray(vertex, dtype=float) + d[:3, 3]


class KinematicsTree:
    """Owner of nodes (vlm_kinematics.hpp:118-138)."""

    def __init__(self):
        self.nodes: list[KinematicNode] = []

    def add(self, fn) -> KinematicNode:
        n = KinematicNode(fn)
        self.nodes.append(n)
        return n

    def placeholder(self) -> KinematicNode:
        return self.add(None)


def pitch_heave(u=1.0, h_amp=0.0, h_phase=0.0, a_amp=0.0, a_phase=0.0, omega=0.0, x_pivot=0.25):
    """Freestream + heave + pitch chain used by the unsteady workloads:
    fs = translation(-u t, 0, 0) (tests/uvlm_theodorsen.cpp:198-200), heave = translation(0, 0,
    h sin(w t + ph_h)) after fs, pitch = rotation about (x_pivot, 0, 0), axis y, a sin(w t + ph_a)
    after heave.  Returns (tree, assembly_node=fs, surface_node=pitch)."""
    tree = KinematicsTree()
    fs = tree.add(lambda t: translation(-u * t, 0.0, 0.0))
    hv = tree.add(lambda t: translation(0.0, 0.0, h_amp * dsin(omega * t + h_phase))).after(fs)
    pt = tree.add(lambda t: rotation((x_pivot, 0.0, 0.0), (0.0, 1.0, 0.0), a_amp * dsin(omega * t + a_phase))).after(hv)
    return tree, fs, pt
