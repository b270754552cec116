"""TEST INFRASTRUCTURE — forward-mode dual numbers through the NumPy restatement of the
dense collocation assembly (oracle/oracle_np.py), at frozen rule selection.

A Dual carries a value and one directional derivative (tangent) through every operation,
so F(X + t dX, k + t dk) and C are differentiated exactly in t (to rounding) without
finite differences: the check the adjoint kernel is held to besides central differences.
The assembly formulas are those of oracle_np.assemble (interpolate_elements / jacobian!,
src/3d/assembly_collocation.jl:27-66, src/3d/jacobian.jl:7-87; freens3d! / freens3dk0!,
src/3d/kernels.jl:25-55), with the rule of every (target, element) pair given.
"""
from __future__ import annotations

import numpy as np

import oracle_np


class Dual:
    """value + t * tangent, elementwise over numpy arrays (complex allowed)."""

    __array_priority__ = 1000

    def __init__(self, v, d=None):
        self.v = np.asarray(v)
        self.d = np.zeros_like(self.v) if d is None else np.asarray(d)

    @staticmethod
    def lift(x):
        return x if isinstance(x, Dual) else Dual(x, np.zeros_like(np.asarray(x)))

    def __add__(self, o):
        o = Dual.lift(o)
        return Dual(self.v + o.v, self.d + o.d)

    __radd__ = __add__

    def __neg__(self):
        return Dual(-self.v, -self.d)

    def __sub__(self, o):
        return self + (-Dual.lift(o))

    def __rsub__(self, o):
        return Dual.lift(o) - self

    def __mul__(self, o):
        o = Dual.lift(o)
        return Dual(self.v * o.v, self.d * o.v + self.v * o.d)

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = Dual.lift(o)
        return Dual(self.v / o.v, (self.d * o.v - self.v * o.d) / (o.v * o.v))

    def __rtruediv__(self, o):
        return Dual.lift(o) / self

    def __pow__(self, p: int):
        return Dual(self.v ** p, p * self.v ** (p - 1) * self.d)

    def __getitem__(self, idx):
        return Dual(self.v[idx], self.d[idx])

    def __matmul__(self, B):  # constant right factor
        return Dual(self.v @ B, self.d @ B)

    def sum(self, axis=None):
        return Dual(self.v.sum(axis=axis), self.d.sum(axis=axis))


def dsqrt(x: Dual) -> Dual:
    s = np.sqrt(x.v)
    return Dual(s, x.d / (2 * s))


def dexp(x: Dual) -> Dual:
    e = np.exp(x.v)
    return Dual(e, e * x.d)


def lin_map(spec, X: Dual, B):
    """einsum(spec, X, B) for a constant B (linear in X)."""
    return Dual(np.einsum(spec, X.v, B), np.einsum(spec, X.d, B))


def cross(a: Dual, b: Dual) -> Dual:
    return Dual(np.cross(a.v, b.v), np.cross(a.d, b.v) + np.cross(a.v, b.d))


def assemble_F_C(mesh, X, dX, k, dk, rules, targets=None, dtargets=None, n=4, m=4):
    """(F, C) as Duals at coordinates X + t dX and wavenumber k + t dk, rule of pair
    (i, e) = rules[i, e].  targets None: collocation at the mesh sources, which move with
    the coordinates (mesh sources must be its coordinate nodes, physics "geometry")."""
    Xd = Dual(np.asarray(X, dtype=np.float64), np.asarray(dX, dtype=np.float64))
    if targets is None:
        T = Xd
    else:
        T = Dual(np.asarray(targets, dtype=np.float64),
                 np.zeros_like(np.asarray(targets, dtype=np.float64)) if dtargets is None
                 else np.asarray(dtargets, dtype=np.float64))
    kd = Dual(np.asarray(complex(k)), np.asarray(complex(dk)))
    M, N = T.v.shape[1], mesh.n_sources
    R = 6 if mesh.geometry_kind == 6 else 3
    Fv = np.zeros((M, N), dtype=np.complex128)
    Fd = np.zeros((M, N), dtype=np.complex128)
    Cv = np.zeros(M, dtype=np.complex128)
    Cd = np.zeros(M, dtype=np.complex128)
    for c in range(1, R + 1):
        ii, ee = np.nonzero(rules == c)
        if ii.size == 0:
            continue
        u, v, w = oracle_np.rule(c, n, m)
        Ng, du, dv = oracle_np.basis(mesh.geometry_kind, u, v)
        phi, _, _ = oracle_np.basis(mesh.physics_kind, u, v)
        Xe = Xd[:, mesh.topology[:, ee]]                       # (3, ng, P)
        y = lin_map("dkp,qk->pqd", Xe, Ng)                     # (P, Q, 3)
        tu = lin_map("dkp,qk->pqd", Xe, du)
        tv = lin_map("dkp,qk->pqd", Xe, dv)
        cr = cross(tu, tv)
        J = dsqrt((cr * cr).sum(axis=2))
        nrm = -cr / Dual(J.v[..., None], J.d[..., None])
        x = T[:, ii]                                           # (3, P)
        xt = Dual(x.v.T[:, None, :], x.d.T[:, None, :])        # (P, 1, 3)
        d = y - xt
        r = dsqrt((d * d).sum(axis=2))
        nd = (nrm * d).sum(axis=2)
        jw = J * w[None, :]
        ikr = Dual(1j * kd.v * r.v, 1j * (kd.d * r.v + kd.v * r.d))
        ekr = dexp(ikr)
        f = ekr * (ikr - 1.0) * nd / (4 * np.pi * r ** 3) * jw
        k0 = -nd / (4 * np.pi * r ** 3) * jw
        fp = f @ phi                                           # (P, np)
        cols = mesh.physics_topology[:, ee].T
        np.add.at(Fv, (ii[:, None], cols), fp.v)
        np.add.at(Fd, (ii[:, None], cols), fp.d)
        ks = k0.sum(axis=1)
        np.add.at(Cv, ii, ks.v)
        np.add.at(Cd, ii, ks.d)
    return Dual(Fv, Fd), Dual(Cv, Cd)
