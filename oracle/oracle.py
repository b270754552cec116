"""TEST INFRASTRUCTURE — ctypes front end of the CPU oracle (oracle/liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module, and only to compare results with.  The oracle restates the reference's
assemble_parallel! (src/3d/assembly_collocation.jl:259-479); see bie_oracle.cpp for the
per-function citations.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i8p = ctypes.POINTER(ctypes.c_int8)


def build(force: bool = False) -> str:
    """Compile liboracle.so from the sources beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_julia_hypot3.restype = ctypes.c_double
        L.orc_julia_hypot3.argtypes = [ctypes.c_double] * 3
        L.orc_gauss_legendre.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_rule.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.orc_basis.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_element_geometry.argtypes = [_dp, _i64p, ctypes.c_int, ctypes.c_int64,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, _dp, _dp, _dp]
        L.orc_select.argtypes = [_dp, ctypes.c_int64, _i64p, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int64, ctypes.c_int, _dp, ctypes.c_int64, _i8p]
        L.orc_assemble.argtypes = [_dp, _i64p, ctypes.c_int, _dp, ctypes.c_int64, _i64p,
                                   ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int64,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                   ctypes.c_int]
        L.orc_assemble_taylor.argtypes = [_dp, _i64p, ctypes.c_int, _dp, ctypes.c_int64,
                                          _i64p, ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_double, _dp,
                                          ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.orc_assemble_frozen.argtypes = [_dp, _i64p, ctypes.c_int, _dp, ctypes.c_int64, _i64p,
                                          ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_double, _dp,
                                          ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, _i8p, _dp, _dp, _dp, ctypes.c_int]
        L.orc_select_diff.restype = ctypes.c_int64
        L.orc_select_diff.argtypes = [_dp, ctypes.c_int64, _i64p, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int64, ctypes.c_int, _dp, ctypes.c_int64, _i8p]
        L.orc_hypot3_scaled.restype = ctypes.c_double
        L.orc_hypot3_scaled.argtypes = [ctypes.c_double] * 3
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_subdivision_rule.argtypes = [_dp, ctypes.c_double, _dp, _dp, _dp, ctypes.c_int]
        L.orc_connected_sources.argtypes = [_i64p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                            ctypes.c_int, ctypes.c_int, _i64p, _i64p, _i64p,
                                            _i64p]
        L.orc_sparse_assemble.argtypes = [_dp, _i64p, ctypes.c_int, _dp, _i64p, ctypes.c_int,
                                          ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp]
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def julia_hypot3(x, y, z) -> float:
    return lib().orc_julia_hypot3(float(x), float(y), float(z))


def gauss_legendre(n: int):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, _p(x), _p(w))
    return x, w


def rule(c: int, n: int, m: int):
    Q = lib().orc_rule(c, n, m, None, None, None)
    u, v, w = np.zeros(Q), np.zeros(Q), np.zeros(Q)
    lib().orc_rule(c, n, m, _p(u), _p(v), _p(w))
    return u, v, w


def basis(kind: int, u, v):
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    N = np.zeros((u.size, kind))
    du = np.zeros((u.size, kind))
    dv = np.zeros((u.size, kind))
    lib().orc_basis(kind, u.size, _p(u), _p(v), _p(N), _p(du), _p(dv))
    return N, du, dv


def _mesh_args(mesh):
    if mesh.physics_kind not in (3, 6):
        raise ValueError("the C restatement covers continuous physics; discontinuous meshes "
                         "are checked with oracle_np")
    coords = np.asfortranarray(mesh.coordinates)
    topo = np.asfortranarray(mesh.topology)
    src = np.asfortranarray(mesh.sources)
    ptop = np.asfortranarray(mesh.physics_topology)
    return coords, topo, src, ptop


def element_geometry(mesh, c: int, n: int, m: int):
    coords, topo, _, _ = _mesh_args(mesh)
    Q = lib().orc_rule(c, n, m, None, None, None)
    E = mesh.n_elements
    y = np.zeros((E, Q, 3))
    nr = np.zeros((E, Q, 3))
    jw = np.zeros((E, Q))
    lib().orc_element_geometry(_p(coords), _p(topo, _i64p), mesh.geometry_kind, E, 0, c, n,
                               m, _p(y), _p(nr), _p(jw))
    return y, nr, jw


def select(mesh, targets=None, hypot="julia1.9"):
    """Rule index (1..R) per (target, element) as int8 [M, E]; hypot "scaled" takes the
    distances of orc_hypot3_scaled instead of orc_julia_hypot3."""
    _, _, src, ptop = _mesh_args(mesh)
    tg = np.asfortranarray(mesh.sources if targets is None else targets, dtype=np.float64)
    M = tg.shape[1]
    out = np.zeros((M, mesh.n_elements), dtype=np.int8)
    lib().orc_select(_p(src), mesh.n_sources, _p(ptop, _i64p), mesh.physics_kind,
                     mesh.geometry_kind, mesh.n_elements, 0, _p(tg), M, _p(out, _i8p))
    if hypot == "scaled":
        _, alt = select_diff(mesh, targets, want_rules=True)
        out = np.where(alt != 0, alt, out)
    elif hypot != "julia1.9":
        raise ValueError("hypot must be 'julia1.9' or 'scaled'")
    return out


def select_diff(mesh, targets=None, want_rules=False):
    """Pairs whose rule differs between the two hypot formulas (orc_select_diff): the count,
    and with want_rules the int8 [M, E] array holding the other formula's rule where they
    differ (0 elsewhere)."""
    _, _, src, ptop = _mesh_args(mesh)
    tg = np.asfortranarray(mesh.sources if targets is None else targets, dtype=np.float64)
    M = tg.shape[1]
    out = np.zeros((M, mesh.n_elements), dtype=np.int8) if want_rules else None
    n = lib().orc_select_diff(_p(src), mesh.n_sources, _p(ptop, _i64p), mesh.physics_kind,
                              mesh.geometry_kind, mesh.n_elements, 0, _p(tg), M,
                              _p(out, _i8p) if want_rules else None)
    return (int(n), out) if want_rules else int(n)


def assemble(mesh, k, targets=None, n=4, m=4, fOn=True, gOn=True, cOn=True, nthreads=0):
    """(F, G, C) as complex128, F and G Fortran-ordered M x N (the reference's layout)."""
    coords, topo, src, ptop = _mesh_args(mesh)
    tg = np.asfortranarray(mesh.sources if targets is None else targets, dtype=np.float64)
    M, N = tg.shape[1], mesh.n_sources
    F = np.zeros((M, N), dtype=np.complex128, order="F")
    G = np.zeros((M, N), dtype=np.complex128, order="F")
    C = np.zeros(M, dtype=np.complex128)
    k = complex(k)
    flags = (1 if fOn else 0) | (2 if gOn else 0) | (4 if cOn else 0)
    rc = lib().orc_assemble(_p(coords), _p(topo, _i64p), mesh.geometry_kind, _p(src), N,
                            _p(ptop, _i64p), mesh.physics_kind, mesh.n_elements, 0, k.real,
                            k.imag, _p(tg), M, n, m, flags, _p(F), _p(G), _p(C), nthreads)
    if rc != 0:
        raise ValueError(f"oracle assemble failed ({rc})")
    return F, G, C


def assemble_taylor(mesh, k, order, targets=None, n=4, m=4, nthreads=0):
    """d^order/dk^order of (F, G) at k (the reference's taylor_assemble! kernels)."""
    coords, topo, src, ptop = _mesh_args(mesh)
    tg = np.asfortranarray(mesh.sources if targets is None else targets, dtype=np.float64)
    M, N = tg.shape[1], mesh.n_sources
    F = np.zeros((M, N), dtype=np.complex128, order="F")
    G = np.zeros((M, N), dtype=np.complex128, order="F")
    k = complex(k)
    rc = lib().orc_assemble_taylor(_p(coords), _p(topo, _i64p), mesh.geometry_kind, _p(src),
                                   N, _p(ptop, _i64p), mesh.physics_kind, mesh.n_elements, 0,
                                   k.real, k.imag, _p(tg), M, n, m, order, _p(F), _p(G),
                                   nthreads)
    if rc != 0:
        raise ValueError(f"oracle taylor assemble failed ({rc})")
    return F, G


def assemble_frozen(mesh, k, rules, targets=None, n=4, m=4, fOn=True, gOn=True, cOn=True,
                    nthreads=0):
    """Assembly with the rule index of every (target, element) pair given (int8 [M, E],
    values 1..R): a smooth function of coordinates and k for finite differences."""
    coords, topo, src, ptop = _mesh_args(mesh)
    tg = np.asfortranarray(mesh.sources if targets is None else targets, dtype=np.float64)
    M, N = tg.shape[1], mesh.n_sources
    rules = np.ascontiguousarray(rules, dtype=np.int8)
    assert rules.shape == (M, mesh.n_elements)
    F = np.zeros((M, N), dtype=np.complex128, order="F")
    G = np.zeros((M, N), dtype=np.complex128, order="F")
    C = np.zeros(M, dtype=np.complex128)
    k = complex(k)
    flags = (1 if fOn else 0) | (2 if gOn else 0) | (4 if cOn else 0)
    rc = lib().orc_assemble_frozen(_p(coords), _p(topo, _i64p), mesh.geometry_kind, _p(src), N,
                                   _p(ptop, _i64p), mesh.physics_kind, mesh.n_elements, 0,
                                   k.real, k.imag, _p(tg), M, n, m, flags, _p(rules, _i8p),
                                   _p(F), _p(G), _p(C), nthreads)
    if rc != 0:
        raise ValueError(f"oracle frozen assemble failed ({rc})")
    return F, G, C


def max_threads() -> int:
    return lib().orc_max_threads()


# ---- sparse near field (sparse_oracle.cpp) ---------------------------------------------

def subdivision_rule(src, tol=1e-6):
    """(u, v, w) of singular_triangle_integration(., 3, src, I, tol)."""
    src = np.ascontiguousarray(src, dtype=np.float64)
    Q = lib().orc_subdivision_rule(_p(src), tol, None, None, None, 0)
    u, v, w = np.zeros(Q), np.zeros(Q), np.zeros(Q)
    lib().orc_subdivision_rule(_p(src), tol, _p(u), _p(v), _p(w), Q)
    return u, v, w


def connected_sources(mesh, depth=1):
    """connected_sources(mesh, depth): (element lists, source lists), 0-based, sorted."""
    ptop = np.asfortranarray(mesh.physics_topology)
    N, E = mesh.n_sources, mesh.n_elements
    rp, ep = np.zeros(N + 1, dtype=np.int64), np.zeros(N + 1, dtype=np.int64)
    if lib().orc_connected_sources(_p(ptop, _i64p), mesh.physics_kind, E, N, 0, depth, _p(rp, _i64p),
                                   None, _p(ep, _i64p), None) != 0:
        raise ValueError("depth must be 1 or 2")
    cols, els = np.zeros(rp[-1], dtype=np.int64), np.zeros(ep[-1], dtype=np.int64)
    lib().orc_connected_sources(_p(ptop, _i64p), mesh.physics_kind, E, N, 0, depth, _p(rp, _i64p),
                                _p(cols, _i64p), _p(ep, _i64p), _p(els, _i64p))
    return ([els[ep[i]:ep[i + 1]] for i in range(N)], [cols[rp[i]:rp[i + 1]] for i in range(N)])


def sparse_assemble(mesh, k, sources=None, depth=1, offset=0.0, fOn=True, gOn=True,
                    hypot="julia1.9"):
    """sparse_assemble_parallel! as (I, J, F, G): 0-based rows and columns of the stored
    entries in the reference's order, complex128 values."""
    coords, topo, src, ptop = _mesh_args(mesh)
    pts = np.asfortranarray(src if sources is None else sources, dtype=np.float64)
    N, E = mesh.n_sources, mesh.n_elements
    if pts.shape != (3, N):
        raise ValueError("one source point per mesh source")
    _, cols = connected_sources(mesh, depth)
    nnz = sum(len(c) for c in cols)
    F = np.zeros(nnz, dtype=np.complex128)
    G = np.zeros(nnz, dtype=np.complex128)
    k = complex(k)
    fg = (1 if fOn else 0) | (2 if gOn else 0)
    variant = {"julia1.9": 0, "scaled": 1}[hypot]
    rc = lib().orc_sparse_assemble(_p(coords), _p(topo, _i64p), mesh.geometry_kind, _p(pts),
                                   _p(ptop, _i64p), mesh.physics_kind, E, N, 0, k.real, k.imag,
                                   depth, float(offset), fg, variant, _p(F), _p(G))
    if rc != 0:
        raise ValueError(f"oracle sparse assemble failed ({rc})")
    I = np.concatenate([np.full(len(c), i, dtype=np.int64) for i, c in enumerate(cols)])
    J = np.concatenate(cols) if nnz else np.zeros(0, dtype=np.int64)
    return I, J, F, G
