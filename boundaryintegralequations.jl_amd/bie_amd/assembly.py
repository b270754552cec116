"""Host-side mirror of the reference's assembly entry point, backed by the HIP library.

``assemble_parallel(mesh, k, in_sources; fOn, gOn, cOn, sparse, m=4, n=4, progress)``
mirrors ``assemble_parallel!`` (src/3d/assembly_collocation.jl:259-268): same keyword
names, defaults and argument meaning, returning ``(F, G, C)`` with F, G complex128
M x N in column-major order and C complex128 of length M.  Errors raise
``ValueError``/``NotImplementedError`` where the reference throws ``ArgumentError`` or
has no method.

``DeviceMesh`` is the device-resident form (bie_dmesh) for callers that keep F, G, C
in HBM: it takes raw device pointers (e.g. ``torch.Tensor.data_ptr()``) and a HIP
stream handle.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .mesh import Mesh3d


def _flags(fOn, gOn, cOn):
    return (_lib.BIE_F if fOn else 0) | (_lib.BIE_G if gOn else 0) | (_lib.BIE_C if cOn else 0)


def assemble_parallel(mesh: Mesh3d, k, in_sources=None, *, fOn=True, gOn=True, cOn=True,
                      sparse=False, m=4, n=4, progress=True, depth=2, offset=None,
                      hypot="julia1.9"):
    """Dense F, G, C on linear/quadratic triangles (drop-in for assemble_parallel!).

    ``hypot`` (an addition to the reference's keywords) names the n-ary hypot whose
    rounding decides the rule of exactly tied (target, element) pairs: "julia1.9" (the
    reference's CI Julia) or "scaled" (later Julia releases)."""
    if sparse:  # (F, G) of the near field, as the reference returns them
        return sparse_assemble_parallel(mesh, k, in_sources, fOn=fOn, gOn=gOn, depth=depth,
                                        offset=offset, hypot=hypot)
    del progress, depth, offset  # accepted for signature parity; unused by dense assembly
    hm = _lib.HostMesh(mesh, _lib.hypot_flag(hypot))
    if in_sources is None or in_sources is mesh.sources:
        targets = None
        M = mesh.n_sources
    else:
        targets = np.asfortranarray(in_sources, dtype=np.float64)
        if targets.ndim != 2 or targets.shape[0] != 3:
            raise ValueError("in_sources must be a 3 x M array")
        M = targets.shape[1]
    N = mesh.n_sources
    F = np.empty((M, N), dtype=np.complex128, order="F")
    G = np.empty((M, N), dtype=np.complex128, order="F")
    C = np.empty(M, dtype=np.complex128)
    k = complex(k)
    rc = _lib.lib().bie_assemble(ctypes.byref(hm.struct), k.real, k.imag,
                                 _lib._ptr(targets), M, int(n), int(m), _flags(fOn, gOn, cOn),
                                 _lib._ptr(F), _lib._ptr(G), _lib._ptr(C))
    if rc == _lib.BIE_EINVAL:
        raise ValueError(_lib.lib().bie_last_error().decode())
    if rc == _lib.BIE_EUNSUPPORTED:
        raise NotImplementedError(_lib.lib().bie_last_error().decode())
    _lib.check(rc)
    return F, G, C


class Graph:
    """An instantiated HIP graph (bie_graph); launch() replays it on a stream."""

    def __init__(self, handle):
        self.handle = handle

    def launch(self, stream: int = 0):
        _lib.check(_lib.lib().bie_graph_launch(self.handle, ctypes.c_void_p(stream or None)))

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().bie_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceMesh:
    """A mesh resident on one GPU (bie_dmesh), with its cached rule tables."""

    def __init__(self, mesh: Mesh3d, device: int = 0, hypot: str = "julia1.9"):
        self.mesh = mesh
        hm = _lib.HostMesh(mesh, _lib.hypot_flag(hypot))
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().bie_dmesh_create(ctypes.byref(hm.struct), int(device),
                                               ctypes.byref(h)))
        self.handle = h
        self.device = int(device)
        self.n_sources = mesh.n_sources
        self.n_elements = mesh.n_elements

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().bie_dmesh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def sources_ptr(self) -> int:
        return _lib.lib().bie_dmesh_sources(self.handle)

    def assemble(self, k, F_ptr: int, G_ptr: int, C_ptr: int, *, targets_ptr: int = 0,
                 n_targets: int | None = None, n=4, m=4, fOn=True, gOn=True, cOn=True,
                 stream: int = 0):
        """Asynchronous device assembly into caller-owned device buffers."""
        k = complex(k)
        M = self.n_sources if n_targets is None else int(n_targets)
        _lib.check(_lib.lib().bie_assemble_device(
            self.handle, k.real, k.imag, ctypes.c_void_p(targets_ptr or None), M, int(n),
            int(m), _flags(fOn, gOn, cOn), ctypes.c_void_p(F_ptr), ctypes.c_void_p(G_ptr),
            ctypes.c_void_p(C_ptr), ctypes.c_void_p(stream or None)))

    def select_rules(self, rules_ptr: int, *, targets_ptr: int = 0, n_targets=None,
                     stream: int = 0):
        M = self.n_sources if n_targets is None else int(n_targets)
        _lib.check(_lib.lib().bie_select_rules_device(
            self.handle, ctypes.c_void_p(targets_ptr or None), M, ctypes.c_void_p(rules_ptr),
            ctypes.c_void_p(stream or None)))

    def assemble_sweep(self, k_ptr: int, nk: int, F_ptr: int, G_ptr: int, stride: int,
                       C_ptr: int, *, targets_ptr: int = 0, n_targets: int | None = None,
                       n=4, m=4, fOn=True, gOn=True, cOn=True, stream: int = 0):
        """F_q, G_q for nk real wavenumbers (device array at k_ptr); F_q at
        F_ptr + 16 * q * stride bytes.  Asynchronous."""
        M = self.n_sources if n_targets is None else int(n_targets)
        _lib.check(_lib.lib().bie_assemble_sweep_device(
            self.handle, ctypes.c_void_p(k_ptr), int(nk), ctypes.c_void_p(targets_ptr or None),
            M, int(n), int(m), _flags(fOn, gOn, cOn), ctypes.c_void_p(F_ptr),
            ctypes.c_void_p(G_ptr), int(stride), ctypes.c_void_p(C_ptr),
            ctypes.c_void_p(stream or None)))

    def sweep_tune(self, k_ptr: int, nk: int, F_ptr: int, G_ptr: int, stride: int, C_ptr: int,
                   *, targets_ptr: int = 0, n_targets: int | None = None, n=4, m=4, fOn=True,
                   gOn=True, cOn=True, stream: int = 0):
        """Tile-shape autotuner of the sweep (bie_sweep_tune): times every candidate
        (elements per tile x LDS slots), keeps the fastest for later sweeps on this mesh;
        returns (elems_per_tile, slots, ms) with 0 meaning whole surface / as many as fit."""
        M = self.n_sources if n_targets is None else int(n_targets)
        e, sl, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        _lib.check(_lib.lib().bie_sweep_tune(
            self.handle, ctypes.c_void_p(k_ptr), int(nk), ctypes.c_void_p(targets_ptr or None),
            M, int(n), int(m), _flags(fOn, gOn, cOn), ctypes.c_void_p(F_ptr),
            ctypes.c_void_p(G_ptr), int(stride), ctypes.c_void_p(C_ptr),
            ctypes.c_void_p(stream or None), ctypes.byref(e), ctypes.byref(sl), ctypes.byref(ms)))
        return e.value, sl.value, ms.value

    def sweep_graph(self, k_ptr: int, nk: int, F_ptr: int, G_ptr: int, stride: int, C_ptr: int,
                    *, targets_ptr: int = 0, n_targets: int | None = None, n=4, m=4,
                    fOn=True, gOn=True, cOn=True) -> "Graph":
        """The same sweep captured into a HIP graph."""
        M = self.n_sources if n_targets is None else int(n_targets)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().bie_sweep_graph_create(
            self.handle, ctypes.c_void_p(k_ptr), int(nk), ctypes.c_void_p(targets_ptr or None),
            M, int(n), int(m), _flags(fOn, gOn, cOn), ctypes.c_void_p(F_ptr),
            ctypes.c_void_p(G_ptr), int(stride), ctypes.c_void_p(C_ptr), ctypes.byref(h)))
        return Graph(h)

    def element_geometry(self, rule: int, n: int, m: int):
        from . import rules as _r
        Q = _r.rule_size(rule, n, m)
        E = self.n_elements
        y = np.zeros((E, Q, 3))
        nr = np.zeros((E, Q, 3))
        jw = np.zeros((E, Q))
        _lib.check(_lib.lib().bie_element_geometry(self.handle, rule, n, m, _lib._ptr(y),
                                                   _lib._ptr(nr), _lib._ptr(jw)))
        return y, nr, jw


class Batch:
    """An uneven batch of device meshes assembled in one persistent launch (bie_batch_*).

    Every mesh shares element kinds and the rule set (n, m); instance i is assembled at its
    own wavenumber with collocation at its own nodes, into caller-owned device buffers
    (F_i, G_i: N_i x N_i column-major complex128; C_i: N_i)."""

    def __init__(self, dmeshes, *, n=4, m=4):
        self.dmeshes = list(dmeshes)           # keeps the meshes alive while the batch is
        arr = (ctypes.c_void_p * len(self.dmeshes))(*[d.handle.value for d in self.dmeshes])
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().bie_batch_create(arr, len(self.dmeshes), int(n), int(m),
                                               ctypes.byref(h)))
        self.handle = h
        self.n, self.m = n, m

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().bie_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.dmeshes)

    def configure(self, elems_per_tile=64, slots=64, order=0):
        _lib.check(_lib.lib().bie_batch_configure(self.handle, int(elems_per_tile), int(slots),
                                                  int(order)))

    def info(self):
        t, s, o = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        nt = ctypes.c_int64()
        _lib.check(_lib.lib().bie_batch_info(self.handle, ctypes.byref(t), ctypes.byref(s),
                                             ctypes.byref(o), ctypes.byref(nt)))
        return dict(elems_per_tile=t.value, slots=s.value, order=o.value, n_tiles=nt.value)

    def _args(self, ks, F_ptrs, G_ptrs, C_ptrs):
        n = len(self.dmeshes)
        ks = [complex(k) for k in ks]
        if not (len(ks) == len(F_ptrs) == len(G_ptrs) == len(C_ptrs) == n):
            raise ValueError("one k and one F, G, C pointer per instance")
        kr = (ctypes.c_double * n)(*[k.real for k in ks])
        ki = (ctypes.c_double * n)(*[k.imag for k in ks])
        P = ctypes.c_void_p * n
        return kr, ki, P(*[p or None for p in F_ptrs]), P(*[p or None for p in G_ptrs]), \
            P(*[p or None for p in C_ptrs])

    def assemble(self, ks, F_ptrs, G_ptrs, C_ptrs, *, fOn=True, gOn=True, cOn=True,
                 stream: int = 0):
        """Asynchronous batched assembly."""
        kr, ki, F, G, C = self._args(ks, F_ptrs, G_ptrs, C_ptrs)
        _lib.check(_lib.lib().bie_batch_assemble(self.handle, kr, ki, _flags(fOn, gOn, cOn), F, G,
                                                 C, ctypes.c_void_p(stream or None)))

    def tune(self, ks, F_ptrs, G_ptrs, C_ptrs, *, fOn=True, gOn=True, cOn=True,
             stream: int = 0) -> float:
        """Times every schedule candidate, keeps the fastest; returns its time in ms."""
        kr, ki, F, G, C = self._args(ks, F_ptrs, G_ptrs, C_ptrs)
        best = ctypes.c_double()
        _lib.check(_lib.lib().bie_batch_tune(self.handle, kr, ki, _flags(fOn, gOn, cOn), F, G, C,
                                             ctypes.c_void_p(stream or None), ctypes.byref(best)))
        return best.value

    def graph(self, ks, F_ptrs, G_ptrs, C_ptrs, *, fOn=True, gOn=True, cOn=True) -> Graph:
        """The batched assembly with these arguments captured into a HIP graph
        (bie_batch_graph_create); Graph.launch(stream) replays it."""
        kr, ki, F, G, C = self._args(ks, F_ptrs, G_ptrs, C_ptrs)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().bie_batch_graph_create(self.handle, kr, ki, _flags(fOn, gOn, cOn),
                                                     F, G, C, ctypes.byref(h)))
        return Graph(h)

    def dfdk(self, ks, alpha_ptrs, p_ptrs, w_ptr: int, *, stream: int = 0):
        """w_i = alpha_i^T (dF_i/dk) p_i for every instance in one launch (bie_batch_dfdk_device):
        alpha_i, p_i device arrays of N_i complex128, w a device array of len(batch) complex128.
        Returns when the launch has finished."""
        n = len(self.dmeshes)
        ks = [complex(k) for k in ks]
        if not (len(ks) == len(alpha_ptrs) == len(p_ptrs) == n):
            raise ValueError("one k and one alpha, p pointer per instance")
        kr = (ctypes.c_double * n)(*[k.real for k in ks])
        ki = (ctypes.c_double * n)(*[k.imag for k in ks])
        P = ctypes.c_void_p * n
        _lib.check(_lib.lib().bie_batch_dfdk_device(
            self.handle, kr, ki, P(*alpha_ptrs), P(*p_ptrs), ctypes.c_void_p(w_ptr),
            ctypes.c_void_p(stream or None)))


# ---- sparse near field -------------------------------------------------------------------

def _sparse_matrices(I, J, F, G, n, fOn, gOn):
    import scipy.sparse as sps  # the reference returns SparseMatrixCSC; scipy is its analogue
    shape = (n, n)
    Fm = sps.csr_matrix((F, (I, J)), shape=shape)
    Gm = sps.csr_matrix((G, (I, J)), shape=shape)
    return Fm, Gm


def sparse_assemble_parallel(mesh: Mesh3d, k, sources=None, *, fOn=True, gOn=True, depth=1,
                             offset=None, progress=True, Ngauss=40, hypot="julia1.9"):
    """Near-field F and G (drop-in for sparse_assemble_parallel!,
    src/3d/assembly_collocation_losses.jl:200-405): the entries (i, j) of the sources
    connected to node i through at most ``depth`` elements (connected_sources, :41-132),
    each element integrated with the rule clustered at its node nearest to source i: the
    subdivision rule for continuous physics, the polar rule (create_polar_elements,
    :409-447, ``Ngauss``; 40 for quadratic physics) around the physics node for
    discontinuous physics.  Returns scipy CSR matrices n_sources x n_sources holding
    exactly the reference's stored entries (a disabled output is all zeros, as in the
    reference).  ``offset``: the subdivision rules' source offset (None: 0)."""
    del progress
    if depth not in (1, 2):
        raise ValueError("depth must be 1 or 2")
    N = mesh.n_sources
    if sources is None or sources is mesh.sources:
        pts = None
    else:
        pts = np.asfortranarray(sources, dtype=np.float64)
        if pts.shape != (3, N):
            raise ValueError("sources must be 3 x n_sources (one point per mesh source)")
    hm = _lib.HostMesh(mesh, _lib.hypot_flag(hypot))
    L = _lib.lib()
    nnz = ctypes.c_int64()
    _lib.check(L.bie_sparse_count(ctypes.byref(hm.struct), int(depth), ctypes.byref(nnz)))
    n = nnz.value
    I = np.empty(n, dtype=np.int64)
    J = np.empty(n, dtype=np.int64)
    F = np.empty(n, dtype=np.complex128)
    G = np.empty(n, dtype=np.complex128)
    k = complex(k)
    off = float("nan") if offset is None else float(offset)
    rc = L.bie_sparse_assemble_ex(ctypes.byref(hm.struct), k.real, k.imag, _lib._ptr(pts),
                                  N if pts is not None else 0, int(depth), off, int(Ngauss),
                                  _flags(fOn, gOn, False), n, _lib._ptr(I), _lib._ptr(J),
                                  _lib._ptr(F), _lib._ptr(G))
    if rc == _lib.BIE_EINVAL:
        raise ValueError(L.bie_last_error().decode())
    _lib.check(rc)
    base = hm.struct.index_base
    return _sparse_matrices(I - base, J - base, F, G, N, fOn, gOn)


def partial_assemble_parallel(mesh: Mesh3d, k, in_sources=None, *, fOn=True, gOn=True, n=3,
                              progress=False, depth=1, hypot="julia1.9"):
    """The partial near field of a fast operator (drop-in for partial_assemble_parallel!,
    src/3d/fast_multipole_method.jl:11-81): the pattern of sparse_assemble_parallel, every
    connected (source, element) pair integrated with the one symmetric rule
    gauss_points_triangle(n) (n = 1, 3, 4, 7, 12, 13).  Returns scipy CSR matrices."""
    del progress
    if depth not in (1, 2):
        raise ValueError("depth must be 1 or 2")
    N = mesh.n_sources
    if in_sources is None or in_sources is mesh.sources:
        pts = None
    else:
        pts = np.asfortranarray(in_sources, dtype=np.float64)
        if pts.shape != (3, N):
            raise ValueError("in_sources must be 3 x n_sources (one point per mesh source)")
    hm = _lib.HostMesh(mesh, _lib.hypot_flag(hypot))
    L = _lib.lib()
    nnz = ctypes.c_int64()
    _lib.check(L.bie_sparse_count(ctypes.byref(hm.struct), int(depth), ctypes.byref(nnz)))
    m = nnz.value
    I = np.empty(m, dtype=np.int64)
    J = np.empty(m, dtype=np.int64)
    F = np.empty(m, dtype=np.complex128)
    G = np.empty(m, dtype=np.complex128)
    k = complex(k)
    rc = L.bie_partial_assemble(ctypes.byref(hm.struct), k.real, k.imag, _lib._ptr(pts),
                                N if pts is not None else 0, int(depth), int(n),
                                _flags(fOn, gOn, False), m, _lib._ptr(I), _lib._ptr(J),
                                _lib._ptr(F), _lib._ptr(G))
    if rc == _lib.BIE_EINVAL:
        raise ValueError(L.bie_last_error().decode())
    _lib.check(rc)
    base = hm.struct.index_base
    return _sparse_matrices(I - base, J - base, F, G, N, fOn, gOn)


class SparsePattern:
    """Device-resident sparse near-field assembly (bie_sparse) on a DeviceMesh: pattern,
    pairs and rules (subdivision, or polar for discontinuous physics) built once, then
    assemblies for any k."""

    def __init__(self, dmesh: DeviceMesh, depth=1, offset=None, Ngauss=40):
        h = ctypes.c_void_p()
        off = float("nan") if offset is None else float(offset)
        _lib.check(_lib.lib().bie_sparse_create_ex(dmesh.handle, int(depth), off, int(Ngauss),
                                                   ctypes.byref(h)))
        self.handle = h
        self.dmesh = dmesh  # the pattern must not outlive its mesh
        self.nnz = int(_lib.lib().bie_sparse_nnz(h))

    def pattern(self, index_base=0):
        I = np.empty(self.nnz, dtype=np.int64)
        J = np.empty(self.nnz, dtype=np.int64)
        _lib.check(_lib.lib().bie_sparse_pattern(self.handle, int(index_base), _lib._ptr(I),
                                                 _lib._ptr(J)))
        return I, J

    def assemble_device(self, k, F_ptr: int, G_ptr: int, *, sources_ptr: int = 0, fOn=True,
                        gOn=True, stream: int = 0):
        """Asynchronous assembly into device arrays of nnz complex128 entries."""
        k = complex(k)
        _lib.check(_lib.lib().bie_sparse_assemble_device(
            self.handle, k.real, k.imag, ctypes.c_void_p(sources_ptr or None),
            _flags(fOn, gOn, False), ctypes.c_void_p(F_ptr or None),
            ctypes.c_void_p(G_ptr or None), ctypes.c_void_p(stream or None)))

    def partial_device(self, k, n_gauss, F_ptr: int, G_ptr: int, *, sources_ptr: int = 0,
                       fOn=True, gOn=True, stream: int = 0):
        """The partial near field (gauss_points_triangle(n_gauss) on every pair) into
        device arrays of nnz complex entries.  Asynchronous."""
        k = complex(k)
        _lib.check(_lib.lib().bie_sparse_partial_device(
            self.handle, k.real, k.imag, int(n_gauss), ctypes.c_void_p(sources_ptr or None),
            _flags(fOn, gOn, False), ctypes.c_void_p(F_ptr or None),
            ctypes.c_void_p(G_ptr or None), ctypes.c_void_p(stream or None)))

    def assemble(self, k, *, fOn=True, gOn=True):
        """Host copies of the nnz values of F and G (pattern order)."""
        dF = _lib.DeviceBuffer.empty(16 * self.nnz)
        dG = _lib.DeviceBuffer.empty(16 * self.nnz)
        self.assemble_device(k, dF.ptr, dG.ptr, fOn=fOn, gOn=gOn)
        _lib.check(_lib.lib().bie_device_synchronize())
        return (dF.get(np.empty(self.nnz, dtype=np.complex128)),
                dG.get(np.empty(self.nnz, dtype=np.complex128)))

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().bie_sparse_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
