"""Triangular surface meshes in the layout of the reference's ``Mesh3d``.

Mirrors ``Mesh3d`` (src/utils/mesh.jl:23-34): ``coordinates`` 3 x Nc, ``topology``
n_g x E, ``sources`` 3 x N, ``physics_topology`` n_p x E, all column-major.  Indices
are stored 0-based here; the C-ABI takes an ``index_base`` so 1-based Julia arrays
pass through unchanged.

Generators:
  * ``icosphere(f, order)``: geodesic sphere, every icosahedron face split f x f,
    vertices projected to the sphere; 10 f^2 + 2 corner nodes, 20 f^2 triangles.
    Quadratic meshes add the 30 f^2 edge mid nodes (projected), corners first.
  * ``ellipsoid(f, axes, order)``: the same sphere scaled per axis.
  * ``load_comsol(path, geometry_order, physics_order)``: reader for the ``.mphtxt``
    files the reference's tests use (src/3d/read_comsol.jl:9-119), including the
    swap of local nodes 5 and 6 of 6-node triangles (read_comsol.jl:49-52).

Element node order follows the reference (src/3d/SurfaceFunction.jl:40-57): corners
1,2,3 then mid(1,2), mid(2,3), mid(3,1); (x2-x1) x (x3-x1) points out of the body,
as in the bundled COMSOL spheres, so the reference's negated normal
(src/3d/jacobian.jl:63-70) points into the body.
"""
from __future__ import annotations

import dataclasses
import os

import numpy as np

TRI3 = 3
TRI6 = 6
# discontinuous physics kinds (physics_kind only): low byte = nodes per element
DTRI1 = 0x101  # DiscontinuousTriangularConstant
DTRI3 = 0x103  # DiscontinuousTriangularLinear
DTRI6 = 0x106  # DiscontinuousTriangularQuadratic
_DISC_ORDERS = {"disctriconstant": DTRI1, "disctrilinear": DTRI3, "disctriquadratic": DTRI6}
# quadrilaterals (geometry and physics kinds; low byte = nodes per element)
QUAD4 = 0x204    # QuadrilateralLinear4 (COMSOL layout)
QUAD4S = 0x404   # QuadrilateralLinearSerendipity
QUAD8 = 0x208    # QuadrilateralQuadratic (8-node serendipity)
QUAD9 = 0x209    # QuadrilateralQuadraticLagrange (COMSOL layout)
DQUAD1 = 0x301   # DiscontinuousQuadrilateralConstant
DQUAD4 = 0x304   # DiscontinuousQuadrilateralLinear4
DQUAD9 = 0x309   # DiscontinuousQuadrilateralQuadraticLagrange
_DISC_QUAD_ORDERS = {"discquadconstant": DQUAD1, "discquadlinear": DQUAD4,
                     "discquadquadratic": DQUAD9}


def physics_nodes(kind: int) -> int:
    """Nodes per element of a physics kind (its low byte)."""
    return int(kind) & 0xFF


def geometry_nodes(kind: int) -> int:
    """Nodes per element of a geometry kind (its low byte)."""
    return int(kind) & 0xFF


def is_quad(kind: int) -> bool:
    return (int(kind) >> 9) != 0


@dataclasses.dataclass
class Mesh3d:
    """Data model of the reference's ``Mesh3d`` restricted to what assembly reads."""

    coordinates: np.ndarray        # (3, Nc) float64
    topology: np.ndarray           # (n_g, E) int64, 0-based
    sources: np.ndarray            # (3, N) float64, collocation nodes
    physics_topology: np.ndarray   # (n_p, E) int64, 0-based, indexes ``sources``
    geometry_kind: int             # TRI3 | TRI6
    physics_kind: int              # TRI3 | TRI6 | DTRI1 | DTRI3 | DTRI6
    physics_beta: float = 0.0      # node offset of DTRI3 / DTRI6

    def __post_init__(self):
        self.coordinates = np.ascontiguousarray(self.coordinates, dtype=np.float64)
        self.sources = np.ascontiguousarray(self.sources, dtype=np.float64)
        self.topology = np.ascontiguousarray(self.topology, dtype=np.int64)
        self.physics_topology = np.ascontiguousarray(self.physics_topology, dtype=np.int64)
        if self.coordinates.shape[0] != 3 or self.sources.shape[0] != 3:
            raise ValueError("coordinates and sources must be 3 x n")
        if self.topology.shape[0] != geometry_nodes(self.geometry_kind):
            raise ValueError("topology rows must equal the geometry node count")
        if self.physics_topology.shape[0] != physics_nodes(self.physics_kind):
            raise ValueError("physics_topology rows must equal the physics node count")
        if self.topology.shape[1] != self.physics_topology.shape[1]:
            raise ValueError("topology and physics_topology disagree on element count")

    @property
    def n_elements(self) -> int:
        return int(self.topology.shape[1])

    @property
    def n_sources(self) -> int:
        return int(self.sources.shape[1])

    @property
    def n_coordinates(self) -> int:
        return int(self.coordinates.shape[1])

    def fortran_views(self):
        """Arrays laid out column-major (Julia order) as flat contiguous buffers."""
        return (np.asfortranarray(self.coordinates), np.asfortranarray(self.topology),
                np.asfortranarray(self.sources), np.asfortranarray(self.physics_topology))


# ---------------------------------------------------------------------------------------
# Geodesic sphere
# ---------------------------------------------------------------------------------------
def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    # orient every face so (b-a) x (c-a) points outward
    for i, (a, b, c) in enumerate(f):
        n = np.cross(v[b] - v[a], v[c] - v[a])
        if np.dot(n, v[a] + v[b] + v[c]) < 0:
            f[i] = [a, c, b]
    return v, f


def _subdivide(f_level: int):
    """Split every icosahedron face into f^2 triangles; returns unprojected vertices."""
    v0, faces0 = _icosahedron()
    key_to_idx: dict = {}
    verts = []

    # Exact integer keys: a point on face (a,b,c) at barycentric (f-i-j, i, j)/f is keyed
    # by the sorted (vertex, weight) pairs, which is identical on shared edges.
    def vid(a, b, c, i, j):
        w = {}
        for node, wt in ((a, f_level - i - j), (b, i), (c, j)):
            if wt:
                w[node] = w.get(node, 0) + wt
        key = tuple(sorted(w.items()))
        idx = key_to_idx.get(key)
        if idx is None:
            idx = len(verts)
            key_to_idx[key] = idx
            p = sum(v0[node] * wt for node, wt in key) / f_level
            verts.append(p)
        return idx

    tris = []
    for a, b, c in faces0:
        grid = {}
        for i in range(f_level + 1):
            for j in range(f_level + 1 - i):
                grid[i, j] = vid(a, b, c, i, j)
        for i in range(f_level):
            for j in range(f_level - i):
                tris.append((grid[i, j], grid[i + 1, j], grid[i, j + 1]))
                if i + j + 2 <= f_level:
                    tris.append((grid[i + 1, j], grid[i + 1, j + 1], grid[i, j + 1]))
    return np.array(verts), np.array(tris, dtype=np.int64)


def _project(p, radius=1.0):
    return radius * p / np.linalg.norm(p, axis=-1, keepdims=True)


def icosphere(f: int, order: str = "linear", radius: float = 1.0,
              physics: str = "geometry", beta_type="legendre") -> Mesh3d:
    """Geodesic sphere mesh (nodes 10 f^2 + 2 linear, 40 f^2 + 2 quadratic)."""
    if f < 1:
        raise ValueError("f must be >= 1")
    verts, tris = _subdivide(f)
    corners = _project(verts, 1.0)
    # orientation: (x2-x1) x (x3-x1) outward
    a, b, c = corners[tris[:, 0]], corners[tris[:, 1]], corners[tris[:, 2]]
    flip = np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) < 0
    tris[flip] = tris[flip][:, [0, 2, 1]]
    if order == "linear":
        coords = corners * radius
        topo = tris.T.copy()
        gk = TRI3
    elif order == "quadratic":
        edge_id: dict = {}
        mids = []
        n_c = corners.shape[0]
        quad = np.empty((tris.shape[0], 6), dtype=np.int64)
        for e, (p, q, r) in enumerate(tris):
            quad[e, :3] = (p, q, r)
            for slot, (u, w) in zip((3, 4, 5), ((p, q), (q, r), (r, p))):
                key = (min(u, w), max(u, w))
                idx = edge_id.get(key)
                if idx is None:
                    idx = n_c + len(mids)
                    edge_id[key] = idx
                    mids.append(_project(0.5 * (corners[u] + corners[w]), 1.0))
                quad[e, slot] = idx
        coords = np.vstack([corners, np.array(mids)]) * radius
        topo = quad.T.copy()
        gk = TRI6
    else:
        raise ValueError("order must be 'linear' or 'quadratic'")
    return _with_physics(coords.T.copy(), topo, gk, physics, beta_type)


def ellipsoid(f: int, axes=(1.0, 1.0, 1.0), order: str = "linear",
              physics: str = "geometry", beta_type="legendre") -> Mesh3d:
    """Icosphere scaled by semi-axes (a, b, c); orientation is preserved."""
    m = icosphere(f, order, 1.0, physics="geometry")
    s = np.asarray(axes, dtype=np.float64).reshape(3, 1)
    return _with_physics(m.coordinates * s, m.topology, m.geometry_kind, physics, beta_type)


def discontinuous_beta(physics: str, beta_type="legendre") -> float:
    """Node offset beta of a discontinuous triangle (get_beta_tri_linear /
    get_beta_tri_quadratic, src/3d/mesh.jl:182-214): half the quadrilateral value, which is
    1 + (the smallest Gauss-Legendre node of 2 or 3 points) for "legendre", 1/2 or 1/3 for
    "equidistant", and the number itself when one is given (so a number is halved too)."""
    if physics in ("disctriconstant", "discquadconstant"):
        return 0.0
    quad_order = physics in ("discquadlinear", "discquadquadratic")
    if physics not in ("disctrilinear", "disctriquadratic") and not quad_order:
        raise ValueError("not a discontinuous physics order")
    lin = physics in ("disctrilinear", "discquadlinear")
    if isinstance(beta_type, (int, float)) and not isinstance(beta_type, bool):
        quad = float(beta_type)
    elif beta_type == "legendre":
        # gausslegendre(2)[1][1] = -1/sqrt(3), gausslegendre(3)[1][1] = -sqrt(3/5)
        quad = 1.0 + (-(1.0 / np.sqrt(3.0)) if lin else -np.sqrt(3.0 / 5.0))
    elif beta_type == "equidistant":
        quad = 0.5 if lin else 1.0 / 3.0
    else:
        raise ValueError("beta_type must be 'legendre', 'equidistant' or a number")
    # get_beta_quad_linear / _quadratic are used as they are on quadrilaterals
    return quad if quad_order else quad / 2


def _lagrange3(x):
    """1-D quadratic Lagrange polynomials on -1, 0, 1."""
    return [0.5 * x * (x - 1.0), (1.0 - x) * (1.0 + x), 0.5 * x * (x + 1.0)]


_QUAD9_IJ = [(0, 0), (2, 0), (0, 2), (2, 2), (1, 0), (0, 1), (1, 1), (2, 1), (1, 2)]


def geometry_basis(kind: int, u, v) -> np.ndarray:
    """Geometry basis (nodes x len(u)) of triangularLinear / triangularQuadratic and the
    quadrilaterals (src/3d/surface_functions.jl:1-95, 191-213)."""
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    if kind in (QUAD4, QUAD4S):
        su = [-1, 1, -1, 1] if kind == QUAD4 else [-1, 1, 1, -1]
        sv = [-1, -1, 1, 1]
        return np.stack([(1 + a * u) * (1 + b * v) / 4 for a, b in zip(su, sv)])
    if kind == QUAD8:
        su, sv = [-1, 1, 1, -1], [-1, -1, 1, 1]
        corners = [(1 + a * u) * (1 + b * v) * (a * u + b * v - 1) / 4 for a, b in zip(su, sv)]
        mids = [(1 - v) * (1 - u * u) / 2, (1 + u) * (1 - v * v) / 2, (1 + v) * (1 - u * u) / 2,
                (1 - u) * (1 - v * v) / 2]
        return np.stack(corners + mids)
    if kind == QUAD9:
        lu, lv = _lagrange3(u), _lagrange3(v)
        return np.stack([lu[i] * lv[j] for i, j in _QUAD9_IJ])
    if kind == TRI3:
        return np.stack([(1.0 - u) - v, u, v])
    w = (1.0 - v) - u
    return np.stack([w * ((1.0 - 2 * v) - 2 * u), u * (2 * u - 1.0), v * (2 * v - 1.0),
                     (4 * u) * w, (4 * u) * v, (4 * v) * w])


def discontinuous_nodes(kind: int, beta: float):
    """Local (u, v) of the nodes of a discontinuous element (get_nodal_nodes_u/_v,
    src/3d/SurfaceFunction.jl:429-470)."""
    b = float(beta)
    if kind == DQUAD1:
        return np.array([0.0]), np.array([0.0])
    if kind == DQUAD4:
        return (np.array([b - 1, 1 - b, b - 1, 1 - b]), np.array([b - 1, b - 1, 1 - b, 1 - b]))
    if kind == DQUAD9:
        lo, hi = b - 1, 1 - b
        return (np.array([lo, hi, lo, hi, 0.0, lo, 0.0, hi, 0.0]),
                np.array([lo, lo, hi, hi, lo, 0.0, 0.0, 0.0, hi]))
    if kind == DTRI1:
        return np.array([1 / 3]), np.array([1 / 3])
    if kind == DTRI3:
        return np.array([b, 1 - 2 * b, b]), np.array([b, b, 1 - 2 * b])
    if kind == DTRI6:
        h = (1 - b) / 2
        return (np.array([b, 1 - 2 * b, b, h, h, b]), np.array([b, b, 1 - 2 * b, b, h, h]))
    raise ValueError("not a discontinuous physics kind")


def discontinuous_sources(coords, topo, gk, kind, beta):
    """compute_sources (src/3d/mesh.jl:116-148): every element gets its own physics nodes,
    numbered element by element, at the geometry interpolated at the local nodes."""
    u, v = discontinuous_nodes(kind, beta)
    N = geometry_basis(gk, u, v)                             # (gk, np)
    X = np.asarray(coords)[:, np.asarray(topo)]              # (3, gk, E)
    src = np.einsum("dke,kp->dpe", X, N)                     # (3, np, E)
    npk, E = N.shape[1], X.shape[2]
    sources = src.reshape(3, npk * E, order="F")
    ptop = np.arange(npk * E, dtype=np.int64).reshape(npk, E, order="F")
    return np.ascontiguousarray(sources), ptop


def with_discontinuous_physics(mesh: "Mesh3d", physics: str, beta_type="legendre") -> "Mesh3d":
    """The same geometry with discontinuous physics ("disctriconstant", "disctrilinear",
    "disctriquadratic" on triangles, "discquadconstant", "discquadlinear",
    "discquadquadratic" on quadrilaterals; set_physics_element + compute_sources,
    src/3d/mesh.jl:116-180)."""
    orders = _DISC_QUAD_ORDERS if is_quad(mesh.geometry_kind) else _DISC_ORDERS
    kind = orders.get(physics)
    if kind is None:
        raise ValueError("physics must be " + ", ".join(orders) + " on this geometry")
    beta = discontinuous_beta(physics, beta_type)
    sources, ptop = discontinuous_sources(mesh.coordinates, mesh.topology, mesh.geometry_kind,
                                          kind, beta)
    return Mesh3d(mesh.coordinates, mesh.topology, sources, ptop, mesh.geometry_kind, kind,
                  beta)


def _with_physics(coords, topo, gk, physics, beta_type="legendre"):
    if physics == "geometry":
        return Mesh3d(coords, topo, coords.copy(), topo.copy(), gk, gk)
    if physics in _DISC_ORDERS:
        return with_discontinuous_physics(Mesh3d(coords, topo, coords.copy(), topo.copy(), gk,
                                                 gk), physics, beta_type)
    if physics == "linear":
        # read_comsol.jl:87-93: corner connectivity; sources are the used corner nodes.
        ptop = topo[:3, :]
        used = np.unique(ptop)
        remap = -np.ones(coords.shape[1], dtype=np.int64)
        remap[used] = np.arange(used.size)
        return Mesh3d(coords, topo, coords[:, used].copy(), remap[ptop], gk, TRI3)
    raise ValueError("physics must be 'geometry', 'linear' or a discontinuous triangular order")


# ---------------------------------------------------------------------------------------
# COMSOL .mphtxt reader
# ---------------------------------------------------------------------------------------
def read_comsol(path: str):
    """Return (coordinates 3xNc, topology n x E 0-based) as in read_comsol.jl:9-54."""
    if not path.endswith(".mphtxt"):
        path = path + ".mphtxt"
    coords, topo = [], []
    with open(path, "r") as fh:
        lines = fh.read().splitlines()
    i = 0
    while i < len(lines):
        line = lines[i]
        if line == "# Mesh vertex coordinates":
            i += 1
            while i < len(lines) and lines[i].strip():
                coords.append([float(x) for x in lines[i].split()])
                i += 1
        elif line == "# Elements":
            i += 1
            while i < len(lines) and lines[i].strip():
                topo.append([int(x) for x in lines[i].split()])
                i += 1
        i += 1
    coords = np.array(coords, dtype=np.float64).T.copy()
    topo = np.array(topo, dtype=np.int64).T.copy()
    if topo.shape[0] == 6:
        topo[[4, 5], :] = topo[[5, 4], :]   # read_comsol.jl:49-52
    return coords, topo


def _remove_unused(topo):
    """Compact renumbering preserving order (src/3d/mesh.jl:6-33)."""
    used = np.unique(topo)
    remap = -np.ones(int(topo.max()) + 1, dtype=np.int64)
    remap[used] = np.arange(used.size)
    return remap[topo], used


def load_comsol(path: str, geometry_order: str = "quadratic",
                physics_order: str = "geometry", beta_type="legendre") -> Mesh3d:
    """Restatement of ``load3dTriangularComsolMesh`` (read_comsol.jl:61-119) for
    geometry in {linear, quadratic} and physics in {geometry, linear, disctriconstant,
    disctrilinear, disctriquadratic} (the discontinuous ones: compute_sources with the
    node offset of beta_type, src/3d/mesh.jl:116-214)."""
    coords0, topo0 = read_comsol(path)
    if geometry_order == "quadratic":
        if topo0.shape[0] != 6:
            raise ValueError("quadratic geometry needs a 6-node triangle mesh")
        gk = TRI6
    elif geometry_order == "linear":
        topo0 = topo0[:3, :]
        gk = TRI3
    else:
        raise ValueError("Only quadratic and linear geometries are supported")
    topo, used = _remove_unused(topo0)
    coords = coords0[:, used]
    if physics_order == "geometry":
        return Mesh3d(coords, topo, coords[:, np.unique(topo)].copy(), topo.copy(), gk, gk)
    if physics_order == "linear":
        ptop = topo[:3, :]
        used_p = np.unique(ptop)
        # The reference indexes ``sources`` with ``topology[1:3,:]`` directly; that is only
        # consistent when corner nodes come first, which holds for the bundled files.
        if used_p.size and int(used_p.max()) != used_p.size - 1:
            raise ValueError("corner nodes are not numbered first; the reference's "
                             "linear-physics indexing would be inconsistent")
        return Mesh3d(coords, topo, coords[:, used_p].copy(), ptop.copy(), gk, TRI3)
    if physics_order in _DISC_ORDERS:
        base = Mesh3d(coords, topo, coords.copy(), topo.copy(), gk, gk)
        return with_discontinuous_physics(base, physics_order, beta_type)
    raise ValueError("physics_order must be geometry, linear, disctriconstant, disctrilinear "
                     "or disctriquadratic")


def _read_obj(path):
    verts, faces = [], []
    with open(path) as fh:
        for line in fh:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "v":
                verts.append([float(t) for t in tok[1:4]])
            elif tok[0] == "f":
                idx = []
                for t in tok[1:]:
                    i = int(t.split("/")[0])          # "v", "v/vt", "v//vn", "v/vt/vn"
                    idx.append(i - 1 if i > 0 else len(verts) + i)   # 1-based or relative
                faces.append(idx)
    return verts, faces


def _read_off(path):
    """OFF (Object File Format): a header line "OFF", then "nv nf ne", nv vertex lines and nf
    face lines "n i_1 .. i_n [colour]".  Comments (#) and blank lines are skipped; the
    colour, normal and texture variants (COFF, NOFF, STOFF, ...) carry extra values per
    vertex, of which the first three (x, y, z) are used; values after a face's n indices
    (a face colour) are ignored.  4OFF / nOFF (other dimensions) are rejected."""
    with open(path) as fh:
        rows = [r for r in (line.split("#")[0].split() for line in fh) if r]
    if not rows:
        raise ValueError(f"{path}: not an OFF file")
    head = rows[0][0]
    if not head.endswith("OFF") or not set(head[:-3]) <= set("STCN"):
        raise ValueError(f"{path}: not an OFF file (header {head!r}; only 3D OFF variants "
                         "such as OFF, COFF, NOFF, STOFF are read)")
    rest = rows[0][1:]
    r = 1
    if not rest:                       # counts on their own line
        rest, r = rows[1], 2
    nv, nf = int(rest[0]), int(rest[1])
    if len(rows) < r + nv + nf:
        raise ValueError(f"{path}: expected {nv} vertices and {nf} faces")
    verts = [[float(t) for t in rows[r + i][:3]] for i in range(nv)]
    faces = []
    for row in rows[r + nv:r + nv + nf]:
        n = int(row[0])
        if len(row) < 1 + n:
            raise ValueError(f"{path}: face with fewer than {n} indices")
        faces.append([int(t) for t in row[1:1 + n]])
    return verts, faces


_PLY_TYPES = {"char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2",
              "int16": "i2", "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4",
              "uint": "u4", "uint32": "u4", "float": "f4", "float32": "f4", "double": "f8",
              "float64": "f8"}


def _read_ply(path):
    """PLY, ASCII or binary (little or big endian): the elements are read in header order;
    x, y, z of "vertex" and the index list of "face" (vertex_indices or vertex_index) are
    kept, every other element and property is skipped."""
    with open(path, "rb") as fh:
        data = fh.read()
    end = data.find(b"end_header")
    if not data.startswith(b"ply") or end < 0:
        raise ValueError(f"{path}: not a PLY file")
    body_at = data.index(b"\n", end) + 1
    header = [ln.split() for ln in data[:end].decode("latin-1").splitlines()]
    fmt, elems = None, []                       # elems: [name, count, [props]]
    for tok in header[1:]:
        if not tok or tok[0] in ("comment", "obj_info"):
            continue
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            elems.append([tok[1], int(tok[2]), []])
        elif tok[0] == "property":
            if not elems:
                raise ValueError(f"{path}: property before any element")
            if tok[1] == "list":                # property list <count type> <item type> name
                elems[-1][2].append(("list", tok[2], tok[3], tok[4]))
            else:
                elems[-1][2].append(("scalar", tok[1], None, tok[2]))
    if fmt not in ("ascii", "binary_little_endian", "binary_big_endian"):
        raise ValueError(f"{path}: unknown PLY format {fmt!r}")
    verts, faces = None, None
    if fmt == "ascii":
        toks = data[body_at:].decode("latin-1").split()
        p = 0
        for name, count, props in elems:
            rows = []
            for _ in range(count):
                vals = {}
                for kind, t1, t2, pname in props:
                    if kind == "scalar":
                        vals[pname] = float(toks[p])
                        p += 1
                    else:
                        n = int(toks[p])
                        vals[pname] = [int(float(t)) for t in toks[p + 1:p + 1 + n]]
                        p += 1 + n
                rows.append(vals)
            if name == "vertex":
                verts = [[r["x"], r["y"], r["z"]] for r in rows]
            elif name == "face":
                key = "vertex_indices" if rows and "vertex_indices" in rows[0] else "vertex_index"
                faces = [r[key] for r in rows]
    else:
        bo = "<" if fmt == "binary_little_endian" else ">"
        p = body_at
        for name, count, props in elems:
            if all(k == "scalar" for k, *_ in props):   # fixed-size records: one read
                dt = np.dtype([(pn, bo + _PLY_TYPES[t1]) for _, t1, _, pn in props])
                rec = np.frombuffer(data, dtype=dt, count=count, offset=p)
                p += dt.itemsize * count
                if name == "vertex":
                    verts = np.stack([rec["x"], rec["y"], rec["z"]], 1).astype(np.float64).tolist()
                continue
            rows = []
            for _ in range(count):
                vals = {}
                for kind, t1, t2, pname in props:
                    if kind == "scalar":
                        dt = np.dtype(bo + _PLY_TYPES[t1])
                        vals[pname] = np.frombuffer(data, dtype=dt, count=1, offset=p)[0]
                        p += dt.itemsize
                    else:
                        dc, di = np.dtype(bo + _PLY_TYPES[t1]), np.dtype(bo + _PLY_TYPES[t2])
                        n = int(np.frombuffer(data, dtype=dc, count=1, offset=p)[0])
                        p += dc.itemsize
                        vals[pname] = np.frombuffer(data, dtype=di, count=n, offset=p).astype(int).tolist()
                        p += di.itemsize * n
                rows.append(vals)
            if name == "face":
                key = "vertex_indices" if rows and "vertex_indices" in rows[0] else "vertex_index"
                faces = [r[key] for r in rows]
            elif name == "vertex":
                verts = [[float(r["x"]), float(r["y"]), float(r["z"])] for r in rows]
    if verts is None or faces is None:
        raise ValueError(f"{path}: PLY file without vertex or face element")
    return verts, faces


def _read_stl(path):
    """STL, ASCII ("solid ... facet ... vertex x y z ... endsolid") or binary (80-byte header,
    uint32 count, 50-byte facets).  STL stores every facet's three corners separately; corners
    with identical coordinates are merged into one node (in order of first appearance), so
    neighbouring facets share nodes as collocation needs."""
    with open(path, "rb") as fh:
        data = fh.read()
    corners = None
    if len(data) >= 84:
        nfac = int(np.frombuffer(data, dtype="<u4", count=1, offset=80)[0])
        if len(data) == 84 + 50 * nfac:
            rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)),
                                                      ("a", "<u2")]), count=nfac, offset=84)
            corners = rec["v"].reshape(-1, 3).astype(np.float64)
    if corners is None:
        text = data.decode("latin-1")
        if not text.lstrip().startswith("solid"):
            raise ValueError(f"{path}: not an STL file")
        tok = text.split()
        corners = np.array([[float(tok[i + 1]), float(tok[i + 2]), float(tok[i + 3])]
                            for i, t in enumerate(tok) if t == "vertex"], dtype=np.float64)
        if corners.shape[0] % 3:
            raise ValueError(f"{path}: facets without three vertices")
    uniq, first, inv = np.unique(corners, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first)                  # nodes numbered by first appearance
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    verts = uniq[order].tolist()
    faces = rank[inv.reshape(-1)].reshape(-1, 3).tolist()
    return verts, faces


def cubed_sphere(f: int, order: str = "linear", radius: float = 1.0,
                 physics: str = "geometry", beta_type="legendre") -> Mesh3d:
    """Quadrilateral sphere: every cube face cut into f x f squares, nodes projected to the
    sphere.  order "linear": QuadrilateralLinear4 (6 f^2 + 2 nodes); "quadratic":
    QuadrilateralQuadraticLagrange (24 f^2 + 2 nodes).  COMSOL node order; (x2 - x1) x
    (x3 - x1) points out of the body, as in the reference's quad_sphere.mphtxt."""
    if f < 1:
        raise ValueError("f must be >= 1")
    q = 2 if order == "quadratic" else 1
    if order not in ("linear", "quadratic"):
        raise ValueError("order must be 'linear' or 'quadratic'")
    g = q * f                                    # grid intervals per face side
    ids: dict = {}
    pts = []

    def node(key):
        i = ids.get(key)
        if i is None:
            i = len(pts)
            ids[key] = i
            pts.append(key)
        return i

    elems = []
    for axis in range(3):
        for sign in (-1, 1):
            a1, a2 = (axis + 1) % 3, (axis + 2) % 3
            if sign < 0:
                a1, a2 = a2, a1                   # e1 x e2 = sign * e_axis: outward
            for j in range(f):
                for i in range(f):
                    def key(a, b):
                        k = [0, 0, 0]
                        k[axis] = sign * g
                        k[a1] = -g + 2 * (q * i + a)
                        k[a2] = -g + 2 * (q * j + b)
                        return tuple(k)
                    if q == 1:
                        elems.append([node(key(0, 0)), node(key(1, 0)), node(key(0, 1)),
                                      node(key(1, 1))])
                    else:
                        elems.append([node(key(a, b)) for a, b in _QUAD9_IJ])
    P = np.array(pts, dtype=np.float64)
    coords = (radius * P / np.linalg.norm(P, axis=1, keepdims=True)).T.copy()
    topo = np.array(elems, dtype=np.int64).T.copy()
    gk = QUAD9 if q == 2 else QUAD4
    base = Mesh3d(coords, topo, coords.copy(), topo.copy(), gk, gk)
    if physics == "geometry":
        return base
    if physics == "linear" and gk == QUAD9:
        return _quad_linear_physics(coords, topo)
    if physics in _DISC_QUAD_ORDERS:
        return with_discontinuous_physics(base, physics, beta_type)
    raise ValueError("physics must be geometry, linear (quadratic geometry) or a discquad order")


def _quad_linear_physics(coords, topo):
    """QuadrilateralLinear4 physics on 9-node geometry: the corner nodes (the first four
    of every element), renumbered in order of appearance."""
    ptop = topo[:4, :]
    used = np.unique(ptop)
    remap = -np.ones(coords.shape[1], dtype=np.int64)
    remap[used] = np.arange(used.size)
    return Mesh3d(coords, topo, coords[:, used].copy(), remap[ptop], QUAD9, QUAD4)


def load_comsol_quad(path: str, geometry_order: str = "quadratic",
                     physics_order: str = "geometry", beta_type="legendre") -> Mesh3d:
    """Restatement of ``load3dQuadComsolMesh`` (src/3d/read_comsol.jl:121-182): 9-node
    COMSOL quadrilaterals as QuadrilateralQuadraticLagrange (quadratic) or their first four
    nodes as QuadrilateralLinear4 (linear); physics geometry, linear (QuadrilateralLinear4
    on the corner nodes) or discquadconstant / discquadlinear / discquadquadratic."""
    coords0, topo0 = read_comsol(path)
    if topo0.shape[0] != 9:
        raise ValueError("a 9-node quadrilateral (quad2) mesh is expected")
    if geometry_order == "quadratic":
        gk = QUAD9
    elif geometry_order == "linear":
        topo0 = topo0[:4, :]
        gk = QUAD4
    else:
        raise ValueError("Only QuadQuadratic and QuadLinear geometry types supported")
    topo, used = _remove_unused(topo0)
    coords = coords0[:, used]
    base = Mesh3d(coords, topo, coords[:, np.unique(topo)].copy(), topo.copy(), gk, gk)
    if physics_order == "geometry":
        return base
    if physics_order == "linear":
        ptop = topo[:4, :]
        used_p = np.unique(ptop)
        # The reference indexes ``sources`` with ``topology[1:4,:]`` directly; that is only
        # consistent when corner nodes come first.
        if used_p.size and int(used_p.max()) != used_p.size - 1:
            raise ValueError("corner nodes are not numbered first; the reference's "
                             "linear-physics indexing would be inconsistent")
        return Mesh3d(coords, topo, coords[:, used_p].copy(), ptop.copy(), gk, QUAD4)
    if physics_order in _DISC_QUAD_ORDERS:
        return with_discontinuous_physics(base, physics_order, beta_type)
    raise ValueError("physics_order must be geometry, linear, discquadconstant, "
                     "discquadlinear or discquadquadratic")


def read_mesh_file(path: str):
    """(coordinates 3 x Nc, topology 3 x E, 0-based) of a triangle mesh in Wavefront OBJ,
    OFF, PLY (ASCII or binary) or STL (ASCII or binary): the formats ``load_mesh_file`` reads
    through MeshIO
    (src/3d/meshing.jl:8-34), restated from the published file formats.  Faces must be
    triangles (the linear-triangle path takes the first three nodes of every face)."""
    ext = os.path.splitext(path)[1].lower()
    readers = {".obj": _read_obj, ".off": _read_off, ".ply": _read_ply, ".stl": _read_stl}
    if ext not in readers:
        raise NotImplementedError(f"{ext} meshes are not read here (OBJ, OFF, PLY, STL are)")
    verts, faces = readers[ext](path)
    if any(len(f) != 3 for f in faces):
        raise ValueError(f"{path}: only triangular faces are supported")
    coords = np.asarray(verts, dtype=np.float64).reshape(-1, 3).T.copy()
    topo = np.asarray(faces, dtype=np.int64).reshape(-1, 3).T.copy()
    if topo.size and (topo.min() < 0 or topo.max() >= coords.shape[1]):
        raise ValueError(f"{path}: face index out of range")
    return coords, topo


def load3d_triangular_mesh(path: str, geometry_order: str = "linear",
                           physics_order: str = "geometry", beta_type="legendre") -> Mesh3d:
    """Restatement of ``load3dTriangularMesh`` (src/3d/meshing.jl:48-106) for linear
    geometry (the only order the reference accepts there): unused nodes removed; geometry or
    linear physics collocate at the used geometry nodes, the discontinuous triangular orders
    at their element nodes (compute_sources, src/3d/mesh.jl:116-148)."""
    if geometry_order != "linear":
        raise ValueError("Currently only linear element supported")
    if physics_order not in ("geometry", "linear") and physics_order not in _DISC_ORDERS:
        raise ValueError("physics_order must be geometry, linear, disctriconstant, "
                         "disctrilinear or disctriquadratic")
    coords0, topo0 = read_mesh_file(path)
    topo, used = _remove_unused(topo0)
    coords = coords0[:, used]
    base = Mesh3d(coords, topo, coords[:, np.unique(topo)].copy(), topo.copy(), TRI3, TRI3)
    if physics_order in _DISC_ORDERS:
        return with_discontinuous_physics(base, physics_order, beta_type)
    return base


def with_coordinates(mesh: Mesh3d, coords) -> Mesh3d:
    """The same mesh with new coordinates; collocation nodes that are coordinates move
    with them (physics = geometry, or the corner nodes for linear physics), discontinuous
    physics nodes are interpolated again."""
    coords = np.asarray(coords, dtype=np.float64)
    if mesh.physics_kind in (DTRI1, DTRI3, DTRI6, DQUAD1, DQUAD4, DQUAD9):
        src, ptop = discontinuous_sources(coords, mesh.topology, mesh.geometry_kind,
                                          mesh.physics_kind, mesh.physics_beta)
        return Mesh3d(coords, mesh.topology, src, ptop, mesh.geometry_kind, mesh.physics_kind,
                      mesh.physics_beta)
    lut = {tuple(c): j for j, c in enumerate(mesh.coordinates.T)}
    idx = np.array([lut[tuple(s)] for s in mesh.sources.T], dtype=np.int64)
    return Mesh3d(coords, mesh.topology, coords[:, idx].copy(), mesh.physics_topology,
                  mesh.geometry_kind, mesh.physics_kind)


def perturb_radial(mesh: Mesh3d, sigma: float, seed: int) -> Mesh3d:
    """Scale every coordinate by (1 + eps), eps ~ N(0, sigma^2) per node (seeded)."""
    rng = np.random.default_rng(seed)
    eps = rng.normal(0.0, sigma, mesh.n_coordinates)
    return with_coordinates(mesh, mesh.coordinates * (1.0 + eps))


def fibonacci_sphere(n: int, radius: float = 1.0) -> np.ndarray:
    """n nearly uniform points on a sphere (Fibonacci lattice), 3 x n."""
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    phi = np.pi * (1.0 + 5.0 ** 0.5) * i
    rho = np.sqrt(1.0 - z * z)
    return radius * np.vstack([rho * np.cos(phi), rho * np.sin(phi), z])


def golden_dir() -> str:
    here = os.path.dirname(os.path.abspath(__file__))
    return os.path.normpath(os.path.join(here, "..", "..", "tests", "golden"))


def uneven_ellipsoid_batch(n_inst: int = 48, n_range=(500, 8000), axes_range=(0.5, 2.0),
                           k_range=(1.0, 20.0), seed: int = 3, order: str = "linear"):
    """Uneven batch of ellipsoids: semi-axes uniform in axes_range, node counts log-uniform
    in n_range rounded to the nearest geodesic count 10 f^2 + 2, wavenumbers uniform in
    k_range.  Returns (meshes, ks)."""
    rng = np.random.default_rng(seed)
    meshes, ks = [], []
    for _ in range(n_inst):
        axes = rng.uniform(axes_range[0], axes_range[1], 3)
        n_target = np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1])))
        f = max(1, int(round(np.sqrt((n_target - 2) / 10.0))))
        k = rng.uniform(k_range[0], k_range[1])
        meshes.append(ellipsoid(f, tuple(axes), order))
        ks.append(float(k))
    return meshes, ks
