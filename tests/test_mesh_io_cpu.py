"""Triangle-mesh files (load3dTriangularMesh, src/3d/meshing.jl:8-106): OBJ, OFF, PLY (ASCII
and binary) and STL (ASCII and binary) written from a generated sphere read back to the same
arrays; OBJ index forms, OFF variants (COFF, face colours, counts on the header line), PLY
element order and extra elements, unused vertices, non-triangular faces and unsupported
formats."""
import struct

import numpy as np
import pytest

from bie_amd import mesh as M


def _write(path, fmt, coords, topo, extra_vertex=False):
    V = coords.T.tolist() + ([[9.0, 9.0, 9.0]] if extra_vertex else [])
    F = topo.T.tolist()
    with open(path, "w") as fh:
        if fmt == "obj":
            fh.write("# sphere\no sphere\n")
            for v in V:
                fh.write("v %.17g %.17g %.17g\n" % tuple(v))
            for k, f in enumerate(F):   # mixed index forms: plain, v/vt/vn, v//vn, relative
                a, b, c = (i + 1 for i in f)
                if k % 4 == 0:
                    fh.write(f"f {a} {b} {c}\n")
                elif k % 4 == 1:
                    fh.write(f"f {a}/1/1 {b}/2/2 {c}/3/3\n")
                elif k % 4 == 2:
                    fh.write(f"f {a}//1 {b}//2 {c}//3\n")
                else:
                    n = len(V)
                    fh.write(f"f {a - 1 - n} {b - 1 - n} {c - 1 - n}\n")
        elif fmt == "off":
            fh.write(f"OFF\n# comment\n{len(V)} {len(F)} 0\n")
            for v in V:
                fh.write("%.17g %.17g %.17g\n" % tuple(v))
            for f in F:
                fh.write("3 %d %d %d\n" % tuple(f))
        else:
            fh.write("ply\nformat ascii 1.0\ncomment sphere\n"
                     f"element vertex {len(V)}\nproperty float x\nproperty float y\n"
                     f"property float z\nelement face {len(F)}\n"
                     "property list uchar int vertex_indices\nend_header\n")
            for v in V:
                fh.write("%.17g %.17g %.17g\n" % tuple(v))
            for f in F:
                fh.write("3 %d %d %d\n" % tuple(f))


@pytest.mark.parametrize("fmt", ["obj", "off", "ply"])
@pytest.mark.parametrize("extra", [False, True])
def test_round_trip_matches_the_generated_mesh(tmp_path, fmt, extra):
    ref = M.icosphere(3, "linear")
    path = str(tmp_path / f"sphere.{fmt}")
    _write(path, fmt, ref.coordinates, ref.topology, extra_vertex=extra)
    mesh = M.load3d_triangular_mesh(path)
    assert np.array_equal(mesh.coordinates, ref.coordinates)   # the unused vertex is dropped
    assert np.array_equal(mesh.topology, ref.topology)
    assert np.array_equal(mesh.sources, ref.sources)
    assert np.array_equal(mesh.physics_topology, ref.physics_topology)
    assert mesh.geometry_kind == M.TRI3 and mesh.physics_kind == M.TRI3
    lin = M.load3d_triangular_mesh(path, physics_order="linear")
    assert np.array_equal(lin.sources, ref.sources)


def test_rejections(tmp_path):
    quad = tmp_path / "quad.obj"
    quad.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n")
    with pytest.raises(ValueError):
        M.load3d_triangular_mesh(str(quad))
    bad = tmp_path / "bad.off"
    bad.write_text("OFF\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n3 0 1 5\n")
    with pytest.raises(ValueError):
        M.load3d_triangular_mesh(str(bad))
    with pytest.raises(NotImplementedError):
        M.load3d_triangular_mesh(str(tmp_path / "x.vtk"))
    four = tmp_path / "four.off"
    four.write_text("4OFF\n3 1 0\n0 0 0 0\n1 0 0 0\n0 1 0 0\n3 0 1 2\n")
    with pytest.raises(ValueError):
        M.load3d_triangular_mesh(str(four))
    with pytest.raises(ValueError):
        M.load3d_triangular_mesh(str(quad), geometry_order="quadratic")


def _check(mesh, ref):
    assert np.array_equal(mesh.coordinates, ref.coordinates)
    assert np.array_equal(mesh.topology, ref.topology)
    assert np.array_equal(mesh.sources, ref.sources)


def test_off_variants(tmp_path):
    ref = M.icosphere(2, "linear")
    V, F = ref.coordinates.T.tolist(), ref.topology.T.tolist()
    coff = tmp_path / "c.off"          # vertex colours, face colours, counts on the header line
    with open(coff, "w") as fh:
        fh.write(f"COFF {len(V)} {len(F)} 0\n")
        for v in V:
            fh.write("%.17g %.17g %.17g 255 0 0 255\n" % tuple(v))
        for f in F:
            fh.write("3 %d %d %d 0.5 0.5 0.5\n" % tuple(f))
    _check(M.load3d_triangular_mesh(str(coff)), ref)
    noff = tmp_path / "n.off"          # normals after the coordinates, blank lines
    with open(noff, "w") as fh:
        fh.write(f"NOFF\n\n{len(V)} {len(F)} 0\n\n")
        for v in V:
            fh.write("%.17g %.17g %.17g 0 0 1\n" % tuple(v))
        for f in F:
            fh.write("3 %d %d %d\n" % tuple(f))
    _check(M.load3d_triangular_mesh(str(noff)), ref)


def _ply_header(fmt, nv, nf, face_first=False):
    vert = (f"element vertex {nv}\nproperty double x\nproperty double y\nproperty double z\n"
            "property uchar red\n")
    face = f"element face {nf}\nproperty list uchar int vertex_indices\nproperty float quality\n"
    other = "element edge 1\nproperty int vertex1\nproperty int vertex2\n"
    body = (face + vert) if face_first else (vert + face)
    return (f"ply\nformat {fmt} 1.0\ncomment generated\n\n{body}{other}end_header\n").encode()


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
@pytest.mark.parametrize("face_first", [False, True])
def test_ply_formats_and_element_order(tmp_path, fmt, face_first):
    ref = M.icosphere(2, "linear")
    V, F = ref.coordinates.T.tolist(), ref.topology.T.tolist()
    out = bytearray(_ply_header(fmt, len(V), len(F), face_first))
    if fmt == "ascii":
        vtxt = "".join("%.17g %.17g %.17g 7\n" % tuple(v) for v in V)
        ftxt = "".join("3 %d %d %d 0.25\n" % tuple(f) for f in F)
        out += ((ftxt + vtxt) if face_first else (vtxt + ftxt)).encode() + b"0 1\n"
    else:
        bo = "<" if fmt.endswith("little_endian") else ">"
        vb = b"".join(struct.pack(bo + "dddB", *v, 7) for v in V)
        fb = b"".join(struct.pack(bo + "Biiif", 3, *f, 0.25) for f in F)
        out += ((fb + vb) if face_first else (vb + fb)) + struct.pack(bo + "ii", 0, 1)
    path = tmp_path / "s.ply"
    path.write_bytes(bytes(out))
    _check(M.load3d_triangular_mesh(str(path)), ref)


@pytest.mark.parametrize("binary", [False, True])
def test_stl_merges_facet_corners(tmp_path, binary):
    ref = M.icosphere(2, "linear")
    X = ref.coordinates.astype(np.float32).astype(np.float64) if binary else ref.coordinates
    tri = X[:, ref.topology].transpose(2, 1, 0)          # E x 3 corners x 3
    path = tmp_path / "s.stl"
    if binary:
        out = bytearray(b"\0" * 80) + struct.pack("<I", tri.shape[0])
        for t in tri:
            out += struct.pack("<12fH", 0, 0, 0, *t.reshape(-1), 0)
        path.write_bytes(bytes(out))
    else:
        with open(path, "w") as fh:
            fh.write("solid s\n")
            for t in tri:
                fh.write("facet normal 0 0 0\nouter loop\n")
                for c in t:
                    fh.write("vertex %.17g %.17g %.17g\n" % tuple(c))
                fh.write("endloop\nendfacet\n")
            fh.write("endsolid s\n")
    mesh = M.load3d_triangular_mesh(str(path))
    assert mesh.n_coordinates == ref.n_coordinates and mesh.n_elements == ref.n_elements
    # same triangles (as corner coordinates), nodes shared between neighbouring facets
    assert np.array_equal(mesh.coordinates[:, mesh.topology], X[:, ref.topology])
    assert np.bincount(mesh.topology.ravel()).min() >= 5
