"""Triangle-mesh files (load3dTriangularMesh, src/3d/meshing.jl:8-106): OBJ, OFF and ASCII
PLY written from a generated sphere read back to the same arrays; OBJ index forms, unused
vertices, non-triangular faces and unsupported formats."""
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
        M.load3d_triangular_mesh(str(tmp_path / "x.stl"))
    with pytest.raises(ValueError):
        M.load3d_triangular_mesh(str(quad), geometry_order="quadratic")
