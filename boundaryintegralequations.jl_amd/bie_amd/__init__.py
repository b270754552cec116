"""bie_amd — MI355X-native dense 3D collocation assembly (drop-in for the reference's
assemble_parallel! on linear and quadratic triangles) behind a C ABI.

The compute path is the HIP library libbie_amd.so (built for gfx950); this package is
its host-side mirror: mesh data model and generators, ctypes binding, and the
``assemble_parallel`` entry point with the reference's keyword interface (its
``sparse=True`` form, the near-field ``sparse_assemble_parallel``, included).
"""
from .mesh import (Mesh3d, TRI3, TRI6, icosphere, ellipsoid, load_comsol, read_comsol,  # noqa: F401
                   load3d_triangular_mesh, read_mesh_file)
from .assembly import (assemble_parallel, Batch, DeviceMesh,  # noqa: F401
                       sparse_assemble_parallel, SparsePattern)

__all__ = ["Mesh3d", "TRI3", "TRI6", "icosphere", "ellipsoid", "load_comsol", "read_comsol",
           "load3d_triangular_mesh", "read_mesh_file", "assemble_parallel", "DeviceMesh",
           "sparse_assemble_parallel", "SparsePattern"]
