"""bie_amd — MI355X-native dense 3D collocation assembly (drop-in for the reference's
assemble_parallel! on triangles and quadrilaterals, continuous and discontinuous physics)
behind a C ABI.

The compute path is the HIP library libbie_amd.so (built for gfx950); this package is
its host-side mirror: mesh data model and generators, ctypes binding, and the
``assemble_parallel`` entry point with the reference's keyword interface (its
``sparse=True`` form, the near-field ``sparse_assemble_parallel``, included), and the
fast operators' partial near field ``partial_assemble_parallel``.
"""
from .mesh import (Mesh3d, TRI3, TRI6, DTRI1, DTRI3, DTRI6, QUAD4, QUAD4S, QUAD8,  # noqa: F401
                   QUAD9, DQUAD1, DQUAD4, DQUAD9, icosphere, ellipsoid, cubed_sphere,
                   load_comsol, load_comsol_quad, read_comsol, load3d_triangular_mesh,
                   read_mesh_file, with_discontinuous_physics)
from .assembly import (assemble_parallel, Batch, DeviceMesh,  # noqa: F401
                       partial_assemble_parallel, sparse_assemble_parallel, SparsePattern)

__all__ = ["Mesh3d", "TRI3", "TRI6", "DTRI1", "DTRI3", "DTRI6", "QUAD4", "QUAD4S", "QUAD8",
           "QUAD9", "DQUAD1", "DQUAD4", "DQUAD9", "icosphere", "ellipsoid", "cubed_sphere",
           "load_comsol", "load_comsol_quad", "read_comsol", "load3d_triangular_mesh",
           "read_mesh_file", "with_discontinuous_physics", "assemble_parallel", "DeviceMesh",
           "Batch", "partial_assemble_parallel", "sparse_assemble_parallel", "SparsePattern"]
