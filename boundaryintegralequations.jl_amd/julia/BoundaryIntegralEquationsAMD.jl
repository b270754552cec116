"""
    BoundaryIntegralEquationsAMD

Drop-in MI355X backend for `BoundaryIntegralEquations.assemble_parallel!`
(src/3d/assembly_collocation.jl:259-268) on linear and quadratic triangles, calling
`libbie_amd.so` (C ABI: include/bie_amd.h) through `ccall`.

    using BoundaryIntegralEquations
    import BoundaryIntegralEquationsAMD: assemble_parallel!
    F, G, C = assemble_parallel!(mesh, k, mesh.sources; n=2, m=2)

Supported meshes (TriangularLinear / TriangularQuadratic geometry with geometry or linear
physics) run on the GPU; everything else (quads, discontinuous physics, `sparse=true`)
is forwarded to the reference method unchanged, so the call stays a drop-in.  Errors from
the library are thrown as `ArgumentError` (bad input) or `ErrorException` (HIP failure).

This file is not executed in the build container (no Julia there); the same ABI is
exercised by the Python ctypes harness.
"""
module BoundaryIntegralEquationsAMD

import BoundaryIntegralEquations
const BIE = BoundaryIntegralEquations

const libbie = Ref{String}(get(ENV, "BIE_AMD_LIB",
                               joinpath(@__DIR__, "..", "bie_amd", "libbie_amd.so")))

# Mirror of `bie_mesh` in include/bie_amd.h (field order and widths must match).
struct BieMesh
    coordinates::Ptr{Float64}
    topology::Ptr{Int64}
    sources::Ptr{Float64}
    physics_topology::Ptr{Int64}
    n_coordinates::Int64
    n_elements::Int64
    n_sources::Int64
    geometry_kind::Int32
    physics_kind::Int32
    index_base::Int32
    reserved::Int32
end

element_kind(::BIE.TriangularLinear) = Int32(3)
element_kind(::BIE.TriangularQuadratic) = Int32(6)
element_kind(::Any) = Int32(0)

last_error() = unsafe_string(ccall((:bie_last_error, libbie[]), Cstring, ()))

function check(rc::Integer)
    rc == 0 && return nothing
    msg = last_error()
    rc == 1 && throw(ArgumentError(msg))
    rc == 2 && throw(ArgumentError("unsupported element family: " * msg))
    throw(ErrorException("bie_amd error $rc: " * msg))
end

supported(mesh) = element_kind(mesh.shape_function) != 0 &&
                  element_kind(mesh.physics_function) != 0 &&
                  !(element_kind(mesh.shape_function) == 3 && element_kind(mesh.physics_function) == 6)

"""
    assemble_hip(mesh, k, in_sources; fOn=true, gOn=true, cOn=true, m=4, n=4)

Dense F, G (M x N ComplexF64) and C (length M) on the GPU.
"""
function assemble_hip(mesh::BIE.Mesh3d, k::Number, in_sources::AbstractMatrix;
                      fOn=true, gOn=true, cOn=true, m=4, n=4)
    coords = Matrix{Float64}(mesh.coordinates)
    topo = Matrix{Int64}(mesh.topology)
    srcs = Matrix{Float64}(mesh.sources)
    ptopo = Matrix{Int64}(mesh.physics_topology)
    collocation = in_sources === mesh.sources
    targets = collocation ? srcs : Matrix{Float64}(in_sources)
    M = size(targets, 2)
    N = size(srcs, 2)
    F = Matrix{ComplexF64}(undef, M, N)
    G = Matrix{ComplexF64}(undef, M, N)
    C = Vector{ComplexF64}(undef, M)
    flags = UInt32((fOn ? 1 : 0) | (gOn ? 2 : 0) | (cOn ? 4 : 0))
    kc = ComplexF64(k)
    GC.@preserve coords topo srcs ptopo targets F G C begin
        bm = BieMesh(pointer(coords), pointer(topo), pointer(srcs), pointer(ptopo),
                     size(coords, 2), size(topo, 2), N,
                     element_kind(mesh.shape_function), element_kind(mesh.physics_function),
                     Int32(1), Int32(0))
        rc = ccall((:bie_assemble, libbie[]), Cint,
                   (Ref{BieMesh}, Cdouble, Cdouble, Ptr{Float64}, Int64, Int32, Int32, UInt32,
                    Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{ComplexF64}),
                   bm, real(kc), imag(kc), pointer(targets), M, Int32(n), Int32(m), flags,
                   F, G, C)
        check(rc)
    end
    return F, G, C
end

"""
    assemble_parallel!(mesh, k, in_sources; fOn, gOn, cOn, sparse=false, m=4, n=4, kwargs...)

Same signature and defaults as the reference; routes supported dense assemblies to the
GPU and everything else to `BoundaryIntegralEquations.assemble_parallel!`.
"""
function assemble_parallel!(mesh::BIE.Mesh3d, k, in_sources; fOn=true, gOn=true, cOn=true,
                            sparse=false, m=4, n=4, progress=true, depth=2, offset=nothing)
    if sparse || !supported(mesh)
        return BIE.assemble_parallel!(mesh, k, in_sources; fOn=fOn, gOn=gOn, cOn=cOn,
                                      sparse=sparse, m=m, n=n, progress=progress,
                                      depth=depth, offset=offset)
    end
    return assemble_hip(mesh, k, in_sources; fOn=fOn, gOn=gOn, cOn=cOn, m=m, n=n)
end

end # module
