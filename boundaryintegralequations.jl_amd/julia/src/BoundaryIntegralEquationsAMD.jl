"""
    BoundaryIntegralEquationsAMD

Drop-in MI355X backend for `BoundaryIntegralEquations.assemble_parallel!`
(src/3d/assembly_collocation.jl:259-268) on linear and quadratic triangles, calling
`libbie_amd.so` (C ABI: include/bie_amd.h) through `ccall`.

    using BoundaryIntegralEquations
    import BoundaryIntegralEquationsAMD: assemble_parallel!
    F, G, C = assemble_parallel!(mesh, k, mesh.sources; n=2, m=2)

Supported meshes run on the GPU: TriangularLinear / TriangularQuadratic geometry with
geometry, linear or discontinuous triangular physics, and the quadrilaterals
(QuadrilateralLinear4, QuadrilateralLinearSerendipity, QuadrilateralQuadratic,
QuadrilateralQuadraticLagrange) with their own, linear or discontinuous quadrilateral
physics; dense, and on triangles also `sparse=true` (the near field,
`sparse_assemble_parallel!`, with the subdivision rules of continuous physics or the polar
rules of discontinuous physics) and the fast operators' `partial_assemble_parallel!`.
Everything else (the near field of quadrilaterals, other element types) is forwarded to the
reference method unchanged, so the call stays a drop-in.  Errors from
the library are thrown as `ArgumentError` (bad input) or `ErrorException` (HIP failure).

This file is not executed in the build container (no Julia there); the same ABI is
exercised by the Python ctypes harness.
"""
module BoundaryIntegralEquationsAMD

import BoundaryIntegralEquations
import LinearAlgebra
import SparseArrays
const BIE = BoundaryIntegralEquations

const libbie = Ref{String}(get(ENV, "BIE_AMD_LIB",
                               joinpath(@__DIR__, "..", "..", "bie_amd", "libbie_amd.so")))

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
    flags::Int32
    physics_beta::Float64
end

# Rule selection compares hypot distances, and Julia releases round the 3-argument hypot
# differently (Julia 1.9: maxabs * sqrt(sum((x/maxabs)^2)); later: power-of-two scaled
# sqrt(x^2 + y^2 + z^2)).  The flag follows the running Julia, found from one argument
# triple on which the two formulas round differently.
const BIE_MESH_SCALED_HYPOT = Int32(1)
hypot_flag() = hypot(0.6732655185893088, 0.3428080423874833, 0.13687617154257523) ==
               0.7678143647933542 ? BIE_MESH_SCALED_HYPOT : Int32(0)

element_kind(::BIE.TriangularLinear) = Int32(3)
element_kind(::BIE.TriangularQuadratic) = Int32(6)
element_kind(::BIE.QuadrilateralLinear4) = Int32(0x204)
element_kind(::BIE.QuadrilateralLinearSerendipity) = Int32(0x404)
element_kind(::BIE.QuadrilateralQuadratic) = Int32(0x208)
element_kind(::BIE.QuadrilateralQuadraticLagrange) = Int32(0x209)
element_kind(::Any) = Int32(0)

# physics_kind of the C ABI: the continuous kinds, or 0x100 | node count for the
# discontinuous triangles (surface_functions.jl:234-318), whose beta field sets the nodes
physics_kind(pf) = element_kind(pf)
physics_kind(::BIE.DiscontinuousTriangularConstant) = Int32(0x101)
physics_kind(::BIE.DiscontinuousTriangularLinear) = Int32(0x103)
physics_kind(::BIE.DiscontinuousTriangularQuadratic) = Int32(0x106)
physics_kind(::BIE.DiscontinuousQuadrilateralConstant) = Int32(0x301)
physics_kind(::BIE.DiscontinuousQuadrilateralLinear4) = Int32(0x304)
physics_kind(::BIE.DiscontinuousQuadrilateralQuadraticLagrange) = Int32(0x309)
physics_beta(pf) = 0.0
physics_beta(pf::BIE.DiscontinuousTriangularLinear) = Float64(pf.beta)
physics_beta(pf::BIE.DiscontinuousTriangularQuadratic) = Float64(pf.beta)
physics_beta(pf::BIE.DiscontinuousQuadrilateralLinear4) = Float64(pf.beta)
physics_beta(pf::BIE.DiscontinuousQuadrilateralQuadraticLagrange) = Float64(pf.beta)
quadrilateral(mesh) = mesh.shape_function isa BIE.Quadrilateral

bie_mesh(mesh, coords, topo, srcs, ptopo) =
    BieMesh(pointer(coords), pointer(topo), pointer(srcs), pointer(ptopo), size(coords, 2),
            size(topo, 2), size(srcs, 2), element_kind(mesh.shape_function),
            physics_kind(mesh.physics_function), Int32(1), hypot_flag(),
            physics_beta(mesh.physics_function))

last_error() = unsafe_string(ccall((:bie_last_error, libbie[]), Cstring, ()))

function check(rc::Integer)
    rc == 0 && return nothing
    msg = last_error()
    rc == 1 && throw(ArgumentError(msg))
    rc == 2 && throw(ArgumentError("unsupported element family: " * msg))
    throw(ErrorException("bie_amd error $rc: " * msg))
end

function supported(mesh)
    g = element_kind(mesh.shape_function)
    p = physics_kind(mesh.physics_function)
    (g == 0 || p == 0) && return false
    if mesh.shape_function isa BIE.Quadrilateral   # the combinations include/bie_amd.h lists
        return p == g || (g == 0x209 && p == 0x204) || p in (0x301, 0x304, 0x309)
    end
    return p in (3, 6, 0x101, 0x103, 0x106) && !(g == 3 && p == 6)
end

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
    # collocation (the sources themselves, or an equal array: the library also recognises
    # value-identical targets) passes NULL, which selects the recorded mid-edge pairs
    collocation = in_sources === mesh.sources || in_sources == mesh.sources
    targets = collocation ? srcs : Matrix{Float64}(in_sources)
    M = size(targets, 2)
    N = size(srcs, 2)
    F = Matrix{ComplexF64}(undef, M, N)
    G = Matrix{ComplexF64}(undef, M, N)
    C = Vector{ComplexF64}(undef, M)
    flags = UInt32((fOn ? 1 : 0) | (gOn ? 2 : 0) | (cOn ? 4 : 0))
    kc = ComplexF64(k)
    GC.@preserve coords topo srcs ptopo targets F G C begin
        bm = bie_mesh(mesh, coords, topo, srcs, ptopo)
        rc = ccall((:bie_assemble, libbie[]), Cint,
                   (Ref{BieMesh}, Cdouble, Cdouble, Ptr{Float64}, Int64, Int32, Int32, UInt32,
                    Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{ComplexF64}),
                   bm, real(kc), imag(kc), collocation ? Ptr{Float64}(C_NULL) : pointer(targets),
                   M, Int32(n), Int32(m), flags, F, G, C)
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
    # the near field of quadrilaterals stays with the reference
    if !supported(mesh) || (sparse && quadrilateral(mesh))
        return BIE.assemble_parallel!(mesh, k, in_sources; fOn=fOn, gOn=gOn, cOn=cOn,
                                      sparse=sparse, m=m, n=n, progress=progress,
                                      depth=depth, offset=offset)
    end
    sparse && return sparse_hip(mesh, k, in_sources; fOn=fOn, gOn=gOn, depth=depth,
                                offset=offset)
    return assemble_hip(mesh, k, in_sources; fOn=fOn, gOn=gOn, cOn=cOn, m=m, n=n)
end

"""
    sparse_hip(mesh, k, sources; fOn=true, gOn=true, depth=1, offset=nothing, Ngauss=40)

Near-field F, G on the GPU as `SparseMatrixCSC`s (n_sources x n_sources) with the
reference's stored entries (`sparse_assemble_parallel!`,
src/3d/assembly_collocation_losses.jl:200-405): `sparse(I, J, F), sparse(I, J, G)`.
`Ngauss`: the polar rules' order for discontinuous physics (40 for quadratic physics,
as the reference forces).
"""
function sparse_hip(mesh::BIE.Mesh3d, k::Number, sources::AbstractMatrix; fOn=true, gOn=true,
                    depth=1, offset=nothing, Ngauss=40)
    coords = Matrix{Float64}(mesh.coordinates)
    topo = Matrix{Int64}(mesh.topology)
    srcs = Matrix{Float64}(mesh.sources)
    ptopo = Matrix{Int64}(mesh.physics_topology)
    N = size(srcs, 2)
    size(sources) == (3, N) || throw(ArgumentError("one source point per mesh source"))
    own = sources === mesh.sources
    pts = own ? srcs : Matrix{Float64}(sources)
    flags = UInt32((fOn ? 1 : 0) | (gOn ? 2 : 0))
    kc = ComplexF64(k)
    off = offset === nothing ? NaN : Float64(offset)
    GC.@preserve coords topo srcs ptopo pts begin
        bm = bie_mesh(mesh, coords, topo, srcs, ptopo)
        nnz = Ref{Int64}(0)
        check(ccall((:bie_sparse_count, libbie[]), Cint, (Ref{BieMesh}, Int32, Ref{Int64}),
                    bm, Int32(depth), nnz))
        I = Vector{Int64}(undef, nnz[]); J = similar(I)
        F = Vector{ComplexF64}(undef, nnz[]); G = similar(F)
        check(ccall((:bie_sparse_assemble_ex, libbie[]), Cint,
                    (Ref{BieMesh}, Cdouble, Cdouble, Ptr{Float64}, Int64, Int32, Cdouble, Int32,
                     UInt32, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{ComplexF64}, Ptr{ComplexF64}),
                    bm, real(kc), imag(kc), own ? Ptr{Float64}(C_NULL) : pointer(pts),
                    own ? 0 : N, Int32(depth), off, Int32(Ngauss), flags, nnz[], I, J, F, G))
        return SparseArrays.sparse(I, J, F, N, N), SparseArrays.sparse(I, J, G, N, N)
    end
end

# ---------------------------------------------------------------------------------------
# Routing hook inside the reference's own generic function
# ---------------------------------------------------------------------------------------
#
# The reference's dispatcher (assembly_collocation.jl:259-268) calls the 4-argument methods
# assemble_parallel!(mesh, k, in_sources, mesh.shape_function; ...) for Triangular (:270) and
# TriangularQuadratic (:355) geometry.  The methods below are more specific in k and
# in_sources, so every caller inside the reference - the examples, the dense lossy branch
# assemble_parallel!(mesh, kₐ, sources; m, n) at LossyGlobalMatrix.jl:119 (complex kₐ) -
# reaches the GPU once this package is loaded.  Unsupported physics and `use_gpu!(false)`
# go back to the reference's own method through `invoke`.

const GPU_ROUTING = Ref(true)

"""
    use_gpu!(on::Bool)

Turn the routing of the reference's `assemble_parallel!` to the GPU on or off (on by
default once the package is loaded).
"""
use_gpu!(on::Bool) = (GPU_ROUTING[] = on; on)

function BIE.assemble_parallel!(mesh::BIE.Mesh3d, k::Number, in_sources::AbstractMatrix{<:Real},
                                shape_function::BIE.TriangularLinear; fOn=true, gOn=true,
                                cOn=true, m=3, n=3, progress=true)
    if !(GPU_ROUTING[] && supported(mesh))
        return invoke(BIE.assemble_parallel!, Tuple{BIE.Mesh3d,Any,Any,BIE.Triangular}, mesh, k,
                      in_sources, shape_function; fOn=fOn, gOn=gOn, cOn=cOn, m=m, n=n,
                      progress=progress)
    end
    return assemble_hip(mesh, k, in_sources; fOn=fOn, gOn=gOn, cOn=cOn, m=m, n=n)
end

function BIE.assemble_parallel!(mesh::BIE.Mesh3d, k::Number, in_sources::AbstractMatrix{<:Real},
                                shape_function::BIE.TriangularQuadratic; fOn=true, gOn=true,
                                cOn=true, m=3, n=3, progress=true)
    if !(GPU_ROUTING[] && supported(mesh))
        return invoke(BIE.assemble_parallel!, Tuple{BIE.Mesh3d,Any,Any,BIE.TriangularQuadratic},
                      mesh, k, in_sources, shape_function; fOn=fOn, gOn=gOn, cOn=cOn, m=m,
                      n=n, progress=progress)
    end
    return assemble_hip(mesh, k, in_sources; fOn=fOn, gOn=gOn, cOn=cOn, m=m, n=n)
end

# Quadrilateral geometry takes the reference's generic method (assembly_collocation.jl:481-538,
# defaults n = m = 4); this method is more specific than it for every Quadrilateral.
function BIE.assemble_parallel!(mesh::BIE.Mesh3d, k::Number, in_sources::AbstractMatrix{<:Real},
                                shape_function::BIE.Quadrilateral; fOn=true, gOn=true,
                                cOn=true, m=4, n=4, progress=true)
    if !(GPU_ROUTING[] && supported(mesh))
        return invoke(BIE.assemble_parallel!, Tuple{BIE.Mesh3d,Any,Any,BIE.SurfaceFunction},
                      mesh, k, in_sources, shape_function; fOn=fOn, gOn=gOn, cOn=cOn, m=m,
                      n=n, progress=progress)
    end
    return assemble_hip(mesh, k, in_sources; fOn=fOn, gOn=gOn, cOn=cOn, m=m, n=n)
end

# The near field: sparse_assemble_parallel!(mesh, k, sources, physics_function; fOn, gOn,
# depth, offset, progress, Ngauss) (assembly_collocation_losses.jl:200-405), called by the
# reference's fast operators (fast_multipole_method.jl:181,187) and lossy systems through
# assemble_parallel!(...; sparse=true).  Continuous (subdivision rules) and discontinuous
# (polar rules, Ngauss) triangular physics run on the GPU when the physics function is the
# mesh's own (the kernel takes the mesh's physics); another physics function, quadrilaterals
# and a source count that differs from the mesh's go to the reference's method.
function BIE.sparse_assemble_parallel!(mesh::BIE.Mesh3d, k::Number,
                                       sources::AbstractMatrix{<:Real},
                                       physics_function::Union{BIE.TriangularLinear,
                                                               BIE.TriangularQuadratic,
                                                               BIE.DiscontinuousTriangular};
                                       fOn=true, gOn=true, depth=1, offset=nothing,
                                       progress=true, Ngauss=40)
    same = typeof(physics_function) == typeof(mesh.physics_function) &&
           physics_beta(physics_function) == physics_beta(mesh.physics_function)
    if !(GPU_ROUTING[] && supported(mesh) && !quadrilateral(mesh) && same &&
         size(sources, 2) == size(mesh.sources, 2))
        return invoke(BIE.sparse_assemble_parallel!, Tuple{BIE.Mesh3d,Any,Any,Any}, mesh, k,
                      sources, physics_function; fOn=fOn, gOn=gOn, depth=depth,
                      offset=offset, progress=progress, Ngauss=Ngauss)
    end
    return sparse_hip(mesh, k, sources; fOn=fOn, gOn=gOn, depth=depth, offset=offset,
                      Ngauss=Ngauss)
end

"""
    partial_hip(mesh, k, sources; fOn=true, gOn=true, n=3, depth=1)

The partial near field of a fast operator on the GPU (`partial_assemble_parallel!`,
src/3d/fast_multipole_method.jl:11-81): the sparse pattern, every pair integrated with
`gauss_points_triangle(n)`; `sparse(I, J, F), sparse(I, J, G)`.
"""
function partial_hip(mesh::BIE.Mesh3d, k::Number, sources::AbstractMatrix; fOn=true, gOn=true,
                     n=3, depth=1)
    coords = Matrix{Float64}(mesh.coordinates)
    topo = Matrix{Int64}(mesh.topology)
    srcs = Matrix{Float64}(mesh.sources)
    ptopo = Matrix{Int64}(mesh.physics_topology)
    N = size(srcs, 2)
    size(sources) == (3, N) || throw(ArgumentError("one source point per mesh source"))
    own = sources === mesh.sources
    pts = own ? srcs : Matrix{Float64}(sources)
    flags = UInt32((fOn ? 1 : 0) | (gOn ? 2 : 0))
    kc = ComplexF64(k)
    GC.@preserve coords topo srcs ptopo pts begin
        bm = bie_mesh(mesh, coords, topo, srcs, ptopo)
        nnz = Ref{Int64}(0)
        check(ccall((:bie_sparse_count, libbie[]), Cint, (Ref{BieMesh}, Int32, Ref{Int64}),
                    bm, Int32(depth), nnz))
        I = Vector{Int64}(undef, nnz[]); J = similar(I)
        F = Vector{ComplexF64}(undef, nnz[]); G = similar(F)
        check(ccall((:bie_partial_assemble, libbie[]), Cint,
                    (Ref{BieMesh}, Cdouble, Cdouble, Ptr{Float64}, Int64, Int32, Int32, UInt32,
                     Int64, Ptr{Int64}, Ptr{Int64}, Ptr{ComplexF64}, Ptr{ComplexF64}),
                    bm, real(kc), imag(kc), own ? Ptr{Float64}(C_NULL) : pointer(pts),
                    own ? 0 : N, Int32(depth), Int32(n), flags, nnz[], I, J, F, G))
        return SparseArrays.sparse(I, J, F, N, N), SparseArrays.sparse(I, J, G, N, N)
    end
end

# The reference's fast operators call partial_assemble_parallel!(mesh, zk, targets,
# shape_function; fOn, gOn, depth) (fast_multipole_method.jl:179,185); continuous and
# discontinuous physics on triangles run on the GPU, the rest goes to the reference's method.
function BIE.partial_assemble_parallel!(mesh::BIE.Mesh3d, k::Number,
                                        in_sources::AbstractMatrix{<:Real},
                                        shape_function::BIE.Triangular; fOn=true, gOn=true,
                                        n=3, progress=false, depth=1)
    if !(GPU_ROUTING[] && supported(mesh) && !quadrilateral(mesh) &&
         n in (1, 3, 4, 7, 12, 13) && size(in_sources, 2) == size(mesh.sources, 2))
        return invoke(BIE.partial_assemble_parallel!, Tuple{BIE.Mesh3d,Any,Any,BIE.Triangular},
                      mesh, k, in_sources, shape_function; fOn=fOn, gOn=gOn, n=n,
                      progress=progress, depth=depth)
    end
    return partial_hip(mesh, k, in_sources; fOn=fOn, gOn=gOn, n=n, depth=depth)
end

# ---------------------------------------------------------------------------------------
# Device-resident meshes and the batched / adjoint / apply entry points
# ---------------------------------------------------------------------------------------

"""Device buffer from the library's allocator (freed by a finalizer)."""
mutable struct DeviceBuffer
    ptr::Ptr{Cvoid}
    nbytes::Int
    function DeviceBuffer(nbytes::Integer)
        p = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:bie_device_malloc, libbie[]), Cint, (Ptr{Ptr{Cvoid}}, Int64), p,
                    max(Int64(nbytes), 1)))
        b = new(p[], Int(nbytes))
        finalizer(x -> ccall((:bie_device_free, libbie[]), Cint, (Ptr{Cvoid},), x.ptr), b)
        return b
    end
end

function DeviceBuffer(a::Array)
    b = DeviceBuffer(sizeof(a))
    GC.@preserve a check(ccall((:bie_memcpy_h2d, libbie[]), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int64),
                               b.ptr, pointer(a), sizeof(a)))
    return b
end

function download!(a::Array, b::DeviceBuffer)
    check(ccall((:bie_device_synchronize, libbie[]), Cint, ()))
    GC.@preserve a check(ccall((:bie_memcpy_d2h, libbie[]), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int64),
                               pointer(a), b.ptr, sizeof(a)))
    return a
end

"""A mesh uploaded once (bie_dmesh), with its cached rule tables and schedules."""
mutable struct DeviceMesh
    handle::Ptr{Cvoid}
    mesh::BIE.Mesh3d
    function DeviceMesh(mesh::BIE.Mesh3d; device::Integer=0)
        supported(mesh) || throw(ArgumentError("unsupported element family"))
        coords = Matrix{Float64}(mesh.coordinates)
        topo = Matrix{Int64}(mesh.topology)
        srcs = Matrix{Float64}(mesh.sources)
        ptopo = Matrix{Int64}(mesh.physics_topology)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        GC.@preserve coords topo srcs ptopo begin
            bm = bie_mesh(mesh, coords, topo, srcs, ptopo)
            check(ccall((:bie_dmesh_create, libbie[]), Cint, (Ref{BieMesh}, Cint, Ptr{Ptr{Cvoid}}),
                        bm, Cint(device), h))
        end
        dm = new(h[], mesh)
        finalizer(x -> ccall((:bie_dmesh_destroy, libbie[]), Cint, (Ptr{Cvoid},), x.handle), dm)
        return dm
    end
end

n_sources(dm::DeviceMesh) = size(dm.mesh.sources, 2)
flagbits(fOn, gOn, cOn) = UInt32((fOn ? 1 : 0) | (gOn ? 2 : 0) | (cOn ? 4 : 0))

"""
    assemble_sweep(dm, ks; n=4, m=4, fOn=true, gOn=true, cOn=true) -> (Fs, Gs, C)

F_k, G_k for every real wavenumber in `ks` in one launch (bie_assemble_sweep_device);
C does not depend on k.
"""
function assemble_sweep(dm::DeviceMesh, ks::AbstractVector{<:Real}; n=4, m=4, fOn=true,
                        gOn=true, cOn=true)
    N = n_sources(dm)
    nk = length(ks)
    kd = DeviceBuffer(Vector{Float64}(ks))
    dF = DeviceBuffer(16 * N * N * nk)
    dG = DeviceBuffer(16 * N * N * nk)
    dC = DeviceBuffer(16 * N)
    check(ccall((:bie_assemble_sweep_device, libbie[]), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Cvoid}, Int64, Int32, Int32, UInt32,
                 Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid}),
                dm.handle, kd.ptr, Int32(nk), C_NULL, N, Int32(n), Int32(m),
                flagbits(fOn, gOn, cOn), dF.ptr, dG.ptr, Int64(N * N), dC.ptr, C_NULL))
    Fall = download!(Array{ComplexF64}(undef, N, N, nk), dF)
    Gall = download!(Array{ComplexF64}(undef, N, N, nk), dG)
    C = download!(Vector{ComplexF64}(undef, N), dC)
    return [Fall[:, :, q] for q in 1:nk], [Gall[:, :, q] for q in 1:nk], C
end

"""
    batch_assemble(meshes, ks; n=4, m=4, tune=false) -> Vector of (F, G, C)

Collocation F, G, C of several meshes (same element kinds), each at its own wavenumber, in
one persistent launch (bie_batch_*); `tune=true` runs the schedule autotuner first.
"""
function batch_assemble(dms::Vector{DeviceMesh}, ks::AbstractVector; n=4, m=4, tune=false)
    length(dms) == length(ks) || throw(ArgumentError("one wavenumber per mesh"))
    handles = [dm.handle for dm in dms]
    b = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:bie_batch_create, libbie[]), Cint, (Ptr{Ptr{Cvoid}}, Int32, Int32, Int32,
                                                      Ptr{Ptr{Cvoid}}),
                handles, Int32(length(dms)), Int32(n), Int32(m), b))
    try
        Ns = [n_sources(dm) for dm in dms]
        bufs = [(DeviceBuffer(16 * N * N), DeviceBuffer(16 * N * N), DeviceBuffer(16 * N)) for N in Ns]
        kr = Float64[real(k) for k in ks]
        ki = Float64[imag(k) for k in ks]
        pF = [x[1].ptr for x in bufs]; pG = [x[2].ptr for x in bufs]; pC = [x[3].ptr for x in bufs]
        if tune
            best = Ref{Float64}(0.0)
            check(ccall((:bie_batch_tune, libbie[]), Cint,
                        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, UInt32, Ptr{Ptr{Cvoid}},
                         Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Ptr{Float64}),
                        b[], kr, ki, UInt32(7), pF, pG, pC, C_NULL, best))
        else
            check(ccall((:bie_batch_assemble, libbie[]), Cint,
                        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, UInt32, Ptr{Ptr{Cvoid}},
                         Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}),
                        b[], kr, ki, UInt32(7), pF, pG, pC, C_NULL))
        end
        return [(download!(Matrix{ComplexF64}(undef, N, N), x[1]),
                 download!(Matrix{ComplexF64}(undef, N, N), x[2]),
                 download!(Vector{ComplexF64}(undef, N), x[3])) for (N, x) in zip(Ns, bufs)]
    finally
        ccall((:bie_batch_destroy, libbie[]), Cint, (Ptr{Cvoid},), b[])
    end
end

"""
    assembly_vjp(dm, k, alpha, p; beta=nothing, targets=nothing, target_node=nothing, n=4, m=4)

Gradient of Re[alpha' * F * p + beta' * C] (plain transposes: alpha, beta are weights, not
conjugated) with respect to the mesh coordinates (3 x Nc) and k, with the rule selection
held fixed (bie_assembly_vjp_device).  `target_node[i]` (1-based) marks target i as sitting
on mesh coordinate `target_node[i]` (it moves with the mesh); `targets === nothing` means
collocation at the sources.
"""
function assembly_vjp(dm::DeviceMesh, k::Real, alpha::AbstractVector, p::AbstractVector;
                      beta=nothing, targets=nothing, target_node=nothing, n=4, m=4)
    mesh = dm.mesh
    M = targets === nothing ? n_sources(dm) : size(targets, 2)
    Nc = size(mesh.coordinates, 2)
    dA = DeviceBuffer(Vector{ComplexF64}(alpha))
    dP = DeviceBuffer(Vector{ComplexF64}(p))
    dB = beta === nothing ? nothing : DeviceBuffer(Vector{ComplexF64}(beta))
    dT = targets === nothing ? nothing : DeviceBuffer(Matrix{Float64}(targets))
    dN = target_node === nothing ? nothing : DeviceBuffer(Vector{Int32}(target_node .- 1))
    gX = DeviceBuffer(zeros(Float64, 3 * Nc))
    gk = DeviceBuffer(zeros(Float64, 1))
    ptr(x) = x === nothing ? C_NULL : x.ptr
    check(ccall((:bie_assembly_vjp_device, libbie[]), Cint,
                (Ptr{Cvoid}, Cdouble, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                 Ptr{Cvoid}, Int32, Int32, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                dm.handle, Float64(k), ptr(dT), Int64(M), ptr(dN), dA.ptr, ptr(dB), dP.ptr,
                Int32(n), Int32(m), gX.ptr, gk.ptr, C_NULL))
    g = download!(Vector{Float64}(undef, 3 * Nc), gX)
    return reshape(g, 3, Nc), download!(Vector{Float64}(undef, 1), gk)[1]
end

plane_wave(x, k, d) = exp.(im * k .* vec(d' * x))

"""
    assemble_device!(dm, k, dF, dG, dC; targets=nothing, M=n_sources(dm), n=4, m=4,
                     fOn=true, gOn=true, cOn=true)

F, G, C into device buffers (bie_assemble_device), collocation or at the M device targets.
"""
function assemble_device!(dm::DeviceMesh, k::Number, dF::DeviceBuffer, dG::DeviceBuffer,
                          dC::DeviceBuffer; targets=nothing, M::Integer=n_sources(dm), n=4, m=4,
                          fOn=true, gOn=true, cOn=true)
    kc = ComplexF64(k)
    check(ccall((:bie_assemble_device, libbie[]), Cint,
                (Ptr{Cvoid}, Cdouble, Cdouble, Ptr{Cvoid}, Int64, Int32, Int32, UInt32,
                 Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                dm.handle, real(kc), imag(kc), targets === nothing ? C_NULL : targets.ptr,
                Int64(M), Int32(n), Int32(m), flagbits(fOn, gOn, cOn), dF.ptr, dG.ptr, dC.ptr,
                C_NULL))
    return nothing
end

lu_factor!(dA::DeviceBuffer, N, ipiv::DeviceBuffer, dC) =
    check(ccall((:bie_lu_factor_device, libbie[]), Cint,
                (Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                N, dA.ptr, N, dC === nothing ? C_NULL : dC.ptr, ipiv.ptr, C_NULL))

lu_solve!(dA::DeviceBuffer, N, ipiv::DeviceBuffer, dB::DeviceBuffer; trans=0) =
    check(ccall((:bie_lu_solve_device, libbie[]), Cint,
                (Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int32, Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}),
                N, dA.ptr, N, ipiv.ptr, Int32(trans), dB.ptr, N, 1, C_NULL))

gemv!(dy::DeviceBuffer, dA::DeviceBuffer, m, n, dx::DeviceBuffer; trans=0) =
    check(ccall((:bie_zgemv_device, libbie[]), Cint,
                (Int32, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                Int32(trans), m, n, dA.ptr, max(m, 1), dx.ptr, dy.ptr, C_NULL))

# The forward problem with F and F_field on the device: A = F + diag(1 - C) factored in
# place, p = A \ b, u = p_inc - F_field p; only vectors of N and P entries are downloaded.
function forward_device(dmesh::DeviceMesh, k::Real, field_points::AbstractMatrix, n, m,
                        direction)
    srcs = Matrix{Float64}(dmesh.mesh.sources)
    pts = Matrix{Float64}(field_points)
    N, P = size(srcs, 2), size(pts, 2)
    dF = DeviceBuffer(16 * N * N)
    dFf = DeviceBuffer(16 * P * N)
    dG = DeviceBuffer(16 * N * max(N, P))      # G is not needed (zero-filled)
    dC = DeviceBuffer(16 * N)
    dCf = DeviceBuffer(16 * P)
    dT = DeviceBuffer(pts)
    assemble_device!(dmesh, k, dF, dG, dC; n=n, m=m, gOn=false)
    assemble_device!(dmesh, k, dFf, dG, dCf; targets=dT, M=P, n=n, m=m, gOn=false, cOn=false)
    ipiv = DeviceBuffer(4 * N)
    lu_factor!(dF, N, ipiv, dC)
    b = plane_wave(srcs, k, direction)
    dp = DeviceBuffer(b)
    lu_solve!(dF, N, ipiv, dp)
    dfp = DeviceBuffer(16 * P)
    gemv!(dfp, dFf, P, N, dp)
    p = download!(Vector{ComplexF64}(undef, N), dp)
    pinc = plane_wave(pts, k, direction)
    u = pinc - download!(Vector{ComplexF64}(undef, P), dfp)
    return (; dF, dFf, ipiv, b, p, pinc, u, pts, srcs, N, P)
end

"""
    scattering_gradient(mesh, k, field_points; n=4, m=4, direction=[0, 0, 1]) -> (J, dJdX, dJdk)

J = sum |u|^2 of the field pressure u = p_inc - F_field p at `field_points`, with p solving
(F + diag(1 - C)) p = p_inc on the rigid body (examples/3d_rigid_sphere.jl:47-52,
examples/3d_rosebem.jl:81), and its gradient with respect to the mesh coordinates and k by
the adjoint method: A' lambda = F_field' u, then two vector-Jacobian products of the
assembly (rule selection held fixed).  F, F_field and the LU factors of A stay on the
device (bie_lu_factor_device, bie_lu_solve_device, bie_zgemv_device); only vectors of N
and P entries and the gradient come back.
"""
function scattering_gradient(mesh::BIE.Mesh3d, k::Real, field_points::AbstractMatrix; n=4,
                             m=4, direction=[0.0, 0.0, 1.0], dm=nothing)
    dmesh = dm === nothing ? DeviceMesh(mesh) : dm
    fw = forward_device(dmesh, k, field_points, n, m, direction)
    srcs, pts, b, p, u, pinc = fw.srcs, fw.pts, fw.b, fw.p, fw.u, fw.pinc
    J = real(sum(abs2, u))
    du = DeviceBuffer(u)
    dlam = DeviceBuffer(16 * fw.N)
    gemv!(dlam, fw.dFf, fw.P, fw.N, du; trans=2)      # F_field' u
    lu_solve!(fw.dF, fw.N, fw.ipiv, dlam; trans=2)     # A' lambda = F_field' u
    lam = download!(Vector{ComplexF64}(undef, fw.N), dlam)
    # surface nodes are mesh coordinates: find each source's coordinate index
    lut = Dict(Tuple(srcs[:, j]) => j for j in 1:size(srcs, 2))
    coord_of = zeros(Int, size(srcs, 2))
    for j in 1:size(mesh.coordinates, 2)
        t = Tuple(mesh.coordinates[:, j])
        haskey(lut, t) && (coord_of[lut[t]] = j)
    end
    all(>(0), coord_of) || throw(ArgumentError("collocation nodes are not mesh coordinates"))
    gs, gks = assembly_vjp(dmesh, k, conj.(lam), p; beta=-conj.(lam) .* p,
                           target_node=coord_of, n=n, m=m)
    gf, gkf = assembly_vjp(dmesh, k, -conj.(u), p; targets=pts, n=n, m=m)
    grad = 2 .* (gs .+ gf)
    dk = 2 * (gks + gkf)
    d = Vector{Float64}(direction)
    coef = real.(-conj.(lam) .* (im * k) .* b)
    for j in 1:size(srcs, 2), c in 1:3
        grad[c, coord_of[j]] += 2 * d[c] * coef[j]
    end
    dk += 2 * real(sum(-conj.(lam) .* im .* vec(d' * srcs) .* b))
    dk += 2 * real(sum(conj.(u) .* im .* vec(d' * pts) .* pinc))
    return J, grad, dk
end

"""
    apply(A, X) -> A * X

Dense complex apply on FP64 MFMA (bie_zgemm_device), e.g. the assembled F or G times a
block of right-hand sides.
"""
function apply(A::AbstractMatrix{ComplexF64}, X::AbstractVecOrMat{ComplexF64})
    Ah = Matrix{ComplexF64}(A)
    Xh = Matrix{ComplexF64}(reshape(X, size(X, 1), :))
    M, K = size(Ah)
    size(Xh, 1) == K || throw(DimensionMismatch("inner dimensions differ"))
    nrhs = size(Xh, 2)
    dA = DeviceBuffer(Ah); dX = DeviceBuffer(Xh); dY = DeviceBuffer(16 * M * nrhs)
    check(ccall((:bie_zgemm_device, libbie[]), Cint,
                (Int64, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64,
                 Ptr{Cvoid}),
                M, nrhs, K, dA.ptr, max(M, 1), dX.ptr, max(K, 1), dY.ptr, max(M, 1), C_NULL))
    Y = download!(Matrix{ComplexF64}(undef, M, nrhs), dY)
    return X isa AbstractVector ? vec(Y) : Y
end

"""
    gemv_batched!(dY, dF, N, dX, ldx, active) -> dY

y_q += F_q x_q on the device (bie_zgemv_batched_device) for the systems `active` (0-based)
of a batched sweep's output `dF` (F_q, N x N column-major, at byte offset 16 q N^2), with
x_q at entry offset q ldx of `dX` and y_q at q N of `dY`: the operator of Krylov solves over
all wavenumbers at once.  Waits for the device before returning.
"""
function gemv_batched!(dY::DeviceBuffer, dF::DeviceBuffer, N::Integer, dX::DeviceBuffer,
                       ldx::Integer, active::AbstractVector{<:Integer})
    dI = DeviceBuffer(Vector{Int32}(active))
    check(ccall((:bie_zgemv_batched_device, libbie[]), Cint,
                (Int64, Int64, Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64,
                 Ptr{Cvoid}, Int32, Ptr{Cvoid}),
                N, N, dF.ptr, Int64(N)^2, N, dX.ptr, ldx, dY.ptr, N, dI.ptr,
                Int32(length(active)), C_NULL))
    check(ccall((:bie_device_synchronize, libbie[]), Cint, ()))
    return dY
end

"""
    scattering_objective(mesh, k, field_points; n=4, m=4, direction=[0, 0, 1]) -> J

The objective of `scattering_gradient` alone.  With ChainRulesCore loaded, its `rrule`
returns the adjoint gradient with respect to the mesh coordinates (as a tangent of
`mesh.coordinates`) and k.
"""
function scattering_objective(mesh::BIE.Mesh3d, k::Real, field_points::AbstractMatrix; n=4,
                              m=4, direction=[0.0, 0.0, 1.0])
    fw = forward_device(DeviceMesh(mesh), k, field_points, n, m, direction)
    return real(sum(abs2, fw.u))
end

# ---------------------------------------------------------------------------------------
# Wavenumber derivatives and their Taylor expansion (src/3d/rosebem.jl)
# ---------------------------------------------------------------------------------------

"""
    taylor_assemble!(mesh, k, in_sources, shape_function; M=0, fOn=true, gOn=true, cOn=true,
                     m=3, n=3, progress=true, U=I) -> (From, Grom, C)

Same signature and defaults as the reference's `taylor_assemble!` (src/3d/rosebem.jl:56-157):
`From[:, :, q+1]` is the q-th k-derivative of F for q = 0..M, all orders from one device call
(bie_assemble_taylor_device), reduced to `U' * F_q * U` on the host when a basis U is given.
Unsupported meshes, quadrilaterals and M > 63 go to the reference method.
"""
function taylor_assemble!(mesh::BIE.Mesh3d, k, in_sources, shape_function; M=0, fOn=true,
                          gOn=true, cOn=true, m=3, n=3, progress=true, U=LinearAlgebra.I)
    if !supported(mesh) || quadrilateral(mesh) || M > 63
        return BIE.taylor_assemble!(mesh, k, in_sources, shape_function; M=M, fOn=fOn,
                                    gOn=gOn, cOn=cOn, m=m, n=n, progress=progress, U=U)
    end
    dm = DeviceMesh(mesh)
    N = n_sources(dm)
    dX = DeviceBuffer(Matrix{Float64}(in_sources))
    nt = size(in_sources, 2)
    nord = M + 1
    dF = fOn ? DeviceBuffer(16 * nt * N * nord) : nothing
    dG = gOn ? DeviceBuffer(16 * nt * N * nord) : nothing
    dC = DeviceBuffer(16 * nt)
    kc = ComplexF64(k)
    check(ccall((:bie_assemble_taylor_device, libbie[]), Cint,
                (Ptr{Cvoid}, Cdouble, Cdouble, Int32, Ptr{Cvoid}, Int64, Int32, Int32, UInt32,
                 Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid}),
                dm.handle, real(kc), imag(kc), Int32(nord), dX.ptr, nt, Int32(n), Int32(m),
                flagbits(fOn, gOn, cOn), fOn ? dF.ptr : C_NULL, gOn ? dG.ptr : C_NULL,
                Int64(nt) * N, dC.ptr, C_NULL))
    F = fOn ? download!(Array{ComplexF64}(undef, nt, N, nord), dF) : zeros(ComplexF64, nt, N, nord)
    G = gOn ? download!(Array{ComplexF64}(undef, nt, N, nord), dG) : zeros(ComplexF64, nt, N, nord)
    C = download!(Vector{ComplexF64}(undef, nt), dC)
    U isa LinearAlgebra.UniformScaling && return F, G, C
    From = cat((U' * F[:, :, q] * U for q in 1:nord)...; dims=3)
    Grom = cat((U' * G[:, :, q] * U for q in 1:nord)...; dims=3)
    return From, Grom, C
end

taylor_weight(k, k0, i) = i > 20 ? ComplexF64((k - k0)^i / factorial(Int128(i))) :
                                   ComplexF64((k - k0)^i / factorial(i))

function expand_device(basis::AbstractArray, k, k0)
    B = Array{ComplexF64}(basis)
    nt = size(B, ndims(B))
    n = length(B) ÷ nt
    coef = ComplexF64[taylor_weight(k, k0, i) for i in 0:nt-1]
    dA = DeviceBuffer(B)
    dY = DeviceBuffer(16 * n)
    GC.@preserve coef check(ccall((:bie_taylor_expand_device, libbie[]), Cint,
                                  (Int32, Int64, Ptr{Cvoid}, Int64, Ptr{ComplexF64}, Ptr{Cvoid},
                                   Ptr{Cvoid}),
                                  Int32(nt), n, dA.ptr, n, coef, dY.ptr, C_NULL))
    return download!(Array{ComplexF64}(undef, size(B)[1:end-1]...), dY)
end

"""
    apply_taylor_expansion(Abasis, k, k0[, bbasis])

The reference's `apply_taylor_expansion` (src/3d/rosebem.jl:164-199) as one streaming pass
on the device (bie_taylor_expand_device): sum over i of (k - k0)^i / i! Abasis[:, :, i+1]
(and the same for the vectors bbasis).  More than 64 terms go to the reference method.
"""
function apply_taylor_expansion(Abasis, k, k0)
    size(Abasis, 3) > 64 && return BIE.apply_taylor_expansion(Abasis, k, k0)
    return expand_device(Abasis, k, k0)
end

function apply_taylor_expansion(Abasis, bbasis, k, k0)
    size(Abasis, 3) > 64 && return BIE.apply_taylor_expansion(Abasis, bbasis, k, k0)
    return expand_device(Abasis, k, k0), expand_device(bbasis, k, k0)
end

end # module
