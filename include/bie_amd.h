/*
 * bie_amd.h — C ABI of the MI355X (gfx950) dense 3D collocation assembly library.
 *
 * Drop-in boundary for BoundaryIntegralEquations.jl's
 *     assemble_parallel!(mesh::Mesh3d, k, in_sources; fOn, gOn, cOn, m=4, n=4) -> (F, G, C)
 * (src/3d/assembly_collocation.jl:259-268 dispatcher, :270-353 linear triangles,
 *  :355-479 quadratic triangles).  All arrays are column-major (Julia order); complex
 * values are interleaved (re, im) float64 pairs, identical to Julia's ComplexF64 and
 * HIP's hipDoubleComplex.  Every entry point returns BIE_OK (0) or an error code and
 * leaves a message for bie_last_error() on the calling thread.
 */
#ifndef BIE_AMD_H
#define BIE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BIE_OK 0
#define BIE_EINVAL 1        /* bad argument: null pointer, size, index, kind  (ArgumentError) */
#define BIE_EUNSUPPORTED 2  /* element family or entry point outside what this library covers */
#define BIE_ENOMEM 3        /* device or host allocation failed                              */
#define BIE_EHIP 4          /* HIP runtime error                                             */

#define BIE_TRI3 3          /* TriangularLinear    (src/3d/SurfaceFunction.jl:31)  */
#define BIE_TRI6 6          /* TriangularQuadratic (src/3d/SurfaceFunction.jl:58)  */
/* Discontinuous physics (physics_kind only; geometry stays BIE_TRI3 or BIE_TRI6): one set
 * of nodes per element inside it, the low byte is the node count.  Nodes and bases
 * (src/3d/SurfaceFunction.jl:429-438, src/3d/surface_functions.jl:234-318), with
 * psi(s) = (s - beta) / (1 - 3 beta):
 *   BIE_DTRI1  DiscontinuousTriangularConstant   N = 1, node (1/3, 1/3)
 *   BIE_DTRI3  DiscontinuousTriangularLinear     N = psi(1-u-v), psi(u), psi(v)
 *   BIE_DTRI6  DiscontinuousTriangularQuadratic  the quadratic Lagrange basis in those three
 * The mesh's sources are the geometry at the nodes (compute_sources, src/3d/mesh.jl:116-148)
 * and physics_topology numbers them element by element; beta is bie_mesh.physics_beta. */
#define BIE_DTRI1 0x101
#define BIE_DTRI3 0x103
#define BIE_DTRI6 0x106
/* Quadrilaterals (geometry and physics; the low byte is the node count).  They take the
 * reference's generic path (assemble_parallel!(mesh, k, sources, ::SurfaceFunction),
 * assembly_collocation.jl:481-538): one tensor Gauss-Legendre rule on [-1,1]^2 for every
 * pair, m points in u and n in v (create_shape_function calls getQuadpoints(sf, m, n)), no
 * rule selection.  Node orders (src/3d/surface_functions.jl:1-149, SurfaceFunction.jl:439-470):
 *   BIE_QUAD4   QuadrilateralLinear4 (COMSOL): (-1,-1) (1,-1) (-1,1) (1,1)
 *   BIE_QUAD4S  QuadrilateralLinearSerendipity: counter-clockwise from (-1,-1)
 *   BIE_QUAD8   QuadrilateralQuadratic: corners counter-clockwise, then (0,-1) (1,0) (0,1) (-1,0)
 *   BIE_QUAD9   QuadrilateralQuadraticLagrange (COMSOL): (-1,-1) (1,-1) (-1,1) (1,1) (0,-1)
 *               (-1,0) (0,0) (1,0) (0,1); odd n and m together are refused as the reference
 *               refuses them ("Gauss Node On Singularity", :497-504)
 * Physics on quadrilateral geometry: the geometry kind itself, BIE_QUAD4 on BIE_QUAD9 (the
 * loader's physics_order = :linear, corner nodes first), or the discontinuous
 *   BIE_DQUAD1  DiscontinuousQuadrilateralConstant, node (0,0)
 *   BIE_DQUAD4  DiscontinuousQuadrilateralLinear4: the BIE_QUAD4 basis of (u, v) / (1 - beta)
 *   BIE_DQUAD9  DiscontinuousQuadrilateralQuadraticLagrange: BIE_QUAD9 of (u, v) / (1 - beta) */
#define BIE_QUAD4 0x204
#define BIE_QUAD4S 0x404
#define BIE_QUAD8 0x208
#define BIE_QUAD9 0x209
#define BIE_DQUAD1 0x301
#define BIE_DQUAD4 0x304
#define BIE_DQUAD9 0x309

#define BIE_F 1u            /* fOn */
#define BIE_G 2u            /* gOn */
#define BIE_C 4u            /* cOn */
#define BIE_FGC 7u

/* bie_mesh.flags: rule selection (find_closest_corner, assembly_collocation.jl:102-115)
 * compares hypot distances, and Julia releases differ in how hypot(x, y, z) rounds:
 * 0 = maxabs * sqrt(sum((x_i/maxabs)^2)) (Julia 1.9, the reference's CI version),
 * BIE_MESH_SCALED_HYPOT = power-of-two scaled sqrt(x^2 + y^2 + z^2) (later releases).
 * The choice decides the rule of exactly tied pairs, which symmetric meshes have. */
#define BIE_MESH_SCALED_HYPOT 1

/* Mesh3d (src/utils/mesh.jl:23-34), the fields assembly reads. */
typedef struct bie_mesh {
  const double* coordinates;        /* 3 x n_coordinates                              */
  const int64_t* topology;          /* nodes x n_elements (geometry_kind & 0xff)       */
  const double* sources;            /* 3 x n_sources (collocation nodes)               */
  const int64_t* physics_topology;  /* nodes x n_elements (nodes = physics_kind & 0xff),
                                       indexes sources                                  */
  int64_t n_coordinates;
  int64_t n_elements;
  int64_t n_sources;
  int32_t geometry_kind;            /* BIE_TRI3 | BIE_TRI6 | BIE_QUAD4 | BIE_QUAD4S | BIE_QUAD8 |
                                       BIE_QUAD9                                        */
  int32_t physics_kind;             /* a geometry kind, BIE_DTRI1 | BIE_DTRI3 | BIE_DTRI6,
                                       BIE_DQUAD1 | BIE_DQUAD4 | BIE_DQUAD9 (see above)   */
  int32_t index_base;               /* 1 for Julia arrays, 0 for C/Python              */
  int32_t flags;                    /* 0 or BIE_MESH_SCALED_HYPOT                      */
  double physics_beta;              /* node offset beta of BIE_DTRI3 / BIE_DTRI6 /
                                       BIE_DQUAD4 / BIE_DQUAD9 (the physics function's
                                       beta field), else ignored                        */
} bie_mesh;

typedef struct bie_dmesh bie_dmesh;  /* device-resident mesh + cached rule tables */
typedef struct bie_graph bie_graph;  /* an instantiated HIP graph of one batched assembly */
typedef struct bie_batch bie_batch;  /* an uneven batch of device meshes, one rule set */
typedef struct bie_sparse bie_sparse;  /* pattern and rules of a sparse near-field assembly */

const char* bie_last_error(void);
int bie_version(void);

/* ---- Host-array drop-in ------------------------------------------------------------
 * Replaces assemble_parallel!(mesh, k, in_sources; fOn, gOn, cOn, m, n)
 * (assembly_collocation.jl:259-268).  k = k_re + i k_im (complex k as in the lossy caller,
 * LossyGlobalMatrix.jl:119).  targets: 3 x M (NULL = mesh sources).  F, G: M x N complex,
 * C: M complex, caller-allocated; disabled outputs are zero-filled (as :308-310).
 * Rule selection for any target reads mesh sources (defined behaviour; the reference
 * indexes the targets there, :325/:436, which is only valid for collocation).
 * The device form of the last mesh (upload, rule tables, schedule, recorded mid-edge pairs)
 * is kept for the next call: a call whose mesh matches it byte for byte skips that setup
 * (a frequency loop calling assemble_parallel! once per k); device memory it holds is
 * released when another mesh replaces it or the process ends.  Outputs of 256 MiB or more
 * are copied back through two pinned 128 MiB host buffers (allocated on first use, kept for
 * the process) with host threads. */
int bie_assemble(const bie_mesh* mesh, double k_re, double k_im, const double* targets,
                 int64_t n_targets, int32_t n, int32_t m, uint32_t flags, double* F,
                 double* G, double* C);

/* State of the host-array entry point's kept mesh (no counterpart in the reference):
 * calls that reused it (hits), meshes uploaded (builds), its n_sources (-1: none kept) and
 * the entries of its mid-edge pair queue (0 when none is held; any pointer may be NULL).
 * Targets equal to the mesh sources value for value are collocation; the pair queue of
 * other targets is released when the call returns. */
int bie_host_cache_info(int64_t* hits, int64_t* builds, int64_t* n_sources,
                        int64_t* queue_entries);
/* Releases the kept mesh (its device memory is freed when no call is using it) and the
 * pinned staging buffers; the next bie_assemble builds them again. */
int bie_host_cache_release(void);

/* ---- Sparse near field -----------------------------------------------------------------
 * Replaces sparse_assemble_parallel!(mesh, k, sources, physics_function; fOn, gOn, depth,
 * offset) (src/3d/assembly_collocation_losses.jl:200-405; assemble_parallel! with
 * sparse=true, src/3d/assembly_collocation.jl:261-263): F and G on the pairs (source i,
 * element e) with e connected to node i through at most `depth` (1 or 2) elements
 * (connected_sources, assembly_collocation_losses.jl:98-132), each integrated with the
 * subdivision rule (singular_triangle_integration, src/3d/adaptive_integration.jl:172-246)
 * clustered at the element's geometry node closest to source i.  The reference returns
 * sparse(I, J, F), sparse(I, J, G): I = the row of every stored entry, J = its column (the
 * connected sources of each row, ascending), in that order; these entry points give the
 * same I, J (index_base of the mesh) and values.  offset: the reference's `offset` keyword
 * (NaN = nothing, i.e. 0 for continuous physics).  No C (free term) is formed there.
 * Triangular geometry; continuous or discontinuous physics.  Discontinuous physics
 * (BIE_DTRI1/3/6) follows the reference's branch for it (:233-257 and connected_sources,
 * :41-96): row r is node r mod n of element r / n (n nodes per element, so n_sources must
 * be n_elements * n), connected at depth 1 to its own element and at depth 2 to every
 * element sharing a corner with it; each pair is integrated with the polar rule
 * (create_polar_elements, :409-447: polar_gaussian(Ngauss), src/3d/quadrature.jl:196-206,
 * on three sub-triangles) around the physics node whose index is that of the closest
 * geometry node (the reference's filler rules for the missing ones: nodes 1, 2, 3; the
 * constant element: node 1).  offset is not used there. */
int bie_sparse_count(const bie_mesh* mesh, int32_t depth, int64_t* nnz);
/* sources: 3 x n_sources (NULL = mesh sources; n_sources must equal the mesh's), flags
 * BIE_F | BIE_G (a disabled output is zero-filled), I, J (nnz int64), F, G (nnz complex)
 * host arrays. */
int bie_sparse_assemble(const bie_mesh* mesh, double k_re, double k_im, const double* sources,
                        int64_t n_sources, int32_t depth, double offset, uint32_t flags,
                        int64_t nnz, int64_t* I, int64_t* J, double* F, double* G);
/* The same with the reference's Ngauss keyword (the polar rules' order for discontinuous
 * physics, 1..400; the reference's default 40, which bie_sparse_assemble uses, and forced to
 * 40 for BIE_DTRI6 as the reference does; not used for continuous physics). */
int bie_sparse_assemble_ex(const bie_mesh* mesh, double k_re, double k_im,
                           const double* sources, int64_t n_sources, int32_t depth,
                           double offset, int32_t n_gauss, uint32_t flags, int64_t nnz,
                           int64_t* I, int64_t* J, double* F, double* G);
/* Device-resident form: the pattern, pairs and rules built once per (mesh, depth, offset),
 * then any number of assemblies into device arrays of nnz complex entries (asynchronous on
 * stream; d_sources NULL = the mesh sources).  Destroy the handle before its dmesh. */
int bie_sparse_create(bie_dmesh* dm, int32_t depth, double offset, bie_sparse** out);
/* The same with Ngauss (as bie_sparse_assemble_ex). */
int bie_sparse_create_ex(bie_dmesh* dm, int32_t depth, double offset, int32_t n_gauss,
                         bie_sparse** out);
int bie_sparse_destroy(bie_sparse* sp);
int64_t bie_sparse_nnz(const bie_sparse* sp);
int bie_sparse_pattern(const bie_sparse* sp, int32_t index_base, int64_t* I, int64_t* J);
int bie_sparse_assemble_device(bie_sparse* sp, double k_re, double k_im, const double* d_sources,
                               uint32_t flags, double* dF, double* dG, void* stream);

/* ---- Partial (fast-operator) near field -------------------------------------------------
 * Replaces partial_assemble_parallel!(mesh, k, in_sources, shape_function::Triangular; fOn,
 * gOn, n=3, depth=1) (src/3d/fast_multipole_method.jl:11-81): the same pattern as the
 * sparse near field (connected_sources), every pair integrated with the one symmetric
 * rule gauss_points_triangle(n_gauss) (src/3d/adaptive_integration.jl:13-64; n_gauss = 1,
 * 3, 4, 7, 12 or 13), no selection and no subdivision - the quadrature a fast operator
 * applies to the near field, which setup_fast_operator (:163-192) subtracts with it.  Same
 * I, J and value layout as bie_sparse_assemble; nnz from bie_sparse_count. */
int bie_partial_assemble(const bie_mesh* mesh, double k_re, double k_im, const double* sources,
                         int64_t n_sources, int32_t depth, int32_t n_gauss, uint32_t flags,
                         int64_t nnz, int64_t* I, int64_t* J, double* F, double* G);
/* Device-resident form on a bie_sparse handle (its pattern; its offset is not used). */
int bie_sparse_partial_device(bie_sparse* sp, double k_re, double k_im, int32_t n_gauss,
                              const double* d_sources, uint32_t flags, double* dF, double* dG,
                              void* stream);

/* ---- Device-resident path ------------------------------------------------------------
 * Uploads the mesh once (the element loop's inputs, Mesh3d src/utils/mesh.jl:23-34). */
int bie_dmesh_create(const bie_mesh* mesh, int device, bie_dmesh** out);
int bie_dmesh_destroy(bie_dmesh* dm);
int64_t bie_dmesh_n_sources(const bie_dmesh* dm);
int64_t bie_dmesh_n_elements(const bie_dmesh* dm);
/* Device copy of the collocation nodes (3 x N), for use as targets. */
const double* bie_dmesh_sources(const bie_dmesh* dm);

/* Same semantics as bie_assemble, with device pointers and an optional hipStream_t
 * (NULL = the legacy default stream).  d_targets NULL = collocation at mesh sources.
 * Asynchronous with respect to the host.  Calls on one handle share its device scratch
 * (element tables, schedules, mid-edge pair lists): each call's work waits for the
 * previous call's on that handle when the two use different streams, so calls may come
 * from several threads and streams.  Work captured into a graph (bie_sweep_graph_create,
 * or a caller's own capture) is not ordered this way: its launches must not overlap other
 * work on the same handle. */
int bie_assemble_device(bie_dmesh* dm, double k_re, double k_im, const double* d_targets,
                        int64_t n_targets, int32_t n, int32_t m, uint32_t flags, double* dF,
                        double* dG, double* dC, void* stream);

/* Rule index 1..R chosen for every (target i, element e), written as int8 to d_rules
 * [i * n_elements + e] (find_closest_corner, assembly_collocation.jl:102-115). */
int bie_select_rules_device(bie_dmesh* dm, const double* d_targets, int64_t n_targets,
                            int8_t* d_rules, void* stream);

/* Element geometry at rule c (1..6) points: y (E x Q x 3), normal (E x Q x 3, the
 * reference's negated normal), J*w (E x Q): interpolate_elements
 * (assembly_collocation.jl:27-66, jacobian.jl:78-87).  Host output arrays. */
int bie_element_geometry(bie_dmesh* dm, int32_t rule, int32_t n, int32_t m, double* y,
                         double* normals, double* jw);

/* Builds the LDS-slot schedule of the assembly kernel for a host mesh and replays it,
 * checking every invariant the kernel relies on (no counterpart in the reference; the
 * reference's threads own whole rows, assembly_collocation.jl:314), for steps of one and of
 * two elements per wave.  stats (6 values, one element per wave): segments, steps, max
 * live slots, shared nodes, zeroed columns, flush entries. */
int bie_schedule_check(const bie_mesh* mesh, int32_t min_segments, int64_t* stats);

/* ---- Batched frequency sweep ------------------------------------------------------------
 * F_q, G_q for nk real wavenumbers d_k[q] (device array) on one device mesh, equal to nk
 * calls of bie_assemble_device(k = d_k[q]) (each the reference's assemble_parallel!,
 * assembly_collocation.jl:259-479, which a frequency sweep calls once per k).  F_q starts
 * at dF + 2*q*stride doubles (stride >= n_targets * n_sources complex elements); likewise
 * G_q.  C is k-independent and written once to dC.  Everything k-independent (rule
 * selection, geometry, table loads, 1/r, n.d) is shared by the wavenumbers of a block, and
 * uniformly spaced blocks replace sincos by a recurrence.  A disabled F / G (flag off) may
 * be NULL.  Asynchronous on stream. */
int bie_assemble_sweep_device(bie_dmesh* dm, const double* d_k, int32_t nk,
                              const double* d_targets, int64_t n_targets, int32_t n, int32_t m,
                              uint32_t flags, double* dF, double* dG, int64_t stride,
                              double* dC, void* stream);
/* Tile-shape autotuner of the sweep: times the sweep with these arguments for every
 * candidate tile, 64 targets (one per lane) x elements per tile in {16, 32, 64, 128, the
 * whole surface}, each with as many LDS slots as fit or with 32, and keeps the fastest for
 * later sweeps and sweep graphs on this handle (the outputs hold the last candidate's
 * result, equal to any other's up to the order of shared-column additions).  Reports the
 * chosen elements per tile (0: whole surface), slots (0: as many as fit) and its time.  No
 * counterpart in the reference.  Synchronous. */
int bie_sweep_tune(bie_dmesh* dm, const double* d_k, int32_t nk, const double* d_targets,
                   int64_t n_targets, int32_t n, int32_t m, uint32_t flags, double* dF,
                   double* dG, int64_t stride, double* dC, void* stream,
                   int32_t* elems_per_tile, int32_t* slots, double* best_ms);
/* The same sweep captured once into a HIP graph (kernels only, no memset nodes); launching
 * the graph replays it on any stream.  The graph keeps the pointers it was built with. */
int bie_sweep_graph_create(bie_dmesh* dm, const double* d_k, int32_t nk,
                           const double* d_targets, int64_t n_targets, int32_t n, int32_t m,
                           uint32_t flags, double* dF, double* dG, int64_t stride, double* dC,
                           bie_graph** out);
int bie_graph_launch(bie_graph* g, void* stream);
int bie_graph_destroy(bie_graph* g);

/* ---- Wavenumber derivatives (Taylor expansion in k) ------------------------------------
 * d^q F / dk^q and d^q G / dk^q for q = 0 .. n_orders-1 (n_orders <= 64) at one complex
 * wavenumber k_re + i k_im: replaces taylor_assemble! (src/3d/rosebem.jl:56-157, kernels
 * taylor_greens3d! / taylor_freens3d!, src/3d/kernels.jl:94-118) with U = I.  Order q goes
 * to dF + 2*q*stride doubles (stride >= n_targets * n_sources complex elements), column-
 * major n_targets x n_sources: the reference's F[:, :, q+1].  Likewise G; C (order
 * independent) once to dC.  Order 0 equals bie_assemble_device.  The reference defines the
 * method for linear triangles; quadratic meshes are accepted as well.  Same rule selection,
 * tables and kernel schedule as the frequency sweep (one e^{ikr} per point serves every
 * order).  A disabled F / G (flag off) may be NULL.  Asynchronous on stream. */
int bie_assemble_taylor_device(bie_dmesh* dm, double k_re, double k_im, int32_t n_orders,
                               const double* d_targets, int64_t n_targets, int32_t n, int32_t m,
                               uint32_t flags, double* dF, double* dG, int64_t stride,
                               double* dC, void* stream);

/* Y = sum_t coef_t A_t for t = 0 .. n_terms-1 (1 <= n_terms <= 64): A_t are n complex128
 * entries at dA + 2*t*strideA doubles (strideA >= n when n_terms > 1), coef 2*n_terms
 * doubles (interleaved complex) in host memory, read during the call.  With coef_t =
 * (k - k0)^t / t! this is apply_taylor_expansion (src/3d/rosebem.jl:164-199) on the
 * derivative matrices of bie_assemble_taylor_device (or on right-hand-side vectors).  dY may
 * equal dA (the order-0 array is then overwritten); no other overlap.  Asynchronous on
 * stream. */
int bie_taylor_expand_device(int32_t n_terms, int64_t n, const double* dA, int64_t strideA,
                             const double* coef, double* dY, void* stream);

/* ---- Uneven batch ---------------------------------------------------------------------
 * Collocation assembly (targets = each mesh's sources) of many meshes of different sizes,
 * each with its own wavenumber, in one persistent kernel launch: every workgroup stages
 * the shared basis table once and takes (instance, 64-row block, element segment) tiles
 * from a device counter.  All meshes share geometry and physics kinds and the rule set
 * (n, m); each instance i writes its own F_i, G_i (N_i x N_i, column-major complex128) and
 * C_i (N_i), same values as bie_assemble_device on that mesh.  Arrays of per-instance
 * values (k_re, k_im, dF, dG, dC) live on the host; the pointers in dF / dG / dC are device
 * pointers.  k_im may be NULL (real wavenumbers).  A disabled F / G (flag off) may be NULL;
 * if given, it is zeroed.  No counterpart in the reference, whose assembly handles one
 * mesh and one k per call (assembly_collocation.jl:259-268); a batch of uneven meshes was
 * handled there by looping that call.  The batch keeps the dmesh pointers: destroy it
 * before them. */
int bie_batch_create(bie_dmesh* const* meshes, int32_t n_inst, int32_t n, int32_t m,
                     bie_batch** out);
int bie_batch_destroy(bie_batch* b);
/* Schedule of the persistent launch: target elements per segment (>= 8), LDS slots per
 * workgroup (32 or 64, capped at what fits beside the basis tables and the staged rows;
 * 0: as many as fit) and tile order (0: longest tiles first, 1: interleaved across
 * instances).  Takes effect at the next call. */
int bie_batch_configure(bie_batch* b, int32_t elems_per_tile, int32_t slots, int32_t order);
int bie_batch_info(bie_batch* b, int32_t* elems_per_tile, int32_t* slots, int32_t* order,
                   int64_t* n_tiles);
int bie_batch_assemble(bie_batch* b, const double* k_re, const double* k_im, uint32_t flags,
                       double* const* dF, double* const* dG, double* const* dC, void* stream);
/* Autotuner: times every schedule in {32, 64, 128} elements per tile x {32, 64} slots x
 * {longest-first, interleaved} on these inputs (synchronising), keeps the fastest and
 * leaves its outputs in dF / dG / dC; *best_ms (may be NULL) receives its time. */
int bie_batch_tune(bie_batch* b, const double* k_re, const double* k_im, uint32_t flags,
                   double* const* dF, double* const* dG, double* const* dC, void* stream,
                   double* best_ms);

/* Batched derivative in k, one launch over the batch's instances: for every instance i
 *   w_i = alpha_i^T (dF_i/dk) p_i     (complex; F_i its collocation F at k_re[i] + i k_im[i])
 * written to d_w[2 i], d_w[2 i + 1] (device array of 2 n_inst doubles, overwritten).  alpha_i
 * (N_i complex) and p_i (N_i complex) are device arrays listed in host arrays; k_im may be
 * NULL.  dF/dk = -k e^{ikr} (n.d) J w / (4 pi r) per quadrature point (the order-1 term of
 * the reference's taylor_freens3d!, src/3d/kernels.jl:104-118), with the assembly's rules.
 * For a real k and Phi_i = Re[alpha_i^T F_i p_i], dPhi_i/dk = Re(w_i): the k-part of an
 * adjoint gradient (alpha_i from the adjoint solve of instance i).  Waits for the launch.
 * No counterpart in the reference (it differentiates by finite differences,
 * examples/2d_optimization.jl:184-216). */
/* The batched assembly (bie_batch_assemble with these arguments) captured into a HIP graph:
 * kernels only (fill kernels for the zeroing, no memset nodes), with the graph's own copies of
 * the instance table, the tile list and the tile counter, and its own mid-edge pair queues
 * (quadratic physics); launching it replays the batch on any stream (bie_graph_launch),
 * destroying it frees those (bie_graph_destroy).  The graph keeps the output pointers and
 * the meshes' rule tables and schedules: replays must not overlap other work on the batch
 * or its meshes, and reconfiguring or tuning the batch, assembling its meshes with another
 * rule set (n, m), or destroying either invalidates it.
 * No counterpart in the reference (a persistent launch captured in a graph is the GPU form
 * of its per-mesh loop). */
int bie_batch_graph_create(bie_batch* b, const double* k_re, const double* k_im,
                           uint32_t flags, double* const* dF, double* const* dG,
                           double* const* dC, bie_graph** out);
int bie_batch_dfdk_device(bie_batch* b, const double* k_re, const double* k_im,
                          const double* const* d_alpha, const double* const* d_p, double* d_w,
                          void* stream);

/* ---- Adjoint building block ------------------------------------------------------------
 * Vector-Jacobian product of the assembly at real k with the rule selection held fixed:
 *   d_grad_coords[3*v + d] += d/dX[v,d]  Re[ sum_i alpha_i sum_j F_ij p_j + sum_i beta_i C_i ]
 *   *d_grad_k             += d/dk      (same functional)
 * X = mesh coordinates (3 x n_coordinates, the array given to bie_dmesh_create).  alpha
 * (n_targets complex), beta (n_targets complex, NULL = 0) and p (n_sources complex) are
 * device arrays.  d_target_node[i] >= 0 marks target i as sitting on mesh coordinate
 * d_target_node[i] (it moves with X); -1 (or a NULL array) keeps it fixed (field points).
 * The pair quadrature is recomputed, nothing is stored.  No counterpart in the reference,
 * which differentiates by finite differences (examples/2d_optimization.jl:184-216); the
 * gradient of a scattering objective through A = F + diag(1 - C) and its solve is
 * composed from two calls (surface and field operators) on the host. */
int bie_assembly_vjp_device(bie_dmesh* dm, double k, const double* d_targets, int64_t n_targets,
                            const int32_t* d_target_node, const double* d_alpha,
                            const double* d_beta, const double* d_p, int32_t n, int32_t m,
                            double* d_grad_coords, double* d_grad_k, void* stream);

/* As bie_assembly_vjp_device at a complex wavenumber k_re + i k_im (the lossy systems'
 * kernels, e.g. assemble_parallel!(mesh, k_a, sources; ...) at
 * src/3d/LossyGlobalMatrix.jl:119 with complex k_a): d_grad_k (2 doubles, accumulated, or
 * NULL) receives d/dRe k and d/dIm k of the same functional.  No counterpart in the
 * reference. */
int bie_assembly_vjp_complex_device(bie_dmesh* dm, double k_re, double k_im,
                                    const double* d_targets, int64_t n_targets,
                                    const int32_t* d_target_node, const double* d_alpha,
                                    const double* d_beta, const double* d_p, int32_t n,
                                    int32_t m, double* d_grad_coords, double* d_grad_k,
                                    void* stream);

/* ---- Dense operator apply --------------------------------------------------------------
 * Y = A X for complex128 column-major A (M x K, leading dimension lda), X (K x nrhs, ldx),
 * Y (M x nrhs, ldy), all device pointers, on FP64 MFMA; asynchronous on stream.
 * Replaces the BLAS ZGEMM behind `F*X` / `G*vs` in the reference's callers
 * (examples/3d_pulsating_sphere.jl:42,62), batched like the FMM operators' 32-RHS path
 * (src/3d/fast_multipole_method.jl:249-287).  Uses a cached per-device workspace of
 * 16 * K * ceil(nrhs/32)*32 bytes, shared by all calls on the device: a call on another
 * stream than the previous one waits for it on the device.  Under graph capture the
 * workspace cannot grow (BIE_EINVAL; make one call of that size outside the capture first),
 * and replays of the graph must not overlap other calls.  The kernel addresses each K split
 * with 32-bit buffer offsets: lda and ceil(nrhs/32)*32 up to 1.6e6 (BIE_EINVAL beyond). */
int bie_zgemm_device(int64_t M, int64_t nrhs, int64_t K, const double* dA, int64_t lda,
                     const double* dX, int64_t ldx, double* dY, int64_t ldy, void* stream);

/* As bie_zgemm_device, allocating nothing: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops).
 * The packed X goes to a 16 MiB device array that the library holds on every GPU it runs
 * on until the process ends, instead of the growable workspace: K * ceil(nrhs/32)*32 must
 * not exceed 1048576 complex entries (e.g. K = 10242 with up to 96 right-hand sides);
 * larger calls return BIE_EINVAL.  Y is cleared by a fill kernel, never by a memset.  All
 * calls on one GPU share that array: calls that overlap on one GPU (other streams, other
 * threads, replays of graphs that captured it) race and give wrong results; order them. */
int bie_zgemm_device_noalloc(int64_t M, int64_t nrhs, int64_t K, const double* dA,
                             int64_t lda, const double* dX, int64_t ldx, double* dY,
                             int64_t ldy, void* stream);

/* Batched matrix-vector product for Krylov solves over a frequency sweep's operators:
 *   y_q += A_q x_q   for q = d_idx[0 .. n_idx)  (device int32 array, n_idx <= 65535)
 * A_q column-major M x K complex128 at dA + 2 q strideA doubles (leading dimension lda),
 * x_q at dX + 2 q ldx doubles (K entries), y_q at dY + 2 q ldy doubles (M entries, added
 * to: initialise y first).  Strides and leading dimensions count complex entries; the
 * batched sweep's F_k (stride N*N) are the intended A_q.  Asynchronous on stream.  No
 * counterpart in the reference, which solves one system at a time
 * (examples/3d_rigid_sphere.jl:47-52). */
int bie_zgemv_batched_device(int64_t M, int64_t K, const double* dA, int64_t strideA,
                             int64_t lda, const double* dX, int64_t ldx, double* dY,
                             int64_t ldy, const int32_t* d_idx, int32_t n_idx, void* stream);

/* ---- Solves on the device ----------------------------------------------------------
 * The reference solves the collocation system A p = b with A = F + diag(1 - C) by Julia's
 * `\` (LAPACK getrf/getrs; examples/3d_rigid_sphere.jl:47-52, test/3d/assembly_test.jl:37)
 * or IterativeSolvers.gmres, and forms the field pressure p_inc - F_field p
 * (examples/3d_rosebem.jl:81).  These keep F where the assembly wrote it.  Matrices are
 * column-major complex128; dimensions and leading dimensions below 2^31. */

/* A := F + diag(1 - C) (dC NULL: A := F) in place at dA (n x n, leading dimension lda),
 * then its LU factorisation with partial pivoting in place (rocSOLVER zgetrf); pivots to
 * d_ipiv (n int32, device).  Waits for the factorisation: a singular matrix returns
 * BIE_EINVAL naming the zero pivot's column. */
int bie_lu_factor_device(int64_t n, double* dA, int64_t lda, const double* dC, int32_t* d_ipiv,
                         void* stream);
/* B := op(A)^-1 B with the factors of bie_lu_factor_device: op(A) = A (trans 0), A^T (1)
 * or A^H (2); B n x nrhs (leading dimension ldb).  Asynchronous (rocSOLVER zgetrs). */
int bie_lu_solve_device(int64_t n, const double* dA, int64_t lda, const int32_t* d_ipiv,
                        int32_t trans, double* dB, int64_t ldb, int64_t nrhs, void* stream);
/* y := op(A) x for A m x n (leading dimension lda): trans 0 (y: m entries), 1 (A^T) or 2
 * (A^H; y: n entries).  Asynchronous (rocBLAS zgemv). */
int bie_zgemv_device(int32_t trans, int64_t m, int64_t n, const double* dA, int64_t lda,
                     const double* dx, double* dy, void* stream);
/* Restarted GMRES (Arnoldi with classical Gram-Schmidt applied twice, Givens rotations,
 * restart after `restart` steps, at most maxiter steps, stop at ||b_q - A_q x_q|| <=
 * tol ||b_q||) for the n_systems systems (F_q + diag(1 - C)) x_q = b_q at once: F_q at
 * dF + 2*q*strideF doubles (leading dimension ldf: a frequency sweep's output), C shared
 * (n complex), b_q at dB + 2*q*ldb doubles, x_q written at dX + 2*q*ldx doubles (x0 = 0).
 * The Arnoldi vectors stay on the device; the small Hessenberg problems are solved on the
 * host, one small transfer per step for all systems.  Synchronous; iterations (int32) and
 * relative residuals (double) per system go to the host arrays. */
int bie_gmres_batched_device(int64_t n, int32_t n_systems, const double* dF, int64_t strideF,
                             int64_t ldf, const double* dC, const double* dB, int64_t ldb,
                             double* dX, int64_t ldx, double tol, int32_t restart,
                             int32_t maxiter, int32_t* iterations, double* residuals,
                             void* stream);

/* ---- Kernel timing ------------------------------------------------------------------
 * While enabled, every device assembly records four HIP events on its launch stream:
 * call start, before the dense assembly kernel, between it and the mid-edge kernel (the
 * pairs that need the split rules 4-6; zero time on linear physics), and after it.
 * bie_timing_collect waits for the recorded events, returns the summed dense-kernel,
 * mid-edge-kernel and whole-call durations (ms) and the number of calls, and clears the
 * record.  No counterpart in the reference (it only wraps calls in @time,
 * examples/sphere.jl:51). */
int bie_timing_enable(int on);
int bie_timing_collect(double* kernel_ms, double* mid_ms, double* call_ms,
                       int64_t* n_calls);

/* ---- Device memory helpers (for callers without their own allocator) -------------- */
int bie_device_malloc(void** ptr, int64_t bytes);
int bie_device_free(void* ptr);
int bie_memcpy_h2d(void* dst, const void* src, int64_t bytes);
int bie_memcpy_d2h(void* dst, const void* src, int64_t bytes);
int bie_device_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif /* BIE_AMD_H */
