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
#define BIE_EUNSUPPORTED 2  /* element family outside linear/quadratic triangles            */
#define BIE_ENOMEM 3        /* device or host allocation failed                              */
#define BIE_EHIP 4          /* HIP runtime error                                             */

#define BIE_TRI3 3          /* TriangularLinear    (src/3d/SurfaceFunction.jl:31)  */
#define BIE_TRI6 6          /* TriangularQuadratic (src/3d/SurfaceFunction.jl:58)  */

#define BIE_F 1u            /* fOn */
#define BIE_G 2u            /* gOn */
#define BIE_C 4u            /* cOn */
#define BIE_FGC 7u

/* Mesh3d (src/utils/mesh.jl:23-34), the fields assembly reads. */
typedef struct bie_mesh {
  const double* coordinates;        /* 3 x n_coordinates                              */
  const int64_t* topology;          /* geometry_kind x n_elements                     */
  const double* sources;            /* 3 x n_sources (collocation nodes)               */
  const int64_t* physics_topology;  /* physics_kind x n_elements, indexes sources      */
  int64_t n_coordinates;
  int64_t n_elements;
  int64_t n_sources;
  int32_t geometry_kind;            /* BIE_TRI3 | BIE_TRI6 */
  int32_t physics_kind;             /* BIE_TRI3 | BIE_TRI6 */
  int32_t index_base;               /* 1 for Julia arrays, 0 for C/Python              */
  int32_t reserved;
} bie_mesh;

typedef struct bie_dmesh bie_dmesh;  /* device-resident mesh + cached rule tables */
typedef struct bie_graph bie_graph;  /* an instantiated HIP graph of one batched assembly */
typedef struct bie_batch bie_batch;  /* an uneven batch of device meshes, one rule set */

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
 * Asynchronous with respect to the host. */
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
 * reference's threads own whole rows, assembly_collocation.jl:314).  stats (6 values):
 * segments, steps, max live slots, shared nodes, zeroed columns, flush entries. */
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

/* ---- Dense operator apply --------------------------------------------------------------
 * Y = A X for complex128 column-major A (M x K, leading dimension lda), X (K x nrhs, ldx),
 * Y (M x nrhs, ldy), all device pointers, on FP64 MFMA; asynchronous on stream.
 * Replaces the BLAS ZGEMM behind `F*X` / `G*vs` in the reference's callers
 * (examples/3d_pulsating_sphere.jl:42,62), batched like the FMM operators' 32-RHS path
 * (src/3d/fast_multipole_method.jl:249-287).  Uses a cached per-device workspace of
 * 16 * K * ceil(nrhs/32)*32 bytes.  The kernel addresses each K split with 32-bit buffer
 * offsets: lda and ceil(nrhs/32)*32 up to 1.6e6 (BIE_EINVAL beyond). */
int bie_zgemm_device(int64_t M, int64_t nrhs, int64_t K, const double* dA, int64_t lda,
                     const double* dX, int64_t ldx, double* dY, int64_t ldy, void* stream);

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
