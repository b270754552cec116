// Shared host-side declarations of the bie_amd library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bie_amd.h"

namespace bie {

void set_error(const std::string& msg);

struct Error {
  int code;
  std::string msg;
};

#define BIE_HIP_CHECK(expr)                                                       \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess)                                                         \
      throw ::bie::Error{BIE_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e)}; \
  } while (0)

#define BIE_REQUIRE(cond, code, msg)                         \
  do {                                                       \
    if (!(cond)) throw ::bie::Error{(code), std::string(msg)}; \
  } while (0)

constexpr double kPi = 3.14159265358979323846;

// Bound on |x| for the branch-free table sincos of the slot kernels (sincos_tab_fast): 2^46,
// so that 64 |x| / pi stays below 2^51 and its rounding by an added 1.5 * 2^52 is exact.
// The host chooses that form only when |k| times a bound on every target-to-element-point
// distance is below it.
constexpr double kTabFastMaxArg = 70368744177664.0;

// Point record of the concatenated rule set: u, v, w, then N, dN/du, dN/dv of the
// geometry basis (6 slots each, unused slots zero).
constexpr int kPtStride = 30;
// offsets in a point record of N, dN/du and dN/dv (9 geometry nodes each)
constexpr int kPtN = 3, kPtDu = 12, kPtDv = 21;

// All quadrature rules one assembly uses, for a geometry and a physics kind.  Triangles:
// rules 1..3 cluster at the corners (Q = n*m each); rules 4..6 at the edge mid points of
// quadratic elements (Q = 2*n*m', m' = m rounded up to even).  Quadrilaterals: one tensor
// Gauss-Legendre rule on [-1, 1]^2 (Q = n*m, R = 1).
struct RuleSet {
  int gk = 3, pk = 3, n = 0, m = 0;  // geometry and physics nodes per element
  int gkind = 3;                     // geometry kind (BIE_TRI3 .. BIE_QUAD9)
  int pkind = 3;                     // physics kind (BIE_TRI3 .. BIE_DTRI6)
  double beta = 0.0;                 // node offset of BIE_DTRI3 / BIE_DTRI6
  int R = 3;          // number of rules
  int Qc = 0, Qm = 0; // corner / mid-edge rule sizes
  int Pe = 0;         // points per element over all rules
  int Qmax = 0;
  int off[6] = {0, 0, 0, 0, 0, 0};
  int len[6] = {0, 0, 0, 0, 0, 0};
  std::vector<double> pts;  // Pe x kPtStride
  std::vector<double> phi;  // R x Qmax x pk physics basis at the rule points
};

RuleSet build_ruleset(int gkind, int pkind, double beta, int n, int m);

// Element kinds: the low byte is the node count; BIE_DTRI* / BIE_DQUAD* are discontinuous,
// BIE_QUAD* / BIE_DQUAD* quadrilaterals (one Gauss rule, no selection).
inline int physics_nodes(int pkind) { return pkind & 0xff; }
inline int geometry_nodes(int gkind) { return gkind & 0xff; }
inline bool physics_discontinuous(int pkind) { return (pkind & 0x100) != 0; }
inline bool kind_is_quad(int kind) { return (kind >> 9) != 0; }
// physics kinds whose nodes sit at the offset bie_mesh.physics_beta
inline bool kind_has_beta(int pkind) {
  return pkind == BIE_DTRI3 || pkind == BIE_DTRI6 || pkind == BIE_DQUAD4 || pkind == BIE_DQUAD9;
}
// The mid-edge split rules 4-6 exist on quadratic geometry (R = 6) and are chosen only when
// the nearest of six physics nodes is an edge node.
inline bool uses_mid_rules(int gk, int pk) { return gk == 6 && pk == 6; }
void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w);
void gauss_legendre_pm1(int n, std::vector<double>& x, std::vector<double>& w);

// Device-side launch parameters for one assembly.
struct AssemblyArgs {
  const double* targets;   // 3 x M
  int64_t M;
  int64_t N;               // columns of F, G
  int E;
  int NP;                  // physics nodes per element
  const double* elem;      // E x Pe x 8: y(3), n(3), J*w/(4 pi), 0
  int Pe;
  const double* selnodes;  // E x NP x 4 selection node coordinates
  const int* ptopo;        // E x NP physics node indices (0-based, element-major)
  const double* phi;       // R x Qmax x NP
  int Qmax;
  int off[6];
  int len[6];
  int R;
  int hyp;                 // rule-selection distance: 0 julia_hypot3, 1 scaled_hypot3
  double kr, ki;
  unsigned flags;
  double* F;
  double* G;
  double* C;
};

// Device view of a slot schedule plus the mid-edge pair queue.
constexpr int kSlotWaves = 8;   // waves per workgroup of the slot kernel
constexpr int kSlotCount = 64;  // LDS slots of the schedule check (64 rows x 4 doubles each)
constexpr int kMidLds = 512;    // per-workgroup LDS queue of mid-edge pairs
constexpr int kXcds = 8;        // XCDs of an MI355X; workgroup b is dispatched to XCD b mod 8
constexpr int kLdsBytes = 160 * 1024;  // LDS per CU (one slot-kernel workgroup per CU)

constexpr int kStageMaxQ = 16;  // vertex rules up to this length are staged in LDS per wave
// Waves of the wide single-k slot kernel (three per SIMD, one workgroup per CU), an option
// beside kSlotWaves for real k on rules of 4, 6 or 16 points.
constexpr int kWideWaves = 12;

// Table of (cos, sin)(j pi / 64), j < kTrigN, for the table-driven sincos of the slot
// kernels.  It follows the basis table in the same device buffer (at phi_pad(nphi)
// doubles) and in LDS.
constexpr int kTrigN = 128;
constexpr int kTrigDoubles = 2 * kTrigN;

// LDS layout of the slot kernels, in doubles: S slots of 4 x 64 accumulators, the basis
// table (padded to 4 doubles), the trig table (kTrigDoubles), then the per-wave staging
// (when Qc <= kStageMaxQ) of the element's three vertex rules' table rows (3 Qc points x
// 8 doubles) and its selection nodes (kStageSel doubles), then the mid-edge queue.
constexpr int kStageSel = 24;  // NP x 4 doubles of selection-node data, NP <= 6
__host__ __device__ inline int phi_pad(int nphi) { return (nphi + 3) & ~3; }
__host__ __device__ inline int stage_wave_doubles(int Qc) { return 3 * Qc * 8 + kStageSel; }
// The sweep kernel's fixed LDS (doubles): basis and trig tables, no staging.
__host__ __device__ inline int sweep_fixed_doubles(int nphi) {
  return phi_pad(nphi) + kTrigDoubles;
}
__host__ __device__ inline int stage_doubles(int Qc, int waves = kSlotWaves) {
  return Qc <= kStageMaxQ ? waves * stage_wave_doubles(Qc) : 0;
}
// Split form of the slot kernel: `split` waves share each element of a step, wave h taking
// the rule points h Qc/split .. (h+1) Qc/split - 1; each stages the rows of its points of
// the three vertex rules and the selection nodes.
__host__ __device__ inline int split_stage_wave_doubles(int Qh) { return 3 * Qh * 8 + kStageSel; }
__host__ __device__ inline int split_stage_doubles(int Qc, int split) {
  return kSlotWaves * split * split_stage_wave_doubles(Qc / split);
}

// LDS slots that fit beside `fixed_doubles` (padded basis table + staging) and the
// mid-edge queue when each slot holds `per_slot` accumulators of 64 rows x 4 doubles.
inline int fitting_slots(int fixed_doubles, int per_slot, long lds_bytes = kLdsBytes) {
  const long fixed = (long)sizeof(double) * fixed_doubles + 8L * kMidLds + 16;
  const long S = (lds_bytes - fixed) / (256L * (long)sizeof(double) * per_slot);
  return (int)(S < 120 ? S : 120);
}

struct SlotArgs {
  int W, S;
  int n_segments;
  int n_rowblocks;
  int seg_major;
  int64_t n_tiles;
  int phi_in_lds;
  const int* seg_step_begin;
  const int* step_elems;
  const int* step_slot;    // (step, wave, node) slot, schedule order
  const double* step_sel;  // (step, wave, node) selection node xyz_, schedule order
  const int* flush_begin;
  const int* flush_node;
  const int* flush_slot;
  unsigned* mid_count;
  int2* mid_queue;
  int64_t mid_cap;
  int push_mid;            // 1: queue mid-edge pairs; 0: they come from a cached sorted list
  int occupancy;           // slot-kernel workgroups per CU the schedule's LDS size allows
  int split;               // waves per element of a step (1, or 2 for the split kernel)
};

// One instance of an uneven batch (assemble.hip, k_assemble_batch): its assembly and
// slot-schedule arguments, and the columns its shared-column atomics need zeroed.
struct BatchInst {
  AssemblyArgs a;
  SlotArgs sa;
  const int* zero_cols;
  int n_zero;
  int pad;
};

// Batched frequency sweep (sweep.hip).
struct SweepArgs {
  const double* k;           // nk real wavenumbers (device)
  int nk;
  const int* block_uniform;  // per k block: 1 when its k are uniformly spaced
  int64_t stride;            // complex elements between consecutive F_k (and G_k)
  int64_t n_tiles;           // k blocks x segments x row blocks
  int taylor;                // 1: the nk matrices are d^q/dk^q, q = 0..nk-1, at k_re + i k_im
  double k_re, k_im;
};

// Vector-Jacobian product of the assembly (adjoint.hip).
struct AdjointArgs {
  const double* targets;   // 3 x M
  int64_t M;
  const int* tvert;        // M: mesh node a target sits on (moves with X), -1 fixed; or NULL
  const double2* alpha;    // M complex weights of F rows
  const double2* beta;     // M complex weights of C (or NULL)
  const double2* p;        // N complex values at the collocation nodes
  const double* coords;    // 3 x Nc
  const int* topo;         // E x NG
  const int* ptopo;        // E x NP
  const double* selnodes;  // E x NP x 4
  const double* pts;       // rule points (kPtStride)
  const double* phi;       // R x Qmax x NP
  int Qmax;
  int off[6];
  int len[6];
  int R;
  int E;
  int hyp;                 // rule-selection distance (as AssemblyArgs::hyp)
  double k;               // Re k
  double ki;              // Im k (complex-k entry point; 0 for real k)
  double* gradX;           // 3 x Nc, accumulated
  double* gradk;           // d/dRe k (and d/dIm k after it for complex k), accumulated, or NULL
};

// One instance of the batched k-derivative w = alpha^T (dF/dk) p (adjoint.hip).
struct DfdkInst {
  const double* targets;   // 3 x M (the instance's collocation nodes)
  int64_t M;
  int E;
  int Pe;
  const double* elem;      // element tables of the assembly (E x Pe x 8)
  const double* selnodes;  // E x NP x 4
  const int* ptopo;        // E x NP
  const double* phi;       // R x Qmax x NP
  int Qmax;
  int R;
  int hyp;
  int off[6];
  int len[6];
  double kr, ki;
  const double2* alpha;    // M
  const double2* p;        // N
  double* w;               // complex result (2 doubles), accumulated
};

// Sparse near-field assembly (sparse.hip).
struct SparseLaunch {
  const double* src;      // 3 x rows source points (device)
  const int2* pairs;      // (row, element) per pair (device)
  const int* pos;         // pairs x NP entry indices (device)
  int64_t n_pairs;
  const double* coords;   // 3 x Nc (device)
  const int* topo;        // E x NG (device)
  const double* gsel;     // E x NG x 4 geometry node coordinates (device)
  const double* rules;    // point records (u, v, w, 0), sparse_record_doubles() each (device)
  const double* trig;     // (cos, sin)(j pi / 64), j < kTrigN, interleaved (device)
  int off[6], len[6];     // per geometry node: first record and point count
  int NG, NP, hyp;
  const int* batches = nullptr;  // pair batches sharing an element (device; 8 ints each)
  int64_t n_batches = 0;
  bool disc;              // discontinuous physics (basis at beta)
  double beta;
  double kr, ki;
  unsigned flags;
  double* F;              // nnz complex (device), zeroed
  double* G;
};
void subdivision_rule(const double src[3], std::vector<double>& u, std::vector<double>& v,
                      std::vector<double>& w);
bool gauss_points_triangle(int n, std::vector<double>& u, std::vector<double>& v,
                           std::vector<double>& w);
void sparse_pattern(const int* ptopo, int NP, int64_t E, int64_t N, int depth,
                    std::vector<int64_t>& row_ptr, std::vector<int>& cols,
                    std::vector<int64_t>& pair_ptr, std::vector<int>& pair_elem);
// connected_sources of discontinuous physics (row r belongs to element r / NP; N == E * NP);
// topo: the geometry topology (E x NG), whose first three nodes are the corners.
void sparse_pattern_disc(const int* topo, int NG, const int* ptopo, int NP, int64_t E,
                         int64_t N, int depth, std::vector<int64_t>& row_ptr,
                         std::vector<int>& cols, std::vector<int64_t>& pair_ptr,
                         std::vector<int>& pair_elem);
// The polar rule of a discontinuous element around its physics node (pu, pv): three
// sub-triangles from the node to the reference triangle's sides, polar_gaussian(n) on each.
void polar_rule(double pu, double pv, int n, std::vector<double>& u, std::vector<double>& v,
                std::vector<double>& w);
// Local (u, v) of node c of a discontinuous triangular physics kind.
void discontinuous_node(int pkind, double beta, int c, double* u, double* v);
// Physics basis of any physics kind at (u, v) (rules.cpp); N holds 9 entries.
void physics_basis(int pkind, double beta, double u, double v, double* N);
void launch_sparse_pairs(const SparseLaunch& L, hipStream_t s);
int sparse_record_doubles();
// Rows per batch of the batched sparse kernel (k_sparse_batch) for NP physics nodes.
int sparse_batch_rows(int NP);
// Geometry basis of a geometry kind and its derivatives at (u, v) (rules.cpp); N, du, dv
// hold 9 entries each (slots beyond the kind's node count are zeroed).
void geometry_basis(int kind, double u, double v, double* N, double* du, double* dv);

// Launchers (assemble.hip)
void launch_fill_zero(double* p, int64_t n_doubles, hipStream_t s);
void launch_fill_zero_u32(unsigned* p, hipStream_t s);
void launch_geometry(const double* coords, const int* topo, int gk, int E,
                     const double* d_pts, int Pe, double* elem, hipStream_t s);
void launch_gather_selnodes(const double* sources, const int* ptopo, int NP, int E,
                            double* selnodes, hipStream_t s);
void launch_assemble(const AssemblyArgs& a, hipStream_t s);
// fast: every |k| r of the launch is below kTabFastMaxArg (kernels_math.hpp), so real-k
// launches may use the branch-free sincos
void launch_assemble_slots(const AssemblyArgs& a, const SlotArgs& sa, bool fast, hipStream_t s);
void launch_assemble_mid(const AssemblyArgs& a, const SlotArgs& sa, hipStream_t s);
void launch_zero_columns(double* X, int64_t M, const int* cols, int ncols, hipStream_t s);
void launch_batch_prep(const BatchInst* d_inst, int n_inst, int64_t max_cols, hipStream_t s);
size_t batch_shmem(int S, int nphi, int nstage);
void launch_assemble_batch(const BatchInst* d_inst, const int4* d_tiles, int n_tiles,
                           unsigned* d_counter, const double* d_phi, int nphi, int nstage,
                           int S, int NP, bool cplx, bool gon, int qc, int grid,
                           hipStream_t s);
void launch_assemble_mid_batch(const BatchInst* d_inst, int n_inst, bool cplx, hipStream_t s);
void launch_mid_recorded(const AssemblyArgs& a, const unsigned long long* keys, int64_t n_pairs,
                         double* vals, const unsigned long long* entries, const int* off,
                         const int* refs, int64_t n_entries, hipStream_t s);
// The two halves of launch_mid_recorded: the pairs' values (reads the element tables, adds
// C), then their addition into F and G (after the slot kernel's column writes).
void launch_mid_compute(const AssemblyArgs& a, const unsigned long long* keys, int64_t n_pairs,
                        double* vals, hipStream_t s);
void launch_mid_scatter(const AssemblyArgs& a, int64_t n_pairs, const double* vals,
                        const unsigned long long* entries, const int* off, const int* refs,
                        int64_t n_entries, hipStream_t s);
int batch_blocks_per_cu(int NP, bool cplx, bool gon, int qc, int S, int nphi, int nstage);
void zgemv_batched_device(int64_t M, int64_t K, const double* dA, int64_t strideA, int64_t lda,
                          const double* dX, int64_t ldx, double* dY, int64_t ldy,
                          const int32_t* d_idx, int32_t n_idx, hipStream_t s);
// Y = sum_t coef_t A_t over n_terms arrays of n complex entries spaced strideA apart
// (apply_taylor_expansion, src/3d/rosebem.jl:164-199); coef on the host.
constexpr int kTaylorMaxOrders = 64;
void taylor_expand_device(int n_terms, int64_t n, const double* dA, int64_t strideA,
                          const double* coef, double* dY, hipStream_t s);
void zgemm_device(int64_t M, int64_t nrhs, int64_t K, const double* dA, int64_t lda,
                  const double* dX, int64_t ldx, double* dY, int64_t ldy, hipStream_t s);
// Capacity of the allocation-free apply's scratch array, in complex128 entries (16 MiB).
constexpr int64_t kApplyScratchComplex = int64_t(1) << 20;
// As zgemm_device with the packed X in the device scratch array instead of an allocated
// workspace: K * ceil(nrhs/32)*32 <= kApplyScratchComplex.
void zgemm_device_noalloc(int64_t M, int64_t nrhs, int64_t K, const double* dA, int64_t lda,
                          const double* dX, int64_t ldx, double* dY, int64_t ldy,
                          hipStream_t s);
void launch_zero_columns_multi(double* X, int64_t stride, int nk, int64_t M, const int* cols,
                               int ncols, hipStream_t s);
int sweep_kb(int NP);  // slot groups (wavenumbers) per k block of the sweep kernel
// k (or derivative-order) matrices per block: twice sweep_kb without G
int sweep_block_size(int NP, bool gon);
void launch_sweep(const AssemblyArgs& a, const SlotArgs& sa, const SweepArgs& sw,
                  hipStream_t s);
void sweep_prepare_attributes(int NP, int nphi, bool taylor, bool gon);
void launch_block_uniform(const double* k, int nk, int KB, int* out, hipStream_t s);
void launch_adjoint(const AdjointArgs& a, int gk, int pk, bool cplx, hipStream_t s);
// tiles: (instance, 256-row block, chunk of ech elements)
void launch_dfdk_batch(const DfdkInst* d_inst, const int4* d_tiles, int n_tiles, int ech, int NP,
                       hipStream_t s);
// Solves on the device (solve.hip)
void lu_factor_device(int64_t n, double* dA, int64_t lda, const double* dC, int32_t* d_ipiv,
                      hipStream_t s);
void lu_solve_device(int64_t n, const double* dA, int64_t lda, const int32_t* d_ipiv, int trans,
                     double* dB, int64_t ldb, int64_t nrhs, hipStream_t s);
void zgemv_device(int trans, int64_t m, int64_t n, const double* dA, int64_t lda,
                  const double* dx, double* dy, hipStream_t s);
void gmres_batched_device(int64_t n, int nsys, const double* dF, int64_t strideF, int64_t ldf,
                          const double* dC, const double* dB, int64_t ldb, double* dX,
                          int64_t ldx, double tol, int restart, int maxiter, int32_t* iters,
                          double* resid, hipStream_t s);
void launch_select(const double* targets, int64_t M, const double* selnodes, int NP, int E,
                   int R, int hyp, int8_t* out, hipStream_t s);

}  // namespace bie

// Device-resident mesh and its cached rule-dependent tables.
struct bie_dmesh {
  int device = 0;
  int gk = 3, pk = 3;           // geometry nodes, physics nodes per element
  int gkind = 3;                // geometry kind (BIE_TRI3 .. BIE_QUAD9)
  int pkind = 3;                // physics kind (BIE_TRI3 .. BIE_DTRI6)
  double beta = 0.0;            // node offset of BIE_DTRI3 / BIE_DTRI6
  int hyp = 0;                  // rule-selection distance (BIE_MESH_SCALED_HYPOT)
  int64_t Nc = 0, E = 0, N = 0;
  double* d_coords = nullptr;
  int* d_topo = nullptr;        // E x gk, element-major
  double* d_sources = nullptr;  // 3 x N
  int* d_ptopo = nullptr;       // E x pk, element-major
  double* d_selnodes = nullptr; // E x pk x 4
  // rule-dependent cache
  int cn = -1, cm = -1;
  bie::RuleSet rs;
  double* d_pts = nullptr;
  double* d_phi = nullptr;
  double* d_elem = nullptr;
  bool geom_valid = false;      // d_elem holds the geometry of the cached rule set
  // bounding sphere of the mesh: centre, radius over the coordinates and over the sources
  double center[3] = {0.0, 0.0, 0.0};
  double coord_radius = 0.0, source_radius = 0.0;
  // host copies for scheduling
  std::vector<double> h_coords, h_sources;
  std::vector<int> h_topo, h_ptopo;
  // slot schedules (mesh-only): [0] single-k kernel, [1] sweep kernel, [2] batch kernel
  struct Sched {
    int S = -1, key = -1, wv = 0;  // wv: elements per step
    int segments = 0, zero_cols = 0, max_live = 0;
    int* d_seg_step_begin = nullptr;
    int* d_step_elems = nullptr;
    int* d_step_slot = nullptr;
    double* d_step_sel = nullptr;
    int* d_flush_begin = nullptr;
    int* d_flush_node = nullptr;
    int* d_flush_slot = nullptr;
    int* d_zero_cols = nullptr;
    std::vector<int> h_seg_step_begin;  // host copy (tile costs of the batch kernel)
  } sched[3];
  int* d_block_uniform = nullptr;
  int block_uniform_cap = 0;
  // tile shape of the frequency sweep (bie_sweep_tune): elements per tile (0: one segment per
  // row block) and LDS slots (0: as many as fit)
  int sweep_elems_per_tile = 0;
  int sweep_slots = 0;
  // mid-edge pair queue
  unsigned* d_mid_count = nullptr;
  int2* d_mid_queue = nullptr;
  int64_t mid_cap = 0;
  // mid-edge pairs of collocation (targets = sources): they depend on the geometry only,
  // so the first collocation assembly records them, sorted by (element, target), with the
  // distinct (column, row) entries they add to (CSR over pair * 6 + node)
  unsigned long long* d_mid_keys = nullptr;
  int64_t n_mid_keys = -1;  // -1: not recorded yet
  double* d_mid_vals = nullptr;  // 24 per pair
  unsigned long long* d_mid_entries = nullptr;
  int* d_mid_off = nullptr;
  int* d_mid_refs = nullptr;
  int64_t n_mid_entries = 0;
  // calls on the handle share the scratch above: each call's work waits for the previous
  // call's (recorded in `done` on `done_stream`) when it runs on another stream
  hipEvent_t done = nullptr;
  hipStream_t done_stream = nullptr;
  bool done_valid = false;
  // side stream of the recorded mid-edge pairs' values, computed while the slot kernel runs
  hipStream_t side = nullptr;
  hipEvent_t side_fork = nullptr, side_join = nullptr;
  std::mutex mu;
};

// An uneven batch: several device meshes (same element kinds) assembled with one rule set
// in one persistent launch (bie_batch_* in include/bie_amd.h).
struct bie_batch {
  int device = 0;
  int gk = 3, pk = 3, n = 0, m = 0;
  int pkind = 3;
  double beta = 0.0;
  std::vector<bie_dmesh*> meshes;
  // schedule configuration (chosen by bie_batch_tune)
  int elems_per_tile = 64;   // target elements per segment
  int slots = 64;            // LDS slots per workgroup (0: as many as fit)
  int slots_used = 64;       // the slot count the tile list was built with
  int order = 0;             // 0: longest tiles first, 1: instances interleaved
  int configured_key = -1;   // configuration the tile list was built for
  // device state
  bie::BatchInst* d_inst = nullptr;
  bie::BatchInst* h_inst = nullptr;  // pinned staging for the per-call table
  hipEvent_t staged = nullptr;       // the last upload of h_inst has completed
  int4* d_tiles = nullptr;
  int n_tiles = 0;
  unsigned* d_counter = nullptr;
  std::vector<int64_t> rowblocks;
  std::vector<int> min_segments;     // per instance: the schedule the tile list indexes
  std::mutex mu;
};
