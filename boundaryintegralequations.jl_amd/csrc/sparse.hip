// Sparse near-field assembly (gfx950, FP64): the reference's
//   sparse_assemble_parallel!(mesh, k, sources, physics_function; fOn, gOn, depth, offset)
// (src/3d/assembly_collocation_losses.jl:200-405), the near-field correction behind its
// fast (FMM / H-matrix) operators and the lossy thermal and viscous systems.
//
// * Pattern and rules: built on the host (near_field.cpp): connected_sources' rows and
//   columns, and the six subdivision rules of singular_triangle_integration in the
//   reference triangle (continuous physics) or the polar rules of create_polar_elements
//   around the physics nodes (discontinuous physics, :233-257 there).
// * Pairs: every (row i, connected element e) is integrated with the rule of the element's
//   geometry node closest to source i (find_closest_corner on the element coordinates,
//   :308-309 there), the kernels of sparse_computing_integrals! (:158-186: freens3d!,
//   greens3d!, kernels.jl:13-37) and the physics basis at the rule points; the results add
//   into the row's entries of the element's physics nodes.
//
// One wave per (row, element) pair: lanes take the rule's points (a few thousand), the
// partial sums are reduced across the wave and added into the entries with FP64 atomics
// (a row's entries receive several elements' contributions).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <type_traits>
#include <vector>

#include "bie_common.hpp"
#include "device_common.hpp"
#include "kernels_math.hpp"

namespace bie {
namespace {

constexpr double kInv4PiS = 0.079577471545947667884;

// ---- device ----------------------------------------------------------------------------

struct SparseArgs {
  const double* src;      // 3 x n_rows source points
  const int2* pairs;      // (row, element)
  const int* pos;         // pairs x NP entry index
  int64_t n_pairs;
  const double* coords;   // 3 x Nc
  const int* topo;        // E x NG
  const double* gsel;     // E x NG x 4 geometry node coordinates (selection)
  const double* rules;    // point records (kSparseRec doubles)
  const double2* trig;    // (cos, sin)(j pi / 64), j < kTrigN
  int off[6], len[6];
  const int* batches;     // 8 ints per batch: element, count, up to 4 pair indices
  int64_t n_batches;
  int hyp;
  bool disc;              // discontinuous physics, nodes at beta
  double beta;
  double kr, ki;
  unsigned flags;
  double* F;              // nnz complex
  double* G;
};

// point record: u, v, w, pad (the bases are evaluated per point from u, v: a 26-double
// record with them precomputed made the kernel wait on L2 for 208 bytes per point)
constexpr int kSparseRec = 4;

// The geometry basis and its derivatives at (u, v), operation for operation as the host
// tables (eval_basis, rules.cpp, compiled without fused multiply-adds) and the CPU
// restatement compute them: the geometry below must round as theirs does.
template <int NG>
__device__ __forceinline__ void sparse_geom_basis(double u, double v, double (&N)[NG],
                                                  double (&du)[NG], double (&dv)[NG]) {
#pragma clang fp contract(off)
  if constexpr (NG == 3) {
    N[0] = (1.0 - u) - v;
    N[1] = u;
    N[2] = v;
    du[0] = -1.0;
    du[1] = 1.0;
    du[2] = 0.0;
    dv[0] = -1.0;
    dv[1] = 0.0;
    dv[2] = 1.0;
  } else {
    const double l = (1.0 - v) - u;
    N[0] = l * ((1.0 - 2.0 * v) - 2.0 * u);
    N[1] = u * (2.0 * u - 1.0);
    N[2] = v * (2.0 * v - 1.0);
    N[3] = 4.0 * u * l;
    N[4] = 4.0 * u * v;
    N[5] = 4.0 * v * l;
    const double d0 = -(2.0 * l - 1.0) - 2.0 * l;
    du[0] = d0;
    du[1] = 4.0 * u - 1.0;
    du[2] = 0.0;
    du[3] = 4.0 * l - 4.0 * u;
    du[4] = 4.0 * v;
    du[5] = -4.0 * v;
    dv[0] = d0;
    dv[1] = 0.0;
    dv[2] = 4.0 * v - 1.0;
    dv[3] = -4.0 * u;
    dv[4] = 4.0 * u;
    dv[5] = 4.0 * l - 4.0 * v;
  }
}

// The physics basis at (u, v): continuous three- or six-node triangles (the geometry
// bases), or the discontinuous ones (surface_functions.jl:234-318) with psi(s) =
// (s - beta) / (1 - 3 beta) of the barycentric coordinates 1 - u - v, u, v.
template <int NP>
__device__ __forceinline__ void sparse_phys_basis(bool disc, double beta, double u, double v,
                                                  double (&N)[NP]) {
  if constexpr (NP == 1) {
    N[0] = 1.0;
  } else {
    if (!disc) {
      double du[NP], dv[NP];
      sparse_geom_basis<NP>(u, v, N, du, dv);
      return;
    }
    const double den = 1.0 - 3.0 * beta;
    const double l1 = (((1.0 - u) - v) - beta) / den;
    const double l2 = (u - beta) / den;
    const double l3 = (v - beta) / den;
    if constexpr (NP == 3) {
      N[0] = l1;
      N[1] = l2;
      N[2] = l3;
    } else {
      N[0] = l1 * (2.0 * l1 - 1.0);
      N[1] = l2 * (2.0 * l2 - 1.0);
      N[2] = l3 * (2.0 * l3 - 1.0);
      N[3] = (4.0 * l2) * l1;
      N[4] = (4.0 * l2) * l3;
      N[5] = (4.0 * l3) * l1;
    }
  }
}

struct SparseGeom {
  double dx, dy, dz, nd, J;
};

// d = y - x, the reference's normal n = -(t_u x t_v)/J and J at one rule point, in the
// oracle's operation order without fused multiply-adds.  On the elements around the source
// node n.d is rounding noise (a flat element contains its own node) divided by r^3 at
// cells ~1e-6 of the element: only bit-identical geometry reproduces that noise, so this
// part is evaluated exactly as the CPU restatement does.
template <int NG>
__device__ __forceinline__ SparseGeom sparse_geometry(const double (&X)[NG][3],
                                                      const double (&N)[NG],
                                                      const double (&Du)[NG],
                                                      const double (&Dv)[NG], double x0,
                                                      double x1, double x2) {
#pragma clang fp contract(off)
  double y[3], tu[3], tv[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double a = 0.0, b = 0.0, c = 0.0;
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      a = a + X[g][d] * N[g];
      b = b + X[g][d] * Du[g];
      c = c + X[g][d] * Dv[g];
    }
    y[d] = a;
    tu[d] = b;
    tv[d] = c;
  }
  const double c0 = tu[1] * tv[2] - tu[2] * tv[1], c1 = tu[2] * tv[0] - tu[0] * tv[2],
               c2 = tu[0] * tv[1] - tu[1] * tv[0];
  SparseGeom sg;
  sg.J = julia_hypot3(c0, c1, c2);
  sg.dx = y[0] - x0;
  sg.dy = y[1] - x1;
  sg.dz = y[2] - x2;
  sg.nd = (-c0 / sg.J) * sg.dx + (-c1 / sg.J) * sg.dy + (-c2 / sg.J) * sg.dz;
  return sg;
}

// The surface at one rule point without the source: y, the reference's normal
// n = -(t_u x t_v) / J and J, operation for operation as sparse_geometry.
template <int NG>
__device__ __forceinline__ void sparse_surface(const double (&X)[NG][3], const double (&N)[NG],
                                               const double (&Du)[NG], const double (&Dv)[NG],
                                               double (&y)[3], double (&n)[3], double& J) {
#pragma clang fp contract(off)
  double tu[3], tv[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double a = 0.0, b = 0.0, c = 0.0;
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      a = a + X[g][d] * N[g];
      b = b + X[g][d] * Du[g];
      c = c + X[g][d] * Dv[g];
    }
    y[d] = a;
    tu[d] = b;
    tv[d] = c;
  }
  const double c0 = tu[1] * tv[2] - tu[2] * tv[1], c1 = tu[2] * tv[0] - tu[0] * tv[2],
               c2 = tu[0] * tv[1] - tu[1] * tv[0];
  J = julia_hypot3(c0, c1, c2);
  n[0] = -c0 / J;
  n[1] = -c1 / J;
  n[2] = -c2 / J;
}

// d = y - x and n.d for one source, in sparse_geometry's order.
__device__ __forceinline__ SparseGeom sparse_toward(const double (&y)[3], const double (&n)[3],
                                                    double J, double x0, double x1,
                                                    double x2) {
#pragma clang fp contract(off)
  SparseGeom sg;
  sg.J = J;
  sg.dx = y[0] - x0;
  sg.dy = y[1] - x1;
  sg.dz = y[2] - x2;
  sg.nd = n[0] * sg.dx + n[1] * sg.dy + n[2] * sg.dz;
  return sg;
}

// The kernels at one rule point: G = e^{ikr} J w / (4 pi r) and
// F = e^{ikr} (ikr - 1) (n.d) J w / (4 pi r^3) (freens3d!, greens3d!, kernels.jl:13-37).
template <bool CPLX>
__device__ __forceinline__ void sparse_point_kernels(const SparseArgs& a, const SparseGeom& sg,
                                                     double wq, const double2* tab, double& Fr,
                                                     double& Fi, double& Gr, double& Gi) {
  const double dx = sg.dx, dy = sg.dy, dz = sg.dz, J = sg.J, nd = sg.nd;
  const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
  const double rinv = rsqrt_nr(r2);
  const double r = r2 * rinv;
  const double g = J * wq * kInv4PiS * rinv;         // J w / (4 pi r)
  const double ak = nd * rinv * rinv * g;             // (n.d) J w / (4 pi r^3)
  double sn, cs;
  sincos_tab(a.kr * r, tab, &sn, &cs);
  if (CPLX) {
    const double ex = exp(-a.ki * r);
    cs *= ex;
    sn *= ex;
    const double ar = fma(-a.ki, r, -1.0), ai = a.kr * r;   // i k r - 1
    Fr = ak * fma(cs, ar, -sn * ai);
    Fi = ak * fma(cs, ai, sn * ar);
  } else {
    const double krr = a.kr * r;
    Fr = -ak * fma(krr, sn, cs);
    Fi = ak * fma(krr, cs, -sn);
  }
  Gr = g * cs;
  Gi = g * sn;
}

// The block's copy of the trig table for sincos_tab (every thread of the block takes part,
// before any leaves).
__device__ __forceinline__ void load_trig(const SparseArgs& a, double2* s_trig) {
  for (int j = threadIdx.x; j < kTrigN; j += blockDim.x) s_trig[j] = a.trig[j];
  __syncthreads();
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

template <int NG, int NP, bool CPLX>
__global__ void __launch_bounds__(256) k_sparse_pairs(const SparseArgs a) {
  const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  __shared__ double2 s_trig[kTrigN];
  load_trig(a, s_trig);
  if (p >= a.n_pairs) return;  // the whole wave leaves together
  const int lane = threadIdx.x & 63;
  const int2 pr = a.pairs[p];
  const double x0 = a.src[3 * (size_t)pr.x], x1 = a.src[3 * (size_t)pr.x + 1],
               x2 = a.src[3 * (size_t)pr.x + 2];
  // no distance below Inf (non-finite coordinates) falls to the sixth rule, the
  // reference's final else branch (assembly_collocation_losses.jl:390-395)
  const int c = select_rule<NG>(x0, x1, x2, a.gsel + (size_t)pr.y * NG * 4, 6, a.hyp) - 1;
  double X[NG][3];
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const int node = a.topo[(size_t)pr.y * NG + g];
#pragma unroll
    for (int d = 0; d < 3; ++d) X[g][d] = a.coords[3 * (size_t)node + d];
  }
  double fr[NP], fi[NP], gr[NP], gi[NP];
#pragma unroll
  for (int s = 0; s < NP; ++s) fr[s] = fi[s] = gr[s] = gi[s] = 0.0;
  const double* R0 = a.rules + (size_t)a.off[c] * kSparseRec;
  const int Q = a.len[c];
  for (int q = lane; q < Q; q += 64) {
    const double2 uv = *reinterpret_cast<const double2*>(R0 + (size_t)q * kSparseRec);
    const double wq = R0[(size_t)q * kSparseRec + 2];
    double Ng[NG], Du[NG], Dv[NG], Np[NP];
    sparse_geom_basis<NG>(uv.x, uv.y, Ng, Du, Dv);
    sparse_phys_basis<NP>(a.disc, a.beta, uv.x, uv.y, Np);
    const SparseGeom sg = sparse_geometry<NG>(X, Ng, Du, Dv, x0, x1, x2);
    double Fr, Fi, Gr, Gi;
    sparse_point_kernels<CPLX>(a, sg, wq, s_trig, Fr, Fi, Gr, Gi);
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      const double ph = Np[s];
      fr[s] = fma(Fr, ph, fr[s]);
      fi[s] = fma(Fi, ph, fi[s]);
      gr[s] = fma(Gr, ph, gr[s]);
      gi[s] = fma(Gi, ph, gi[s]);
    }
  }
  const bool doF = (a.flags & BIE_F) != 0, doG = (a.flags & BIE_G) != 0;
#pragma unroll
  for (int s = 0; s < NP; ++s) {
    const double vfr = wave_sum_d(fr[s]), vfi = wave_sum_d(fi[s]);
    const double vgr = wave_sum_d(gr[s]), vgi = wave_sum_d(gi[s]);
    if (lane == 0) {
      const size_t e = 2 * (size_t)a.pos[p * NP + s];
      if (doF) {
        unsafeAtomicAdd(a.F + e, vfr);
        unsafeAtomicAdd(a.F + e + 1, vfi);
      }
      if (doG) {
        unsafeAtomicAdd(a.G + e, vgr);
        unsafeAtomicAdd(a.G + e + 1, vgi);
      }
    }
  }
}

// Short rules (the partial near field's gauss_points_triangle(n), n <= 13): one thread per
// pair instead of one wave, the same per-point evaluation; the thread adds its sums into
// the entries with FP64 atomics.
template <int NG, int NP, bool CPLX>
__global__ void __launch_bounds__(256) k_sparse_pairs_short(const SparseArgs a) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ double2 s_trig[kTrigN];
  load_trig(a, s_trig);
  if (p >= a.n_pairs) return;
  const int2 pr = a.pairs[p];
  const double x0 = a.src[3 * (size_t)pr.x], x1 = a.src[3 * (size_t)pr.x + 1],
               x2 = a.src[3 * (size_t)pr.x + 2];
  const int c = select_rule<NG>(x0, x1, x2, a.gsel + (size_t)pr.y * NG * 4, 6, a.hyp) - 1;
  double X[NG][3];
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const int node = a.topo[(size_t)pr.y * NG + g];
#pragma unroll
    for (int d = 0; d < 3; ++d) X[g][d] = a.coords[3 * (size_t)node + d];
  }
  double fr[NP], fi[NP], gr[NP], gi[NP];
#pragma unroll
  for (int s = 0; s < NP; ++s) fr[s] = fi[s] = gr[s] = gi[s] = 0.0;
  const double* R0 = a.rules + (size_t)a.off[c] * kSparseRec;
  const int Q = a.len[c];
  for (int q = 0; q < Q; ++q) {
    const double2 uv = *reinterpret_cast<const double2*>(R0 + (size_t)q * kSparseRec);
    const double wq = R0[(size_t)q * kSparseRec + 2];
    double Ng[NG], Du[NG], Dv[NG], Np[NP];
    sparse_geom_basis<NG>(uv.x, uv.y, Ng, Du, Dv);
    sparse_phys_basis<NP>(a.disc, a.beta, uv.x, uv.y, Np);
    const SparseGeom sg = sparse_geometry<NG>(X, Ng, Du, Dv, x0, x1, x2);
    double Fr, Fi, Gr, Gi;
    sparse_point_kernels<CPLX>(a, sg, wq, s_trig, Fr, Fi, Gr, Gi);
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      fr[s] = fma(Fr, Np[s], fr[s]);
      fi[s] = fma(Fi, Np[s], fi[s]);
      gr[s] = fma(Gr, Np[s], gr[s]);
      gi[s] = fma(Gi, Np[s], gi[s]);
    }
  }
  const bool doF = (a.flags & BIE_F) != 0, doG = (a.flags & BIE_G) != 0;
#pragma unroll
  for (int s = 0; s < NP; ++s) {
    const size_t e = 2 * (size_t)a.pos[p * NP + s];
    if (doF) {
      unsafeAtomicAdd(a.F + e, fr[s]);
      unsafeAtomicAdd(a.F + e + 1, fi[s]);
    }
    if (doG) {
      unsafeAtomicAdd(a.G + e, gr[s]);
      unsafeAtomicAdd(a.G + e + 1, gi[s]);
    }
  }
}

// Batches of up to B pairs that share an element (built on the host, grouped by the rule
// the mesh sources select): one wave per batch evaluates the surface at each point of a rule
// once and the kernels for every row of the batch that selected that rule - the surface
// (~2/3 of the per-point work) is shared.  Each row's rule is selected here exactly as in
// k_sparse_pairs, so explicit sources stay correct (they only share less).
template <int NG, int NP, int B, bool CPLX>
__global__ void __launch_bounds__(256) k_sparse_batch(const SparseArgs a) {
  const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  __shared__ double2 s_trig[kTrigN];
  load_trig(a, s_trig);
  if (b >= a.n_batches) return;  // the whole wave leaves together
  const int lane = threadIdx.x & 63;
  const int* rec = a.batches + (size_t)b * 8;
  const int e = rec[0], cnt = rec[1];
  double X[NG][3];
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const int node = a.topo[(size_t)e * NG + g];
#pragma unroll
    for (int d = 0; d < 3; ++d) X[g][d] = a.coords[3 * (size_t)node + d];
  }
  double xs[B][3];
  int cr[B], pj[B];
#pragma unroll
  for (int j = 0; j < B; ++j) {
    pj[j] = j < cnt ? rec[2 + j] : -1;
    cr[j] = -1;
    xs[j][0] = xs[j][1] = xs[j][2] = 0.0;
    if (pj[j] >= 0) {
      const int row = a.pairs[pj[j]].x;
      xs[j][0] = a.src[3 * (size_t)row];
      xs[j][1] = a.src[3 * (size_t)row + 1];
      xs[j][2] = a.src[3 * (size_t)row + 2];
      cr[j] = select_rule<NG>(xs[j][0], xs[j][1], xs[j][2], a.gsel + (size_t)e * NG * 4, 6,
                              a.hyp) - 1;
    }
  }
  double acc[B][NP][4];
#pragma unroll
  for (int j = 0; j < B; ++j)
#pragma unroll
    for (int s = 0; s < NP; ++s) acc[j][s][0] = acc[j][s][1] = acc[j][s][2] = acc[j][s][3] = 0.0;
  unsigned done = 0u;
  for (int j0 = 0; j0 < B; ++j0) {
    if (cr[j0] < 0 || ((done >> j0) & 1u)) continue;
    const int c = cr[j0];
    unsigned m = 0u;  // the rows whose rule has these records (filler slots share them)
#pragma unroll
    for (int j = 0; j < B; ++j)
      if (cr[j] >= 0 && a.off[cr[j]] == a.off[c] && a.len[cr[j]] == a.len[c]) m |= 1u << j;
    done |= m;
    const double* R0 = a.rules + (size_t)a.off[c] * kSparseRec;
    const int Q = a.len[c];
    for (int q = lane; q < Q; q += 64) {
      const double2 uv = *reinterpret_cast<const double2*>(R0 + (size_t)q * kSparseRec);
      const double wq = R0[(size_t)q * kSparseRec + 2];
      double Ng[NG], Du[NG], Dv[NG], Np[NP], y[3], n[3], J;
      sparse_geom_basis<NG>(uv.x, uv.y, Ng, Du, Dv);
      sparse_phys_basis<NP>(a.disc, a.beta, uv.x, uv.y, Np);
      sparse_surface<NG>(X, Ng, Du, Dv, y, n, J);
#pragma unroll
      for (int j = 0; j < B; ++j) {
        if (!((m >> j) & 1u)) continue;
        const SparseGeom sg = sparse_toward(y, n, J, xs[j][0], xs[j][1], xs[j][2]);
        double Fr, Fi, Gr, Gi;
        sparse_point_kernels<CPLX>(a, sg, wq, s_trig, Fr, Fi, Gr, Gi);
#pragma unroll
        for (int s = 0; s < NP; ++s) {
          acc[j][s][0] = fma(Fr, Np[s], acc[j][s][0]);
          acc[j][s][1] = fma(Fi, Np[s], acc[j][s][1]);
          acc[j][s][2] = fma(Gr, Np[s], acc[j][s][2]);
          acc[j][s][3] = fma(Gi, Np[s], acc[j][s][3]);
        }
      }
    }
  }
  const bool doF = (a.flags & BIE_F) != 0, doG = (a.flags & BIE_G) != 0;
#pragma unroll
  for (int j = 0; j < B; ++j) {
    if (pj[j] < 0) continue;
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      const double vfr = wave_sum_d(acc[j][s][0]), vfi = wave_sum_d(acc[j][s][1]);
      const double vgr = wave_sum_d(acc[j][s][2]), vgi = wave_sum_d(acc[j][s][3]);
      if (lane == 0) {
        const size_t ent = 2 * (size_t)a.pos[(size_t)pj[j] * NP + s];
        if (doF) {
          unsafeAtomicAdd(a.F + ent, vfr);
          unsafeAtomicAdd(a.F + ent + 1, vfi);
        }
        if (doG) {
          unsafeAtomicAdd(a.G + ent, vgr);
          unsafeAtomicAdd(a.G + ent + 1, vgi);
        }
      }
    }
  }
}

}  // namespace

void launch_sparse_pairs(const SparseLaunch& L, hipStream_t s) {
  if (L.n_pairs == 0) return;
  SparseArgs a{};
  a.src = L.src;
  a.pairs = L.pairs;
  a.pos = L.pos;
  a.n_pairs = L.n_pairs;
  a.coords = L.coords;
  a.topo = L.topo;
  a.gsel = L.gsel;
  a.rules = L.rules;
  a.trig = reinterpret_cast<const double2*>(L.trig);
  for (int c = 0; c < 6; ++c) {
    a.off[c] = L.off[c];
    a.len[c] = L.len[c];
  }
  a.hyp = L.hyp;
  a.batches = L.batches;
  a.n_batches = L.n_batches;
  a.disc = L.disc;
  a.beta = L.beta;
  a.kr = L.kr;
  a.ki = L.ki;
  a.flags = L.flags;
  a.F = L.F;
  a.G = L.G;
  // every slot's rule short (the partial near field): one thread per pair
  int qmax = 0;
  for (int c = 0; c < 6; ++c) qmax = std::max(qmax, L.len[c]);
  const bool shortr = qmax <= 16;
  const bool batched = !shortr && L.batches != nullptr && L.n_batches > 0;
  const int64_t blocks = shortr    ? (L.n_pairs + 255) / 256
                         : batched ? (L.n_batches + 3) / 4
                                   : (L.n_pairs + 3) / 4;
  BIE_REQUIRE(blocks < 2147483647LL, BIE_EINVAL, "too many sparse pairs");
  const dim3 grid((unsigned)blocks);
  const bool cplx = L.ki != 0.0;
  auto go = [&](auto kern) { hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, a); };
  // geometry nodes 3 or 6; physics nodes 1 (discontinuous constant), 3 or 6
  auto pick = [&](auto ng, auto np, auto cx) {
    constexpr int NG = decltype(ng)::value, NP = decltype(np)::value;
    constexpr bool C = decltype(cx)::value;
    constexpr int B = NP == 1 ? 4 : 3;
    if (shortr)
      go(k_sparse_pairs_short<NG, NP, C>);
    else if (batched)
      go(k_sparse_batch<NG, NP, B, C>);
    else
      go(k_sparse_pairs<NG, NP, C>);
  };
  auto by_cx = [&](auto ng, auto np) {
    cplx ? pick(ng, np, std::true_type()) : pick(ng, np, std::false_type());
  };
  auto by_np = [&](auto ng) {
    if (L.NP == 1)
      by_cx(ng, std::integral_constant<int, 1>());
    else if (L.NP == 3)
      by_cx(ng, std::integral_constant<int, 3>());
    else
      by_cx(ng, std::integral_constant<int, 6>());
  };
  BIE_REQUIRE((L.NG == 3 || L.NG == 6) && (L.NP == 1 || L.NP == 3 || L.NP == 6), BIE_EINVAL,
              "internal: sparse kernel node counts");
  if (L.NG == 3)
    by_np(std::integral_constant<int, 3>());
  else
    by_np(std::integral_constant<int, 6>());
  BIE_HIP_CHECK(hipGetLastError());
}

int sparse_record_doubles() { return kSparseRec; }

int sparse_batch_rows(int NP) { return NP == 1 ? 4 : 3; }

}  // namespace bie
