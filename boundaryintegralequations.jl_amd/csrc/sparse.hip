// Sparse near-field assembly (gfx950, FP64): the reference's
//   sparse_assemble_parallel!(mesh, k, sources, physics_function; fOn, gOn, depth, offset)
// (src/3d/assembly_collocation_losses.jl:200-405), the near-field correction behind its
// fast (FMM / H-matrix) operators and the lossy thermal and viscous systems.
//
// * Pattern: connected_sources (assembly_collocation_losses.jl:98-132): row i holds the
//   elements with node i among their physics nodes and, as columns, those elements'
//   physics nodes (depth 1); depth 2 takes the same lists of every column of depth 1.
//   Columns ascending per row; the reference's I, J = the row of each entry, its column.
// * Rules: six subdivision rules computed once per call in the reference triangle
//   (singular_triangle_integration, adaptive_integration.jl:172-246, called with the
//   identity as element coordinates and the source at a vertex or an edge midpoint,
//   assembly_collocation_losses.jl:226-232): the triangle is split recursively into four
//   while a sub-triangle's largest side exceeds a quarter of its distance to the source
//   (sides below 1e-6 of the element size count as zero), then each sub-triangle gets the
//   3-point rule of gauss_points_triangle (adaptive_integration.jl:13-19), shifted to its
//   centre and scaled by its shortest side, as the reference places it.
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
#include <vector>

#include "bie_common.hpp"
#include "device_common.hpp"
#include "kernels_math.hpp"

namespace bie {
namespace {

constexpr double kInv4PiS = 0.079577471545947667884;

// ---- subdivision rules (host) ---------------------------------------------------------
// Restated operation by operation (no fused multiply-add: the subdivision decisions are
// floating-point comparisons, and the recursion must reach the same sub-triangles).
#pragma clang fp contract(off)

struct P2 {
  double u, v;
};
struct V3 {
  double x, y, z;
};

P2 mid(P2 a, P2 b) { return {(a.u + b.u) / 2, (a.v + b.v) / 2}; }

// The linear basis [(1 - u) - v, u, v] of the element the reference subdivides (a
// discontinuous linear triangle with beta = 0), times the identity coordinates.
V3 lin(P2 p) { return {(1.0 - p.u) - p.v, p.u, p.v}; }

double dist3(V3 a, V3 b) {
  const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return std::sqrt((dx * dx + dy * dy) + dz * dz);
}

// maximum_side_length / minimum_side_length (adaptive_integration.jl:73-86): sides of the
// closed polygon c1 -> c2 -> c3 -> c1.
double max_side3(V3 a, V3 b, V3 c) {
  return std::max(std::max(dist3(b, a), dist3(c, b)), dist3(a, c));
}
double side2(P2 a, P2 b) {
  const double du = a.u - b.u, dv = a.v - b.v;
  return std::sqrt(du * du + dv * dv);
}

struct Tri {
  P2 a, b, c;
};

void subdivide(const std::vector<Tri>& in, V3 src, double esize, double tol,
               std::vector<Tri>& out, int depth) {
  BIE_REQUIRE(depth < 64, BIE_EINVAL, "subdivision does not terminate");
  for (const Tri& t : in) {
    const P2 p12 = mid(t.a, t.b), p23 = mid(t.b, t.c), p31 = mid(t.c, t.a);
    const P2 cen = {((t.a.u + t.b.u) + t.c.u) / 3.0, ((t.a.v + t.b.v) + t.c.v) / 3.0};
    const V3 s0 = lin(cen), s1 = lin(t.a), s2 = lin(t.b), s3 = lin(t.c);
    const double dsub = std::min(std::min(std::min(dist3(s0, src), dist3(s1, src)),
                                          dist3(s2, src)),
                                 dist3(s3, src));
    double size = max_side3(s1, s2, s3);
    if (size < esize * tol) size = 0.0;
    if (size > dsub / 4) {
      const std::vector<Tri> nd = {{t.a, p12, p31}, {p12, t.b, p23}, {p31, p23, t.c},
                                   {p12, p23, p31}};
      subdivide(nd, src, esize, tol, out, depth + 1);
    } else {
      out.push_back(t);
    }
  }
}

#pragma clang fp contract(on)

}  // namespace

// Points (u, v) and weights of the subdivision rule for a source at barycentric-like
// position src = (1 - u - v, u, v) of the reference triangle.
void subdivision_rule(const double src[3], std::vector<double>& u, std::vector<double>& v,
                      std::vector<double>& w) {
#pragma clang fp contract(off)
  const V3 s = {src[0], src[1], src[2]};
  const double esize = max_side3({1, 0, 0}, {0, 1, 0}, {0, 0, 1});
  std::vector<Tri> out;
  subdivide({{{0.0, 0.0}, {1.0, 0.0}, {0.0, 1.0}}}, s, esize, 1e-6, out, 0);
  const double gx[3] = {1.0 / 6, 2.0 / 3, 1.0 / 6}, gy[3] = {1.0 / 6, 1.0 / 6, 2.0 / 3};
  const double gw = 1.0 / 6, a3 = 1.0 / 3.0;
  u.clear();
  v.clear();
  w.clear();
  for (const Tri& t : out) {
    const double dim = std::min(std::min(side2(t.b, t.a), side2(t.c, t.b)), side2(t.a, t.c));
    const double cu = ((t.a.u + t.b.u) + t.c.u) / 3, cv = ((t.a.v + t.b.v) + t.c.v) / 3;
    for (int g = 0; g < 3; ++g) {
      u.push_back((gx[g] - a3) * dim + cu);
      v.push_back((gy[g] - a3) * dim + cv);
      w.push_back(gw * (dim * dim));
    }
  }
}

namespace {

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
  int off[6], len[6];
  int hyp;
  double kr, ki;
  unsigned flags;
  double* F;              // nnz complex
  double* G;
};

// point record: w, N (6), dN/du (6), dN/dv (6), physics basis (6), pad
constexpr int kSparseRec = 26;

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

template <int NG, int NP, bool CPLX>
__global__ void __launch_bounds__(256) k_sparse_pairs(const SparseArgs a) {
  const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= a.n_pairs) return;  // the whole wave leaves together
  const int lane = threadIdx.x & 63;
  const int2 pr = a.pairs[p];
  const double x0 = a.src[3 * (size_t)pr.x], x1 = a.src[3 * (size_t)pr.x + 1],
               x2 = a.src[3 * (size_t)pr.x + 2];
  const int c = select_rule<NG>(x0, x1, x2, a.gsel + (size_t)pr.y * NG * 4, NG, a.hyp) - 1;
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
    const double* P = R0 + (size_t)q * kSparseRec;
    double y[3] = {0, 0, 0}, tu[3] = {0, 0, 0}, tv[3] = {0, 0, 0};
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      const double N = P[1 + g], du = P[7 + g], dv = P[13 + g];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        y[d] = fma(X[g][d], N, y[d]);
        tu[d] = fma(X[g][d], du, tu[d]);
        tv[d] = fma(X[g][d], dv, tv[d]);
      }
    }
    const double c0 = tu[1] * tv[2] - tu[2] * tv[1], c1 = tu[2] * tv[0] - tu[0] * tv[2],
                 c2 = tu[0] * tv[1] - tu[1] * tv[0];
    const double J = sqrt(fma(c0, c0, fma(c1, c1, c2 * c2)));
    // the reference's normal n = -(t_u x t_v)/J; d = y - x
    const double dx = y[0] - x0, dy = y[1] - x1, dz = y[2] - x2;
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double rinv = rsqrt_nr(r2);
    const double r = r2 * rinv;
    const double nd = -(c0 * dx + c1 * dy + c2 * dz) / J;
    const double g = J * P[0] * kInv4PiS * rinv;       // J w / (4 pi r)
    const double ak = nd * rinv * rinv * g;             // (n.d) J w / (4 pi r^3)
    double sn, cs;
    sincos_f64(a.kr * r, &sn, &cs);
    double Fr, Fi;
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
    const double Gr = g * cs, Gi = g * sn;
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      const double ph = P[19 + s];
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

}  // namespace

// Pattern and pairs of the sparse assembly (connected_sources, depth 1 or 2).
void sparse_pattern(const int* ptopo, int NP, int64_t E, int64_t N, int depth,
                    std::vector<int64_t>& row_ptr, std::vector<int>& cols,
                    std::vector<int64_t>& pair_ptr, std::vector<int>& pair_elem) {
  BIE_REQUIRE(depth == 1 || depth == 2, BIE_EINVAL, "depth must be 1 or 2");
  std::vector<std::vector<int>> el(N), sc(N);
  for (int64_t e = 0; e < E; ++e)
    for (int i = 0; i < NP; ++i) {
      const int node = ptopo[e * NP + i];
      el[node].push_back((int)e);
      for (int j = 0; j < NP; ++j) sc[node].push_back(ptopo[e * NP + j]);
    }
  auto uniq = [](std::vector<int>& v) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  };
  for (int64_t i = 0; i < N; ++i) {
    uniq(el[i]);
    uniq(sc[i]);
  }
  if (depth == 2) {
    std::vector<std::vector<int>> el2(N), sc2(N);
    for (int64_t i = 0; i < N; ++i) {
      for (int j : sc[i]) {
        el2[i].insert(el2[i].end(), el[j].begin(), el[j].end());
        sc2[i].insert(sc2[i].end(), sc[j].begin(), sc[j].end());
      }
      uniq(el2[i]);
      uniq(sc2[i]);
    }
    el.swap(el2);
    sc.swap(sc2);
  }
  row_ptr.assign(1, 0);
  pair_ptr.assign(1, 0);
  cols.clear();
  pair_elem.clear();
  for (int64_t i = 0; i < N; ++i) {
    cols.insert(cols.end(), sc[i].begin(), sc[i].end());
    row_ptr.push_back((int64_t)cols.size());
    pair_elem.insert(pair_elem.end(), el[i].begin(), el[i].end());
    pair_ptr.push_back((int64_t)pair_elem.size());
  }
}

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
  for (int c = 0; c < 6; ++c) {
    a.off[c] = L.off[c];
    a.len[c] = L.len[c];
  }
  a.hyp = L.hyp;
  a.kr = L.kr;
  a.ki = L.ki;
  a.flags = L.flags;
  a.F = L.F;
  a.G = L.G;
  const int64_t blocks = (L.n_pairs + 3) / 4;
  BIE_REQUIRE(blocks < 2147483647LL, BIE_EINVAL, "too many sparse pairs");
  const dim3 grid((unsigned)blocks);
  const bool cplx = L.ki != 0.0;
  auto go = [&](auto kern) { hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, a); };
  if (L.NG == 3 && L.NP == 3)
    cplx ? go(k_sparse_pairs<3, 3, true>) : go(k_sparse_pairs<3, 3, false>);
  else if (L.NG == 6 && L.NP == 6)
    cplx ? go(k_sparse_pairs<6, 6, true>) : go(k_sparse_pairs<6, 6, false>);
  else
    cplx ? go(k_sparse_pairs<6, 3, true>) : go(k_sparse_pairs<6, 3, false>);
  BIE_HIP_CHECK(hipGetLastError());
}

int sparse_record_doubles() { return kSparseRec; }

}  // namespace bie
