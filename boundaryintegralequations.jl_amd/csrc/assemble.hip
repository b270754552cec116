// Device kernels of the dense 3D collocation assembly (gfx950, FP64).
//
// k_geometry        interpolate_elements (src/3d/assembly_collocation.jl:27-66) and
//                   jacobian! (src/3d/jacobian.jl:78-87) for every rule point of every
//                   element: y = X N, n = -(X dN/du x X dN/dv)/J, J w / (4 pi).
// select_rule       find_closest_corner (assembly_collocation.jl:102-115), bit-exact.
// k_assemble        the row loop (assembly_collocation.jl:314-349 / 424-475) with
//                   computing_integrals! (:213-252) fused: one pass over the rule points
//                   evaluates G, F and the C integrand (kernels.jl:13-55) from a single
//                   e^{ikr}, contracts with the physics basis (mygemm_vec!,
//                   src/2d/assembly_collocation.jl:142-150) and scatters into F, G.
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "bie_common.hpp"
#include "device_common.hpp"
#include "kernels_math.hpp"
#include "schedule.hpp"

namespace bie {
namespace {

constexpr double kInv4Pi = 0.079577471545947667884;  // 1 / (4 pi)
// Points per stage of the software-pipelined real-k point loop (pair_quadrature_pipelined).
#ifndef BIE_POINT_GROUP
#define BIE_POINT_GROUP 1
#endif
constexpr int kPointGroup = BIE_POINT_GROUP;

__global__ void k_fill_zero(double2* __restrict__ p, int64_t n2) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2; i += stride)
    p[i] = make_double2(0.0, 0.0);
}

__global__ void k_fill_zero_tail(double* __restrict__ p) { p[0] = 0.0; }

__global__ void __launch_bounds__(256)
k_geometry(const double* __restrict__ coords, const int* __restrict__ topo, int gk, int E,
           const double* __restrict__ pts, int Pe, double* __restrict__ elem) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)E * Pe) return;
  const int e = (int)(t / Pe);
  const int p = (int)(t - (int64_t)e * Pe);
  const double* P = pts + (size_t)p * kPtStride;
  double y0 = 0, y1 = 0, y2 = 0, u0 = 0, u1 = 0, u2 = 0, v0 = 0, v1 = 0, v2 = 0;
  for (int k = 0; k < gk; ++k) {
    const int node = topo[(size_t)e * gk + k];
    const double X0 = coords[3 * (size_t)node], X1 = coords[3 * (size_t)node + 1],
                 X2 = coords[3 * (size_t)node + 2];
    const double N = P[kPtN + k], du = P[kPtDu + k], dv = P[kPtDv + k];
    y0 += X0 * N;
    y1 += X1 * N;
    y2 += X2 * N;
    u0 += X0 * du;
    u1 += X1 * du;
    u2 += X2 * du;
    v0 += X0 * dv;
    v1 += X1 * dv;
    v2 += X2 * dv;
  }
  const double c0 = u1 * v2 - u2 * v1, c1 = u2 * v0 - u0 * v2, c2 = u0 * v1 - u1 * v0;
  const double J = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  double* o = elem + (size_t)t * 8;
  const double4 A = make_double4(y0, y1, y2, -c0 / J);
  const double4 B = make_double4(-c1 / J, -c2 / J, J * P[2] * kInv4Pi, 0.0);
  *reinterpret_cast<double4*>(o) = A;
  *reinterpret_cast<double4*>(o + 4) = B;
}

__global__ void k_gather_selnodes(const double* __restrict__ sources,
                                  const int* __restrict__ ptopo, int NP, int E,
                                  double* __restrict__ sel) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E * NP) return;
  const int node = ptopo[t];
  double* o = sel + (size_t)t * 4;
  o[0] = sources[3 * (size_t)node];
  o[1] = sources[3 * (size_t)node + 1];
  o[2] = sources[3 * (size_t)node + 2];
  o[3] = 0.0;
}

template <int NP>
__global__ void __launch_bounds__(256)
k_select(const double* __restrict__ targets, int64_t M, const double* __restrict__ sel,
         int E, int R, int hyp, int8_t* __restrict__ out) {
  const int64_t total = M * (int64_t)E;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += stride) {
    const int64_t i = t / E;
    const int e = (int)(t - i * E);
    out[t] = (int8_t)select_rule<NP>(targets[3 * i], targets[3 * i + 1], targets[3 * i + 2],
                                      sel + (size_t)e * NP * 4, R, hyp);
  }
}

// Quadrature of one (target, element) pair with rule points tab[0..Q) and physics basis
// ph[q * NP + s]: accumulates the F and G contributions of the NP element nodes and the
// C integrand.  Single-layer g = e^{ikr}/(4 pi r), double-layer
// f = e^{ikr}(ikr - 1)(n.d)/(4 pi r^3) and k0 = -(n.d)/(4 pi r^3) (kernels.jl:13-55) share
// one e^{ikr} per point; 1/(4 pi) and J w are folded into tab[6].  With TAB, e^{ikr}
// comes from the LDS trig table trig (sincos_tab; sincos_tab_fast with FAST, for launches
// whose arguments k r are known to stay below its bound), else from sincos_f64.
template <int NP, bool CPLX, bool TAB = false, bool GON = true, bool FAST = false>
__device__ __forceinline__ void point_kernel(double x0, double x1, double x2, const double4 A,
                                             const double4 B, const double* __restrict__ p,
                                             double kr, double ki, double* fr, double* fi,
                                             double* gr, double* gi, double& csum,
                                             const double2* trig = nullptr) {
  const double dx = A.x - x0, dy = A.y - x1, dz = A.z - x2;
  const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
  const double rinv = rsqrt_nr(r2);
  const double r = r2 * rinv;
  const double nd = fma(A.w, dx, fma(B.x, dy, B.y * dz));
  const double g = B.z * rinv;             // J w / (4 pi r)
  const double ak = nd * rinv * rinv * g;  // (n.d) J w / (4 pi r^3)
  csum -= ak;
  double sn, cs;
  if (FAST)
    sincos_tab_fast(kr * r, trig, &sn, &cs);
  else if (TAB)
    sincos_tab(kr * r, trig, &sn, &cs);
  else
    sincos_f64(kr * r, &sn, &cs);
  double Fr, Fi;
  if (CPLX) {
    const double ex = exp(-ki * r);
    cs *= ex;
    sn *= ex;
    const double ar = fma(-ki, r, -1.0), ai = kr * r;
    Fr = ak * fma(cs, ar, -sn * ai);
    Fi = ak * fma(cs, ai, sn * ar);
  } else {
    const double krr = kr * r;
    Fr = -ak * fma(krr, sn, cs);
    Fi = ak * fma(krr, cs, -sn);
  }
  const double Gr = g * cs, Gi = g * sn;
#pragma unroll
  for (int s = 0; s < NP; ++s) {
    const double ps = p[s];
    fr[s] = fma(Fr, ps, fr[s]);
    fi[s] = fma(Fi, ps, fi[s]);
    if constexpr (GON) {  // without G (gOn = false) its accumulation is left out
      gr[s] = fma(Gr, ps, gr[s]);
      gi[s] = fma(Gi, ps, gi[s]);
    }
  }
}

// Quadrature of one (target, element) pair over rule points tab[0..Q) with physics basis
// ph[q * NP + s]: accumulates the F and G contributions of the NP element nodes and the
// C integrand.  Single-layer g = e^{ikr}/(4 pi r), double-layer
// f = e^{ikr}(ikr - 1)(n.d)/(4 pi r^3) and k0 = -(n.d)/(4 pi r^3) (kernels.jl:13-55) share
// one e^{ikr} per point; 1/(4 pi) and J w are folded into tab[6].  Points are taken two at
// a time from two register sets so the next table row is always in flight.
template <int NP, bool CPLX>
__device__ __forceinline__ void pair_quadrature(double x0, double x1, double x2,
                                                const double* __restrict__ tab,
                                                const double* __restrict__ ph, int Q,
                                                double kr, double ki, double* fr, double* fi,
                                                double* gr, double* gi, double& csum) {
  double4 A = *reinterpret_cast<const double4*>(tab);
  double4 B = *reinterpret_cast<const double4*>(tab + 4);
  int q = 0;
  for (; q + 1 < Q; q += 2) {
    const double4 A2 = *reinterpret_cast<const double4*>(tab + 8 * (q + 1));
    const double4 B2 = *reinterpret_cast<const double4*>(tab + 8 * (q + 1) + 4);
    point_kernel<NP, CPLX>(x0, x1, x2, A, B, ph + q * NP, kr, ki, fr, fi, gr, gi, csum);
    const int qn = (q + 2 < Q) ? q + 2 : q + 1;
    A = *reinterpret_cast<const double4*>(tab + 8 * qn);
    B = *reinterpret_cast<const double4*>(tab + 8 * qn + 4);
    point_kernel<NP, CPLX>(x0, x1, x2, A2, B2, ph + (q + 1) * NP, kr, ki, fr, fi, gr, gi, csum);
  }
  if (q < Q) point_kernel<NP, CPLX>(x0, x1, x2, A, B, ph + q * NP, kr, ki, fr, fi, gr, gi, csum);
}

// As pair_quadrature, with the first two table rows (points 0 and 1, or 0 twice when
// Q == 1) already loaded by the caller, so their latency overlaps earlier work.
template <int NP, bool CPLX, bool TAB = false, bool GON = true>
__device__ __forceinline__ void pair_quadrature_pre(double x0, double x1, double x2,
                                                    const double* __restrict__ tab,
                                                    const double* __restrict__ ph, int Q,
                                                    double kr, double ki, double4 A, double4 B,
                                                    double4 A2, double4 B2, double* fr,
                                                    double* fi, double* gr, double* gi,
                                                    double& csum, const double2* trig = nullptr) {
  int q = 0;
  for (; q + 1 < Q; q += 2) {
    point_kernel<NP, CPLX, TAB, GON>(x0, x1, x2, A, B, ph + q * NP, kr, ki, fr, fi, gr, gi,
                                     csum, trig);
    const int q2 = (q + 2 < Q) ? q + 2 : q + 1;
    A = *reinterpret_cast<const double4*>(tab + 8 * q2);
    B = *reinterpret_cast<const double4*>(tab + 8 * q2 + 4);
    point_kernel<NP, CPLX, TAB, GON>(x0, x1, x2, A2, B2, ph + (q + 1) * NP, kr, ki, fr, fi, gr,
                                     gi, csum, trig);
    const int q3 = (q + 3 < Q) ? q + 3 : q + 1;
    A2 = *reinterpret_cast<const double4*>(tab + 8 * q3);
    B2 = *reinterpret_cast<const double4*>(tab + 8 * q3 + 4);
  }
  if (q < Q)
    point_kernel<NP, CPLX, TAB, GON>(x0, x1, x2, A, B, ph + q * NP, kr, ki, fr, fi, gr, gi,
                                     csum, trig);
}

// The real-k point of point_kernel<NP, false, true, GON, true> in two stages, so that a
// software-pipelined loop can put other points' work between them.  front: distance,
// 1/r, n.d, the kernel scale factors and the argument reduction of sincos_tab_fast, ending
// with the issue of the table read; back: the sincos polynomials, the kernel values and
// their accumulation into the NP node sums.  Same operations, in the same order per point,
// as point_kernel.
struct PointFront {
  double krr, d, gg, ak;
  double2 t;
};

__device__ __forceinline__ PointFront point_front(double x0, double x1, double x2,
                                                  const double4 A, const double4 B, double kr,
                                                  double& csum,
                                                  const double2* __restrict__ trig) {
  constexpr double kInvStep = 20.371832715762602978;       // 64 / pi
  constexpr double kStepHi = 1.5707963267948965580 / 32;   // pi/2 rounded, / 32 (exact)
  constexpr double kStepLo = 6.1232339957367658e-17 / 32;  // its remainder, / 32 (exact)
  constexpr double kShift = 6755399441055744.0;            // 1.5 * 2^52
  PointFront f;
  const double dx = A.x - x0, dy = A.y - x1, dz = A.z - x2;
  const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
  const double rinv = rsqrt_nr(r2);
  const double r = r2 * rinv;
  const double nd = fma(A.w, dx, fma(B.x, dy, B.y * dz));
  f.gg = B.z * rinv;                  // J w / (4 pi r)
  f.ak = nd * rinv * rinv * f.gg;     // (n.d) J w / (4 pi r^3)
  csum -= f.ak;
  f.krr = kr * r;
  // sincos_tab_fast(k r): reduction and table read
  const double sh = fma(f.krr, kInvStep, kShift);
  const double n = sh - kShift;
  double d = fma(-n, kStepHi, f.krr);
  f.d = fma(-n, kStepLo, d);
  f.t = trig[__double2loint(sh) & (kTrigN - 1)];
  return f;
}

template <int NP, bool GON>
__device__ __forceinline__ void point_back(const PointFront& f, const double* p, double* fr,
                                           double* fi, double* gr, double* gi) {
  const double d = f.d, d2 = d * d;
  const double sp = fma(d2, fma(d2, -1.0 / 5040.0, 1.0 / 120.0), -1.0 / 6.0);
  const double sd = fma(d * d2, sp, d);
  const double cp = fma(d2, fma(d2, -1.0 / 720.0, 1.0 / 24.0), -0.5);
  const double cd = fma(d2, cp, 1.0);
  const double sn = fma(f.t.y, cd, f.t.x * sd);
  const double cs = fma(f.t.x, cd, -(f.t.y * sd));
  const double Fr = -f.ak * fma(f.krr, sn, cs);
  const double Fi = f.ak * fma(f.krr, cs, -sn);
  const double Gr = f.gg * cs, Gi = f.gg * sn;
#pragma unroll
  for (int s = 0; s < NP; ++s) {
    const double ps = p[s];
    fr[s] = fma(Fr, ps, fr[s]);
    fi[s] = fma(Fi, ps, fi[s]);
    if constexpr (GON) {
      gr[s] = fma(Gr, ps, gr[s]);
      gi[s] = fma(Gi, ps, gi[s]);
    }
  }
}

// Real-k quadrature of QC staged points (rows[8 q .. 8 q + 8)), software-pipelined over
// groups of G points: while group q's back stage runs, group q + G's table reads (issued
// by its front stage) and group q + 2G's row reads are in flight.  With two waves per
// SIMD (the slot kernel's LDS use) a plain point-by-point loop leaves the FP64 pipe idle
// through every LDS latency.
template <int NP, bool GON, int QC, int G>
__device__ __forceinline__ void pair_quadrature_pipelined(double x0, double x1, double x2,
                                                          const double* rows,
                                                          const double* __restrict__ ph,
                                                          double kr, double* fr, double* fi,
                                                          double* gr, double* gi,
                                                          double& csum,
                                                          const double2* trig) {
  static_assert(QC % G == 0, "the rule length must be a multiple of the group size");
  constexpr int NG = QC / G;
  double4 A[G], B[G];
  PointFront cur[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    A[g] = *reinterpret_cast<const double4*>(rows + 8 * g);
    B[g] = *reinterpret_cast<const double4*>(rows + 8 * g + 4);
  }
#pragma unroll
  for (int g = 0; g < G; ++g) cur[g] = point_front(x0, x1, x2, A[g], B[g], kr, csum, trig);
  if constexpr (NG > 1) {
#pragma unroll
    for (int g = 0; g < G; ++g) {
      A[g] = *reinterpret_cast<const double4*>(rows + 8 * (G + g));
      B[g] = *reinterpret_cast<const double4*>(rows + 8 * (G + g) + 4);
    }
  }
#pragma unroll
  for (int k = 0; k < NG; ++k) {
    // group k's basis values first: LDS reads complete in issue order, so reads issued
    // after the next group's table reads would wait for those as well
    double pv[G * NP];
#pragma unroll
    for (int t = 0; t < G * NP; ++t) pv[t] = ph[k * G * NP + t];
    __builtin_amdgcn_sched_barrier(0);
    PointFront nxt[G];
    if (k + 1 < NG) {
#pragma unroll
      for (int g = 0; g < G; ++g) nxt[g] = point_front(x0, x1, x2, A[g], B[g], kr, csum, trig);
      if (k + 2 < NG) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
          A[g] = *reinterpret_cast<const double4*>(rows + 8 * ((k + 2) * G + g));
          B[g] = *reinterpret_cast<const double4*>(rows + 8 * ((k + 2) * G + g) + 4);
        }
      }
    }
    // keep the stages in this order: the scheduler's LDS latency model would otherwise
    // move the next group's reads down to their first use
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int g = 0; g < G; ++g) point_back<NP, GON>(cur[g], pv + g * NP, fr, fi, gr, gi);
    __builtin_amdgcn_sched_barrier(0);
    if (k + 1 < NG) {
#pragma unroll
      for (int g = 0; g < G; ++g) cur[g] = nxt[g];
    }
  }
}

// As pair_quadrature over table rows that this wave staged in LDS (rows[8 q .. 8 q + 8)).
// QC > 0: the rule length is the compile-time QC (== Q), and the loop is unrolled whole, so
// every LDS read has a constant offset; with FAST (real k) the loop is the software-
// pipelined pair_quadrature_pipelined over groups of kPointGroup points.
template <int NP, bool CPLX, bool TAB = false, bool GON = true, int QC = 0, bool FAST = false,
          bool ROLLED = false>
__device__ __forceinline__ void pair_quadrature_lds(double x0, double x1, double x2,
                                                    const double* rows, const double* ph, int Q,
                                                    double kr, double ki, double* fr, double* fi,
                                                    double* gr, double* gi, double& csum,
                                                    const double2* trig = nullptr) {
  if constexpr (QC > 0 && ROLLED) {  // one point at a time: fewest registers live
#pragma unroll 1
    for (int q = 0; q < QC; ++q) {
      const double4 A = *reinterpret_cast<const double4*>(rows + 8 * q);
      const double4 B = *reinterpret_cast<const double4*>(rows + 8 * q + 4);
      point_kernel<NP, CPLX, TAB, GON, FAST>(x0, x1, x2, A, B, ph + q * NP, kr, ki, fr, fi, gr,
                                             gi, csum, trig);
    }
  } else if constexpr (QC > 0 && FAST && !CPLX && QC % kPointGroup == 0) {
    pair_quadrature_pipelined<NP, GON, QC, kPointGroup>(x0, x1, x2, rows, ph, kr, fr, fi, gr,
                                                        gi, csum, trig);
  } else if constexpr (QC > 0) {
#pragma unroll
    for (int q = 0; q < QC; ++q) {
      const double4 A = *reinterpret_cast<const double4*>(rows + 8 * q);
      const double4 B = *reinterpret_cast<const double4*>(rows + 8 * q + 4);
      point_kernel<NP, CPLX, TAB, GON, FAST>(x0, x1, x2, A, B, ph + q * NP, kr, ki, fr, fi, gr,
                                             gi, csum, trig);
    }
  } else {
#pragma unroll 2
    for (int q = 0; q < Q; ++q) {
      const double4 A = *reinterpret_cast<const double4*>(rows + 8 * q);
      const double4 B = *reinterpret_cast<const double4*>(rows + 8 * q + 4);
      point_kernel<NP, CPLX, TAB, GON, FAST>(x0, x1, x2, A, B, ph + q * NP, kr, ki, fr, fi, gr,
                                             gi, csum, trig);
    }
  }
}

// One thread per target row; blockIdx.y selects a contiguous chunk of elements.
// Every (target, element) contribution is added into zeroed F, G with hardware FP64
// atomics; C is accumulated per thread and added once per chunk.  Used for small
// problems, where the slot schedule does not pay.
template <int NP, bool CPLX>
__global__ void __launch_bounds__(256)
k_assemble(const AssemblyArgs a, int ech) {
  const double* __restrict__ phi = a.phi;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.M) return;
  const double x0 = a.targets[3 * i], x1 = a.targets[3 * i + 1], x2 = a.targets[3 * i + 2];
  const int e0 = blockIdx.y * ech;
  const int e1 = min(a.E, e0 + ech);
  const bool doF = (a.flags & BIE_F) != 0, doG = (a.flags & BIE_G) != 0;
  double csum = 0.0;
  for (int e = e0; e < e1; ++e) {
    const int c =
        select_rule<NP>(x0, x1, x2, a.selnodes + (size_t)e * NP * 4, a.R, a.hyp) - 1;
    double fr[NP], fi[NP], gr[NP], gi[NP];
#pragma unroll
    for (int s = 0; s < NP; ++s) fr[s] = fi[s] = gr[s] = gi[s] = 0.0;
    pair_quadrature<NP, CPLX>(x0, x1, x2, a.elem + ((size_t)e * a.Pe + a.off[c]) * 8,
                              phi + c * a.Qmax * NP, a.len[c], a.kr, a.ki, fr, fi, gr, gi,
                              csum);
    const int* cols = a.ptopo + (size_t)e * NP;
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      const size_t idx = 2 * ((size_t)i + (size_t)cols[s] * (size_t)a.M);
      if (doF) {
        unsafeAtomicAdd(a.F + idx, fr[s]);
        unsafeAtomicAdd(a.F + idx + 1, fi[s]);
      }
      if (doG) {
        unsafeAtomicAdd(a.G + idx, gr[s]);
        unsafeAtomicAdd(a.G + idx + 1, gi[s]);
      }
    }
  }
  if (a.flags & BIE_C) unsafeAtomicAdd(a.C + 2 * i, csum);
}

// LDS-slot kernel.  Workgroup = W waves sharing 64 target rows (lane = row) and one
// segment of elements; each step the waves take one element each, add the element's
// node contributions into LDS slots (ds_add_f64: waves may hit the same slot), and after
// a barrier the nodes whose last element was in this step are flushed: owned columns
// with one coalesced store per row block, shared columns with FP64 atomics.
//
// Latency: all waves leave a barrier together, so nothing would hide a load issued
// there.  Each wave therefore selects the rule of its *next* element and starts loading
// that element's first table rows before the barrier; selection data is laid out in
// step order and read one step ahead.
//
// Pairs whose rule is a mid-edge split rule (quadratic physics only, ~1% of pairs) are
// queued in LDS, moved to the global queue once per segment, and integrated by
// k_assemble_mid.
// Flush of one schedule step's finished nodes: each wave takes every W-th entry, reads the
// node's slot (F re, F im, G re, G im for 64 rows), zeroes it and writes the column
// segment of F and G: a plain (streaming) store for owned columns, atomics for shared
// ones.  Entries go two at a time, both slots read before either is written out, so one
// LDS wait serves two nodes; an owned column's (re, im) pairs leave as one 16-byte store
// per lane (a wave writes 1 KB of the column in one instruction).  Runs concurrently with
// the next step's quadrature, which never touches these slots.
__device__ __forceinline__ void flush_write(const AssemblyArgs& a, int node, int ss, int64_t i,
                                            bool doF, bool doG, double vr, double vi,
                                            double wr, double wi) {
  const size_t idx = 2 * ((size_t)i + (size_t)node * (size_t)a.M);
  if (ss & kSharedBit) {
    if (doF) {
      unsafeAtomicAdd(a.F + idx, vr);
      unsafeAtomicAdd(a.F + idx + 1, vi);
    }
    if (doG) {
      unsafeAtomicAdd(a.G + idx, wr);
      unsafeAtomicAdd(a.G + idx + 1, wi);
    }
  } else {
    // the output is written once and not read back here: keep it out of the caches
    if (doF) __builtin_nontemporal_store(d2v{vr, vi}, reinterpret_cast<d2v*>(a.F + idx));
    if (doG) __builtin_nontemporal_store(d2v{wr, wi}, reinterpret_cast<d2v*>(a.G + idx));
  }
}

template <int W>
__device__ __forceinline__ void flush_step(const AssemblyArgs& a, const SlotArgs& sa, int st,
                                           int wave, int lane, int64_t i, bool active,
                                           double* acc) {
  const bool doF = (a.flags & BIE_F) != 0, doG = (a.flags & BIE_G) != 0;
  const int f1 = sload(sa.flush_begin + st + 1);
  int f = sload(sa.flush_begin + st) + wave;
  for (; f + W < f1; f += 2 * W) {
    const int n0 = sload(sa.flush_node + f), s0 = sload(sa.flush_slot + f);
    const int n1 = sload(sa.flush_node + f + W), s1 = sload(sa.flush_slot + f + W);
    double* p0 = acc + (size_t)(s0 & 0xffff) * 256 + lane;
    double* p1 = acc + (size_t)(s1 & 0xffff) * 256 + lane;
    const double a0 = p0[0], b0 = p0[64], c0 = p0[128], d0 = p0[192];
    const double a1 = p1[0], b1 = p1[64], c1 = p1[128], d1 = p1[192];
    p0[0] = 0.0;
    p0[64] = 0.0;
    p0[128] = 0.0;
    p0[192] = 0.0;
    p1[0] = 0.0;
    p1[64] = 0.0;
    p1[128] = 0.0;
    p1[192] = 0.0;
    if (active) {
      flush_write(a, n0, s0, i, doF, doG, a0, b0, c0, d0);
      flush_write(a, n1, s1, i, doF, doG, a1, b1, c1, d1);
    }
  }
  if (f < f1) {
    const int node = sload(sa.flush_node + f);
    const int ss = sload(sa.flush_slot + f);
    double* p = acc + (size_t)(ss & 0xffff) * 256 + lane;
    const double vr = p[0], vi = p[64], wr = p[128], wi = p[192];
    p[0] = 0.0;
    p[64] = 0.0;
    p[128] = 0.0;
    p[192] = 0.0;
    if (active) flush_write(a, node, ss, i, doF, doG, vr, vi, wr, wi);
  }
}

// One (row block, segment) tile of the slot kernel: 64 targets (one per lane) against the
// segment's elements, W elements per step (one per wave).  acc holds S zeroed slots of
// 4 x 64 doubles and is zero again on return (the schedule flushes every slot it uses by
// the end of the segment); phi is the basis table in LDS; s_q / s_qn the LDS mid-edge
// queue, whose count must be zero on entry (it is zero again on return when NP == 6).
// Step st: quadrature of the step's pairs into their slots, then the flush of the nodes
// whose last element was at step st - 1, then one barrier.  With STAGED, each wave keeps
// its element's vertex-rule rows in its LDS rows (s_rows + wave * 24 Qc) and loads the
// next element's rows into registers a whole pair ahead: the global loads never wait
// behind the flush's stores (vector-memory counters are in order).  Ends on a barrier.
// HOLD (with STAGED): the next element's rows are loaded before the pair's quadrature and
// held in registers through it; without HOLD they are loaded after it, which frees those
// registers where more waves per SIMD cover the load latency instead (the wide instance).
template <int NP, bool CPLX, int W, bool STAGED, bool GON = true, int QC = 0, bool FAST = false,
          int V = 1, bool HOLD = true>
__device__ __forceinline__ void slot_tile(const AssemblyArgs& a, const SlotArgs& sa, int rb,
                                          int seg, double* acc, const double* phi,
                                          const double2* trig, double* s_rows, int2* s_q,
                                          unsigned* s_qn) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t i = (int64_t)rb * 64 + lane;
  const bool active = i < a.M;
  const int64_t ic = active ? i : (a.M - 1);
  const double x0 = a.targets[3 * ic], x1 = a.targets[3 * ic + 1], x2 = a.targets[3 * ic + 2];
  const double kr = a.kr, ki = a.ki;
  double csum = 0.0;
  const int Qc = QC > 0 ? QC : a.len[0];
  // the rows of the element's vertex rules (three on triangles, the one rule of a
  // quadrilateral) and, when there is a rule to select, its selection nodes
  // (three- and six-node physics are triangles only: three vertex rules, compile-time sizes)
  constexpr bool kTri = NP == 3 || NP == 6;
  const int nchunk = kTri ? 12 * Qc : 4 * Qc * (a.R < 3 ? a.R : 3);
  double* my_rows = STAGED ? s_rows + (size_t)wave * stage_wave_doubles(Qc) : nullptr;
  double* my_sel = STAGED ? my_rows + 24 * Qc : nullptr;  // the element's selection nodes
  const int nsel = (kTri || a.R != 1) ? NP * 2 : 0;      // 16-byte chunks of them
  double2 stg[3], stg_sel;
  constexpr int WV = W * V;  // elements per step; wave w takes elements v * W + w, v < V
  const int st0 = sload(sa.seg_step_begin + seg), st1 = sload(sa.seg_step_begin + seg + 1);
  int e = (st0 < st1) ? sload(sa.step_elems + (size_t)st0 * WV + wave) : -1;
  int c = 0;
  double4 A0 = make_double4(0, 0, 0, 0), B0 = A0, A1 = A0, B1 = A0;
  if (e >= 0) {
    const double* sel = sa.step_sel + ((size_t)st0 * WV + wave) * NP * 4;
    if (STAGED) {
      stage_load(a.elem + (size_t)e * a.Pe * 8, nchunk, sel, nsel, lane, stg, stg_sel);
      stage_store(my_rows, nchunk, my_sel, nsel, lane, stg, stg_sel);
      c = select_rule<NP>(x0, x1, x2, my_sel, a.R, a.hyp) - 1;
    } else {
      c = select_rule<NP>(x0, x1, x2, sel, a.R, a.hyp) - 1;
      if (!(NP == 6 && c >= 3))
        load_rows01(a.elem + ((size_t)e * a.Pe + a.off[c]) * 8, a.len[c], A0, B0, A1, B1);
    }
  }
  for (int st = st0; st < st1; ++st) {
#pragma unroll 1
    for (int v = 0; v < V; ++v) {
    const size_t cidx = (size_t)st * WV + v * W + wave;
    // the wave's next element: the step's next sub-element, else the next step's first
    const bool last_v = v + 1 == V;
    const size_t nidx = last_v ? (size_t)(st + 1) * WV + wave : cidx + W;
    const int e_nxt = (!last_v || st + 1 < st1) ? sload(sa.step_elems + nidx) : -1;
    // unconditional (element 0 stands in at the segment's end): a path that skipped the
    // LDS store below would leave these loads outstanding at the loop head, where the
    // compiler would wait for them, and with them for the previous step's column stores
    if (STAGED && HOLD)
      stage_load(a.elem + (size_t)max(e_nxt, 0) * a.Pe * 8, nchunk,
                 sa.step_sel + (e_nxt >= 0 ? nidx : 0) * NP * 4, e_nxt >= 0 ? nsel : 0, lane,
                 stg, stg_sel);
    if (e >= 0) {
      const bool mid = NP == 6 && c >= 3;
      if (active && !mid) {
        double fr[NP], fi[NP], gr[NP], gi[NP];
#pragma unroll
        for (int s = 0; s < NP; ++s) fr[s] = fi[s] = gr[s] = gi[s] = 0.0;
        if (STAGED)
          pair_quadrature_lds<NP, CPLX, true, GON, QC, FAST>(
              x0, x1, x2, my_rows + (size_t)c * Qc * 8, phi + c * a.Qmax * NP, Qc, kr, ki, fr,
              fi, gr, gi, csum, trig);
        else
          pair_quadrature_pre<NP, CPLX, true, GON>(
              x0, x1, x2, a.elem + ((size_t)e * a.Pe + a.off[c]) * 8, phi + c * a.Qmax * NP,
              a.len[c], kr, ki, A0, B0, A1, B1, fr, fi, gr, gi, csum, trig);
        const int* so = sa.step_slot + cidx * NP;
#pragma unroll
        for (int s = 0; s < NP; ++s) {
          double* p = acc + (size_t)sload(so + s) * 256 + lane;
          atomicAdd(p, fr[s]);
          atomicAdd(p + 64, fi[s]);
          if constexpr (GON) {
            atomicAdd(p + 128, gr[s]);
            atomicAdd(p + 192, gi[s]);
          }
        }
      }
      if (NP == 6 && sa.push_mid) {
        const bool push = mid && active;
        const unsigned long long mask = __ballot(push);
        if (mask) {
          const int leader = __ffsll((long long)mask) - 1;
          unsigned base = 0;
          if (lane == leader) base = atomicAdd(s_qn, (unsigned)__popcll(mask));
          base = __shfl(base, leader);
          if (push) {
            const unsigned pos = base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
            const int2 entry = make_int2((int)i, e | ((c - 3) << 28));
            if (pos < (unsigned)kMidLds) {
              s_q[pos] = entry;
            } else {  // LDS queue full: go straight to the global queue
              const unsigned g = atomicAdd(sa.mid_count, 1u);
              if ((int64_t)g < sa.mid_cap) sa.mid_queue[g] = entry;
            }
          }
        }
      }
    }
    // next element: rule selection, and its rows (LDS: after this wave's reads above)
    int c_nxt = 0;
    if (STAGED) {  // only this wave reads my_rows and my_sel, and its reads above are done
      if (!HOLD)
        stage_load(a.elem + (size_t)max(e_nxt, 0) * a.Pe * 8, nchunk,
                   sa.step_sel + (e_nxt >= 0 ? nidx : 0) * NP * 4, e_nxt >= 0 ? nsel : 0,
                   lane, stg, stg_sel);
      stage_store(my_rows, nchunk, my_sel, nsel, lane, stg, stg_sel);
      if (e_nxt >= 0) c_nxt = select_rule<NP>(x0, x1, x2, my_sel, a.R, a.hyp) - 1;
    } else if (e_nxt >= 0) {
      c_nxt = select_rule<NP>(x0, x1, x2, sa.step_sel + nidx * NP * 4, a.R, a.hyp) - 1;
      if (!(NP == 6 && c_nxt >= 3))
        load_rows01(a.elem + ((size_t)e_nxt * a.Pe + a.off[c_nxt]) * 8, a.len[c_nxt], A0, B0,
                    A1, B1);
    }
    e = e_nxt;
    c = c_nxt;
    }
    // the previous step's finished nodes: their slots were complete at the last barrier
    if (st > st0) flush_step<W>(a, sa, st - 1, wave, lane, i, active, acc);
    lds_barrier();
  }
  if (st1 > st0) {
    flush_step<W>(a, sa, st1 - 1, wave, lane, i, active, acc);
    lds_barrier();
  }
  if (NP == 6 && wave == 0) {  // after a barrier: the LDS queue is complete
    const unsigned n = min(*s_qn, (unsigned)kMidLds);
    if (n) {
      unsigned base = 0;
      if (lane == 0) base = atomicAdd(sa.mid_count, n);
      base = __shfl(base, 0);
      for (unsigned t = lane; t < n; t += 64)
        if ((int64_t)(base + t) < sa.mid_cap) sa.mid_queue[base + t] = s_q[t];
    }
    if (lane == 0) *s_qn = 0u;
  }
  if (active && (a.flags & BIE_C)) unsafeAtomicAdd(a.C + 2 * i, csum);
  lds_barrier();
}

// Split form of slot_tile (real k, vertex rules of QC points staged in LDS, the branch-free
// table sincos): SPLIT waves share each element of a step, so a workgroup of W * SPLIT
// waves keeps W elements per step - the same slot schedule and LDS slots - with SPLIT
// times the waves resident to hide each other's latency.  Wave h of an element integrates
// the points h QH .. h QH + QH - 1 (QH = QC / SPLIT) of the pair's rule from its own staged
// rows and adds its partial node sums to the slots (LDS atomics, as the waves of different
// elements already do); every wave selects the rule, the first one queues mid-edge pairs;
// the flush is spread over all waves.
template <int NP, int W, bool GON, int QC, int SPLIT>
__device__ __forceinline__ void slot_tile_split(const AssemblyArgs& a, const SlotArgs& sa,
                                                int rb, int seg, double* acc, const double* phi,
                                                const double2* trig, double* s_rows, int2* s_q,
                                                unsigned* s_qn) {
  constexpr int QH = QC / SPLIT;
  static_assert(QH * SPLIT == QC, "the split must divide the rule length");
  constexpr int NV = (12 * QH + 63) / 64;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int we = wave / SPLIT, half = wave - we * SPLIT;
  const int64_t i = (int64_t)rb * 64 + lane;
  const bool active = i < a.M;
  const int64_t ic = active ? i : (a.M - 1);
  const double x0 = a.targets[3 * ic], x1 = a.targets[3 * ic + 1], x2 = a.targets[3 * ic + 2];
  const double kr = a.kr;
  double csum = 0.0;
  double* my_rows = s_rows + (size_t)wave * split_stage_wave_doubles(QH);
  double* my_sel = my_rows + 24 * QH;
  constexpr int nsel = NP * 2;
  const double* my_phi = phi + half * QH * NP;
  double2 stg[NV], stg_sel;
  const int st0 = sload(sa.seg_step_begin + seg), st1 = sload(sa.seg_step_begin + seg + 1);
  int e = (st0 < st1) ? sload(sa.step_elems + (size_t)st0 * W + we) : -1;
  int c = 0;
  if (e >= 0) {
    stage_load_half<QC, QH, NV>(a.elem + (size_t)e * a.Pe * 8, half,
                                sa.step_sel + ((size_t)st0 * W + we) * NP * 4, nsel, lane, stg,
                                stg_sel);
    stage_store_half<QH, NV>(my_rows, my_sel, nsel, lane, stg, stg_sel);
    c = select_rule<NP>(x0, x1, x2, my_sel, a.R, a.hyp) - 1;
  }
  for (int st = st0; st < st1; ++st) {
    const size_t nidx = (size_t)(st + 1) * W + we;
    const int e_nxt = (st + 1 < st1) ? sload(sa.step_elems + nidx) : -1;
    // unconditional, as in slot_tile
    stage_load_half<QC, QH, NV>(a.elem + (size_t)max(e_nxt, 0) * a.Pe * 8, half,
                                sa.step_sel + (e_nxt >= 0 ? nidx : 0) * NP * 4,
                                e_nxt >= 0 ? nsel : 0, lane, stg, stg_sel);
    if (e >= 0) {
      const bool mid = NP == 6 && c >= 3;
      if (active && !mid) {
        double fr[NP], fi[NP], gr[NP], gi[NP];
#pragma unroll
        for (int s = 0; s < NP; ++s) fr[s] = fi[s] = gr[s] = gi[s] = 0.0;
        pair_quadrature_lds<NP, false, true, GON, QH, true, (NP == 6 && (GON || QH > 3))>(
            x0, x1, x2, my_rows + (size_t)c * QH * 8, my_phi + c * a.Qmax * NP, QH, kr, 0.0, fr,
            fi, gr, gi, csum, trig);
        const int* so = sa.step_slot + ((size_t)st * W + we) * NP;
#pragma unroll
        for (int s = 0; s < NP; ++s) {
          double* p = acc + (size_t)sload(so + s) * 256 + lane;
          atomicAdd(p, fr[s]);
          atomicAdd(p + 64, fi[s]);
          if constexpr (GON) {
            atomicAdd(p + 128, gr[s]);
            atomicAdd(p + 192, gi[s]);
          }
        }
      }
      if (NP == 6 && sa.push_mid && half == 0) {
        const bool push = mid && active;
        const unsigned long long mask = __ballot(push);
        if (mask) {
          const int leader = __ffsll((long long)mask) - 1;
          unsigned base = 0;
          if (lane == leader) base = atomicAdd(s_qn, (unsigned)__popcll(mask));
          base = __shfl(base, leader);
          if (push) {
            const unsigned pos = base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
            const int2 entry = make_int2((int)i, e | ((c - 3) << 28));
            if (pos < (unsigned)kMidLds) {
              s_q[pos] = entry;
            } else {
              const unsigned g = atomicAdd(sa.mid_count, 1u);
              if ((int64_t)g < sa.mid_cap) sa.mid_queue[g] = entry;
            }
          }
        }
      }
    }
    int c_nxt = 0;
    stage_store_half<QH, NV>(my_rows, my_sel, nsel, lane, stg, stg_sel);
    if (e_nxt >= 0) c_nxt = select_rule<NP>(x0, x1, x2, my_sel, a.R, a.hyp) - 1;
    if (st > st0) flush_step<W * SPLIT>(a, sa, st - 1, wave, lane, i, active, acc);
    lds_barrier();
    e = e_nxt;
    c = c_nxt;
  }
  if (st1 > st0) {
    flush_step<W * SPLIT>(a, sa, st1 - 1, wave, lane, i, active, acc);
    lds_barrier();
  }
  if (NP == 6 && wave == 0) {  // after a barrier: the LDS queue is complete
    const unsigned n = min(*s_qn, (unsigned)kMidLds);
    if (n) {
      unsigned base = 0;
      if (lane == 0) base = atomicAdd(sa.mid_count, n);
      base = __shfl(base, 0);
      for (unsigned t = lane; t < n; t += 64)
        if ((int64_t)(base + t) < sa.mid_cap) sa.mid_queue[base + t] = s_q[t];
    }
    if (lane == 0) *s_qn = 0u;
  }
  if (active && (a.flags & BIE_C)) unsafeAtomicAdd(a.C + 2 * i, csum);
  lds_barrier();
}

// The split slot kernel: W * SPLIT waves per workgroup (slot_tile_split).
template <int NP, int W, bool GON, int QC, int SPLIT>
__global__ void __launch_bounds__(64 * W * SPLIT)
k_assemble_slots_split(const AssemblyArgs a, const SlotArgs sa) {
  extern __shared__ double lds[];
  double* acc = lds;
  double* s_phi = lds + (size_t)sa.S * 256;
  const int nphi = a.R * a.Qmax * NP;
  const double2* s_trig = reinterpret_cast<const double2*>(s_phi + phi_pad(nphi));
  double* s_rows = s_phi + phi_pad(nphi) + kTrigDoubles;
  int2* s_q = reinterpret_cast<int2*>(s_rows + split_stage_doubles(QC, SPLIT));
  unsigned* s_qn = reinterpret_cast<unsigned*>(s_q + kMidLds);
  const int b = blockIdx.x;
  const int per_xcd = (int)((sa.n_tiles + 7) / 8);
  const int64_t tile = (int64_t)(b & 7) * per_xcd + (b >> 3);
  if ((b >> 3) >= per_xcd || tile >= sa.n_tiles) return;
  int rb, seg;
  if (sa.seg_major) {
    seg = (int)(tile / sa.n_rowblocks);
    rb = (int)(tile - (int64_t)seg * sa.n_rowblocks);
  } else {
    rb = (int)(tile / sa.n_segments);
    seg = (int)(tile - (int64_t)rb * sa.n_segments);
  }
  for (int t = threadIdx.x; t < phi_pad(nphi) + kTrigDoubles; t += blockDim.x) s_phi[t] = a.phi[t];
  for (int t = threadIdx.x; t < sa.S * 256; t += blockDim.x) acc[t] = 0.0;
  if (threadIdx.x == 0) *s_qn = 0u;
  lds_barrier();
  slot_tile_split<NP, W, GON, QC, SPLIT>(a, sa, rb, seg, acc, s_phi, s_trig, s_rows, s_q, s_qn);
}

// GON is whether G is formed (the BIE_G flag): F-only launches use their own instance.
// QC > 0: instance for vertex rules of exactly QC points (staged in LDS, QC <= kStageMaxQ)
// with the e^{ikr} of sincos_tab_fast (FAST: real k, |k| r below its bound); QC == 0: any
// rule length, staged or not.
template <int NP, bool CPLX, int W, bool GON, int QC = 0, bool FAST = false, int OCC = 1,
          int V = 1>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(2 * OCC)))
k_assemble_slots(const AssemblyArgs a, const SlotArgs sa) {
  extern __shared__ double lds[];
  double* acc = lds;                              // [slot][4][64]
  double* s_phi = lds + (size_t)sa.S * 256;
  const int nphi = a.R * a.Qmax * NP;
  const int nstage = stage_doubles(a.len[0], W);
  const double2* s_trig = reinterpret_cast<const double2*>(s_phi + phi_pad(nphi));
  double* s_rows = s_phi + phi_pad(nphi) + kTrigDoubles;
  int2* s_q = reinterpret_cast<int2*>(s_rows + nstage);
  unsigned* s_qn = reinterpret_cast<unsigned*>(s_q + kMidLds);
  // XCD-aware tile order: consecutive block ids go to different XCDs, so give each XCD
  // a contiguous run of (row block, segment) tiles.
  const int b = blockIdx.x;
  const int per_xcd = (int)((sa.n_tiles + 7) / 8);
  const int64_t tile = (int64_t)(b & 7) * per_xcd + (b >> 3);
  if ((b >> 3) >= per_xcd || tile >= sa.n_tiles) return;
  int rb, seg;
  if (sa.seg_major) {  // an XCD's concurrent workgroups share one segment's tables in L2
    seg = (int)(tile / sa.n_rowblocks);
    rb = (int)(tile - (int64_t)seg * sa.n_rowblocks);
  } else {             // an XCD's concurrent workgroups share row blocks (shared-column atomics)
    rb = (int)(tile / sa.n_segments);
    seg = (int)(tile - (int64_t)rb * sa.n_segments);
  }
  // basis table and trig table: one buffer of phi_pad(nphi) + kTrigDoubles doubles
  for (int t = threadIdx.x; t < phi_pad(nphi) + kTrigDoubles; t += blockDim.x) s_phi[t] = a.phi[t];
  for (int t = threadIdx.x; t < sa.S * 256; t += blockDim.x) acc[t] = 0.0;
  if (threadIdx.x == 0) *s_qn = 0u;
  lds_barrier();
  if constexpr (QC > 0) {
    static_assert(QC <= kStageMaxQ, "compile-time rule lengths are staged in LDS");
    slot_tile<NP, CPLX, W, true, GON, QC, FAST, V, W != kWideWaves>(a, sa, rb, seg, acc, s_phi,
                                                                    s_trig, s_rows, s_q, s_qn);
  } else if (nstage) {
    slot_tile<NP, CPLX, W, true, GON>(a, sa, rb, seg, acc, s_phi, s_trig, s_rows, s_q, s_qn);
  } else {
    slot_tile<NP, CPLX, W, false, GON>(a, sa, rb, seg, acc, s_phi, s_trig, s_rows, s_q, s_qn);
  }
}

// Uneven batch: many meshes (same element kinds and rule set) in one persistent launch.
// Each workgroup stages the shared basis table once, then takes (instance, row block,
// segment) tiles until the tile list is exhausted.  The list is cut into kXcds contiguous
// queues, one per XCD (workgroup b runs on XCD b mod 8), so an XCD's concurrent tiles
// share a segment's element tables in its L2; a workgroup whose queue is empty takes tiles
// from the following queues.  Every queue's counter only grows, so each workgroup leaves
// after finding all queues exhausted once.
// QC > 0: the real-k instance for vertex rules of exactly QC points with the e^{ikr} of
// sincos_tab_fast (every instance's |k| r below its bound), as k_assemble_slots.
template <int NP, bool CPLX, int W, bool GON = true, int QC = 0>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(2)))
k_assemble_batch(const BatchInst* __restrict__ inst, const int4* __restrict__ tiles,
                 int n_tiles, unsigned* __restrict__ counters, const double* __restrict__ phi,
                 int nphi, int nstage, int S) {
  extern __shared__ double lds[];
  double* acc = lds;
  double* s_phi = lds + (size_t)S * 256;
  const double2* s_trig = reinterpret_cast<const double2*>(s_phi + phi_pad(nphi));
  double* s_rows = s_phi + phi_pad(nphi) + kTrigDoubles;
  int2* s_q = reinterpret_cast<int2*>(s_rows + nstage);
  unsigned* s_qn = reinterpret_cast<unsigned*>(s_q + kMidLds);
  int* s_tile = reinterpret_cast<int*>(s_qn + 1);
  for (int t = threadIdx.x; t < phi_pad(nphi) + kTrigDoubles; t += blockDim.x) s_phi[t] = phi[t];
  for (int t = threadIdx.x; t < S * 256; t += blockDim.x) acc[t] = 0.0;
  if (threadIdx.x == 0) *s_qn = 0u;
  const int x0 = blockIdx.x & (kXcds - 1);
  for (int q = 0; q < kXcds; ++q) {
    const int x = (x0 + q) & (kXcds - 1);
    const int begin = (int)(((int64_t)n_tiles * x) / kXcds);
    const int end = (int)(((int64_t)n_tiles * (x + 1)) / kXcds);
    for (;;) {
      if (threadIdx.x == 0) {
        const unsigned got = atomicAdd(counters + x, 1u);
        *s_tile = (got < (unsigned)(end - begin)) ? begin + (int)got : -1;
      }
      lds_barrier();
      const int t = __builtin_amdgcn_readfirstlane(*s_tile);
      lds_barrier();  // every wave has read s_tile before thread 0 may write it again
      if (t < 0) break;
      const int4 td = tiles[t];
      const int ii = __builtin_amdgcn_readfirstlane(td.x);
      const int rb = __builtin_amdgcn_readfirstlane(td.y);
      const int seg = __builtin_amdgcn_readfirstlane(td.z);
      if constexpr (QC > 0)
        slot_tile<NP, false, W, true, GON, QC, true>(inst[ii].a, inst[ii].sa, rb, seg, acc,
                                                     s_phi, s_trig, s_rows, s_q, s_qn);
      else if (nstage)
        slot_tile<NP, CPLX, W, true>(inst[ii].a, inst[ii].sa, rb, seg, acc, s_phi, s_trig,
                                     s_rows, s_q, s_qn);
      else
        slot_tile<NP, CPLX, W, false>(inst[ii].a, inst[ii].sa, rb, seg, acc, s_phi, s_trig,
                                      s_rows, s_q, s_qn);
    }
  }
}

// One mid-edge split-rule pair (rotated_midpoint_triangular_quadpoints,
// quadrature.jl:117-144): target i, element e, rule c (3..5): the F and G contributions of
// the element's 6 nodes and the C integrand.
__device__ __forceinline__ void mid_pair_values(const AssemblyArgs& a, int64_t i, int e, int c,
                                                int cplx, double (&fr)[6], double (&fi)[6],
                                                double (&gr)[6], double (&gi)[6],
                                                double& csum) {
  const double x0 = a.targets[3 * i], x1 = a.targets[3 * i + 1], x2 = a.targets[3 * i + 2];
  csum = 0.0;
#pragma unroll
  for (int s = 0; s < 6; ++s) fr[s] = fi[s] = gr[s] = gi[s] = 0.0;
  const double* tab = a.elem + ((size_t)e * a.Pe + a.off[c]) * 8;
  const double* ph = a.phi + c * a.Qmax * 6;
  if (cplx)
    pair_quadrature<6, true>(x0, x1, x2, tab, ph, a.len[c], a.kr, a.ki, fr, fi, gr, gi, csum);
  else
    pair_quadrature<6, false>(x0, x1, x2, tab, ph, a.len[c], a.kr, a.ki, fr, fi, gr, gi, csum);
}

// A mid-edge pair added atomically after the slot kernel has written its columns.
__device__ __forceinline__ void mid_pair(const AssemblyArgs& a, int64_t i, int e, int c,
                                         int cplx) {
  const bool doF = (a.flags & BIE_F) != 0, doG = (a.flags & BIE_G) != 0;
  double fr[6], fi[6], gr[6], gi[6], csum;
  mid_pair_values(a, i, e, c, cplx, fr, fi, gr, gi, csum);
  const int* cols = a.ptopo + (size_t)e * 6;
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const size_t idx = 2 * ((size_t)i + (size_t)cols[s] * (size_t)a.M);
    if (doF) {
      unsafeAtomicAdd(a.F + idx, fr[s]);
      unsafeAtomicAdd(a.F + idx + 1, fi[s]);
    }
    if (doG) {
      unsafeAtomicAdd(a.G + idx, gr[s]);
      unsafeAtomicAdd(a.G + idx + 1, gi[s]);
    }
  }
  if (a.flags & BIE_C) unsafeAtomicAdd(a.C + 2 * i, csum);
}

// Mid-edge pairs from the queue the slot kernel filled: one thread per pair.
__device__ __forceinline__ void mid_pairs(const AssemblyArgs& a, const unsigned* count,
                                          const int2* queue, int64_t cap, int cplx,
                                          int64_t first, int64_t stride) {
  const int64_t n = min((int64_t)*count, cap);
  for (int64_t t = first; t < n; t += stride) {
    const int2 en = queue[t];
    mid_pair(a, en.x, en.y & 0x0fffffff, 3 + (en.y >> 28), cplx);
  }
}

// Recorded collocation mid-edge pairs, part 1: one thread per pair (key: element bits
// 34..61, target 2..33, rule - 3 0..1; sorted, so a wave mostly shares one element) writes
// its 6 x {F re, F im, G re, G im} to vals[24 p ..) with plain stores; C is added here.
__global__ void __launch_bounds__(256)
k_mid_compute(const AssemblyArgs a, const unsigned long long* __restrict__ keys, int64_t n,
              double* __restrict__ vals, int cplx) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += stride) {
    const unsigned long long key = keys[p];
    const int64_t i = (int64_t)((key >> 2) & 0xffffffffull);
    double fr[6], fi[6], gr[6], gi[6], csum;
    mid_pair_values(a, i, (int)(key >> 34), 3 + (int)(key & 3ull), cplx, fr, fi, gr, gi, csum);
    double4* out = reinterpret_cast<double4*>(vals + (size_t)p * 24);
#pragma unroll
    for (int s = 0; s < 6; ++s) out[s] = make_double4(fr[s], fi[s], gr[s], gi[s]);
    if (a.flags & BIE_C) unsafeAtomicAdd(a.C + 2 * i, csum);
  }
}

// Part 2: one thread per distinct (column, row) entry (sorted by column, then row) sums its
// pairs' contributions (refs[off[t] .. off[t+1]) = pair * 6 + node) and adds them to F and G
// with a plain read-modify-write: no two threads share an entry, and the slot kernel's
// writes precede this kernel on the stream.
__global__ void __launch_bounds__(256)
k_mid_scatter(const AssemblyArgs a, const unsigned long long* __restrict__ entries,
              const int* __restrict__ off, const int* __restrict__ refs, int64_t n,
              const double* __restrict__ vals) {
  const bool doF = (a.flags & BIE_F) != 0, doG = (a.flags & BIE_G) != 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += stride) {
    const unsigned long long u = entries[t];
    const size_t col = (size_t)(u >> 32), row = (size_t)(u & 0xffffffffull);
    double fr = 0.0, fi = 0.0, gr = 0.0, gi = 0.0;
    for (int r = off[t]; r < off[t + 1]; ++r) {
      const double4 v = *reinterpret_cast<const double4*>(vals + (size_t)refs[r] * 4);
      fr += v.x;
      fi += v.y;
      gr += v.z;
      gi += v.w;
    }
    const size_t idx = row + col * (size_t)a.M;
    if (doF) {
      double2* f = reinterpret_cast<double2*>(a.F) + idx;
      const double2 o = *f;
      *f = make_double2(o.x + fr, o.y + fi);
    }
    if (doG) {
      double2* g = reinterpret_cast<double2*>(a.G) + idx;
      const double2 o = *g;
      *g = make_double2(o.x + gr, o.y + gi);
    }
  }
}

__global__ void __launch_bounds__(256)
k_assemble_mid(const AssemblyArgs a, const unsigned* __restrict__ count,
               const int2* __restrict__ queue, int64_t cap, int cplx) {
  mid_pairs(a, count, queue, cap, cplx, blockIdx.x * (int64_t)blockDim.x + threadIdx.x,
            (int64_t)gridDim.x * blockDim.x);
}

// Mid-edge pairs of every quadratic instance of a batch (blockIdx.y = instance).
__global__ void __launch_bounds__(256)
k_assemble_mid_batch(const BatchInst* __restrict__ inst, int cplx) {
  const BatchInst& b = inst[blockIdx.y];
  if (b.a.NP != 6 || b.a.M == 0) return;
  mid_pairs(b.a, b.sa.mid_count, b.sa.mid_queue, b.sa.mid_cap, cplx,
            blockIdx.x * (int64_t)blockDim.x + threadIdx.x, (int64_t)gridDim.x * blockDim.x);
}

// Per-call preparation of every instance of a batch (blockIdx.y = instance): C zeroed,
// the shared columns of enabled F / G zeroed (the batch kernel adds into them), disabled
// non-NULL outputs zeroed whole, and the mid-edge queue count reset.
__global__ void __launch_bounds__(256) k_batch_prep(const BatchInst* __restrict__ inst) {
  const BatchInst& b = inst[blockIdx.y];
  const AssemblyArgs& a = b.a;
  const double2 z = make_double2(0.0, 0.0);
  if (blockIdx.x == 0 && threadIdx.x == 0 && b.sa.mid_count) *b.sa.mid_count = 0u;
  if (a.C && blockIdx.x == 0)
    for (int64_t r = threadIdx.x; r < a.M; r += blockDim.x) reinterpret_cast<double2*>(a.C)[r] = z;
  double* outs[2] = {a.F, a.G};
  const unsigned bits[2] = {BIE_F, BIE_G};
  for (int o = 0; o < 2; ++o) {
    double2* X = reinterpret_cast<double2*>(outs[o]);
    if (!X) continue;
    // workgroups stride over the columns to zero, threads over the rows of a column
    const bool whole = !(a.flags & bits[o]);
    const int64_t ncols = whole ? a.N : b.n_zero;
    for (int64_t c = blockIdx.x; c < ncols; c += gridDim.x) {
      double2* col = X + (size_t)(whole ? c : b.zero_cols[c]) * a.M;
      for (int64_t r = threadIdx.x; r < a.M; r += blockDim.x) col[r] = z;
    }
  }
}

// Zero whole columns j of a column-major M x N complex matrix, j from a list: workgroups
// stride over the listed columns, threads over the rows of a column.
__global__ void k_zero_columns(double2* __restrict__ X, int64_t M, const int* __restrict__ cols,
                               int ncols) {
  const double2 z = make_double2(0.0, 0.0);
  for (int c = blockIdx.x; c < ncols; c += gridDim.x) {
    double2* col = X + (size_t)cols[c] * M;
    for (int64_t r = threadIdx.x; r < M; r += blockDim.x) col[r] = z;
  }
}

}  // namespace

__global__ void k_zero_u32(unsigned* p) { *p = 0u; }

__global__ void k_zero_u32_n(unsigned* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

void launch_fill_zero_u32(unsigned* p, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, p);
  BIE_HIP_CHECK(hipGetLastError());
}

void launch_fill_zero(double* p, int64_t n_doubles, hipStream_t s) {
  if (n_doubles <= 0) return;
  const int64_t n2 = n_doubles / 2;
  if (n2 > 0) {
    int64_t blocks = (n2 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, s,
                       reinterpret_cast<double2*>(p), n2);
  }
  if (n_doubles & 1) hipLaunchKernelGGL(k_fill_zero_tail, dim3(1), dim3(1), 0, s, p + n_doubles - 1);
  BIE_HIP_CHECK(hipGetLastError());
}

void launch_geometry(const double* coords, const int* topo, int gk, int E, const double* pts,
                     int Pe, double* elem, hipStream_t s) {
  const int64_t n = (int64_t)E * Pe;
  if (n == 0) return;
  hipLaunchKernelGGL(k_geometry, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, coords,
                     topo, gk, E, pts, Pe, elem);
  BIE_HIP_CHECK(hipGetLastError());
}

void launch_gather_selnodes(const double* sources, const int* ptopo, int NP, int E,
                            double* sel, hipStream_t s) {
  const int n = E * NP;
  if (n == 0) return;
  hipLaunchKernelGGL(k_gather_selnodes, dim3((n + 255) / 256), dim3(256), 0, s, sources, ptopo,
                     NP, E, sel);
  BIE_HIP_CHECK(hipGetLastError());
}

void launch_select(const double* targets, int64_t M, const double* sel, int NP, int E, int R,
                   int hyp, int8_t* out, hipStream_t s) {
  if (M == 0 || E == 0) return;
  int64_t blocks = (M * (int64_t)E + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  dim3 grid((unsigned)blocks);
  if (R == 1)  // one rule (quadrilaterals): the selection is rule 1 whatever the nodes
    hipLaunchKernelGGL(k_select<1>, grid, dim3(256), 0, s, targets, M, sel, E, R, hyp, out);
  else if (NP == 1)
    hipLaunchKernelGGL(k_select<1>, grid, dim3(256), 0, s, targets, M, sel, E, R, hyp, out);
  else if (NP == 3)
    hipLaunchKernelGGL(k_select<3>, grid, dim3(256), 0, s, targets, M, sel, E, R, hyp, out);
  else
    hipLaunchKernelGGL(k_select<6>, grid, dim3(256), 0, s, targets, M, sel, E, R, hyp, out);
  BIE_HIP_CHECK(hipGetLastError());
}

void launch_zero_columns(double* X, int64_t M, const int* cols, int ncols, hipStream_t s) {
  if (ncols <= 0 || M <= 0) return;
  const int blocks = ncols < 8192 ? ncols : 8192;
  hipLaunchKernelGGL(k_zero_columns, dim3((unsigned)blocks), dim3(256), 0, s,
                     reinterpret_cast<double2*>(X), M, cols, ncols);
  BIE_HIP_CHECK(hipGetLastError());
}

void launch_assemble_slots(const AssemblyArgs& a, const SlotArgs& sa, bool fast, hipStream_t s) {
  if (a.M == 0 || a.E == 0) return;
  const bool tuned_ok = fast && a.ki == 0.0 && (a.NP == 3 || a.NP == 6) && sa.occupancy == 1 &&
                        sa.split == 1 && (a.len[0] == 4 || a.len[0] == 6 || a.len[0] == 16);
  BIE_REQUIRE(sa.W == kSlotWaves || ((sa.W == 2 * kSlotWaves || sa.W == kWideWaves) && tuned_ok),
              BIE_EINVAL, "slot schedule built for another wave count");
  // waves of the workgroup: kWideWaves for the wide instance, else kSlotWaves (the two-element
  // instance keeps 8 waves with 16 elements per step)
  const int kw = sa.W == kWideWaves ? kWideWaves : kSlotWaves;
  const int64_t per_xcd = (sa.n_tiles + 7) / 8;
  BIE_REQUIRE(per_xcd * 8 <= 2147483647LL, BIE_EINVAL, "too many tiles");
  const dim3 grid((unsigned)(per_xcd * 8));
  BIE_REQUIRE(sa.phi_in_lds, BIE_EINVAL, "slot kernel needs the basis table in LDS");
  // (only three- and six-node physics have split instances)
  const int split =
      (sa.split == 2 && fast && a.ki == 0.0 && (a.NP == 3 || a.NP == 6)) ? 2 : 1;
  const size_t shmem = sizeof(double) * ((size_t)sa.S * 256 + phi_pad(a.R * a.Qmax * a.NP) +
                                         kTrigDoubles +
                                         (split == 2 ? split_stage_doubles(a.len[0], 2)
                                                     : stage_doubles(a.len[0], kw))) +
                       sizeof(int2) * kMidLds + 16;
  BIE_REQUIRE(shmem <= 160 * 1024, BIE_EINVAL, "slot kernel LDS request exceeds 160 KiB");
  const bool cplx = a.ki != 0.0;
  constexpr int W = kSlotWaves;
  auto go = [&](auto kern) {
    BIE_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    hipLaunchKernelGGL(kern, grid, dim3(64 * kw * split), shmem, s, a, sa);
  };
  const bool gon = (a.flags & BIE_G) != 0;
  if (split == 2) {
    auto split_go = [&](auto np, auto g) -> bool {
      constexpr int NPc = decltype(np)::value;
      constexpr bool Gc = decltype(g)::value;
      switch (a.len[0]) {
        case 4: go(k_assemble_slots_split<NPc, W, Gc, 4, 2>); return true;
        case 6: go(k_assemble_slots_split<NPc, W, Gc, 6, 2>); return true;
        case 16: go(k_assemble_slots_split<NPc, W, Gc, 16, 2>); return true;
        default: return false;
      }
    };
    using N3 = std::integral_constant<int, 3>;
    using N6 = std::integral_constant<int, 6>;
    const bool ok = a.NP == 3 ? (gon ? split_go(N3{}, std::true_type{}) : split_go(N3{}, std::false_type{}))
                              : (gon ? split_go(N6{}, std::true_type{}) : split_go(N6{}, std::false_type{}));
    BIE_REQUIRE(ok, BIE_EINVAL, "internal: split slot kernel for an unsupported rule length");
    BIE_HIP_CHECK(hipGetLastError());
    return;
  }
  // real k with k r below the fast sincos bound and a vertex-rule length with its own
  // instance: the unrolled, branch-free point loop
  const int qc = fast && !cplx ? a.len[0] : 0;
  const bool occ2 = sa.occupancy == 2;
  auto fast_go = [&](auto np, auto g) -> bool {
    constexpr int NPc = decltype(np)::value;
    constexpr bool Gc = decltype(g)::value;
    if constexpr (NPc != 3 && NPc != 6) {  // other node counts: the plain instance only
      switch (qc) {
        case 4: go(k_assemble_slots<NPc, false, W, Gc, 4, true>); return true;
        case 6: go(k_assemble_slots<NPc, false, W, Gc, 6, true>); return true;
        case 16: go(k_assemble_slots<NPc, false, W, Gc, 16, true>); return true;
        default: return false;
      }
    } else {
      if (occ2) {
        switch (qc) {
          case 4: go(k_assemble_slots<NPc, false, W, Gc, 4, true, 2>); return true;
          case 6: go(k_assemble_slots<NPc, false, W, Gc, 6, true, 2>); return true;
          case 16: go(k_assemble_slots<NPc, false, W, Gc, 16, true, 2>); return true;
          default: return false;
        }
      }
      if (sa.W == kWideWaves) {
        switch (qc) {
          case 4: go(k_assemble_slots<NPc, false, kWideWaves, Gc, 4, true>); return true;
          case 6: go(k_assemble_slots<NPc, false, kWideWaves, Gc, 6, true>); return true;
          case 16: go(k_assemble_slots<NPc, false, kWideWaves, Gc, 16, true>); return true;
          default: return false;
        }
      }
      if (sa.W == 2 * W) {
        switch (qc) {
          case 4: go(k_assemble_slots<NPc, false, W, Gc, 4, true, 1, 2>); return true;
          case 6: go(k_assemble_slots<NPc, false, W, Gc, 6, true, 1, 2>); return true;
          case 16: go(k_assemble_slots<NPc, false, W, Gc, 16, true, 1, 2>); return true;
          default: return false;
        }
      }
      switch (qc) {
        case 4: go(k_assemble_slots<NPc, false, W, Gc, 4, true>); return true;
        case 6: go(k_assemble_slots<NPc, false, W, Gc, 6, true>); return true;
        case 16: go(k_assemble_slots<NPc, false, W, Gc, 16, true>); return true;
        default: return false;
      }
    }
  };
  using GT = std::true_type;
  using GF = std::false_type;
  // fast instance, else the generic one, for the node count NPc
  auto run = [&](auto np) {
    constexpr int NPc = decltype(np)::value;
    if (gon ? fast_go(np, GT{}) : fast_go(np, GF{})) return;
    if (cplx)
      gon ? go(k_assemble_slots<NPc, true, W, true>) : go(k_assemble_slots<NPc, true, W, false>);
    else
      gon ? go(k_assemble_slots<NPc, false, W, true>) : go(k_assemble_slots<NPc, false, W, false>);
  };
  bool done = true;
  switch (a.NP) {
    case 1: run(std::integral_constant<int, 1>{}); break;
    case 4: run(std::integral_constant<int, 4>{}); break;
    case 8: run(std::integral_constant<int, 8>{}); break;
    case 9: run(std::integral_constant<int, 9>{}); break;
    default: done = false;
  }
  using T3 = std::integral_constant<int, 3>;
  using T6 = std::integral_constant<int, 6>;
  if (!done && a.NP == 3)
    done = gon ? fast_go(T3{}, GT{}) : fast_go(T3{}, GF{});
  else if (!done)
    done = gon ? fast_go(T6{}, GT{}) : fast_go(T6{}, GF{});
  if (!done) {
    BIE_REQUIRE(a.NP == 3 || a.NP == 6, BIE_EINVAL, "internal: no slot kernel for this node count");
    if (a.NP == 3) {
      if (cplx)
        gon ? go(k_assemble_slots<3, true, W, true>) : go(k_assemble_slots<3, true, W, false>);
      else
        gon ? go(k_assemble_slots<3, false, W, true>) : go(k_assemble_slots<3, false, W, false>);
    } else {
      if (cplx)
        gon ? go(k_assemble_slots<6, true, W, true>) : go(k_assemble_slots<6, true, W, false>);
      else
        gon ? go(k_assemble_slots<6, false, W, true>) : go(k_assemble_slots<6, false, W, false>);
    }
  }
  BIE_HIP_CHECK(hipGetLastError());
}

void launch_assemble_mid(const AssemblyArgs& a, const SlotArgs& sa, hipStream_t s) {
  if (!uses_mid_rules(a.R, a.NP) || a.M == 0 || a.E == 0) return;
  hipLaunchKernelGGL(k_assemble_mid, dim3(2048), dim3(256), 0, s, a, sa.mid_count,
                     sa.mid_queue, sa.mid_cap, a.ki != 0.0 ? 1 : 0);
  BIE_HIP_CHECK(hipGetLastError());
}

void launch_batch_prep(const BatchInst* d_inst, int n_inst, int64_t max_cols, hipStream_t s) {
  if (n_inst <= 0) return;
  BIE_REQUIRE(n_inst <= 65535, BIE_EINVAL, "too many batch instances");
  int64_t blocks = max_cols;
  if (blocks < 1) blocks = 1;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_batch_prep, dim3((unsigned)blocks, (unsigned)n_inst), dim3(256), 0, s,
                     d_inst);
  BIE_HIP_CHECK(hipGetLastError());
}

size_t batch_shmem(int S, int nphi, int nstage) {
  return sizeof(double) * ((size_t)S * 256 + phi_pad(nphi) + kTrigDoubles + nstage) +
         sizeof(int2) * kMidLds + 16;
}

// Calls fn with the batch kernel instance for (NP, cplx, gon, qc): qc > 0 (real k only)
// selects the unrolled fast-trig instance for rules of qc points, qc == 0 the generic one.
template <class Fn>
static void with_batch_kernel(int NP, bool cplx, bool gon, int qc, Fn&& fn) {
  constexpr int W = kSlotWaves;
  auto pick = [&](auto np) {
    constexpr int N = decltype(np)::value;
    if (qc > 0 && !cplx) {
      auto fast = [&](auto g) -> bool {
        constexpr bool G = decltype(g)::value;
        switch (qc) {
          case 4: fn(k_assemble_batch<N, false, W, G, 4>); return true;
          case 6: fn(k_assemble_batch<N, false, W, G, 6>); return true;
          case 16: fn(k_assemble_batch<N, false, W, G, 16>); return true;
          default: return false;
        }
      };
      if (gon ? fast(std::true_type{}) : fast(std::false_type{})) return;
    }
    if (cplx)
      fn(k_assemble_batch<N, true, W>);
    else
      fn(k_assemble_batch<N, false, W>);
  };
  if (NP == 1)
    pick(std::integral_constant<int, 1>{});
  else if (NP == 3)
    pick(std::integral_constant<int, 3>{});
  else
    pick(std::integral_constant<int, 6>{});
}

void launch_assemble_batch(const BatchInst* d_inst, const int4* d_tiles, int n_tiles,
                           unsigned* d_counter, const double* d_phi, int nphi, int nstage,
                           int S, int NP, bool cplx, bool gon, int qc, int grid,
                           hipStream_t s) {
  if (n_tiles <= 0) return;
  const size_t shmem = batch_shmem(S, nphi, nstage);
  BIE_REQUIRE(shmem <= 160 * 1024, BIE_EINVAL, "batch kernel LDS request exceeds 160 KiB");
  BIE_REQUIRE(grid >= 1, BIE_EINVAL, "empty persistent grid");
  BIE_REQUIRE(qc == 0 || (!cplx && stage_doubles(qc) == nstage), BIE_EINVAL,
              "internal: fast batch instance for another rule length");
  hipLaunchKernelGGL(k_zero_u32_n, dim3(1), dim3(64), 0, s, d_counter, kXcds);
  BIE_HIP_CHECK(hipGetLastError());
  constexpr int W = kSlotWaves;
  with_batch_kernel(NP, cplx, gon, qc, [&](auto kern) {
    BIE_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * W), shmem, s, d_inst, d_tiles,
                       n_tiles, d_counter, d_phi, nphi, nstage, S);
  });
  BIE_HIP_CHECK(hipGetLastError());
}

// Resident workgroups of the batch kernel per CU for slot count S (LDS-bound).
int batch_blocks_per_cu(int NP, bool cplx, bool gon, int qc, int S, int nphi, int nstage) {
  const size_t shmem = batch_shmem(S, nphi, nstage);
  BIE_REQUIRE(shmem <= 160 * 1024, BIE_EINVAL, "batch kernel LDS request exceeds 160 KiB");
  constexpr int W = kSlotWaves;
  int nb = 0;
  with_batch_kernel(NP, cplx, gon, qc, [&](auto kern) {
    BIE_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    BIE_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &nb, reinterpret_cast<const void*>(kern), 64 * W, shmem));
  });
  return nb < 1 ? 1 : nb;
}

void launch_mid_compute(const AssemblyArgs& a, const unsigned long long* keys, int64_t n_pairs,
                        double* vals, hipStream_t s) {
  if (n_pairs <= 0) return;
  int64_t blocks = (n_pairs + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_mid_compute, dim3((unsigned)blocks), dim3(256), 0, s, a, keys, n_pairs,
                     vals, a.ki != 0.0 ? 1 : 0);
  BIE_HIP_CHECK(hipGetLastError());
}

void launch_mid_scatter(const AssemblyArgs& a, int64_t n_pairs, const double* vals,
                        const unsigned long long* entries, const int* off, const int* refs,
                        int64_t n_entries, hipStream_t s) {
  if (n_pairs <= 0 || n_entries <= 0 || !(a.flags & (BIE_F | BIE_G))) return;
  const int64_t blocks = std::min<int64_t>((n_entries + 255) / 256, 16384);
  hipLaunchKernelGGL(k_mid_scatter, dim3((unsigned)blocks), dim3(256), 0, s, a, entries, off, refs,
                     n_entries, vals);
  BIE_HIP_CHECK(hipGetLastError());
}

void launch_mid_recorded(const AssemblyArgs& a, const unsigned long long* keys, int64_t n_pairs,
                         double* vals, const unsigned long long* entries, const int* off,
                         const int* refs, int64_t n_entries, hipStream_t s) {
  if (n_pairs <= 0) return;
  launch_mid_compute(a, keys, n_pairs, vals, s);
  if (n_entries <= 0 || !(a.flags & (BIE_F | BIE_G))) return;
  int64_t blocks = (n_entries + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_mid_scatter, dim3((unsigned)blocks), dim3(256), 0, s, a, entries, off, refs,
                     n_entries, vals);
  BIE_HIP_CHECK(hipGetLastError());
}

void launch_assemble_mid_batch(const BatchInst* d_inst, int n_inst, bool cplx, hipStream_t s) {
  if (n_inst <= 0) return;
  hipLaunchKernelGGL(k_assemble_mid_batch, dim3(256, (unsigned)n_inst), dim3(256), 0, s, d_inst,
                     cplx ? 1 : 0);
  BIE_HIP_CHECK(hipGetLastError());
}

void launch_assemble(const AssemblyArgs& a, hipStream_t s) {
  if (a.M == 0 || a.E == 0) return;
  const int64_t rowblocks = (a.M + 255) / 256;
  BIE_REQUIRE(rowblocks <= 2147483647LL, BIE_EINVAL, "too many targets");
  int64_t chunks = (2048 + rowblocks - 1) / rowblocks;
  if (chunks > a.E) chunks = a.E;
  if (chunks > 65535) chunks = 65535;
  if (chunks < 1) chunks = 1;
  const int ech = (int)((a.E + chunks - 1) / chunks);
  chunks = (a.E + ech - 1) / ech;
  dim3 grid((unsigned)rowblocks, (unsigned)chunks);
  const bool cplx = a.ki != 0.0;
  auto go = [&](auto np) {
    constexpr int NPc = decltype(np)::value;
    if (cplx)
      hipLaunchKernelGGL((k_assemble<NPc, true>), grid, dim3(256), 0, s, a, ech);
    else
      hipLaunchKernelGGL((k_assemble<NPc, false>), grid, dim3(256), 0, s, a, ech);
  };
  if (a.NP == 1) {
    go(std::integral_constant<int, 1>{});
  } else if (a.NP == 4) {
    go(std::integral_constant<int, 4>{});
  } else if (a.NP == 8) {
    go(std::integral_constant<int, 8>{});
  } else if (a.NP == 9) {
    go(std::integral_constant<int, 9>{});
  } else if (a.NP == 3) {
    if (cplx)
      hipLaunchKernelGGL((k_assemble<3, true>), grid, dim3(256), 0, s, a, ech);
    else
      hipLaunchKernelGGL((k_assemble<3, false>), grid, dim3(256), 0, s, a, ech);
  } else {
    if (cplx)
      hipLaunchKernelGGL((k_assemble<6, true>), grid, dim3(256), 0, s, a, ech);
    else
      hipLaunchKernelGGL((k_assemble<6, false>), grid, dim3(256), 0, s, a, ech);
  }
  BIE_HIP_CHECK(hipGetLastError());
}

}  // namespace bie
