// Vector-Jacobian product of the dense collocation assembly (gfx950, FP64):
//
//   grad_X, grad_k of  Phi = Re[ sum_i alpha_i sum_j F_ij(X, k) p_j + sum_i beta_i C_i(X) ]
//
// for complex weights alpha (per target), beta (per target) and p (per collocation node),
// with the rule selection held fixed (it is piecewise constant in X).  X are the mesh
// coordinates (3 x Nc); targets that are mesh nodes (tvert[i] >= 0) move with them, field
// targets stay fixed.  This is the building block of the adjoint gradient of objectives
// that go through F, C and a solve; the reference has no counterpart (its shape
// optimisation example differentiates by finite differences,
// examples/2d_optimization.jl:184-216).
//
// Quadrature is recomputed per pair instead of stored: at each rule point
// y = X N, t_u = X dN/du, t_v = X dN/dv, c = t_u x t_v.  Since the reference's normal is
// n = -c/|c| and J = |c|, the integrands times J are
//   f J w = -(c.d) w h(r),   h = e^{ikr}(ikr - 1)/(4 pi r^3)      (kernels.jl:29-37)
//   k0 J w = (c.d) w / (4 pi r^3)                                 (kernels.jl:48-55)
// with d = y - x, so Phi's integrand is (c.d) H with H = -A1 h + A2/(4 pi r^3),
// A1 = alpha p_h w (p_h = physics-basis interpolant of p), A2 = beta w, and
//   dPsi/dd = H c + (c.d) H_r d / r,  dPsi/dc = H d,  dPsi/dk = (c.d) A1 k e^{ikr}/(4 pi r),
//   H_r = -A1 h' + A2 (-3/(4 pi r^4)),  h' = -k^2 e^{ikr}/(4 pi r^2) - 3 h / r.
// The same formulas hold for complex k (the lossy systems' wavenumbers) in complex
// arithmetic; then Phi's derivatives in Re k and Im k are Re(D) and -Im(D) of
// D = sum (c.d) A1 k e^{ikr} / (4 pi r) (F is holomorphic in k).
#include <hip/hip_runtime.h>

#include "bie_common.hpp"
#include "device_common.hpp"
#include "kernels_math.hpp"

namespace bie {
namespace {

constexpr double kInv4Pi = 0.079577471545947667884;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

template <int NG, int NP, bool CPLX>
__global__ void __launch_bounds__(256) k_adjoint(const AdjointArgs a, int ech) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool active = i < a.M;
  const int64_t ic = active ? i : (a.M - 1);
  const double x0 = a.targets[3 * ic], x1 = a.targets[3 * ic + 1], x2 = a.targets[3 * ic + 2];
  double2 al = active ? a.alpha[ic] : make_double2(0.0, 0.0);
  double2 be = (active && a.beta) ? a.beta[ic] : make_double2(0.0, 0.0);
  const double k = a.k, ki = CPLX ? a.ki : 0.0;
  double gx0 = 0, gx1 = 0, gx2 = 0, gk = 0, gki = 0;
  const int e0 = blockIdx.y * ech;
  const int e1 = min(a.E, e0 + ech);
  for (int e = e0; e < e1; ++e) {
    const int c =
        select_rule<NP>(x0, x1, x2, a.selnodes + (size_t)e * NP * 4, a.R, a.hyp) - 1;
    double X[NG][3];
#pragma unroll
    for (int q = 0; q < NG; ++q) {
      const int node = a.topo[(size_t)e * NG + q];
#pragma unroll
      for (int d = 0; d < 3; ++d) X[q][d] = a.coords[3 * (size_t)node + d];
    }
    double pr[NP], pim[NP];
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      const double2 v = a.p[a.ptopo[(size_t)e * NP + s]];
      pr[s] = v.x;
      pim[s] = v.y;
    }
    double gX[NG][3];
#pragma unroll
    for (int q = 0; q < NG; ++q) gX[q][0] = gX[q][1] = gX[q][2] = 0.0;
    const int Q = a.len[c];
    const double* P0 = a.pts + (size_t)a.off[c] * kPtStride;
    const double* ph0 = a.phi + (size_t)c * a.Qmax * NP;
    for (int q = 0; q < Q; ++q) {
      const double* P = P0 + (size_t)q * kPtStride;
      const double w = P[2];
      double y[3] = {0, 0, 0}, tu[3] = {0, 0, 0}, tv[3] = {0, 0, 0};
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        const double N = P[kPtN + g], du = P[kPtDu + g], dv = P[kPtDv + g];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          y[d] = fma(X[g][d], N, y[d]);
          tu[d] = fma(X[g][d], du, tu[d]);
          tv[d] = fma(X[g][d], dv, tv[d]);
        }
      }
      const double cv[3] = {tu[1] * tv[2] - tu[2] * tv[1], tu[2] * tv[0] - tu[0] * tv[2],
                            tu[0] * tv[1] - tu[1] * tv[0]};
      const double d[3] = {y[0] - x0, y[1] - x1, y[2] - x2};
      const double r2 = fma(d[0], d[0], fma(d[1], d[1], d[2] * d[2]));
      const double rinv = rsqrt_nr(r2);
      const double r = r2 * rinv;
      const double cd = fma(cv[0], d[0], fma(cv[1], d[1], cv[2] * d[2]));
      // p interpolated at the point, times alpha and w
      double phr = 0, phi_ = 0;
#pragma unroll
      for (int s = 0; s < NP; ++s) {
        const double b = ph0[q * NP + s];
        phr = fma(b, pr[s], phr);
        phi_ = fma(b, pim[s], phi_);
      }
      const double A1r = w * (al.x * phr - al.y * phi_), A1i = w * (al.x * phi_ + al.y * phr);
      const double A2r = w * be.x;
      double sn, cs;
      sincos_f64(k * r, &sn, &cs);
      if (CPLX) {  // e^{ikr} = e^{-Im k r} e^{i Re k r}
        const double ex = exp(-ki * r);
        cs *= ex;
        sn *= ex;
      }
      const double i4r = kInv4Pi * rinv;            // 1/(4 pi r)
      const double i4r3 = i4r * rinv * rinv;        // 1/(4 pi r^3)
      const double kr = k * r;
      double hr, hi, hpr, hpi;
      if (CPLX) {
        // h = e^{ikr} (ikr - 1)/(4 pi r^3), ikr - 1 = (-Im k r - 1) + i Re k r
        const double ar = fma(-ki, r, -1.0), ai = kr;
        hr = (cs * ar - sn * ai) * i4r3;
        hi = (cs * ai + sn * ar) * i4r3;
        // h' = -k^2 e^{ikr}/(4 pi r^2) - 3 h / r, k^2 = (Re k^2 - Im k^2) + 2 i Re k Im k
        const double k2r = k * k - ki * ki, k2i = 2.0 * k * ki, s2 = i4r * rinv;
        hpr = -(k2r * cs - k2i * sn) * s2 - 3.0 * hr * rinv;
        hpi = -(k2r * sn + k2i * cs) * s2 - 3.0 * hi * rinv;
      } else {
        hr = (-cs - kr * sn) * i4r3;
        hi = (kr * cs - sn) * i4r3;
        const double k2i4r2 = k * k * i4r * rinv;   // k^2/(4 pi r^2)
        hpr = -k2i4r2 * cs - 3.0 * hr * rinv;
        hpi = -k2i4r2 * sn - 3.0 * hi * rinv;
      }
      // Re(H) = -Re(A1 h) + A2r/(4 pi r^3);  Re(H_r) = -Re(A1 h') - 3 A2r/(4 pi r^4)
      const double ReH = -(A1r * hr - A1i * hi) + A2r * i4r3;
      const double ReHr = -(A1r * hpr - A1i * hpi) - 3.0 * A2r * i4r3 * rinv;
      if (CPLX) {
        // D += cd A1 k e^{ikr} / (4 pi r): d/dRe k = Re D, d/dIm k = -Im D
        const double er = k * cs - ki * sn, ei = k * sn + ki * cs;   // k e^{ikr}
        gk += cd * i4r * (A1r * er - A1i * ei);
        gki -= cd * i4r * (A1r * ei + A1i * er);
      } else {
        // dPsi/dk = cd * Re(A1 k e^{ikr}) / (4 pi r)
        gk += cd * k * i4r * (A1r * cs - A1i * sn);
      }
      const double t = ReHr * cd * rinv;
      const double gd[3] = {fma(t, d[0], ReH * cv[0]), fma(t, d[1], ReH * cv[1]),
                            fma(t, d[2], ReH * cv[2])};
      const double gc[3] = {ReH * d[0], ReH * d[1], ReH * d[2]};
      gx0 -= gd[0];
      gx1 -= gd[1];
      gx2 -= gd[2];
      // d(g.(t_u x t_v))/dt_u = t_v x g,  /dt_v = g x t_u
      const double gtu[3] = {tv[1] * gc[2] - tv[2] * gc[1], tv[2] * gc[0] - tv[0] * gc[2],
                             tv[0] * gc[1] - tv[1] * gc[0]};
      const double gtv[3] = {gc[1] * tu[2] - gc[2] * tu[1], gc[2] * tu[0] - gc[0] * tu[2],
                             gc[0] * tu[1] - gc[1] * tu[0]};
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        const double N = P[kPtN + g], du = P[kPtDu + g], dv = P[kPtDv + g];
#pragma unroll
        for (int dd = 0; dd < 3; ++dd)
          gX[g][dd] = fma(N, gd[dd], fma(du, gtu[dd], fma(dv, gtv[dd], gX[g][dd])));
      }
    }
    // the wave shares element e: sum its node gradients over the lanes, one atomic each
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      const int node = a.topo[(size_t)e * NG + g];
#pragma unroll
      for (int dd = 0; dd < 3; ++dd) {
        const double v = wave_sum(gX[g][dd]);
        if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(a.gradX + 3 * (size_t)node + dd, v);
      }
    }
  }
  if (active && a.tvert) {
    const int v = a.tvert[i];
    if (v >= 0) {
      unsafeAtomicAdd(a.gradX + 3 * (size_t)v, gx0);
      unsafeAtomicAdd(a.gradX + 3 * (size_t)v + 1, gx1);
      unsafeAtomicAdd(a.gradX + 3 * (size_t)v + 2, gx2);
    }
  }
  const double gks = wave_sum(gk);
  if ((threadIdx.x & 63) == 0 && a.gradk) unsafeAtomicAdd(a.gradk, gks);
  if (CPLX) {
    const double gkis = wave_sum(gki);
    if ((threadIdx.x & 63) == 0 && a.gradk) unsafeAtomicAdd(a.gradk + 1, gkis);
  }
}

// Derivative in k of the assembly, contracted: w = alpha^T (dF/dk) p for one instance of an
// uneven batch (blockIdx.x = tile: instance, 256-row block, element chunk).  With the
// element tables of the assembly (y, the reference's normal n, J w / (4 pi) at every rule
// point), d/dk of the double-layer integrand e^{ikr}(ikr - 1)(n.d) J w / (4 pi r^3)
// (kernels.jl:29-37; its k-derivative is taylor_freens3d! at order 1, kernels.jl:104-118)
// is -k e^{ikr} (n.d) J w / (4 pi r), holomorphic in k, so complex wavenumbers (the lossy
// callers) are covered.  Every (target, element) pair is included, mid-edge rules too, with
// the assembly's rule selection.  One thread per target row; the wave's partial sums are
// added into w[instance] with one atomic per wave.
template <int NP>
__global__ void __launch_bounds__(256)
k_dfdk_batch(const DfdkInst* __restrict__ inst, const int4* __restrict__ tiles, int ech) {
  const int4 td = tiles[blockIdx.x];
  const DfdkInst& a = inst[td.x];
  const int64_t i = (int64_t)td.y * 256 + threadIdx.x;
  const bool active = i < a.M;
  const int64_t ic = active ? i : (a.M - 1);
  const double x0 = a.targets[3 * ic], x1 = a.targets[3 * ic + 1], x2 = a.targets[3 * ic + 2];
  const double kr = a.kr, ki = a.ki;
  double sr = 0.0, si = 0.0;  // sum over pairs and points of (n.d) Jw/(4 pi r) e^{ikr} p_h
  const int e0 = td.z * ech, e1 = min(a.E, e0 + ech);
  for (int e = e0; e < e1; ++e) {
    const int c = select_rule<NP>(x0, x1, x2, a.selnodes + (size_t)e * NP * 4, a.R, a.hyp) - 1;
    double pr[NP], pim[NP];
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      const double2 v = a.p[a.ptopo[(size_t)e * NP + s]];
      pr[s] = v.x;
      pim[s] = v.y;
    }
    const double* tab = a.elem + ((size_t)e * a.Pe + a.off[c]) * 8;
    const double* ph = a.phi + (size_t)c * a.Qmax * NP;
    const int Q = a.len[c];
    for (int q = 0; q < Q; ++q) {
      const double4 A = *reinterpret_cast<const double4*>(tab + 8 * q);
      const double4 B = *reinterpret_cast<const double4*>(tab + 8 * q + 4);
      const double dx = A.x - x0, dy = A.y - x1, dz = A.z - x2;
      const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
      const double rinv = rsqrt_nr(r2);
      const double r = r2 * rinv;
      const double nd = fma(A.w, dx, fma(B.x, dy, B.y * dz));
      double sn, cs;
      sincos_f64(kr * r, &sn, &cs);
      double amp = nd * B.z * rinv;  // (n.d) J w / (4 pi r)
      if (ki != 0.0) amp *= exp(-ki * r);
      double phr = 0.0, phi_ = 0.0;
#pragma unroll
      for (int s = 0; s < NP; ++s) {
        const double b = ph[q * NP + s];
        phr = fma(b, pr[s], phr);
        phi_ = fma(b, pim[s], phi_);
      }
      // amp e^{i kr r} (phr + i phi)
      sr = fma(amp, cs * phr - sn * phi_, sr);
      si = fma(amp, cs * phi_ + sn * phr, si);
    }
  }
  // w += alpha (-k) S
  const double2 al = active ? a.alpha[ic] : make_double2(0.0, 0.0);
  const double tr = -(kr * sr - ki * si), ti = -(kr * si + ki * sr);
  const double wr = wave_sum(al.x * tr - al.y * ti), wi = wave_sum(al.x * ti + al.y * tr);
  if ((threadIdx.x & 63) == 0) {
    unsafeAtomicAdd(a.w, wr);
    unsafeAtomicAdd(a.w + 1, wi);
  }
}

}  // namespace

void launch_dfdk_batch(const DfdkInst* d_inst, const int4* d_tiles, int n_tiles, int ech, int NP,
                       hipStream_t s) {
  if (n_tiles <= 0) return;
  if (NP == 1)
    hipLaunchKernelGGL(k_dfdk_batch<1>, dim3((unsigned)n_tiles), dim3(256), 0, s, d_inst, d_tiles,
                       ech);
  else if (NP == 3)
    hipLaunchKernelGGL(k_dfdk_batch<3>, dim3((unsigned)n_tiles), dim3(256), 0, s, d_inst, d_tiles,
                       ech);
  else
    hipLaunchKernelGGL(k_dfdk_batch<6>, dim3((unsigned)n_tiles), dim3(256), 0, s, d_inst, d_tiles,
                       ech);
  BIE_HIP_CHECK(hipGetLastError());
}

void launch_adjoint(const AdjointArgs& a, int gk, int pk, bool cplx, hipStream_t s) {
  if (a.M == 0 || a.E == 0) return;
  const int64_t rowblocks = (a.M + 255) / 256;
  BIE_REQUIRE(rowblocks < 2147483647LL, BIE_EINVAL, "too many targets");
  int64_t chunks = (4096 + rowblocks - 1) / rowblocks;
  if (chunks > a.E) chunks = a.E;
  if (chunks > 65535) chunks = 65535;
  if (chunks < 1) chunks = 1;
  const int ech = (int)((a.E + chunks - 1) / chunks);
  chunks = (a.E + ech - 1) / ech;
  const dim3 grid((unsigned)rowblocks, (unsigned)chunks);
  auto go = [&](auto kern) { hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, a, ech); };
  if (gk == 3 && pk == 3)
    cplx ? go(k_adjoint<3, 3, true>) : go(k_adjoint<3, 3, false>);
  else if (gk == 6 && pk == 6)
    cplx ? go(k_adjoint<6, 6, true>) : go(k_adjoint<6, 6, false>);
  else
    cplx ? go(k_adjoint<6, 3, true>) : go(k_adjoint<6, 3, false>);
  BIE_HIP_CHECK(hipGetLastError());
}

}  // namespace bie
