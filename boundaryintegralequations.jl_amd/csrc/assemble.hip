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

#include "bie_common.hpp"

namespace bie {
namespace {

constexpr double kInv4Pi = 0.079577471545947667884;  // 1 / (4 pi)

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
    const double N = P[3 + k], du = P[9 + k], dv = P[15 + k];
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

// Julia Base n-ary hypot for Float64 (maxabs * sqrt(sum((x_i/maxabs)^2)), left to
// right), with no fused multiply-add so that the result is bit-identical to the CPU.
__device__ __noinline__ double julia_hypot3(double x, double y, double z) {
#pragma clang fp contract(off)
  const double ax = fabs(x), ay = fabs(y), az = fabs(z);
  if (isinf(ax) || isinf(ay) || isinf(az)) return INFINITY;
  if (isnan(ax) || isnan(ay) || isnan(az)) return NAN;
  double mx = ax > ay ? ax : ay;
  mx = mx > az ? mx : az;
  if (mx == 0.0) return mx;
  const double a = ax / mx, b = ay / mx, c = az / mx;
  double s = a * a;
  s = s + b * b;
  s = s + c * c;
  return mx * sqrt(s);
}

// 1-based rule index of the node nearest to x (first strict minimum of the Julia hypot
// distance).  A squared-distance filter decides every pair whose two nearest nodes are
// separated by more than 1e-12 relative (far beyond the few-ulp error of either
// distance, so the exact argmin is the same); near ties, zero, inf and NaN distances go
// through the exact sequence.  Index 0 (no distance below Inf) maps to the last rule,
// as the reference's if/elseif chain does.
template <int NP>
__device__ __forceinline__ int select_rule(double x0, double x1, double x2,
                                           const double* __restrict__ P, int R) {
  double d2[NP];
#pragma unroll
  for (int s = 0; s < NP; ++s) {
    const double dx = P[4 * s] - x0, dy = P[4 * s + 1] - x1, dz = P[4 * s + 2] - x2;
    d2[s] = fma(dx, dx, fma(dy, dy, dz * dz));
  }
  int jb = 0;
  double b = d2[0];
#pragma unroll
  for (int s = 1; s < NP; ++s)
    if (d2[s] < b) {
      b = d2[s];
      jb = s;
    }
  const double lim = b * (1.0 + 1e-12);
  bool ok = (b <= 1.0e300);
#pragma unroll
  for (int s = 0; s < NP; ++s)
    if (s != jb && !(d2[s] > lim)) ok = false;
  int c;
  if (ok) {
    c = jb + 1;
  } else {
    double best = INFINITY;
    int idx = 0;
    for (int s = 0; s < NP; ++s) {
      const double d = julia_hypot3(P[4 * s] - x0, P[4 * s + 1] - x1, P[4 * s + 2] - x2);
      if (d < best) {
        best = d;
        idx = s + 1;
      }
    }
    c = idx;
  }
  if (c < 1 || c > R) c = R;
  return c;
}

// 1/sqrt(x) from the hardware estimate refined by two Newton steps.
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * fma(-h, y * y, 1.5);
  y = y * fma(-h, y * y, 1.5);
  return y;
}

template <int NP>
__global__ void __launch_bounds__(256)
k_select(const double* __restrict__ targets, int64_t M, const double* __restrict__ sel,
         int E, int R, int8_t* __restrict__ out) {
  const int64_t total = M * (int64_t)E;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += stride) {
    const int64_t i = t / E;
    const int e = (int)(t - i * E);
    out[t] = (int8_t)select_rule<NP>(targets[3 * i], targets[3 * i + 1], targets[3 * i + 2],
                                      sel + (size_t)e * NP * 4, R);
  }
}

// One thread per target row; blockIdx.y selects a contiguous chunk of elements.
// Every (target, element) contribution is added into zeroed F, G with hardware FP64
// atomics; C is accumulated per thread and added once per chunk.
template <int NP, bool CPLX>
__global__ void __launch_bounds__(256)
k_assemble(const AssemblyArgs a, int ech, int phi_in_lds) {
  extern __shared__ double s_phi[];
  const int nphi = a.R * a.Qmax * NP;
  if (phi_in_lds)
    for (int t = threadIdx.x; t < nphi; t += blockDim.x) s_phi[t] = a.phi[t];
  __syncthreads();
  const double* phi = phi_in_lds ? s_phi : a.phi;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.M) return;
  const double x0 = a.targets[3 * i], x1 = a.targets[3 * i + 1], x2 = a.targets[3 * i + 2];
  const int e0 = blockIdx.y * ech;
  const int e1 = min(a.E, e0 + ech);
  const bool doF = (a.flags & BIE_F) != 0, doG = (a.flags & BIE_G) != 0;
  const double kr = a.kr, ki = a.ki;
  double csum = 0.0;
  for (int e = e0; e < e1; ++e) {
    const int c = select_rule<NP>(x0, x1, x2, a.selnodes + (size_t)e * NP * 4, a.R) - 1;
    const double* tab = a.elem + ((size_t)e * a.Pe + a.off[c]) * 8;
    const double* ph = phi + c * a.Qmax * NP;
    const int Q = a.len[c];
    double fr[NP], fi[NP], gr[NP], gi[NP];
#pragma unroll
    for (int s = 0; s < NP; ++s) fr[s] = fi[s] = gr[s] = gi[s] = 0.0;
    for (int q = 0; q < Q; ++q) {
      const double4 A = *reinterpret_cast<const double4*>(tab + 8 * q);
      const double4 B = *reinterpret_cast<const double4*>(tab + 8 * q + 4);
      const double dx = A.x - x0, dy = A.y - x1, dz = A.z - x2;
      const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
      const double rinv = rsqrt_nr(r2);
      const double r = r2 * rinv;
      const double nd = fma(A.w, dx, fma(B.x, dy, B.y * dz));
      const double g = B.z * rinv;             // J w / (4 pi r)
      const double ak = nd * rinv * rinv * g;  // (n.d) J w / (4 pi r^3)
      csum -= ak;
      double sn, cs;
      sincos(kr * r, &sn, &cs);
      double Gr, Gi, Fr, Fi;
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
      Gr = g * cs;
      Gi = g * sn;
#pragma unroll
      for (int s = 0; s < NP; ++s) {
        const double p = ph[q * NP + s];
        fr[s] = fma(Fr, p, fr[s]);
        fi[s] = fma(Fi, p, fi[s]);
        gr[s] = fma(Gr, p, gr[s]);
        gi[s] = fma(Gi, p, gi[s]);
      }
    }
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

}  // namespace

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
                   int8_t* out, hipStream_t s) {
  if (M == 0 || E == 0) return;
  int64_t blocks = (M * (int64_t)E + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  dim3 grid((unsigned)blocks);
  if (NP == 3)
    hipLaunchKernelGGL(k_select<3>, grid, dim3(256), 0, s, targets, M, sel, E, R, out);
  else
    hipLaunchKernelGGL(k_select<6>, grid, dim3(256), 0, s, targets, M, sel, E, R, out);
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
  size_t shmem = sizeof(double) * (size_t)a.R * a.Qmax * a.NP;
  int phi_in_lds = 1;
  if (shmem > 48 * 1024) {  // very high orders: read the basis table from global memory
    shmem = 0;
    phi_in_lds = 0;
  }
  const bool cplx = a.ki != 0.0;
  if (a.NP == 3) {
    if (cplx)
      hipLaunchKernelGGL((k_assemble<3, true>), grid, dim3(256), shmem, s, a, ech, phi_in_lds);
    else
      hipLaunchKernelGGL((k_assemble<3, false>), grid, dim3(256), shmem, s, a, ech, phi_in_lds);
  } else {
    if (cplx)
      hipLaunchKernelGGL((k_assemble<6, true>), grid, dim3(256), shmem, s, a, ech, phi_in_lds);
    else
      hipLaunchKernelGGL((k_assemble<6, false>), grid, dim3(256), shmem, s, a, ech, phi_in_lds);
  }
  BIE_HIP_CHECK(hipGetLastError());
}

}  // namespace bie
