// Device helpers shared by the assembly kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace bie {
namespace {

// Workgroup barrier for LDS hand-offs only: waits for this wave's LDS (and scalar-load)
// operations, then s_barrier.  Unlike __syncthreads() it does not wait for outstanding
// global loads and stores, so prefetched table rows and the column stores of F and G stay
// in flight across it; nothing in global memory is exchanged between the waves of a
// workgroup through these barriers.  The memory clobber keeps the compiler from moving
// memory accesses across it.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Wave-uniform read of read-only data (the host-built slot schedules) through the scalar
// data cache: the load goes to an SGPR and is counted by lgkmcnt.  A vector load of the
// same value would be counted by vmcnt, which on gfx9-family GPUs retires loads and stores
// in issue order, so waiting for it would also wait for every column store issued before
// it.  Only for addresses that are the same in all active lanes and memory the kernel
// never writes.
template <class T>
__device__ __forceinline__ T sload(const T* p) {
  return *reinterpret_cast<const __attribute__((address_space(4))) T*>(
      reinterpret_cast<uintptr_t>(p));
}

// Two doubles as one 16-byte vector: a complex entry's (re, im) written by one store.
typedef double d2v __attribute__((ext_vector_type(2)));

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

// The n-ary hypot that later Julia releases use for IEEE floats: the arguments scaled by
// 2^-E with E the exponent of maxabs (clamped to the normal range), squared and summed left
// to right, square-rooted and scaled back by 2^E.  The scalings are exact, so away from overflow and underflow this is
// sqrt((x^2 + y^2) + z^2) with every operation rounded once; no fused multiply-add, so the
// result is bit-identical to the CPU.
__device__ __noinline__ double scaled_hypot3(double x, double y, double z) {
#pragma clang fp contract(off)
  const double ax = fabs(x), ay = fabs(y), az = fabs(z);
  if (isinf(ax) || isinf(ay) || isinf(az)) return INFINITY;
  if (isnan(ax) || isnan(ay) || isnan(az)) return NAN;
  double mx = ax > ay ? ax : ay;
  mx = mx > az ? mx : az;
  if (mx == 0.0) return mx;
  int e = 0;
  (void)frexp(mx, &e);  // 2^E <= mx < 2^(E+1) with E = e - 1, clamped to the normal range
  const int E = max(-1022, min(1022, e - 1));
  const double inv = ldexp(1.0, -E), scale = ldexp(1.0, E);
  const double a = ax * inv, b = ay * inv, c = az * inv;
  double s = a * a;
  s = s + b * b;
  s = s + c * c;
  return scale * sqrt(s);
}

// 1-based rule index of the node nearest to x (first strict minimum of the Julia hypot
// distance; hyp 0: julia_hypot3, 1: scaled_hypot3).  A squared-distance filter decides
// every pair whose two nearest nodes are separated by more than 1e-12 relative (far beyond
// the few-ulp error of either distance, so the exact argmin is the same under either
// formula); near ties, zero, inf and NaN distances go through the exact sequence.  Index 0
// (no distance below Inf) maps to the last rule, as the reference's if/elseif chain does.
template <int NP>
__device__ __forceinline__ int select_rule(double x0, double x1, double x2,
                                           const double* __restrict__ P, int R, int hyp) {
  // quadrilaterals (node counts other than 3 and 6): one rule, nothing to select
  if (NP != 3 && NP != 6 && R == 1) return 1;
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
      const double dx = P[4 * s] - x0, dy = P[4 * s + 1] - x1, dz = P[4 * s + 2] - x2;
      const double d = hyp ? scaled_hypot3(dx, dy, dz) : julia_hypot3(dx, dy, dz);
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

__device__ __forceinline__ void load_rows01(const double* __restrict__ tab, int Q, double4& A,
                                            double4& B, double4& A2, double4& B2) {
  A = *reinterpret_cast<const double4*>(tab);
  B = *reinterpret_cast<const double4*>(tab + 4);
  const int q1 = Q > 1 ? 1 : 0;
  A2 = *reinterpret_cast<const double4*>(tab + 8 * q1);
  B2 = *reinterpret_cast<const double4*>(tab + 8 * q1 + 4);
}

// Per-wave staging of an element's three vertex-rule tables (3 Qc rows of 64 bytes,
// nchunk = 12 Qc chunks of 16 bytes, at most 3 per lane) and of its nsel chunks of
// selection-node data (one per lane, nsel <= 12): global -> registers early, registers ->
// the wave's LDS rows later, so no global load is pending in the point loop or in the
// rule selection.  The loads are buffer loads whose record counts are nchunk and nsel
// chunks, so lanes past the end read zeros from the bounds check (nsel = 0 reads nothing):
// no exec mask and no register zeroing around them, which would otherwise make the
// compiler wait for every older vector-memory operation (the previous step's column
// stores included) before issuing them.
__device__ __forceinline__ void stage_load(const double* __restrict__ elem_e, int nchunk,
                                           const double* __restrict__ sel, int nsel, int lane,
                                           double2 (&v)[3], double2& vs) {
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<double*>(elem_e), (short)0, nchunk * 16, 0x00020000);
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const u4 w = __builtin_amdgcn_raw_buffer_load_b128(rs, (lane + 64 * t) * 16, 0, 0);
    v[t] = __builtin_bit_cast(double2, w);
  }
  const __amdgpu_buffer_rsrc_t rsel = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<double*>(sel), (short)0, nsel * 16, 0x00020000);
  vs = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rsel, lane * 16, 0, 0));
}

__device__ __forceinline__ void stage_store(double* rows, int nchunk, double* sel_rows,
                                            int nsel, int lane, const double2 (&v)[3],
                                            const double2& vs) {
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const int j = lane + 64 * t;
    if (j < nchunk) reinterpret_cast<double2*>(rows)[j] = v[t];
  }
  if (lane < nsel) reinterpret_cast<double2*>(sel_rows)[lane] = vs;
}

// Split-kernel staging: a wave takes the rows of points half*QH .. half*QH + QH - 1 of
// each of the element's three vertex rules (Qc points each, 64 bytes per row): 12 QH chunks
// of 16 bytes, chunk t from rule t / (4 QH), row chunk t % (4 QH), at the element's
// table chunk rule * 4 Qc + half * 4 QH + t % (4 QH).  Bounds-checked buffer loads as in
// stage_load (lanes past 12 QH read zeros), plus the nsel selection-node chunks.
template <int QC, int QH, int NV>
__device__ __forceinline__ void stage_load_half(const double* __restrict__ elem_e, int half,
                                                const double* __restrict__ sel, int nsel,
                                                int lane, double2 (&v)[NV], double2& vs) {
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<double*>(elem_e), (short)0, 12 * QC * 16, 0x00020000);
#pragma unroll
  for (int u = 0; u < NV; ++u) {
    const int t = lane + 64 * u;
    const int c = t / (4 * QH);
    const int src = c * 4 * QC + half * 4 * QH + (t - c * 4 * QH);
    const int off = t < 12 * QH ? src * 16 : 12 * QC * 16;
    v[u] = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
  }
  const __amdgpu_buffer_rsrc_t rsel = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<double*>(sel), (short)0, nsel * 16, 0x00020000);
  vs = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rsel, lane * 16, 0, 0));
}

template <int QH, int NV>
__device__ __forceinline__ void stage_store_half(double* rows, double* sel_rows, int nsel,
                                                 int lane, const double2 (&v)[NV],
                                                 const double2& vs) {
#pragma unroll
  for (int u = 0; u < NV; ++u) {
    const int t = lane + 64 * u;
    if (t < 12 * QH) reinterpret_cast<double2*>(rows)[t] = v[u];
  }
  if (lane < nsel) reinterpret_cast<double2*>(sel_rows)[lane] = vs;
}

}  // namespace
}  // namespace bie
