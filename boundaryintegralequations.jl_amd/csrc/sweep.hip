// Batched frequency sweep: dense F_k, G_k for a list of real wavenumbers k_0..k_{nk-1}
// on one mesh, in one launch per kernel (graph-capturable: no host synchronisation, no
// memset nodes).
//
// Each wave carries KB wavenumbers at once.  Everything that does not depend on k is
// shared between them: rule selection (find_closest_corner,
// src/3d/assembly_collocation.jl:102-115), the table loads, d = y - x, 1/r, n.d, the
// C integrand (C is k-independent, kernels.jl:48-55, and is formed once by the first
// k block).  For a block of uniformly spaced k, e^{i k_b r} = e^{i k_0 r} (e^{i dk r})^b
// replaces KB sincos evaluations by two plus KB-1 complex products; other blocks evaluate
// each e^{i k_b r} directly, so any list of k is exact.  Accumulation uses the LDS-slot
// schedule of the single-k kernel with KB accumulators per slot (the schedule is built
// for as many slots of KB accumulators as fit in LDS).  The reference has no batched path: it calls
// assemble_parallel! once per k (e.g. the Taylor sweep of src/3d/rosebem.jl:56-157
// shares geometry across derivative orders in the same way).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bie_common.hpp"
#include "device_common.hpp"
#include "kernels_math.hpp"
#include "schedule.hpp"

namespace bie {
namespace {

// With TAB, e^{ikr} comes from the LDS trig table trig (sincos_tab), else from sincos_f64.
template <int NP, int KB, bool TAB>
__device__ __forceinline__ void sweep_point(double x0, double x1, double x2, const double4 A,
                                            const double4 B, const double* __restrict__ p,
                                            const double* kv, double dk, bool rec, bool doG,
                                            double (&fr)[KB][NP], double (&fi)[KB][NP],
                                            double (&gr)[KB][NP], double (&gi)[KB][NP],
                                            double& csum, const double2* trig) {
  const double dx = A.x - x0, dy = A.y - x1, dz = A.z - x2;
  const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
  const double rinv = rsqrt_nr(r2);
  const double r = r2 * rinv;
  const double nd = fma(A.w, dx, fma(B.x, dy, B.y * dz));
  const double g = B.z * rinv;          // J w / (4 pi r)
  const double t = nd * rinv * rinv;     // (n.d) / r^2: f = t (ikr - 1) g e^{ikr}
  csum -= t * g;
  // the recurrence runs on g e^{ik_b r} (= G_b): one complex product per further k
  double sn, cs, zs = 0.0, zc = 1.0;
  if (TAB) {
    sincos_tab(kv[0] * r, trig, &sn, &cs);
    if (rec) sincos_tab(dk * r, trig, &zs, &zc);
  } else {
    sincos_f64(kv[0] * r, &sn, &cs);
    if (rec) sincos_f64(dk * r, &zs, &zc);
  }
  double Gr = g * cs, Gi = g * sn;
  double ps[NP];
#pragma unroll
  for (int s = 0; s < NP; ++s) ps[s] = p[s];
#pragma unroll
  for (int b = 0; b < KB; ++b) {
    if (b > 0) {
      if (rec) {
        const double c2 = fma(Gr, zc, -Gi * zs);
        Gi = fma(Gr, zs, Gi * zc);
        Gr = c2;
      } else {
        if (TAB)
          sincos_tab(kv[b] * r, trig, &sn, &cs);
        else
          sincos_f64(kv[b] * r, &sn, &cs);
        Gr = g * cs;
        Gi = g * sn;
      }
    }
    const double krr = kv[b] * r;
    const double Fr = -t * fma(krr, Gi, Gr);
    const double Fi = t * fma(krr, Gr, -Gi);
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      fr[b][s] = fma(Fr, ps[s], fr[b][s]);
      fi[b][s] = fma(Fi, ps[s], fi[b][s]);
    }
    if (doG) {  // the same for every lane: F-only sweeps (gOn = false) skip this work
#pragma unroll
      for (int s = 0; s < NP; ++s) {
        gr[b][s] = fma(Gr, ps[s], gr[b][s]);
        gi[b][s] = fma(Gi, ps[s], gi[b][s]);
      }
    }
  }
}

// k-derivatives of orders m0 .. m0+KB-1 at one complex k = kr + i ki (taylor_greens3d!,
// taylor_freens3d!, src/3d/kernels.jl:94-118, as taylor_assemble! integrates them,
// src/3d/rosebem.jl:7-44):
//   G_m = (i r)^m e^{ikr} J w / (4 pi r),   F_m = (i k r - 1 + m) (n.d / r^2) G_m,
// so one e^{ikr} serves every order and each further order is one product by i r.
template <int NP, int KB, bool TAB>
__device__ __forceinline__ void taylor_point(double x0, double x1, double x2, const double4 A,
                                             const double4 B, const double* __restrict__ p,
                                             double kr, double ki, int m0, bool doG,
                                             double (&fr)[KB][NP], double (&fi)[KB][NP],
                                             double (&gr)[KB][NP], double (&gi)[KB][NP],
                                             double& csum, const double2* trig) {
  const double dx = A.x - x0, dy = A.y - x1, dz = A.z - x2;
  const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
  const double rinv = rsqrt_nr(r2);
  const double r = r2 * rinv;
  const double nd = fma(A.w, dx, fma(B.x, dy, B.y * dz));
  const double g = B.z * rinv;       // J w / (4 pi r)
  const double t = nd * rinv * rinv;  // (n.d) / r^2
  csum -= t * g;
  double sn, cs;
  if (TAB)
    sincos_tab(kr * r, trig, &sn, &cs);
  else
    sincos_f64(kr * r, &sn, &cs);
  double gm = g;
  if (ki != 0.0) gm *= exp(-ki * r);
  // r^m0 by binary powering (m0 is the same for the whole wave, m0 < 64)
  double rb = r;
  for (int e = m0; e != 0; e >>= 1) {
    if (e & 1) gm *= rb;
    rb *= rb;
  }
  double Gr = gm * cs, Gi = gm * sn;
  if (m0 & 1) {  // times i^m0
    const double t2 = -Gi;
    Gi = Gr;
    Gr = t2;
  }
  if (m0 & 2) {
    Gr = -Gr;
    Gi = -Gi;
  }
  const double krr = kr * r;
  const double a0 = (double)(m0 - 1) - ki * r;  // Re(i k r - 1 + m0)
  double ps[NP];
#pragma unroll
  for (int s = 0; s < NP; ++s) ps[s] = p[s];
#pragma unroll
  for (int b = 0; b < KB; ++b) {
    if (b > 0) {  // G_{m+1} = i r G_m
      const double t2 = -r * Gi;
      Gi = r * Gr;
      Gr = t2;
    }
    const double am = a0 + (double)b;
    const double Fr = t * fma(am, Gr, -krr * Gi);
    const double Fi = t * fma(am, Gi, krr * Gr);
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      fr[b][s] = fma(Fr, ps[s], fr[b][s]);
      fi[b][s] = fma(Fi, ps[s], fi[b][s]);
    }
    if (doG) {  // the same for every lane: F-only stacks (gOn = false) skip this work
#pragma unroll
      for (int s = 0; s < NP; ++s) {
        gr[b][s] = fma(Gr, ps[s], gr[b][s]);
        gi[b][s] = fma(Gi, ps[s], gi[b][s]);
      }
    }
  }
}

// Flush of one step's finished nodes for the KBS slot groups of a block (see the
// single-k kernel): slot contents to F_k, G_k columns, zeroing the slots.  With PACK a
// group holds F of two consecutive matrices (k0 + 2 q2 at +0/+64, k0 + 2 q2 + 1 at
// +128/+192) and no G.
template <int KBS, int W, bool PACK>
__device__ __forceinline__ void sweep_flush(const AssemblyArgs& a, const SlotArgs& sa,
                                            const SweepArgs& sw, int st, int wave, int lane,
                                            int64_t i, bool active, int k0, double* acc) {
  const bool doF = (a.flags & BIE_F) != 0, doG = (a.flags & BIE_G) != 0;
  const int f1 = sload(sa.flush_begin + st + 1);
  for (int f = sload(sa.flush_begin + st) + wave; f < f1; f += W) {
    const int node = sload(sa.flush_node + f);
    const int ss = sload(sa.flush_slot + f);
#pragma unroll
    for (int q2 = 0; q2 < KBS; ++q2) {
      double* p = acc + ((size_t)(ss & 0xffff) * KBS + q2) * 256 + lane;
      const double vr = p[0], vi = p[64], wr = p[128], wi = p[192];
      p[0] = 0.0;
      p[64] = 0.0;
      p[128] = 0.0;
      p[192] = 0.0;
      if constexpr (PACK) {
        const int ka = k0 + 2 * q2;
        if (active && doF) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            if (ka + h >= sw.nk) break;
            const double xr = h ? wr : vr, xi = h ? wi : vi;
            const size_t idx =
                2 * ((size_t)(ka + h) * sw.stride + (size_t)i + (size_t)node * a.M);
            if (ss & kSharedBit) {
              unsafeAtomicAdd(a.F + idx, xr);
              unsafeAtomicAdd(a.F + idx + 1, xi);
            } else {
              __builtin_nontemporal_store(d2v{xr, xi}, reinterpret_cast<d2v*>(a.F + idx));
            }
          }
        }
        continue;
      }
      const int kidx = k0 + q2;
      if (active && kidx < sw.nk) {
        const size_t idx = 2 * ((size_t)kidx * sw.stride + (size_t)i + (size_t)node * a.M);
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
          if (doF) __builtin_nontemporal_store(d2v{vr, vi}, reinterpret_cast<d2v*>(a.F + idx));
          if (doG) __builtin_nontemporal_store(d2v{wr, wi}, reinterpret_cast<d2v*>(a.G + idx));
        }
      }
    }
  }
}

// With TAY the KB accumulator sets of a block hold the k-derivative orders
// kb*KB .. kb*KB+KB-1 at the one wavenumber sw.k_re + i sw.k_im instead of KB wavenumbers.
// GON is whether G is formed (the BIE_G flag): F-only launches skip the G accumulation and
// its LDS atomics in a separate instance, leaving the F-and-G code as it is.
// Without G (PACK) the KB wavenumbers or orders of a block fill KB / 2 slot groups, two
// matrices' F per group (see sweep_flush), so a block carries twice the matrices that the
// slot schedule built for KB / 2 groups would otherwise allow.
template <int NP, int KB, int W, bool TAY, bool GON>
__global__ void __launch_bounds__(64 * W)
k_sweep_slots(const AssemblyArgs a, const SlotArgs sa, const SweepArgs sw) {
  constexpr bool PACK = !GON;
  constexpr int KBS = PACK ? KB / 2 : KB;  // slot groups per slot
  static_assert(!PACK || KB % 2 == 0, "packed blocks hold an even number of orders");
  extern __shared__ double lds[];
  double* acc = lds;  // [slot][group][4][64]
  const int nphi = a.R * a.Qmax * NP;
  double* s_phi = lds + (size_t)sa.S * KBS * 256;
  const double2* s_trig = reinterpret_cast<const double2*>(s_phi + phi_pad(nphi));
  int2* s_q = reinterpret_cast<int2*>(s_phi + phi_pad(nphi) + kTrigDoubles);
  unsigned* s_qn = reinterpret_cast<unsigned*>(s_q + kMidLds);
  const int b = blockIdx.x;
  const int64_t per_xcd = (sw.n_tiles + 7) / 8;
  const int64_t tile = (int64_t)(b & 7) * per_xcd + (b >> 3);
  if ((b >> 3) >= per_xcd || tile >= sw.n_tiles) return;
  // k block outermost, then segment, then row block
  const int64_t per_kb = (int64_t)sa.n_segments * sa.n_rowblocks;
  const int kb = (int)(tile / per_kb);
  const int64_t rest = tile - (int64_t)kb * per_kb;
  const int seg = (int)(rest / sa.n_rowblocks);
  const int rb = (int)(rest - (int64_t)seg * sa.n_rowblocks);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t i = (int64_t)rb * 64 + lane;
  const bool active = i < a.M;
  // basis table and trig table: one buffer of phi_pad(nphi) + kTrigDoubles doubles
  for (int t = threadIdx.x; t < phi_pad(nphi) + kTrigDoubles; t += blockDim.x) s_phi[t] = a.phi[t];
  for (int t = threadIdx.x; t < sa.S * KBS * 256; t += blockDim.x) acc[t] = 0.0;
  if (threadIdx.x == 0) *s_qn = 0u;
  lds_barrier();
  const int64_t ic = active ? i : (a.M - 1);
  const double x0 = a.targets[3 * ic], x1 = a.targets[3 * ic + 1], x2 = a.targets[3 * ic + 2];
  const bool first_kb = kb == 0;
  constexpr bool doG = GON;
  double kv[KB];
  const int k0 = kb * KB;
  bool rec = false;
  double dk = 0.0;
  if constexpr (!TAY) {
#pragma unroll
    for (int q = 0; q < KB; ++q) kv[q] = sw.k[min(k0 + q, sw.nk - 1)];
    rec = sw.block_uniform[kb] != 0;
    dk = KB > 1 ? kv[1] - kv[0] : 0.0;
  }
  double csum = 0.0;
  const int st0 = sload(sa.seg_step_begin + seg), st1 = sload(sa.seg_step_begin + seg + 1);
  int e = (st0 < st1) ? sload(sa.step_elems + (size_t)st0 * W + wave) : -1;
  int c = 0;
  double4 A0 = make_double4(0, 0, 0, 0), B0 = A0;
  if (e >= 0) {
    c = select_rule<NP>(x0, x1, x2, sa.step_sel + ((size_t)st0 * W + wave) * NP * 4, a.R,
                        a.hyp) - 1;
    if (!(NP == 6 && c >= 3)) {
      const double* t0 = a.elem + ((size_t)e * a.Pe + a.off[c]) * 8;
      A0 = *reinterpret_cast<const double4*>(t0);
      B0 = *reinterpret_cast<const double4*>(t0 + 4);
    }
  }
  for (int st = st0; st < st1; ++st) {
    const size_t nidx = (size_t)(st + 1) * W + wave;
    const int e_nxt =
        (st + 1 < st1) ? sload(sa.step_elems + nidx) : -1;
    if (e >= 0) {
      const bool mid = NP == 6 && c >= 3;
      if (active && !mid) {
        double fr[KB][NP], fi[KB][NP], gr[KB][NP], gi[KB][NP];
#pragma unroll
        for (int q = 0; q < KB; ++q)
#pragma unroll
          for (int s = 0; s < NP; ++s) fr[q][s] = fi[q][s] = gr[q][s] = gi[q][s] = 0.0;
        const double* tab = a.elem + ((size_t)e * a.Pe + a.off[c]) * 8;
        const double* ph = s_phi + c * a.Qmax * NP;
        const int Q = a.len[c];
        double cs_local = 0.0;
        // point 0 was fetched before the barrier; later rows are loaded in turn
        if constexpr (TAY) {
          taylor_point<NP, KB, true>(x0, x1, x2, A0, B0, ph, sw.k_re, sw.k_im, k0, doG, fr, fi,
                                     gr, gi, cs_local, s_trig);
          for (int q = 1; q < Q; ++q) {
            const double4 A = *reinterpret_cast<const double4*>(tab + 8 * q);
            const double4 B = *reinterpret_cast<const double4*>(tab + 8 * q + 4);
            taylor_point<NP, KB, true>(x0, x1, x2, A, B, ph + q * NP, sw.k_re, sw.k_im, k0, doG,
                                       fr, fi, gr, gi, cs_local, s_trig);
          }
        } else {
          sweep_point<NP, KB, true>(x0, x1, x2, A0, B0, ph, kv, dk, rec, doG, fr, fi, gr, gi,
                                    cs_local, s_trig);
          for (int q = 1; q < Q; ++q) {
            const double4 A = *reinterpret_cast<const double4*>(tab + 8 * q);
            const double4 B = *reinterpret_cast<const double4*>(tab + 8 * q + 4);
            sweep_point<NP, KB, true>(x0, x1, x2, A, B, ph + q * NP, kv, dk, rec, doG, fr, fi,
                                      gr, gi, cs_local, s_trig);
          }
        }
        csum += cs_local;
        const int* so = sa.step_slot + ((size_t)st * W + wave) * NP;
#pragma unroll
        for (int s = 0; s < NP; ++s)
#pragma unroll
          for (int q2 = 0; q2 < KB; ++q2) {
            if constexpr (PACK) {
              double* p = acc + ((size_t)sload(so + s) * KBS + q2 / 2) * 256 + lane +
                          (q2 & 1) * 128;
              atomicAdd(p, fr[q2][s]);
              atomicAdd(p + 64, fi[q2][s]);
            } else {
              double* p = acc + ((size_t)sload(so + s) * KB + q2) * 256 + lane;
              atomicAdd(p, fr[q2][s]);
              atomicAdd(p + 64, fi[q2][s]);
              if (doG) {
                atomicAdd(p + 128, gr[q2][s]);
                atomicAdd(p + 192, gi[q2][s]);
              }
            }
          }
      }
      if (NP == 6 && first_kb) {  // mid-edge pairs are queued once, for every k
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
    if (e_nxt >= 0) {
      c_nxt = select_rule<NP>(x0, x1, x2, sa.step_sel + nidx * NP * 4, a.R, a.hyp) - 1;
      if (!(NP == 6 && c_nxt >= 3)) {
        const double* t0 = a.elem + ((size_t)e_nxt * a.Pe + a.off[c_nxt]) * 8;
        A0 = *reinterpret_cast<const double4*>(t0);
        B0 = *reinterpret_cast<const double4*>(t0 + 4);
      }
    }
    // the previous step's finished nodes (their slots were complete at the last barrier)
    if (st > st0) sweep_flush<KBS, W, PACK>(a, sa, sw, st - 1, wave, lane, i, active, k0, acc);
    lds_barrier();
    e = e_nxt;
    c = c_nxt;
  }
  if (st1 > st0) {
    sweep_flush<KBS, W, PACK>(a, sa, sw, st1 - 1, wave, lane, i, active, k0, acc);
    lds_barrier();
  }
  if (NP == 6 && first_kb && wave == 0) {
    const unsigned n = min(*s_qn, (unsigned)kMidLds);
    if (n) {
      unsigned base = 0;
      if (lane == 0) base = atomicAdd(sa.mid_count, n);
      base = __shfl(base, 0);
      for (unsigned t = lane; t < n; t += 64)
        if ((int64_t)(base + t) < sa.mid_cap) sa.mid_queue[base + t] = s_q[t];
    }
  }
  if (first_kb && active && (a.flags & BIE_C)) unsafeAtomicAdd(a.C + 2 * i, csum);
}

// Mid-edge pairs for every k (with TAY: every derivative order): one thread per (pair, k).
template <bool TAY>
__global__ void __launch_bounds__(256)
k_sweep_mid(const AssemblyArgs a, const SlotArgs sa, const SweepArgs sw) {
  const int64_t n = min((int64_t)*sa.mid_count, sa.mid_cap);
  const int64_t total = n * sw.nk;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const bool doF = (a.flags & BIE_F) != 0, doG = (a.flags & BIE_G) != 0;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += stride) {
    const int kidx = (int)(t / n);
    const int2 en = sa.mid_queue[t - (int64_t)kidx * n];
    const int64_t i = en.x;
    const int e = en.y & 0x0fffffff;
    const int c = 3 + (en.y >> 28);
    const double x0 = a.targets[3 * i], x1 = a.targets[3 * i + 1], x2 = a.targets[3 * i + 2];
    double fr[1][6], fi[1][6], gr[1][6], gi[1][6], csum = 0.0;
#pragma unroll
    for (int s = 0; s < 6; ++s) fr[0][s] = fi[0][s] = gr[0][s] = gi[0][s] = 0.0;
    const double* tab = a.elem + ((size_t)e * a.Pe + a.off[c]) * 8;
    const double* ph = a.phi + c * a.Qmax * 6;
    const double kk = TAY ? sw.k_re : sw.k[kidx];
    for (int q = 0; q < a.len[c]; ++q) {
      const double4 A = *reinterpret_cast<const double4*>(tab + 8 * q);
      const double4 B = *reinterpret_cast<const double4*>(tab + 8 * q + 4);
      if constexpr (TAY)
        taylor_point<6, 1, false>(x0, x1, x2, A, B, ph + q * 6, kk, sw.k_im, kidx, doG, fr, fi,
                                  gr, gi, csum, nullptr);
      else
        sweep_point<6, 1, false>(x0, x1, x2, A, B, ph + q * 6, &kk, 0.0, false, doG, fr, fi,
                                 gr, gi, csum, nullptr);
    }
    const int* cols = a.ptopo + (size_t)e * 6;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const size_t idx = 2 * ((size_t)kidx * sw.stride + (size_t)i + (size_t)cols[s] * a.M);
      if (doF) {
        unsafeAtomicAdd(a.F + idx, fr[0][s]);
        unsafeAtomicAdd(a.F + idx + 1, fi[0][s]);
      }
      if (doG) {
        unsafeAtomicAdd(a.G + idx, gr[0][s]);
        unsafeAtomicAdd(a.G + idx + 1, gi[0][s]);
      }
    }
    if (kidx == 0 && (a.flags & BIE_C)) unsafeAtomicAdd(a.C + 2 * i, csum);
  }
}

// Zero listed columns of nk column-major M x N complex matrices spaced stride apart:
// workgroups stride over the (matrix, column) pairs, threads over the rows of a column.
__global__ void k_zero_columns_multi(double2* __restrict__ X, int64_t stride, int nk, int64_t M,
                                     const int* __restrict__ cols, int ncols) {
  const double2 z = make_double2(0.0, 0.0);
  const int64_t pairs = (int64_t)nk * ncols;
  for (int64_t pc = blockIdx.x; pc < pairs; pc += gridDim.x) {
    const int kq = (int)(pc / ncols);
    const int c = (int)(pc - (int64_t)kq * ncols);
    double2* col = X + (size_t)kq * stride + (size_t)cols[c] * M;
    for (int64_t r = threadIdx.x; r < M; r += blockDim.x) col[r] = z;
  }
}

}  // namespace

void launch_zero_columns_multi(double* X, int64_t stride, int nk, int64_t M, const int* cols,
                               int ncols, hipStream_t s) {
  if (ncols <= 0 || M <= 0 || nk <= 0) return;
  int64_t blocks = (int64_t)ncols * nk;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_zero_columns_multi, dim3((unsigned)blocks), dim3(256), 0, s,
                     reinterpret_cast<double2*>(X), stride, nk, M, cols, ncols);
  BIE_HIP_CHECK(hipGetLastError());
}

int sweep_kb(int NP) { return NP == 6 ? 2 : 4; }

int sweep_block_size(int NP, bool gon) { return gon ? sweep_kb(NP) : 2 * sweep_kb(NP); }

namespace {

__global__ void k_block_uniform(const double* __restrict__ k, int nk, int KB, int* out) {
  const int kb = blockIdx.x * blockDim.x + threadIdx.x;
  if (kb * KB >= nk) return;
  const double k0 = k[kb * KB];
  const double dk = (kb * KB + 1 < nk) ? k[kb * KB + 1] - k0 : 0.0;
  int ok = (kb * KB + KB <= nk) ? 1 : 0;  // a partial block evaluates directly
  for (int b = 2; b < KB && ok; ++b) {
    const double kb_ = k[kb * KB + b];
    // the recurrence produces e^{i (k0 + b dk) r}; accept it when it differs from
    // e^{i k_b r} by at most ~2 ulp of k_b r (a rounded linspace passes)
    if (fabs(kb_ - (k0 + b * dk)) > 4.5e-16 * fabs(kb_)) ok = 0;
  }
  out[kb] = ok;
}

size_t g_sweep_attr[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

template <int NP, int KB, int W>
const void* sweep_kernel_np(bool taylor, bool gon) {
  if (taylor)
    return gon ? reinterpret_cast<const void*>(k_sweep_slots<NP, KB, W, true, true>)
               : reinterpret_cast<const void*>(k_sweep_slots<NP, 2 * KB, W, true, false>);
  return gon ? reinterpret_cast<const void*>(k_sweep_slots<NP, KB, W, false, true>)
             : reinterpret_cast<const void*>(k_sweep_slots<NP, 2 * KB, W, false, false>);
}

template <int W>
const void* sweep_kernel(int NP, bool taylor, bool gon) {
  if (NP == 1) return sweep_kernel_np<1, 4, W>(taylor, gon);
  return NP == 3 ? sweep_kernel_np<3, 4, W>(taylor, gon) : sweep_kernel_np<6, 2, W>(taylor, gon);
}

}  // namespace

void launch_block_uniform(const double* k, int nk, int KB, int* out, hipStream_t s) {
  const int nkb = (nk + KB - 1) / KB;
  if (nkb <= 0) return;
  hipLaunchKernelGGL(k_block_uniform, dim3((nkb + 63) / 64), dim3(64), 0, s, k, nk, KB, out);
  BIE_HIP_CHECK(hipGetLastError());
}

size_t sweep_shmem(int NP, int nphi, int S) {
  const int KB = sweep_kb(NP);
  return sizeof(double) * ((size_t)S * KB * 256 + sweep_fixed_doubles(nphi)) +
         sizeof(int2) * kMidLds + 16;
}

void sweep_prepare_attributes(int NP, int nphi, bool taylor, bool gon) {
  const size_t shmem =
      sweep_shmem(NP, nphi, fitting_slots(sweep_fixed_doubles(nphi), sweep_kb(NP)));
  constexpr int W = kSlotWaves;
  const void* kern = sweep_kernel<W>(NP, taylor, gon);
  size_t& done = g_sweep_attr[(NP == 1 ? 0 : NP == 3 ? 1 : 2) + (taylor ? 3 : 0) + (gon ? 6 : 0)];
  if (shmem > done) {
    BIE_HIP_CHECK(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)std::max(shmem, (size_t)(144 * 1024))));
    done = std::max(shmem, (size_t)(144 * 1024));
  }
}

void launch_sweep(const AssemblyArgs& a, const SlotArgs& sa, const SweepArgs& sw,
                  hipStream_t s) {
  if (a.M == 0 || a.E == 0 || sw.nk == 0) return;
  const int KB = sweep_kb(a.NP);
  const int nphi = a.R * a.Qmax * a.NP;
  BIE_REQUIRE(sa.W == kSlotWaves && sa.S <= fitting_slots(sweep_fixed_doubles(nphi), KB),
              BIE_EINVAL, "sweep schedule does not match the kernel");
  const int64_t per_xcd = (sw.n_tiles + 7) / 8;
  BIE_REQUIRE(per_xcd * 8 <= 2147483647LL, BIE_EINVAL, "too many sweep tiles");
  const size_t shmem = sweep_shmem(a.NP, nphi, sa.S);
  BIE_REQUIRE(shmem <= 160 * 1024, BIE_EINVAL, "sweep kernel LDS request exceeds 160 KiB");
  const bool tay = sw.taylor != 0;
  const bool gon = (a.flags & BIE_G) != 0;
  sweep_prepare_attributes(a.NP, nphi, tay, gon);
  constexpr int W = kSlotWaves;
  const void* kern = sweep_kernel<W>(a.NP, tay, gon);
  AssemblyArgs a_ = a;
  SlotArgs sa_ = sa;
  SweepArgs sw_ = sw;
  void* args[] = {&a_, &sa_, &sw_};
  BIE_HIP_CHECK(hipLaunchKernel(kern, dim3((unsigned)(per_xcd * 8)), dim3(64 * W), args, shmem,
                                s));
  if (uses_mid_rules(a.R, a.NP)) {
    if (tay)
      hipLaunchKernelGGL(k_sweep_mid<true>, dim3(4096), dim3(256), 0, s, a, sa, sw);
    else
      hipLaunchKernelGGL(k_sweep_mid<false>, dim3(4096), dim3(256), 0, s, a, sa, sw);
    BIE_HIP_CHECK(hipGetLastError());
  }
}

}  // namespace bie
