// TEST INFRASTRUCTURE — CPU oracle for the dense 3D collocation assembly.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
// library, and only to compare results with (or as the timed CPU baseline).  The product
// path never calls it.
//
// A restatement, in C++ with OpenMP threads over target rows, of the reference's
// assemble_parallel! on linear and quadratic triangles:
//   * Gauss-Legendre on [0,1]            src/2d/quadrature.jl:5-11 (third-party
//     FastGaussQuadrature.gausslegendre; restated here by Newton on P_n)
//   * vertex-clustered Duffy rules        src/3d/quadrature.jl:70-82, 99-113
//   * mid-edge split rules                src/3d/quadrature.jl:117-144
//   * rule wiring per closest node        src/3d/assembly_collocation.jl:122-136, 385-399
//   * bases and derivatives               src/3d/surface_functions.jl:191-233
//   * element geometry (y, -n, J*w)       src/3d/assembly_collocation.jl:27-66,
//                                         src/3d/jacobian.jl:7-87
//   * rule selection                      src/3d/assembly_collocation.jl:102-115
//     (select_exact.cpp, compiled without FMA contraction)
//   * kernels G, F, C                     src/3d/kernels.jl:13-55
//   * accumulation                        src/3d/assembly_collocation.jl:213-252,
//                                         src/2d/assembly_collocation.jl:142-150
//   * the row loop                        src/3d/assembly_collocation.jl:270-479
//   * Taylor-in-k kernels (d^m/dk^m)      src/3d/kernels.jl:94-118, src/3d/rosebem.jl:56-157
//
// The reference's loop structure is kept (threads over targets, per-row scratch, G then
// C then F per pair with e^{ikr} evaluated separately for G and F, row-slice writes into
// column-major F and G) so that it doubles as the CPU baseline.
//
// Defined deviation: selection reads mesh.sources[:, physics_topology[:, e]] for every
// target.  The reference indexes the *target* array there (assembly_collocation.jl:325,
// 436), which coincides for collocation and is out of bounds for field points.
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" double orc_julia_hypot3(double x, double y, double z);
extern "C" int orc_find_closest(const double* x, const double* nodes, int n_nodes);
extern "C" int orc_find_closest_variant(const double* x, const double* nodes, int n_nodes,
                                        int variant);

namespace {

constexpr double kPi = 3.14159265358979323846;

void gauss_legendre_m11(int n, double* x, double* w) {
  // Newton iteration on P_n from Chebyshev-like initial guesses; ascending nodes.
  for (int i = 0; i < n; ++i) {
    double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
    double dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int j = 2; j <= n; ++j) {
        double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) p0 = 1.0;
      double pn = (n == 1) ? z : p1;
      double pm = (n == 1) ? 1.0 : p0;
      dp = n * (z * pn - pm) / (z * z - 1.0);
      double dz = pn / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-17) break;
    }
    // final derivative at converged z
    double p0 = 1.0, p1 = z;
    for (int j = 2; j <= n; ++j) {
      double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
      p0 = p1;
      p1 = p2;
    }
    double pn = (n == 1) ? z : p1, pm = (n == 1) ? 1.0 : p0;
    dp = n * (z * pn - pm) / (z * z - 1.0);
    x[n - 1 - i] = z;
    w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

struct Rule {
  std::vector<double> u, v, w;
  int size() const { return (int)u.size(); }
};

// curveLinearQuadpoints: nodes 0.5*(x+1), weights 0.5*w.
void gl01(int n, std::vector<double>& x, std::vector<double>& w) {
  x.resize(n);
  w.resize(n);
  gauss_legendre_m11(n, x.data(), w.data());
  for (int i = 0; i < n; ++i) {
    x[i] = 0.5 * (x[i] + 1.0);
    w[i] = 0.5 * w[i];
  }
}

// Rule for clustering index c (1..6) with n points in t and m points in s.
Rule make_rule(int c, int n, int m) {
  Rule r;
  if (c <= 3) {
    std::vector<double> nx, wx, ny, wy;
    gl01(n, nx, wx);
    gl01(m, ny, wy);
    for (int b = 0; b < m; ++b)
      for (int a = 0; a < n; ++a) {
        double u = nx[a], v = ny[b], w = wy[b] * wx[a];
        double xi, eta, W;
        if (c == 1) {  // rotated_triangular_quadpoints
          xi = u * v;
          eta = (1.0 - u) * v;
          W = w * v;
        } else {  // triangularQuadpoints (c==2), coordinates swapped for c==3
          xi = u;
          eta = v * (1.0 - u);
          W = w * std::fabs(1.0 - u);
          if (c == 3) std::swap(xi, eta);
        }
        r.u.push_back(xi);
        r.v.push_back(eta);
        r.w.push_back(W);
      }
    return r;
  }
  int mm = (int)std::ceil(m / 2.0) * 2;
  std::vector<double> nx, wx, ny, wy;
  gl01(n, nx, wx);
  gl01(mm, ny, wy);
  std::vector<double> X, Y, W;
  for (int blk = 0; blk < 2; ++blk)
    for (int b = 0; b < mm; ++b)
      for (int a = 0; a < n; ++a) {
        double u = nx[a], v = ny[b], w = wy[b] * wx[a];
        if (blk == 0) {  // duffyBasis2 / duffyJacobian2
          X.push_back(0.5 * u);
          Y.push_back(v * (1.0 - u));
          W.push_back(w * (0.5 * (1.0 - u)));
        } else {  // duffyBasis3 / duffyJacobian3
          X.push_back((u - 0.5) * v + 0.5);
          Y.push_back(v * (1.0 - u));
          W.push_back(w * (0.5 * v));
        }
      }
  for (size_t q = 0; q < X.size(); ++q) {
    if (c == 4) {
      r.u.push_back(X[q]);
      r.v.push_back(Y[q]);
    } else if (c == 5) {
      r.u.push_back((1.0 - X[q]) - Y[q]);
      r.v.push_back(X[q]);
    } else {
      r.u.push_back(Y[q]);
      r.v.push_back((1.0 - X[q]) - Y[q]);
    }
    r.w.push_back(W[q]);
  }
  return r;
}

void basis(int kind, double u, double v, double* N, double* du, double* dv) {
  if (kind == 3) {
    N[0] = (1.0 - u) - v;
    N[1] = u;
    N[2] = v;
    if (du) {
      du[0] = -1.0; du[1] = 1.0; du[2] = 0.0;
      dv[0] = -1.0; dv[1] = 0.0; dv[2] = 1.0;
    }
  } else {
    double l = (1.0 - v) - u;
    N[0] = l * ((1.0 - 2.0 * v) - 2.0 * u);
    N[1] = u * (2.0 * u - 1.0);
    N[2] = v * (2.0 * v - 1.0);
    N[3] = 4.0 * u * l;
    N[4] = 4.0 * u * v;
    N[5] = 4.0 * v * l;
    if (du) {
      du[0] = -(2.0 * l - 1.0) - 2.0 * l;
      du[1] = 4.0 * u - 1.0;
      du[2] = 0.0;
      du[3] = 4.0 * l - 4.0 * u;
      du[4] = 4.0 * v;
      du[5] = -4.0 * v;
      dv[0] = -(2.0 * l - 1.0) - 2.0 * l;
      dv[1] = 0.0;
      dv[2] = 4.0 * v - 1.0;
      dv[3] = -4.0 * u;
      dv[4] = 4.0 * u;
      dv[5] = 4.0 * l - 4.0 * v;
    }
  }
}

struct MeshView {
  const double* coords;
  const int64_t* topo;
  int gk;
  const double* sources;
  const int64_t* ptopo;
  int pk;
  int64_t E;
  int64_t N;
  int base;
};

// Per element, per rule point: y (3), normal (3), J*w.
struct ElemTable {
  int Q = 0;
  std::vector<double> y, nrm, jw;  // [e][q][3], [e][q][3], [e][q]
  std::vector<double> phi;         // [q][s]
};

ElemTable interpolate(const MeshView& mv, const Rule& r) {
  ElemTable t;
  t.Q = r.size();
  const int Q = t.Q, ng = mv.gk, np = mv.pk;
  std::vector<double> N(ng * Q), Du(ng * Q), Dv(ng * Q);
  for (int q = 0; q < Q; ++q) basis(ng, r.u[q], r.v[q], &N[q * ng], &Du[q * ng], &Dv[q * ng]);
  t.phi.resize((size_t)Q * np);
  for (int q = 0; q < Q; ++q) basis(np, r.u[q], r.v[q], &t.phi[q * np], nullptr, nullptr);
  t.y.resize((size_t)mv.E * Q * 3);
  t.nrm.resize((size_t)mv.E * Q * 3);
  t.jw.resize((size_t)mv.E * Q);
  for (int64_t e = 0; e < mv.E; ++e) {
    double X[3][6];
    for (int k = 0; k < ng; ++k) {
      int64_t node = mv.topo[k + e * ng] - mv.base;
      for (int d = 0; d < 3; ++d) X[d][k] = mv.coords[d + 3 * node];
    }
    for (int q = 0; q < Q; ++q) {
      double y[3], tu[3], tv[3];
      for (int d = 0; d < 3; ++d) {
        double a = 0, b = 0, c = 0;
        for (int k = 0; k < ng; ++k) {
          a += X[d][k] * N[q * ng + k];
          b += X[d][k] * Du[q * ng + k];
          c += X[d][k] * Dv[q * ng + k];
        }
        y[d] = a;
        tu[d] = b;
        tv[d] = c;
      }
      double cr[3] = {tu[1] * tv[2] - tu[2] * tv[1], tu[2] * tv[0] - tu[0] * tv[2],
                      tu[0] * tv[1] - tu[1] * tv[0]};
      double J = orc_julia_hypot3(cr[0], cr[1], cr[2]);
      size_t o = ((size_t)e * Q + q);
      for (int d = 0; d < 3; ++d) {
        t.y[3 * o + d] = y[d];
        t.nrm[3 * o + d] = -cr[d] / J;
      }
      t.jw[o] = J * r.w[q];
    }
  }
  return t;
}

struct Tables {
  int R = 0;
  ElemTable tab[6];
};

Tables build_tables(const MeshView& mv, int n, int m) {
  Tables T;
  T.R = (mv.gk == 6) ? 6 : 3;
  for (int c = 1; c <= T.R; ++c) T.tab[c - 1] = interpolate(mv, make_rule(c, n, m));
  return T;
}

// Rule index 1..R used for (target x, element e), with the reference's fallthrough:
// index 0 (no distance below Inf) lands on the last branch.
inline int rule_for(const MeshView& mv, const double* x, int64_t e, int R, int variant = 0) {
  double P[3 * 6];
  for (int s = 0; s < mv.pk; ++s) {
    int64_t node = mv.ptopo[s + e * mv.pk] - mv.base;
    for (int d = 0; d < 3; ++d) P[3 * s + d] = mv.sources[d + 3 * node];
  }
  int c = orc_find_closest_variant(x, P, mv.pk, variant);
  if (c < 1 || c > R) c = R;
  return c;
}

struct Flags {
  bool f, g, c;
};

// Taylor order mo = 0 gives the plain kernels; mo >= 1 gives d^mo/dk^mo of G and F
// (kernels.jl:94-118).  C is only formed for mo == 0.
void assemble_impl(const MeshView& mv, double kr, double ki, const double* targets,
                   int64_t M, int n, int m, Flags fl, int mo, double* F, double* G,
                   double* Cc, int nthreads, const int8_t* frozen = nullptr) {
  Tables T = build_tables(mv, n, m);
  const int64_t N = mv.N;
  std::memset(F, 0, sizeof(double) * 2 * M * N);
  std::memset(G, 0, sizeof(double) * 2 * M * N);
  std::memset(Cc, 0, sizeof(double) * 2 * M);
  int Qmax = 0;
  for (int c = 0; c < T.R; ++c) Qmax = std::max(Qmax, T.tab[c].Q);
  const double four_pi = 4.0 * kPi;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
  {
    std::vector<double> rq(Qmax), ire(Qmax), iim(Qmax);
#pragma omp for schedule(dynamic, 4)
    for (int64_t i = 0; i < M; ++i) {
      const double* x = targets + 3 * i;
      double csum = 0.0;
      for (int64_t e = 0; e < mv.E; ++e) {
        int c = frozen ? (int)frozen[i * mv.E + e] : rule_for(mv, x, e, T.R);
        if (c < 1 || c > T.R) c = T.R;
        const ElemTable& t = T.tab[c - 1];
        const int Q = t.Q, np = mv.pk;
        const double* y = &t.y[(size_t)e * Q * 3];
        const double* nr = &t.nrm[(size_t)e * Q * 3];
        const double* jw = &t.jw[(size_t)e * Q];
        // compute_distances!
        for (int q = 0; q < Q; ++q)
          rq[q] = orc_julia_hypot3(y[3 * q] - x[0], y[3 * q + 1] - x[1], y[3 * q + 2] - x[2]);
        const int64_t* pn = mv.ptopo + e * np;
        if (fl.g) {  // greens3d! (or taylor_greens3d!) * jw, then mygemm_vec!
          for (int q = 0; q < Q; ++q) {
            double r = rq[q];
            double ex = std::exp(-ki * r), cs = std::cos(kr * r) * ex, sn = std::sin(kr * r) * ex;
            double re = cs / (four_pi * r), im = sn / (four_pi * r);
            for (int j = 0; j < mo; ++j) {  // multiply by (i r)^mo
              double t2 = -im * r;
              im = re * r;
              re = t2;
            }
            ire[q] = re * jw[q];
            iim[q] = im * jw[q];
          }
          for (int s = 0; s < np; ++s) {
            double ar = 0, ai = 0;
            for (int q = 0; q < Q; ++q) {
              ar += ire[q] * t.phi[q * np + s];
              ai += iim[q] * t.phi[q * np + s];
            }
            int64_t col = pn[s] - mv.base;
            G[2 * (i + col * M)] += ar;
            G[2 * (i + col * M) + 1] += ai;
          }
        }
        if (fl.c && mo == 0) {  // freens3dk0! * jw, then sum_to_c!
          for (int q = 0; q < Q; ++q) {
            double r = rq[q];
            double nd = nr[3 * q] * (y[3 * q] - x[0]) + nr[3 * q + 1] * (y[3 * q + 1] - x[1]) +
                        nr[3 * q + 2] * (y[3 * q + 2] - x[2]);
            csum += (-nd / (four_pi * (r * r * r))) * jw[q];
          }
        }
        if (fl.f) {  // freens3d! (or taylor_freens3d!) * jw, then mygemm_vec!
          for (int q = 0; q < Q; ++q) {
            double r = rq[q];
            double nd = nr[3 * q] * (y[3 * q] - x[0]) + nr[3 * q + 1] * (y[3 * q + 1] - x[1]) +
                        nr[3 * q + 2] * (y[3 * q + 2] - x[2]);
            double ex = std::exp(-ki * r), cs = std::cos(kr * r) * ex, sn = std::sin(kr * r) * ex;
            // (i k r - 1 + mo) with complex k = kr + i ki: (-ki r - 1 + mo) + i kr r
            double ar = -ki * r - 1.0 + mo, ai = kr * r;
            double re = cs * ar - sn * ai, im = cs * ai + sn * ar;
            for (int j = 0; j < mo; ++j) {
              double t2 = -im * r;
              im = re * r;
              re = t2;
            }
            double sc = nd / (four_pi * (r * r * r));
            ire[q] = re * sc * jw[q];
            iim[q] = im * sc * jw[q];
          }
          for (int s = 0; s < np; ++s) {
            double ar = 0, ai = 0;
            for (int q = 0; q < Q; ++q) {
              ar += ire[q] * t.phi[q * np + s];
              ai += iim[q] * t.phi[q * np + s];
            }
            int64_t col = pn[s] - mv.base;
            F[2 * (i + col * M)] += ar;
            F[2 * (i + col * M) + 1] += ai;
          }
        }
      }
      Cc[2 * i] = csum;
      Cc[2 * i + 1] = 0.0;
    }
  }
}

MeshView view(const double* coords, const int64_t* topo, int gk, const double* sources,
              int64_t N, const int64_t* ptopo, int pk, int64_t E, int base) {
  MeshView mv{coords, topo, gk, sources, ptopo, pk, E, N, base};
  return mv;
}

}  // namespace

extern "C" {

int orc_gauss_legendre(int n, double* x, double* w) {
  if (n < 1) return 1;
  gauss_legendre_m11(n, x, w);
  return 0;
}

// Writes the rule for clustering index c (1..6); returns the number of points.
int orc_rule(int c, int n, int m, double* u, double* v, double* w) {
  if (c < 1 || c > 6 || n < 1 || m < 1) return -1;
  Rule r = make_rule(c, n, m);
  if (u) {
    for (int q = 0; q < r.size(); ++q) {
      u[q] = r.u[q];
      v[q] = r.v[q];
      w[q] = r.w[q];
    }
  }
  return r.size();
}

int orc_basis(int kind, int npts, const double* u, const double* v, double* N, double* du,
              double* dv) {
  if (kind != 3 && kind != 6) return 1;
  for (int q = 0; q < npts; ++q)
    basis(kind, u[q], v[q], N + q * kind, du ? du + q * kind : nullptr,
          dv ? dv + q * kind : nullptr);
  return 0;
}

// Element geometry for rule c: y[e][q][3], nrm[e][q][3], jw[e][q].
int orc_element_geometry(const double* coords, const int64_t* topo, int gk, int64_t E,
                         int base, int c, int n, int m, double* y, double* nrm, double* jw) {
  MeshView mv = view(coords, topo, gk, nullptr, 0, topo, gk, E, base);
  ElemTable t = interpolate(mv, make_rule(c, n, m));
  std::memcpy(y, t.y.data(), sizeof(double) * t.y.size());
  std::memcpy(nrm, t.nrm.data(), sizeof(double) * t.nrm.size());
  std::memcpy(jw, t.jw.data(), sizeof(double) * t.jw.size());
  return 0;
}

// Rule indices (1..R) for every (target i, element e), row-major [i][e].
int orc_select(const double* sources, int64_t N, const int64_t* ptopo, int pk, int gk,
               int64_t E, int base, const double* targets, int64_t M, int8_t* out) {
  MeshView mv = view(nullptr, nullptr, gk, sources, N, ptopo, pk, E, base);
  int R = (gk == 6) ? 6 : 3;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < M; ++i)
    for (int64_t e = 0; e < E; ++e) out[i * E + e] = (int8_t)rule_for(mv, targets + 3 * i, e, R);
  return 0;
}

// Number of (target, element) pairs whose rule differs between the two hypot formulas
// (variant 0 and 1 of orc_find_closest_variant); diff (M x E int8, may be NULL) receives the
// variant-1 rule where they differ and 0 elsewhere.
int64_t orc_select_diff(const double* sources, int64_t N, const int64_t* ptopo, int pk, int gk,
                        int64_t E, int base, const double* targets, int64_t M, int8_t* diff) {
  MeshView mv = view(nullptr, nullptr, gk, sources, N, ptopo, pk, E, base);
  int R = (gk == 6) ? 6 : 3;
  int64_t count = 0;
#pragma omp parallel for schedule(static) reduction(+ : count)
  for (int64_t i = 0; i < M; ++i)
    for (int64_t e = 0; e < E; ++e) {
      const int a = rule_for(mv, targets + 3 * i, e, R, 0);
      const int b = rule_for(mv, targets + 3 * i, e, R, 1);
      if (diff) diff[i * E + e] = (int8_t)(a != b ? b : 0);
      count += a != b;
    }
  return count;
}

// Dense assembly; F, G: interleaved complex, column-major M x N; C: complex length M.
// flags: bit0 F, bit1 G, bit2 C.  Returns 0 on success.
int orc_assemble(const double* coords, const int64_t* topo, int gk, const double* sources,
                 int64_t N, const int64_t* ptopo, int pk, int64_t E, int base, double k_re,
                 double k_im, const double* targets, int64_t M, int n, int m, int flags,
                 double* F, double* G, double* C, int nthreads) {
  if ((gk != 3 && gk != 6) || (pk != 3 && pk != 6) || n < 1 || m < 1) return 1;
  MeshView mv = view(coords, topo, gk, sources, N, ptopo, pk, E, base);
  Flags fl{(flags & 1) != 0, (flags & 2) != 0, (flags & 4) != 0};
  assemble_impl(mv, k_re, k_im, targets, M, n, m, fl, 0, F, G, C, nthreads);
  return 0;
}

// d^mo/dk^mo of F and G at k0 (taylor_assemble!, rosebem.jl:56-157).
int orc_assemble_taylor(const double* coords, const int64_t* topo, int gk,
                        const double* sources, int64_t N, const int64_t* ptopo, int pk,
                        int64_t E, int base, double k_re, double k_im, const double* targets,
                        int64_t M, int n, int m, int mo, double* F, double* G, int nthreads) {
  if ((gk != 3 && gk != 6) || (pk != 3 && pk != 6) || mo < 0) return 1;
  MeshView mv = view(coords, topo, gk, sources, N, ptopo, pk, E, base);
  std::vector<double> C(2 * M);
  assemble_impl(mv, k_re, k_im, targets, M, n, m, Flags{true, true, false}, mo, F, G,
                C.data(), nthreads);
  return 0;
}

// As orc_assemble with the rule index of every (target, element) pair given
// (rules[i * E + e] in 1..R): the assembly as a smooth function of the geometry and k at
// a frozen rule selection, for finite-difference checks of derivatives.
int orc_assemble_frozen(const double* coords, const int64_t* topo, int gk,
                        const double* sources, int64_t N, const int64_t* ptopo, int pk,
                        int64_t E, int base, double k_re, double k_im, const double* targets,
                        int64_t M, int n, int m, int flags, const int8_t* rules, double* F,
                        double* G, double* C, int nthreads) {
  if ((gk != 3 && gk != 6) || (pk != 3 && pk != 6) || n < 1 || m < 1 || !rules) return 1;
  MeshView mv = view(coords, topo, gk, sources, N, ptopo, pk, E, base);
  Flags fl{(flags & 1) != 0, (flags & 2) != 0, (flags & 4) != 0};
  assemble_impl(mv, k_re, k_im, targets, M, n, m, fl, 0, F, G, C, nthreads, rules);
  return 0;
}

int orc_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
