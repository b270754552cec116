// Host construction of the quadrature rules of the dense collocation assembly.
//
// Rules follow the reference exactly (same points, weights and order):
//   curveLinearQuadpoints           src/2d/quadrature.jl:5-11 (Gauss-Legendre on [0,1])
//   rotated_triangular_quadpoints   src/3d/quadrature.jl:99-113   (rule 1, node 1)
//   triangularQuadpoints            src/3d/quadrature.jl:70-82    (rule 2, node 2)
//   u/v swapped                     src/3d/assembly_collocation.jl:127-128 (rule 3, node 3)
//   rotated_midpoint_triangular_quadpoints src/3d/quadrature.jl:117-144 (rules 4-6)
//   bases / derivatives             src/3d/surface_functions.jl:191-233
//   quadrilateralQuadpoints         src/3d/quadrature.jl:9-20 (called as getQuadpoints(sf, m, n)
//                                   by create_shape_function, assembly_collocation.jl:12-20)
//   quadrilateral bases             src/3d/surface_functions.jl:1-149
#include <cmath>
#include <stdexcept>

#include "bie_common.hpp"

namespace bie {

// Gauss-Legendre nodes (ascending) and weights on [-1, 1].
void gauss_legendre_pm1(int n, std::vector<double>& x, std::vector<double>& w) {
  // Roots of P_n by Newton's method from cos(pi (4i+3) / (4n+2)), evaluated with the
  // three-term recurrence.
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(kPi * (4.0 * i + 3.0) / (4.0 * n + 2.0));
    double pn = 0, dpn = 0;
    for (int it = 0; it < 64; ++it) {
      double pkm1 = 1.0, pk = z;
      if (n == 1) pkm1 = 1.0;
      for (int k = 2; k <= n; ++k) {
        double pkp1 = ((2.0 * k - 1.0) * z * pk - (k - 1.0) * pkm1) / k;
        pkm1 = pk;
        pk = pkp1;
      }
      pn = (n == 1) ? z : pk;
      double pnm1 = (n == 1) ? 1.0 : pkm1;
      dpn = n * (pnm1 - z * pn) / (1.0 - z * z);
      double dz = pn / dpn;
      z -= dz;
      if (std::fabs(dz) <= 1e-16 * std::fabs(z) + 1e-300) {
        // refresh the derivative at the converged root
        pkm1 = 1.0;
        pk = z;
        for (int k = 2; k <= n; ++k) {
          double pkp1 = ((2.0 * k - 1.0) * z * pk - (k - 1.0) * pkm1) / k;
          pkm1 = pk;
          pk = pkp1;
        }
        pn = (n == 1) ? z : pk;
        pnm1 = (n == 1) ? 1.0 : pkm1;
        dpn = n * (pnm1 - z * pn) / (1.0 - z * z);
        break;
      }
    }
    // i = 0 is the largest root; store ascending.
    x[n - 1 - i] = z;
    w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dpn * dpn);
  }
}

// The same mapped from [-1,1] to [0,1] as the reference does (curveLinearQuadpoints).
void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w) {
  gauss_legendre_pm1(n, x, w);
  for (int i = 0; i < n; ++i) {
    x[i] = 0.5 * (x[i] + 1.0);
    w[i] = 0.5 * w[i];
  }
}

namespace {

// Quadratic Lagrange polynomials on the nodes -1, 0, 1 and their derivatives.
void lagrange3(double x, double* l, double* d) {
  l[0] = 0.5 * x * (x - 1.0);
  l[1] = (1.0 - x) * (1.0 + x);
  l[2] = 0.5 * x * (x + 1.0);
  d[0] = x - 0.5;
  d[1] = -2.0 * x;
  d[2] = x + 0.5;
}

// Quadrilateral geometry bases (surface_functions.jl:1-95) and their derivatives.  Node
// orders: BIE_QUAD4 (-1,-1), (1,-1), (-1,1), (1,1); BIE_QUAD4S counter-clockwise from
// (-1,-1); BIE_QUAD8 the four corners counter-clockwise, then the edge mid points (0,-1),
// (1,0), (0,1), (-1,0); BIE_QUAD9 (COMSOL) (-1,-1), (1,-1), (-1,1), (1,1), (0,-1), (-1,0),
// (0,0), (1,0), (0,1).
void eval_quad_basis(int kind, double u, double v, double* N, double* du, double* dv) {
  if (kind == BIE_QUAD4 || kind == BIE_QUAD4S) {
    const double um = 1.0 - u, up = 1.0 + u, vm = 1.0 - v, vp = 1.0 + v;
    // corner (su, sv) in the node order of the kind
    static const int cu[2][4] = {{-1, 1, -1, 1}, {-1, 1, 1, -1}};
    static const int cv[2][4] = {{-1, -1, 1, 1}, {-1, -1, 1, 1}};
    const int o = kind == BIE_QUAD4S ? 1 : 0;
    for (int i = 0; i < 4; ++i) {
      const double fu = cu[o][i] < 0 ? um : up, fv = cv[o][i] < 0 ? vm : vp;
      N[i] = fu * fv / 4.0;
      du[i] = cu[o][i] * fv / 4.0;
      dv[i] = cv[o][i] * fu / 4.0;
    }
    return;
  }
  if (kind == BIE_QUAD8) {
    const double su[4] = {-1, 1, 1, -1}, sv[4] = {-1, -1, 1, 1};
    for (int i = 0; i < 4; ++i) {  // corners: (1 + su u)(1 + sv v)(su u + sv v - 1) / 4
      const double a = 1.0 + su[i] * u, b = 1.0 + sv[i] * v, c = su[i] * u + sv[i] * v - 1.0;
      N[i] = a * b * c / 4.0;
      du[i] = su[i] * b * (c + a) / 4.0;
      dv[i] = sv[i] * a * (c + b) / 4.0;
    }
    N[4] = (1.0 - v) * (1.0 - u * u) / 2.0;
    du[4] = -(1.0 - v) * u;
    dv[4] = -(1.0 - u * u) / 2.0;
    N[5] = (1.0 + u) * (1.0 - v * v) / 2.0;
    du[5] = (1.0 - v * v) / 2.0;
    dv[5] = -(1.0 + u) * v;
    N[6] = (1.0 + v) * (1.0 - u * u) / 2.0;
    du[6] = -(1.0 + v) * u;
    dv[6] = (1.0 - u * u) / 2.0;
    N[7] = (1.0 - u) * (1.0 - v * v) / 2.0;
    du[7] = -(1.0 - v * v) / 2.0;
    dv[7] = -(1.0 - u) * v;
    return;
  }
  // BIE_QUAD9: tensor products of the 1-D quadratic Lagrange polynomials
  double lu[3], du1[3], lv[3], dv1[3];
  lagrange3(u, lu, du1);
  lagrange3(v, lv, dv1);
  static const int iu[9] = {0, 2, 0, 2, 1, 0, 1, 2, 1}, iv[9] = {0, 0, 2, 2, 0, 1, 1, 1, 2};
  for (int i = 0; i < 9; ++i) {
    N[i] = lu[iu[i]] * lv[iv[i]];
    du[i] = du1[iu[i]] * lv[iv[i]];
    dv[i] = lu[iu[i]] * dv1[iv[i]];
  }
}

void eval_basis(int kind, double u, double v, double* N, double* du, double* dv) {
  for (int i = 0; i < 9; ++i) N[i] = du[i] = dv[i] = 0.0;
  if (kind_is_quad(kind)) {
    eval_quad_basis(kind, u, v, N, du, dv);
    return;
  }
  if (kind == 3) {
    N[0] = (1.0 - u) - v;
    N[1] = u;
    N[2] = v;
    du[0] = -1.0;
    du[1] = 1.0;
    dv[0] = -1.0;
    dv[2] = 1.0;
    return;
  }
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
  du[3] = 4.0 * l - 4.0 * u;
  du[4] = 4.0 * v;
  du[5] = -4.0 * v;
  dv[0] = d0;
  dv[2] = 4.0 * v - 1.0;
  dv[3] = -4.0 * u;
  dv[4] = 4.0 * u;
  dv[5] = 4.0 * l - 4.0 * v;
}

struct Pts {
  std::vector<double> u, v, w;
};

// Tensor grid u = kron(ones(mm), x_n), v = kron(x_mm, ones(n)), w = kron(w_mm, w_n).
void tensor(int n, int mm, Pts& t) {
  std::vector<double> xn, wn, xm, wm;
  gauss_legendre01(n, xn, wn);
  gauss_legendre01(mm, xm, wm);
  for (int b = 0; b < mm; ++b)
    for (int a = 0; a < n; ++a) {
      t.u.push_back(xn[a]);
      t.v.push_back(xm[b]);
      t.w.push_back(wm[b] * wn[a]);
    }
}

Pts rule_points(int c, int n, int m) {
  Pts out;
  if (c <= 3) {
    Pts t;
    tensor(n, m, t);
    for (size_t q = 0; q < t.u.size(); ++q) {
      const double u = t.u[q], v = t.v[q], w = t.w[q];
      double xi, eta, W;
      if (c == 1) {
        xi = u * v;
        eta = (1.0 - u) * v;
        W = w * v;
      } else {
        xi = u;
        eta = v * (1.0 - u);
        W = w * std::fabs(1.0 - u);
        if (c == 3) {
          double tmp = xi;
          xi = eta;
          eta = tmp;
        }
      }
      out.u.push_back(xi);
      out.v.push_back(eta);
      out.w.push_back(W);
    }
    return out;
  }
  const int mm = static_cast<int>(std::ceil(m / 2.0)) * 2;
  Pts t;
  tensor(n, mm, t);
  std::vector<double> X, Y, W;
  for (int blk = 0; blk < 2; ++blk)
    for (size_t q = 0; q < t.u.size(); ++q) {
      const double u = t.u[q], v = t.v[q], w = t.w[q];
      if (blk == 0) {
        X.push_back(0.5 * u);
        Y.push_back(v * (1.0 - u));
        W.push_back(w * (0.5 * (1.0 - u)));
      } else {
        X.push_back((u - 0.5) * v + 0.5);
        Y.push_back(v * (1.0 - u));
        W.push_back(w * (0.5 * v));
      }
    }
  for (size_t q = 0; q < X.size(); ++q) {
    if (c == 4) {
      out.u.push_back(X[q]);
      out.v.push_back(Y[q]);
    } else if (c == 5) {
      out.u.push_back((1.0 - X[q]) - Y[q]);
      out.v.push_back(X[q]);
    } else {
      out.u.push_back(Y[q]);
      out.v.push_back((1.0 - X[q]) - Y[q]);
    }
    out.w.push_back(W[q]);
  }
  return out;
}

}  // namespace

// N, du, dv: 9 entries each (the slots beyond the kind's node count are zeroed)
void geometry_basis(int kind, double u, double v, double* N, double* du, double* dv) {
  eval_basis(kind, u, v, N, du, dv);
}

// Physics basis at (u, v).  Continuous kinds are the geometry bases; the discontinuous
// ones (surface_functions.jl:234-318) use psi(s) = (s - beta) / (1 - 3 beta) of the three
// barycentric coordinates 1-u-v, u, v, each operation in the reference's order.
void physics_basis(int pkind, double beta, double u, double v, double* N) {
  double du[9], dv[9];
  if (pkind == BIE_TRI3 || pkind == BIE_TRI6 || (kind_is_quad(pkind) &&
                                                 !physics_discontinuous(pkind))) {
    eval_basis(pkind, u, v, N, du, dv);
    return;
  }
  if (pkind == BIE_DTRI1 || pkind == BIE_DQUAD1) {
    N[0] = 1.0;
    return;
  }
  if (pkind == BIE_DQUAD4 || pkind == BIE_DQUAD9) {
    // the continuous basis on the element shrunk by 1 - beta
    // (discontinuousQuadrilateral*, surface_functions.jl:115-143)
    eval_basis(pkind == BIE_DQUAD4 ? BIE_QUAD4 : BIE_QUAD9, u / (1.0 - beta), v / (1.0 - beta),
               N, du, dv);
    return;
  }
  const double den = 1.0 - 3.0 * beta;
  const double l1 = (((1.0 - u) - v) - beta) / den;
  const double l2 = (u - beta) / den;
  const double l3 = (v - beta) / den;
  if (pkind == BIE_DTRI3) {
    N[0] = l1;
    N[1] = l2;
    N[2] = l3;
    return;
  }
  N[0] = l1 * (2.0 * l1 - 1.0);
  N[1] = l2 * (2.0 * l2 - 1.0);
  N[2] = l3 * (2.0 * l3 - 1.0);
  N[3] = (4.0 * l2) * l1;
  N[4] = (4.0 * l2) * l3;
  N[5] = (4.0 * l3) * l1;
}

// Tensor Gauss-Legendre rule on [-1, 1]^2 of a quadrilateral: create_shape_function
// (assembly_collocation.jl:12-20) calls getQuadpoints(sf, m, n), i.e. quadrilateralQuadpoints
// with its n and m exchanged, so u takes m points and v takes n:
// u = kron(ones(n), x_m), v = kron(x_n, ones(m)), w = kron(w_n, w_m).
static Pts quad_rule_points(int n, int m) {
  std::vector<double> xm, wm, xn, wn;
  gauss_legendre_pm1(m, xm, wm);
  gauss_legendre_pm1(n, xn, wn);
  Pts t;
  for (int b = 0; b < n; ++b)
    for (int a = 0; a < m; ++a) {
      t.u.push_back(xm[a]);
      t.v.push_back(xn[b]);
      t.w.push_back(wn[b] * wm[a]);
    }
  return t;
}

static bool kind_known(int k) {
  return k == BIE_TRI3 || k == BIE_TRI6 || k == BIE_DTRI1 || k == BIE_DTRI3 || k == BIE_DTRI6 ||
         k == BIE_QUAD4 || k == BIE_QUAD4S || k == BIE_QUAD8 || k == BIE_QUAD9 ||
         k == BIE_DQUAD1 || k == BIE_DQUAD4 || k == BIE_DQUAD9;
}

RuleSet build_ruleset(int gkind, int pkind, double beta, int n, int m) {
  const bool g_ok = gkind == BIE_TRI3 || gkind == BIE_TRI6 || gkind == BIE_QUAD4 ||
                    gkind == BIE_QUAD4S || gkind == BIE_QUAD8 || gkind == BIE_QUAD9;
  if (!g_ok || !kind_known(pkind) || kind_is_quad(gkind) != kind_is_quad(pkind))
    throw Error{BIE_EUNSUPPORTED, "unsupported geometry/physics element pair"};
  if (n < 1 || m < 1 || n > 64 || m > 64)
    throw Error{BIE_EINVAL, "quadrature orders n, m must be in 1..64"};
  const bool quad = kind_is_quad(gkind);
  // assembly_collocation.jl:497-504: a Gauss node at the centre node of a 9-node element
  if (gkind == BIE_QUAD9 && (n % 2 == 1) && (m % 2 == 1))
    throw Error{BIE_EINVAL, "Gauss Node On Singularity. (QuadrilateralQuadraticLagrange "
                            "geometry with odd n and m puts a Gauss point on its centre node)"};
  const int pk = physics_nodes(pkind);
  RuleSet rs;
  rs.gk = geometry_nodes(gkind);
  rs.gkind = gkind;
  rs.pk = pk;
  rs.pkind = pkind;
  const bool has_beta = pkind == BIE_DTRI3 || pkind == BIE_DTRI6 || pkind == BIE_DQUAD4 ||
                        pkind == BIE_DQUAD9;
  rs.beta = has_beta ? beta : 0.0;
  rs.n = n;
  rs.m = m;
  rs.R = quad ? 1 : (gkind == BIE_TRI6 ? 6 : 3);
  std::vector<Pts> rules;
  if (quad)
    rules.push_back(quad_rule_points(n, m));
  else
    for (int c = 1; c <= rs.R; ++c) rules.push_back(rule_points(c, n, m));
  rs.Qc = (int)rules[0].u.size();
  rs.Qm = (rs.R == 6) ? (int)rules[3].u.size() : 0;
  rs.Pe = 0;
  rs.Qmax = 0;
  for (int c = 0; c < rs.R; ++c) {
    rs.off[c] = rs.Pe;
    rs.len[c] = (int)rules[c].u.size();
    rs.Pe += rs.len[c];
    if (rs.len[c] > rs.Qmax) rs.Qmax = rs.len[c];
  }
  rs.pts.assign((size_t)rs.Pe * kPtStride, 0.0);
  rs.phi.assign((size_t)rs.R * rs.Qmax * pk, 0.0);
  for (int c = 0; c < rs.R; ++c) {
    for (int q = 0; q < rs.len[c]; ++q) {
      double* p = &rs.pts[(size_t)(rs.off[c] + q) * kPtStride];
      const double u = rules[c].u[q], v = rules[c].v[q];
      p[0] = u;
      p[1] = v;
      p[2] = rules[c].w[q];
      eval_basis(gkind, u, v, p + kPtN, p + kPtDu, p + kPtDv);
      double N[9];
      physics_basis(pkind, rs.beta, u, v, N);
      for (int s = 0; s < pk; ++s) rs.phi[((size_t)c * rs.Qmax + q) * pk + s] = N[s];
    }
  }
  return rs;
}

}  // namespace bie
