// Host construction of the quadrature rules of the dense collocation assembly.
//
// Rules follow the reference exactly (same points, weights and order):
//   curveLinearQuadpoints           src/2d/quadrature.jl:5-11 (Gauss-Legendre on [0,1])
//   rotated_triangular_quadpoints   src/3d/quadrature.jl:99-113   (rule 1, node 1)
//   triangularQuadpoints            src/3d/quadrature.jl:70-82    (rule 2, node 2)
//   u/v swapped                     src/3d/assembly_collocation.jl:127-128 (rule 3, node 3)
//   rotated_midpoint_triangular_quadpoints src/3d/quadrature.jl:117-144 (rules 4-6)
//   bases / derivatives             src/3d/surface_functions.jl:191-233
#include <cmath>
#include <stdexcept>

#include "bie_common.hpp"

namespace bie {

void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w) {
  // Roots of P_n by Newton's method from cos(pi (4i+3) / (4n+2)), evaluated with the
  // three-term recurrence; mapped from [-1,1] to [0,1] as the reference does.
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
    x[n - 1 - i] = 0.5 * (z + 1.0);
    w[n - 1 - i] = 0.5 * (2.0 / ((1.0 - z * z) * dpn * dpn));
  }
}

namespace {

void eval_basis(int kind, double u, double v, double* N, double* du, double* dv) {
  for (int i = 0; i < 6; ++i) N[i] = du[i] = dv[i] = 0.0;
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

void geometry_basis(int kind, double u, double v, double* N, double* du, double* dv) {
  eval_basis(kind, u, v, N, du, dv);
}

// Physics basis at (u, v).  Continuous kinds are the geometry bases; the discontinuous
// ones (surface_functions.jl:234-318) use psi(s) = (s - beta) / (1 - 3 beta) of the three
// barycentric coordinates 1-u-v, u, v, each operation in the reference's order.
void physics_basis(int pkind, double beta, double u, double v, double* N) {
  if (pkind == BIE_TRI3 || pkind == BIE_TRI6) {
    double du[6], dv[6];
    eval_basis(pkind, u, v, N, du, dv);
    return;
  }
  if (pkind == BIE_DTRI1) {
    N[0] = 1.0;
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

RuleSet build_ruleset(int gk, int pkind, double beta, int n, int m) {
  const bool pk_ok = pkind == BIE_TRI3 || pkind == BIE_TRI6 || pkind == BIE_DTRI1 ||
                     pkind == BIE_DTRI3 || pkind == BIE_DTRI6;
  if ((gk != 3 && gk != 6) || !pk_ok)
    throw Error{BIE_EUNSUPPORTED, "only linear (3-node) and quadratic (6-node) triangles"};
  const int pk = physics_nodes(pkind);
  if (n < 1 || m < 1 || n > 64 || m > 64)
    throw Error{BIE_EINVAL, "quadrature orders n, m must be in 1..64"};
  RuleSet rs;
  rs.gk = gk;
  rs.pk = pk;
  rs.pkind = pkind;
  rs.beta = physics_discontinuous(pkind) && pkind != BIE_DTRI1 ? beta : 0.0;
  rs.n = n;
  rs.m = m;
  rs.R = (gk == 6) ? 6 : 3;
  std::vector<Pts> rules;
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
      eval_basis(gk, u, v, p + 3, p + 9, p + 15);
      double N[6];
      physics_basis(pkind, rs.beta, u, v, N);
      for (int s = 0; s < pk; ++s) rs.phi[((size_t)c * rs.Qmax + q) * pk + s] = N[s];
    }
  }
  return rs;
}

}  // namespace bie
