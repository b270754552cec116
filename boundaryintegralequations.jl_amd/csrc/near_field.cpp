// Host half of the sparse near-field assembly (the kernel is in sparse.hip): the
// reference's subdivision rule and connected_sources, restated for the GPU path.
//
// * subdivision_rule: singular_triangle_integration (src/3d/adaptive_integration.jl:215-246)
//   with subdivide_triangle (:172-213), called with the identity as element coordinates
//   and the source at a vertex or an edge midpoint (assembly_collocation_losses.jl:226-232):
//   the reference triangle is split into four while a sub-triangle's largest side exceeds
//   a quarter of its distance to the source (sides below 1e-6 of the element size count as
//   zero); each sub-triangle then gets gauss_points_triangle(3) (adaptive_integration.jl:
//   13-19), shifted to its centre and scaled by its shortest side, as the reference places
//   it.
// * sparse_pattern: connected_sources (assembly_collocation_losses.jl:98-132).
// * gauss_points_triangle: the symmetric triangle rules of the partial assembly
//   (partial_assemble_parallel!, fast_multipole_method.jl:11-81).
// * polar_rule: create_polar_elements (assembly_collocation_losses.jl:409-447) with
//   polar_gaussian (quadrature.jl:196-206) and polar_quadrature (:163-190), the rules of
//   discontinuous physics; sparse_pattern_disc: connected_sources of discontinuous physics
//   (assembly_collocation_losses.jl:41-96).
// Plain C++ (no device code), so that the CPU sanitizer build covers it.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "bie_common.hpp"

namespace bie {
namespace {

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


// The symmetric triangle rules of gauss_points_triangle (adaptive_integration.jl:13-64,
// OpenBEM's table): n = 1, 3, 4, 7, 12 or 13 points on the reference triangle, weights
// summing to its area 1/2.  Points are (u, v, w) rows; the table values are data that the
// rules are defined by, kept as the reference lists them.
bool gauss_points_triangle(int n, std::vector<double>& u, std::vector<double>& v,
                           std::vector<double>& w) {
  static const double t1[][3] = {{1.0 / 3, 1.0 / 3, 0.5}};
  static const double t3[][3] = {{1.0 / 6, 1.0 / 6, 1.0 / 6},
                                 {2.0 / 3, 1.0 / 6, 1.0 / 6},
                                 {1.0 / 6, 2.0 / 3, 1.0 / 6}};
  static const double t4[][3] = {{1.0 / 3, 1.0 / 3, -9.0 / 32},
                                 {0.2, 0.6, 25.0 / 96},
                                 {0.2, 0.2, 25.0 / 96},
                                 {0.6, 0.2, 25.0 / 96}};
  static const double t7[][3] = {{1.0 / 3, 1.0 / 3, 0.1125},
                                 {0.797426985353087, 0.101286507323456, 0.0629695902724135},
                                 {0.101286507323456, 0.797426985353087, 0.0629695902724135},
                                 {0.101286507323456, 0.101286507323456, 0.0629695902724135},
                                 {0.059715871789770, 0.470142064105115, 0.066197076394253},
                                 {0.470142064105115, 0.059715871789770, 0.066197076394253},
                                 {0.470142064105115, 0.470142064105115, 0.066197076394253}};
  static const double t12[][3] = {{0.873821971016996, 0.063089014491502, 0.0254224531851035},
                                  {0.063089014491502, 0.873821971016996, 0.0254224531851035},
                                  {0.063089014491502, 0.063089014491502, 0.0254224531851035},
                                  {0.501426509658279, 0.249286745170910, 0.0583931378631895},
                                  {0.249286745170910, 0.501426509658279, 0.0583931378631895},
                                  {0.249286745170910, 0.249286745170910, 0.0583931378631895},
                                  {0.636502499121399, 0.310352451033785, 0.041425537809187},
                                  {0.636502499121399, 0.053145049844816, 0.041425537809187},
                                  {0.310352451033785, 0.636502499121399, 0.041425537809187},
                                  {0.310352451033785, 0.053145049844816, 0.041425537809187},
                                  {0.053145049844816, 0.636502499121399, 0.041425537809187},
                                  {0.053145049844816, 0.310352451033785, 0.041425537809187}};
  static const double t13[][3] = {{1.0 / 3, 1.0 / 3, -0.074785022233835},
                                  {0.479308067841923, 0.260345966079038, 0.087807628716602},
                                  {0.260345966079038, 0.479308067841923, 0.087807628716602},
                                  {0.260345966079038, 0.260345966079038, 0.087807628716602},
                                  {0.869739794195568, 0.065130102902216, 0.0266736178044195},
                                  {0.065130102902216, 0.869739794195568, 0.0266736178044195},
                                  {0.065130102902216, 0.065130102902216, 0.0266736178044195},
                                  {0.638444188569809, 0.312865496004875, 0.0385568804451285},
                                  {0.638444188569809, 0.048690315425316, 0.0385568804451285},
                                  {0.312865496004875, 0.638444188569809, 0.0385568804451285},
                                  {0.312865496004875, 0.048690315425316, 0.0385568804451285},
                                  {0.048690315425316, 0.638444188569809, 0.0385568804451285},
                                  {0.048690315425316, 0.312865496004875, 0.0385568804451285}};
  const double(*t)[3] = nullptr;
  int q = 0;
  switch (n) {
    case 1: t = t1; q = 1; break;
    case 3: t = t3; q = 3; break;
    case 4: t = t4; q = 4; break;
    case 7: t = t7; q = 7; break;
    case 12: t = t12; q = 12; break;
    case 13: t = t13; q = 13; break;
    default: return false;
  }
  u.clear();
  v.clear();
  w.clear();
  for (int i = 0; i < q; ++i) {
    u.push_back(t[i][0]);
    v.push_back(t[i][1]);
    w.push_back(t[i][2]);
  }
  return true;
}

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

// Local (u, v) of node c (0-based) of a discontinuous triangular physics kind
// (get_nodal_nodes_u/_v, SurfaceFunction.jl:429-438): the centroid of the constant element;
// the corners moved inwards by beta, then (quadratic) the edge nodes 1-2, 2-3, 3-1.
void discontinuous_node(int pkind, double beta, int c, double* u, double* v) {
  const double b = beta, h = (1.0 - b) / 2.0;
  const double lu[6] = {b, 1.0 - 2.0 * b, b, h, h, b};
  const double lv[6] = {b, b, 1.0 - 2.0 * b, b, h, h};
  if (pkind == BIE_DTRI1) {
    *u = 1.0 / 3.0;
    *v = 1.0 / 3.0;
    return;
  }
  BIE_REQUIRE((pkind == BIE_DTRI3 && c >= 0 && c < 3) || (pkind == BIE_DTRI6 && c >= 0 && c < 6),
              BIE_EINVAL, "internal: discontinuous node out of range");
  *u = lu[c];
  *v = lv[c];
}

// create_polar_elements(physics_function, n, c) for the node (pu, pv): polar_gaussian(n)
// (n rounded up to even radii and as many angles; Gauss-Legendre angles on [0, pi/2], for
// each angle Gauss-Legendre radii on [0, 1 / (cos + sin)], weight w_angle w_radius r) on
// the reference triangle, whose vertex (0, 0) carries the singularity, mapped by the linear
// basis [1 - X - Y, X, Y] onto the sub-triangles (P, (0,1), (0,0)), (P, (1,0), (0,1)) and
// (P, (0,0), (1,0)), weights times |det| of each map; points of the three in that order.
void polar_rule(double pu, double pv, int n, std::vector<double>& u, std::vector<double>& v,
                std::vector<double>& w) {
  BIE_REQUIRE(n >= 1 && n <= 400, BIE_EINVAL, "Ngauss must be between 1 and 400");
  const int nr = ((n + 1) / 2) * 2, na = nr;
  std::vector<double> ga, wa, gr, wr;
  gauss_legendre_pm1(na, ga, wa);
  gauss_legendre_pm1(nr, gr, wr);
  std::vector<double> X, Y, W;
  const double a0 = 0.0, a1 = kPi / 2.0;
  for (int i = 0; i < na; ++i) {
    const double th = (a0 + a1) / 2.0 + (a1 - a0) / 2.0 * ga[i];
    const double thw = (a1 - a0) / 2.0 * wa[i];
    const double rmax = 1.0 / (std::cos(th) + std::sin(th));
    for (int j = 0; j < nr; ++j) {
      const double rad = (0.0 + rmax) / 2.0 + (rmax - 0.0) / 2.0 * gr[j];
      const double radw = (rmax - 0.0) / 2.0 * wr[j] * rad;
      X.push_back(rad * std::cos(th));
      Y.push_back(rad * std::sin(th));
      W.push_back(thw * radw);
    }
  }
  const double A[3][2] = {{0.0, 1.0}, {1.0, 0.0}, {0.0, 0.0}};
  const double B[3][2] = {{0.0, 0.0}, {0.0, 1.0}, {1.0, 0.0}};
  u.clear();
  v.clear();
  w.clear();
  for (int t = 0; t < 3; ++t) {
    const double du0 = A[t][0] - pu, du1 = A[t][1] - pv;  // T * d[1-X-Y, X, Y]/dX
    const double dv0 = B[t][0] - pu, dv1 = B[t][1] - pv;  // T * d[...]/dY
    const double jac = std::fabs(du0 * dv1 - du1 * dv0);
    for (size_t q = 0; q < X.size(); ++q) {
      const double l1 = (1.0 - X[q]) - Y[q];
      u.push_back(pu * l1 + A[t][0] * X[q] + B[t][0] * Y[q]);
      v.push_back(pv * l1 + A[t][1] * X[q] + B[t][1] * Y[q]);
      w.push_back(jac * W[q]);
    }
  }
}

// connected_sources for discontinuous physics: row r is node r mod NP of element r / NP
// (inverse_rle of the element lists, NP rows each).  Depth 1: the element itself and its
// own physics nodes; depth 2: every element sharing a corner (the first three geometry
// nodes) with it, and their physics nodes; lists sorted and unique.
void sparse_pattern_disc(const int* topo, int NG, const int* ptopo, int NP, int64_t E,
                         int64_t N, int depth, std::vector<int64_t>& row_ptr,
                         std::vector<int>& cols, std::vector<int64_t>& pair_ptr,
                         std::vector<int>& pair_elem) {
  BIE_REQUIRE(depth == 1 || depth == 2, BIE_EINVAL, "depth must be 1 or 2");
  BIE_REQUIRE(N == E * (int64_t)NP, BIE_EINVAL,
              "discontinuous near field: one row per element node (n_sources = "
              "n_elements * nodes per element)");
  auto uniq = [](std::vector<int>& v) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  };
  std::vector<std::vector<int>> el(E), sc(E);
  if (depth == 1) {
    for (int64_t e = 0; e < E; ++e) {
      el[e].push_back((int)e);
      sc[e].assign(ptopo + e * NP, ptopo + (e + 1) * NP);
      uniq(sc[e]);
    }
  } else {
    int nc = 0;
    for (int64_t e = 0; e < E; ++e)
      for (int g = 0; g < 3; ++g) nc = std::max(nc, topo[e * NG + g] + 1);
    std::vector<std::vector<int>> corner(nc);
    for (int64_t e = 0; e < E; ++e)
      for (int g = 0; g < 3; ++g) corner[topo[e * NG + g]].push_back((int)e);
    for (int64_t e = 0; e < E; ++e) {
      for (int g = 0; g < 3; ++g) {
        const std::vector<int>& c = corner[topo[e * NG + g]];
        el[e].insert(el[e].end(), c.begin(), c.end());
      }
      uniq(el[e]);
      for (int f : el[e]) {
        const int* p = ptopo + (int64_t)f * NP;
        sc[e].insert(sc[e].end(), p, p + NP);
      }
      uniq(sc[e]);
    }
  }
  row_ptr.assign(1, 0);
  pair_ptr.assign(1, 0);
  cols.clear();
  pair_elem.clear();
  for (int64_t r = 0; r < N; ++r) {
    const int64_t e = r / NP;
    cols.insert(cols.end(), sc[e].begin(), sc[e].end());
    row_ptr.push_back((int64_t)cols.size());
    pair_elem.insert(pair_elem.end(), el[e].begin(), el[e].end());
    pair_ptr.push_back((int64_t)pair_elem.size());
  }
}

}  // namespace bie
