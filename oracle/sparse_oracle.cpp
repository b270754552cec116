// TEST INFRASTRUCTURE — part of the CPU oracle.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it, and only to compare results with.
//
// The reference's sparse near-field assembly restated on the CPU:
//   connected_sources            src/3d/assembly_collocation_losses.jl:98-132
//   subdivide_triangle           src/3d/adaptive_integration.jl:172-213
//   singular_triangle_integration src/3d/adaptive_integration.jl:215-246
//   gauss_points_triangle(3)     src/3d/adaptive_integration.jl:13-19
//   sparse_assemble_parallel!    src/3d/assembly_collocation_losses.jl:200-405
//   sparse_computing_integrals!  src/3d/assembly_collocation_losses.jl:158-186
//   maximum_side_length, minimum_side_length, minimum_distance, compute_center
//                                src/3d/adaptive_integration.jl:73-104
// Compiled with -ffp-contract=off: the subdivision decisions are comparisons of rounded
// distances and must reach the reference's sub-triangles.
//
// Parity status: the subdivision rule is pinned by its moments (weights sum to the
// reference triangle's area, linear functions integrate exactly) and by the closed form
// of the 1/r integral from a vertex; the pattern by an incidence-matrix formulation; the
// assembled values by the dense oracle on the same pairs and by the closed form of the
// single-layer integral of a flat triangle from a vertex.  Julia itself is not available,
// so agreement with the reference's own rounding of the subdivision tests is pinned only
// through those checks.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

extern "C" double orc_julia_hypot3(double x, double y, double z);
extern "C" int orc_find_closest_variant(const double* x, const double* nodes, int n_nodes,
                                        int variant);
extern "C" int orc_basis(int kind, int npts, const double* u, const double* v, double* N,
                         double* du, double* dv);

namespace {

struct Pt {
  double u, v;
};
struct Tri {
  Pt p1, p2, p3;
};

// Column j of element_coordinates * shape_function(u, v) with element_coordinates the
// 3 x 3 identity and shape_function the beta = 0 discontinuous linear triangle
// (basis [1 - u - v, u, v]).
void ref_point(Pt p, double out[3]) {
  out[0] = (1.0 - p.u) - p.v;
  out[1] = p.u;
  out[2] = p.v;
}

// sqrt(sum((a - b).^2)) over three rows, summed top to bottom.
double norm3(const double a[3], const double b[3]) {
  double s = (a[0] - b[0]) * (a[0] - b[0]);
  s = s + (a[1] - b[1]) * (a[1] - b[1]);
  s = s + (a[2] - b[2]) * (a[2] - b[2]);
  return std::sqrt(s);
}

// maximum_side_length of the three columns c1, c2, c3 (cycle c1 c2 c3 c1, diff = next -
// current).
double max_side(const double c1[3], const double c2[3], const double c3[3]) {
  const double s1 = norm3(c2, c1), s2 = norm3(c3, c2), s3 = norm3(c1, c3);
  double m = s1;
  if (s2 > m) m = s2;
  if (s3 > m) m = s3;
  return m;
}

// minimum_side_length of 2 x 3 parameter coordinates.
double min_side_2d(Pt a, Pt b, Pt c) {
  auto len = [](Pt p, Pt q) {
    double s = (p.u - q.u) * (p.u - q.u);
    s = s + (p.v - q.v) * (p.v - q.v);
    return std::sqrt(s);
  };
  const double s1 = len(b, a), s2 = len(c, b), s3 = len(a, c);
  double m = s1;
  if (s2 < m) m = s2;
  if (s3 < m) m = s3;
  return m;
}

void subdivide(const std::vector<Tri>& divs_in, const double src[3], double element_size,
               double tol, std::vector<Tri>& divs_out) {
  for (const Tri& t : divs_in) {
    const Pt p12 = {(t.p1.u + t.p2.u) / 2, (t.p1.v + t.p2.v) / 2};
    const Pt p23 = {(t.p2.u + t.p3.u) / 2, (t.p2.v + t.p3.v) / 2};
    const Pt p31 = {(t.p3.u + t.p1.u) / 2, (t.p3.v + t.p1.v) / 2};
    const Pt center = {((t.p1.u + t.p2.u) + t.p3.u) / 3.0, ((t.p1.v + t.p2.v) + t.p3.v) / 3.0};
    double sub[4][3];
    ref_point(center, sub[0]);
    ref_point(t.p1, sub[1]);
    ref_point(t.p2, sub[2]);
    ref_point(t.p3, sub[3]);
    double dist = norm3(sub[0], src);
    for (int j = 1; j < 4; ++j) {
      const double d = norm3(sub[j], src);
      if (d < dist) dist = d;
    }
    double size = max_side(sub[1], sub[2], sub[3]);
    if (size < element_size * tol) size = 0;
    if (size > dist / 4) {
      // the four children in the reference's column order
      const std::vector<Tri> kids = {{t.p1, p12, p31}, {p12, t.p2, p23}, {p31, p23, t.p3},
                                     {p12, p23, p31}};
      subdivide(kids, src, element_size, tol, divs_out);
    } else {
      divs_out.push_back(t);
    }
  }
}

std::vector<double> rule_points(const double src[3], double tol) {
  const double c1[3] = {1, 0, 0}, c2[3] = {0, 1, 0}, c3[3] = {0, 0, 1};
  const double element_size = max_side(c1, c2, c3);
  std::vector<Tri> out;
  subdivide({{{0.0, 0.0}, {1.0, 0.0}, {0.0, 1.0}}}, src, element_size, tol, out);
  // gauss_points_triangle(3): rows (x, y, w)
  const double gx[3] = {1.0 / 6, 2.0 / 3, 1.0 / 6};
  const double gy[3] = {1.0 / 6, 1.0 / 6, 2.0 / 3};
  const double gw[3] = {1.0 / 6, 1.0 / 6, 1.0 / 6};
  const double a_third = 1.0 / 3.0;
  std::vector<double> uvw;
  for (const Tri& t : out) {
    const double dim = min_side_2d(t.p1, t.p2, t.p3);
    const double cu = ((t.p1.u + t.p2.u) + t.p3.u) / 3;
    const double cv = ((t.p1.v + t.p2.v) + t.p3.v) / 3;
    for (int g = 0; g < 3; ++g) {
      uvw.push_back((gx[g] - a_third) * dim + cu);
      uvw.push_back((gy[g] - a_third) * dim + cv);
      uvw.push_back(gw[g] * (dim * dim));
    }
  }
  return uvw;
}

// connected_sources: per source, sorted unique element and source lists.
void connections(const int64_t* ptopo, int pk, int64_t E, int64_t N, int base, int depth,
                 std::vector<std::vector<int64_t>>& el, std::vector<std::vector<int64_t>>& sc) {
  el.assign(N, {});
  sc.assign(N, {});
  for (int64_t e = 0; e < E; ++e)
    for (int i = 0; i < pk; ++i) {
      const int64_t node = ptopo[i + e * pk] - base;
      el[node].push_back(e);
      for (int j = 0; j < pk; ++j) sc[node].push_back(ptopo[j + e * pk] - base);
    }
  if (depth == 2) {
    std::vector<std::vector<int64_t>> el2(N), sc2(N);
    for (int64_t s = 0; s < N; ++s)
      for (int64_t idx : sc[s]) {  // the lists of depth 1 before sort/unique
        el2[s].insert(el2[s].end(), el[idx].begin(), el[idx].end());
        sc2[s].insert(sc2[s].end(), sc[idx].begin(), sc[idx].end());
      }
    el.swap(el2);
    sc.swap(sc2);
  }
  for (int64_t s = 0; s < N; ++s) {
    std::sort(el[s].begin(), el[s].end());
    el[s].erase(std::unique(el[s].begin(), el[s].end()), el[s].end());
    std::sort(sc[s].begin(), sc[s].end());
    sc[s].erase(std::unique(sc[s].begin(), sc[s].end()), sc[s].end());
  }
}

}  // namespace

extern "C" {

// Subdivision rule for the source src (3 values; the reference's [1; offr; offr] etc.),
// TolE tol: writes u, v, w when u is non-NULL and cap is large enough; returns the number
// of points.
int orc_subdivision_rule(const double* src, double tol, double* u, double* v, double* w,
                         int cap) {
  const std::vector<double> uvw = rule_points(src, tol);
  const int Q = (int)(uvw.size() / 3);
  if (u && cap >= Q)
    for (int q = 0; q < Q; ++q) {
      u[q] = uvw[3 * q];
      v[q] = uvw[3 * q + 1];
      w[q] = uvw[3 * q + 2];
    }
  return Q;
}

// connected_sources flattened: row_ptr (N + 1) and cols (nnz, 0-based) of the source
// lists, el_ptr (N + 1) and els (0-based) of the element lists.  With cols / els NULL
// only the counts are written (row_ptr[N], el_ptr[N]).
int orc_connected_sources(const int64_t* ptopo, int pk, int64_t E, int64_t N, int base,
                          int depth, int64_t* row_ptr, int64_t* cols, int64_t* el_ptr,
                          int64_t* els) {
  if (depth != 1 && depth != 2) return 1;
  std::vector<std::vector<int64_t>> el, sc;
  connections(ptopo, pk, E, N, base, depth, el, sc);
  row_ptr[0] = el_ptr[0] = 0;
  for (int64_t s = 0; s < N; ++s) {
    if (cols) std::copy(sc[s].begin(), sc[s].end(), cols + row_ptr[s]);
    if (els) std::copy(el[s].begin(), el[s].end(), els + el_ptr[s]);
    row_ptr[s + 1] = row_ptr[s] + (int64_t)sc[s].size();
    el_ptr[s + 1] = el_ptr[s] + (int64_t)el[s].size();
  }
  return 0;
}

// sparse_assemble_parallel! for continuous triangular physics: F and G (nnz complex,
// interleaved) in the order of the source lists (row by row, columns ascending).
// sources: 3 x N points (one per mesh source), offset: the six rules' offr.
// fg bit 0 = fOn, bit 1 = gOn; variant selects the hypot formula of find_closest_corner.
int orc_sparse_assemble(const double* coords, const int64_t* topo, int gk,
                        const double* sources, const int64_t* ptopo, int pk, int64_t E,
                        int64_t N, int base, double kr, double ki, int depth, double offr,
                        int fg, int variant, double* F, double* G) {
  if (depth != 1 && depth != 2) return 1;
  std::vector<std::vector<int64_t>> el, sc;
  connections(ptopo, pk, E, N, base, depth, el, sc);
  std::vector<int64_t> start(N + 1, 0);
  for (int64_t s = 0; s < N; ++s) start[s + 1] = start[s] + (int64_t)sc[s].size();
  for (int64_t t = 0; t < 2 * start[N]; ++t) F[t] = G[t] = 0.0;
  // the six rules: sources at the vertices, then the edge midpoints (losses.jl:227-232)
  const double srcs[6][3] = {{1.0, offr, offr}, {offr, 1.0, offr}, {offr, offr, 1.0},
                             {0.5, 0.5, offr},  {offr, 0.5, 0.5},  {0.5, offr, 0.5}};
  struct R {
    int Q;
    std::vector<double> w, Ng, Du, Dv, Np;
  } rules[6];
  for (int c = 0; c < 6; ++c) {
    const std::vector<double> uvw = rule_points(srcs[c], 1e-6);
    R& r = rules[c];
    r.Q = (int)(uvw.size() / 3);
    std::vector<double> u(r.Q), v(r.Q);
    r.w.resize(r.Q);
    for (int q = 0; q < r.Q; ++q) {
      u[q] = uvw[3 * q];
      v[q] = uvw[3 * q + 1];
      r.w[q] = uvw[3 * q + 2];
    }
    r.Ng.resize((size_t)r.Q * gk);
    r.Du.resize((size_t)r.Q * gk);
    r.Dv.resize((size_t)r.Q * gk);
    r.Np.resize((size_t)r.Q * pk);
    std::vector<double> d1((size_t)r.Q * pk), d2((size_t)r.Q * pk);
    orc_basis(gk, r.Q, u.data(), v.data(), r.Ng.data(), r.Du.data(), r.Dv.data());
    orc_basis(pk, r.Q, u.data(), v.data(), r.Np.data(), d1.data(), d2.data());
  }
  const double four_pi = 4.0 * M_PI;
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t s = 0; s < N; ++s) {
    const double* x = sources + 3 * s;
    for (int64_t e : el[s]) {
      double X[6][3];
      for (int g = 0; g < gk; ++g) {
        const int64_t node = topo[g + e * gk] - base;
        for (int d = 0; d < 3; ++d) X[g][d] = coords[d + 3 * node];
      }
      int c = orc_find_closest_variant(x, &X[0][0], gk, variant);
      if (c < 1 || c > 5) c = 6;  // the reference's final else branch
      const R& r = rules[c - 1];
      int64_t pos[6];
      for (int j = 0; j < pk; ++j) {
        const int64_t col = ptopo[j + e * pk] - base;
        pos[j] = start[s] + (std::lower_bound(sc[s].begin(), sc[s].end(), col) - sc[s].begin());
      }
      double fre[6] = {0}, fim[6] = {0}, gre[6] = {0}, gim[6] = {0};
      for (int q = 0; q < r.Q; ++q) {
        double y[3], tu[3], tv[3];
        for (int d = 0; d < 3; ++d) {
          double a = 0, b = 0, cc = 0;
          for (int g = 0; g < gk; ++g) {
            a += X[g][d] * r.Ng[(size_t)q * gk + g];
            b += X[g][d] * r.Du[(size_t)q * gk + g];
            cc += X[g][d] * r.Dv[(size_t)q * gk + g];
          }
          y[d] = a;
          tu[d] = b;
          tv[d] = cc;
        }
        const double cr[3] = {tu[1] * tv[2] - tu[2] * tv[1], tu[2] * tv[0] - tu[0] * tv[2],
                              tu[0] * tv[1] - tu[1] * tv[0]};
        const double J = orc_julia_hypot3(cr[0], cr[1], cr[2]);
        const double jw = J * r.w[q];
        const double dx = y[0] - x[0], dy = y[1] - x[1], dz = y[2] - x[2];
        const double rr = orc_julia_hypot3(dx, dy, dz);
        const double ex = std::exp(-ki * rr);
        const double cs = std::cos(kr * rr) * ex, sn = std::sin(kr * rr) * ex;
        if (fg & 1) {  // freens3d!: e^{ikr}(ikr - 1) (n.d) / (4 pi r^3), n = -(tu x tv)/J
          const double nd = (-cr[0] / J) * dx + (-cr[1] / J) * dy + (-cr[2] / J) * dz;
          const double ar = -ki * rr - 1.0, ai = kr * rr;
          const double sc3 = nd / (four_pi * (rr * rr * rr)) * jw;
          const double re = (cs * ar - sn * ai) * sc3, im = (cs * ai + sn * ar) * sc3;
          for (int j = 0; j < pk; ++j) {
            fre[j] += re * r.Np[(size_t)q * pk + j];
            fim[j] += im * r.Np[(size_t)q * pk + j];
          }
        }
        if (fg & 2) {  // greens3d!: e^{ikr} / (4 pi r)
          const double re = cs / (four_pi * rr) * jw, im = sn / (four_pi * rr) * jw;
          for (int j = 0; j < pk; ++j) {
            gre[j] += re * r.Np[(size_t)q * pk + j];
            gim[j] += im * r.Np[(size_t)q * pk + j];
          }
        }
      }
      for (int j = 0; j < pk; ++j) {
        F[2 * pos[j]] += fre[j];
        F[2 * pos[j] + 1] += fim[j];
        G[2 * pos[j]] += gre[j];
        G[2 * pos[j] + 1] += gim[j];
      }
    }
  }
  return 0;
}

}  // extern "C"
