// TEST INFRASTRUCTURE — part of the CPU oracle.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it, and only to compare results with.
//
// Rule selection of the reference's dense collocation assembly, restated in IEEE-754
// double arithmetic with a fixed operation order.  This translation unit is compiled
// with -ffp-contract=off so that no multiply-add is fused: the GPU kernel evaluates the
// same sequence and must reach bit-identical distances, because symmetric meshes put
// many (target, element) pairs on exact ties.
//
// Parity status: the operation order of Julia's n-argument hypot is restated from
// Julia Base (maxabs * sqrt(sum((x_i/maxabs)^2)), summed left to right).  Julia is not
// available here, so agreement with Julia on exact ties is *parity unpinned*; agreement
// between this oracle and the HIP kernel is pinned bit-exactly by the tests.
#include <algorithm>
#include <cmath>
#include <cstdint>

extern "C" {

// Julia Base n-ary hypot for three Float64 arguments.
double orc_julia_hypot3(double x, double y, double z) {
  double ax = std::fabs(x), ay = std::fabs(y), az = std::fabs(z);
  // Julia: an infinite argument wins over NaN; otherwise NaN propagates through max.
  if (std::isinf(ax) || std::isinf(ay) || std::isinf(az)) return INFINITY;
  if (std::isnan(ax) || std::isnan(ay) || std::isnan(az)) return NAN;
  double m = ax > ay ? ax : ay;
  m = m > az ? m : az;
  if (m == 0.0) return m;
  double a = ax / m, b = ay / m, c = az / m;
  double s = a * a;
  s = s + b * b;
  s = s + c * c;
  return m * std::sqrt(s);
}

// The other n-ary hypot formula Julia releases have used for IEEE floats (restated from
// Julia Base's _hypot / scaleinv): the arguments are scaled by the power of two 2^-E with
// E the exponent of maxabs, clamped to the normal range (exact), squared and summed (left
// to right here; that formula lets the compiler reassociate the sum), square-rooted and
// scaled back by 2^E (exact).  Away from overflow and underflow the scalings cancel exactly, so the
// result is sqrt((x^2 + y^2) + z^2) with every operation rounded once.  Only used to
// measure how many rule choices depend on which formula the reference runs with.
double orc_hypot3_scaled(double x, double y, double z) {
  double ax = std::fabs(x), ay = std::fabs(y), az = std::fabs(z);
  if (std::isinf(ax) || std::isinf(ay) || std::isinf(az)) return INFINITY;
  if (std::isnan(ax) || std::isnan(ay) || std::isnan(az)) return NAN;
  double m = ax > ay ? ax : ay;
  m = m > az ? m : az;
  if (m == 0.0) return m;
  // scale = 2^E with E the exponent of m (m in [2^E, 2^(E+1))), clamped to [-1022, 1022]
  // so that neither the scale nor its inverse leaves the normal range (subnormal m)
  int e = 0;
  std::frexp(m, &e);  // m = f 2^e, f in [0.5, 1)
  const int E = std::max(-1022, std::min(1022, e - 1));
  const double inv = std::ldexp(1.0, -E), scale = std::ldexp(1.0, E);
  const double a = ax * inv, b = ay * inv, c = az * inv;
  double s = a * a;
  s = s + b * b;
  s = s + c * c;
  return scale * std::sqrt(s);
}

// find_closest_corner (src/3d/assembly_collocation.jl:102-115): 1-based index of the
// first node with strictly smallest distance; 0 when no distance compares below Inf.
// variant 0: orc_julia_hypot3 distances, 1: orc_hypot3_scaled.
int orc_find_closest_variant(const double* x, const double* nodes, int n_nodes, int variant) {
  double best = INFINITY;
  int idx = 0;
  for (int i = 0; i < n_nodes; ++i) {
    const double dx = nodes[3 * i + 0] - x[0], dy = nodes[3 * i + 1] - x[1],
                 dz = nodes[3 * i + 2] - x[2];
    double d = variant == 1 ? orc_hypot3_scaled(dx, dy, dz) : orc_julia_hypot3(dx, dy, dz);
    if (d < best) {
      best = d;
      idx = i + 1;
    }
  }
  return idx;
}

int orc_find_closest(const double* x, const double* nodes, int n_nodes) {
  return orc_find_closest_variant(x, nodes, n_nodes, 0);
}

}  // extern "C"
