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

// find_closest_corner (src/3d/assembly_collocation.jl:102-115): 1-based index of the
// first node with strictly smallest distance; 0 when no distance compares below Inf.
int orc_find_closest(const double* x, const double* nodes, int n_nodes) {
  double best = INFINITY;
  int idx = 0;
  for (int i = 0; i < n_nodes; ++i) {
    double d = orc_julia_hypot3(nodes[3 * i + 0] - x[0], nodes[3 * i + 1] - x[1],
                                nodes[3 * i + 2] - x[2]);
    if (d < best) {
      best = d;
      idx = i + 1;
    }
  }
  return idx;
}

}  // extern "C"
