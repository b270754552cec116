// TEST INFRASTRUCTURE: exercises the library's host-side C++ (slot scheduler, rule
// tables, sparse pattern and subdivision rules) and the CPU oracle under AddressSanitizer
// and UndefinedBehaviorSanitizer, on meshes written by tests/test_sanitize_cpu.py.
//
// usage: driver <mesh.bin>...   (each file: int32 E, N, gk, pk, Nc; then double coords
// [3 Nc], int32 topo [gk E], double sources [3 N], int32 ptopo [pk E], 0-based)
//
// Besides running every routine on the meshes, it checks that the library's subdivision
// rules are bit-identical to the oracle's restatement (both are evaluated without fused
// multiply-adds) and that the library's sparse pattern equals the oracle's
// connected_sources.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "bie_common.hpp"
#include "schedule.hpp"

extern "C" {
int orc_subdivision_rule(const double* src, double tol, double* u, double* v, double* w, int cap);
int orc_basis(int kind, int npts, const double* u, const double* v, double* N, double* du,
              double* dv);
int orc_connected_sources(const int64_t* ptopo, int pk, int64_t E, int64_t N, int base,
                          int depth, int64_t* row_ptr, int64_t* cols, int64_t* el_ptr,
                          int64_t* els);
int orc_sparse_assemble(const double* coords, const int64_t* topo, int gk, const double* sources,
                        const int64_t* ptopo, int pk, int64_t E, int64_t N, int base, double kr,
                        double ki, int depth, double offr, int fg, int variant, double* F,
                        double* G);
int orc_select(const double* sources, int64_t N, const int64_t* ptopo, int pk, int gk, int64_t E,
               int base, const double* targets, int64_t M, int8_t* out);
int orc_assemble(const double* coords, const int64_t* topo, int gk, const double* sources,
                 int64_t N, const int64_t* ptopo, int pk, int64_t E, int base, double kr,
                 double ki, const double* targets, int64_t M, int n, int m, int flags, double* F,
                 double* G, double* C, int nthreads);
}

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAIL: %s\n", what);
    ++failures;
  }
}

struct Mesh {
  int E = 0, N = 0, gk = 0, pk = 0, Nc = 0;
  std::vector<double> coords, sources;
  std::vector<int> topo, ptopo;
};

bool read_mesh(const char* path, Mesh& m) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  int hdr[5];
  bool ok = std::fread(hdr, sizeof(int), 5, f) == 5;
  if (ok) {
    m.E = hdr[0], m.N = hdr[1], m.gk = hdr[2], m.pk = hdr[3], m.Nc = hdr[4];
    m.coords.resize(3 * (size_t)m.Nc);
    m.topo.resize((size_t)m.gk * m.E);
    m.sources.resize(3 * (size_t)m.N);
    m.ptopo.resize((size_t)m.pk * m.E);
    ok = std::fread(m.coords.data(), 8, m.coords.size(), f) == m.coords.size() &&
         std::fread(m.topo.data(), 4, m.topo.size(), f) == m.topo.size() &&
         std::fread(m.sources.data(), 8, m.sources.size(), f) == m.sources.size() &&
         std::fread(m.ptopo.data(), 4, m.ptopo.size(), f) == m.ptopo.size();
  }
  std::fclose(f);
  return ok;
}

void check_rules() {
  const int pairs[][2] = {{3, BIE_TRI3},      {3, BIE_DTRI1},     {3, BIE_DTRI3},
                          {3, BIE_DTRI6},     {6, BIE_TRI3},      {6, BIE_TRI6},
                          {6, BIE_DTRI1},     {6, BIE_DTRI3},     {6, BIE_DTRI6},
                          {BIE_QUAD4, BIE_QUAD4}, {BIE_QUAD4S, BIE_QUAD4S},
                          {BIE_QUAD8, BIE_QUAD8}, {BIE_QUAD9, BIE_QUAD9},
                          {BIE_QUAD9, BIE_QUAD4}, {BIE_QUAD4, BIE_DQUAD1},
                          {BIE_QUAD4, BIE_DQUAD4}, {BIE_QUAD9, BIE_DQUAD9}};
  for (const auto& pr : pairs) {
    const int gk = pr[0], pkind = pr[1];
    {
      const int pk = bie::physics_nodes(pkind);
      const bool quad = bie::kind_is_quad(gk);
      for (int n : {1, 2, 3, 4, 6})
        for (int m : {1, 2, 3, 5}) {
          if (gk == BIE_QUAD9 && n % 2 == 1 && m % 2 == 1) {  // refused like the reference
            bool threw = false;
            try {
              bie::build_ruleset(gk, pkind, 0.2113, n, m);
            } catch (const bie::Error&) {
              threw = true;
            }
            expect(threw, "odd n and m on 9-node quadrilaterals are refused");
            continue;
          }
          bie::RuleSet rs = bie::build_ruleset(gk, pkind, 0.2113, n, m);
          expect(rs.R == (quad ? 1 : gk == 6 ? 6 : 3), "rule count");
          expect((int)rs.phi.size() == rs.R * rs.Qmax * pk, "phi size");
          for (int t = 0; t < rs.R * rs.Qmax; ++t) {  // every basis sums to one
            double s = 0;
            for (int j = 0; j < pk; ++j) s += rs.phi[(size_t)t * pk + j];
            const int c = t / rs.Qmax, q = t % rs.Qmax;
            if (q < rs.len[c]) expect(std::fabs(s - 1) < 1e-13, "physics partition of unity");
          }
          for (int c = 0; c < rs.R; ++c) {
            expect(rs.off[c] + rs.len[c] <= rs.Pe, "rule offsets");
            double wsum = 0;
            for (int q = 0; q < rs.len[c]; ++q) wsum += rs.pts[(size_t)(rs.off[c] + q) * bie::kPtStride + 2];
            expect(std::fabs(wsum - (quad ? 4.0 : 0.5)) < 1e-13, "rule weights sum to the area");
          }
        }
    }
  }
  double N[9], du[9], dv[9];
  for (int kind : {3, 6}) {
    bie::geometry_basis(kind, 0.2, 0.3, N, du, dv);
    double s = 0;
    for (int i = 0; i < kind; ++i) s += N[i];
    expect(std::fabs(s - 1) < 1e-15, "partition of unity");
  }
}

void check_subdivision() {
  const double srcs[8][3] = {{1, 0, 0},  {0, 1, 0},     {0, 0, 1},     {0.5, 0.5, 0},
                             {0, 0.5, 0.5}, {0.5, 0, 0.5}, {1, 0.1, 0.1}, {0.5, 0.5, 0.15}};
  for (const auto& s : srcs) {
    std::vector<double> u, v, w;
    bie::subdivision_rule(s, u, v, w);
    const int Q = orc_subdivision_rule(s, 1e-6, nullptr, nullptr, nullptr, 0);
    std::vector<double> uo(Q), vo(Q), wo(Q);
    orc_subdivision_rule(s, 1e-6, uo.data(), vo.data(), wo.data(), Q);
    expect((int)u.size() == Q, "subdivision rule size matches the oracle");
    if ((int)u.size() == Q)
      expect(std::memcmp(u.data(), uo.data(), 8 * Q) == 0 &&
                 std::memcmp(v.data(), vo.data(), 8 * Q) == 0 &&
                 std::memcmp(w.data(), wo.data(), 8 * Q) == 0,
             "subdivision rule bit-identical to the oracle");
    // the sparse kernel's records hold the library's basis values at these points; the
    // kernel reproduces the oracle's geometry bit for bit only if they are identical
    for (int kind : {3, 6}) {
      std::vector<double> No(6 * Q), duo(6 * Q), dvo(6 * Q);
      orc_basis(kind, Q, uo.data(), vo.data(), No.data(), duo.data(), dvo.data());
      bool same = true;
      for (int q = 0; q < Q && same; ++q) {
        double N[9], du[9], dv[9];
        bie::geometry_basis(kind, uo[q], vo[q], N, du, dv);
        same = std::memcmp(N, &No[kind * q], 8 * kind) == 0 &&
               std::memcmp(du, &duo[kind * q], 8 * kind) == 0 &&
               std::memcmp(dv, &dvo[kind * q], 8 * kind) == 0;
      }
      expect(same, "library basis bit-identical to the oracle's at the subdivision points");
    }
  }
}

void check_mesh(const Mesh& m) {
  // slot schedules over a range of step widths, slot budgets and segment counts
  std::vector<double> cent(3 * (size_t)m.E);
  for (int e = 0; e < m.E; ++e)
    for (int d = 0; d < 3; ++d) {
      double a = 0;
      for (int k = 0; k < 3; ++k) a += m.coords[3 * (size_t)m.topo[(size_t)e * m.gk + k] + d];
      cent[3 * (size_t)e + d] = a / 3.0;
    }
  for (int W : {8, 16})
    for (int S : {3 * m.pk + 2, 24, 64})
      for (int segs : {1, 3, 17}) {
        if (S < 2 * m.pk) continue;
        bie::SlotSchedule sc = bie::build_slot_schedule(m.ptopo.data(), m.pk, m.E, m.N,
                                                        cent.data(), W, S, segs);
        bie::validate_slot_schedule(sc, m.ptopo.data(), m.pk, m.E, m.N);
        expect(sc.max_live <= S, "schedule fits its slots");
      }
  // sparse pattern against the oracle's connected_sources
  std::vector<int64_t> pt64(m.ptopo.begin(), m.ptopo.end()), tp64(m.topo.begin(), m.topo.end());
  for (int depth : {1, 2}) {
    std::vector<int64_t> row_ptr, pair_ptr;
    std::vector<int> cols, pair_elem;
    bie::sparse_pattern(m.ptopo.data(), m.pk, m.E, m.N, depth, row_ptr, cols, pair_ptr, pair_elem);
    std::vector<int64_t> rp(m.N + 1), ep(m.N + 1);
    orc_connected_sources(pt64.data(), m.pk, m.E, m.N, 0, depth, rp.data(), nullptr, ep.data(),
                          nullptr);
    std::vector<int64_t> oc(rp[m.N]), oe(ep[m.N]);
    orc_connected_sources(pt64.data(), m.pk, m.E, m.N, 0, depth, rp.data(), oc.data(), ep.data(),
                          oe.data());
    bool same = (int64_t)cols.size() == rp[m.N] && (int64_t)pair_elem.size() == ep[m.N];
    for (size_t t = 0; same && t < cols.size(); ++t) same = cols[t] == oc[t];
    for (size_t t = 0; same && t < pair_elem.size(); ++t) same = pair_elem[t] == oe[t];
    for (int i = 0; same && i <= m.N; ++i) same = row_ptr[i] == rp[i] && pair_ptr[i] == ep[i];
    expect(same, "sparse pattern equals the oracle's connected_sources");
    std::vector<double> F(2 * (size_t)rp[m.N]), G(F.size());
    expect(orc_sparse_assemble(m.coords.data(), tp64.data(), m.gk, m.sources.data(), pt64.data(),
                               m.pk, m.E, m.N, 0, 1.5, 0.2, depth, 0.0, 3, 0, F.data(),
                               G.data()) == 0,
           "oracle sparse assembly");
  }
  // the discontinuous pattern on the element-wise renumbering (three nodes per element):
  // row r belongs to element r / 3; depth 1 its own nodes, depth 2 a superset of them
  {
    const int np = 3;
    std::vector<int> dtopo((size_t)m.E * np);
    for (int64_t t = 0; t < (int64_t)dtopo.size(); ++t) dtopo[t] = (int)t;
    for (int depth : {1, 2}) {
      std::vector<int64_t> row_ptr, pair_ptr;
      std::vector<int> cols, pair_elem;
      bie::sparse_pattern_disc(m.topo.data(), m.gk, dtopo.data(), np, m.E, m.E * np, depth,
                               row_ptr, cols, pair_ptr, pair_elem);
      bool ok = (int64_t)row_ptr.size() == m.E * np + 1;
      for (int64_t r = 0; ok && r < m.E * np; ++r) {
        const int64_t e = r / np;
        int own = 0;
        for (int64_t t = row_ptr[r]; t < row_ptr[r + 1]; ++t) {
          own += cols[t] / np == e;
          ok = ok && (t == row_ptr[r] || cols[t - 1] < cols[t]);
        }
        ok = ok && own == np && (depth == 2 || row_ptr[r + 1] - row_ptr[r] == np);
      }
      expect(ok, "discontinuous sparse pattern");
    }
  }
  // oracle selection and a dense assembly on a few targets
  const int64_t M = m.N < 16 ? m.N : 16;
  std::vector<int8_t> rules((size_t)M * m.E);
  orc_select(m.sources.data(), m.N, pt64.data(), m.pk, m.gk, m.E, 0, m.sources.data(), M,
             rules.data());
  std::vector<double> F(2 * (size_t)M * m.N), G(F.size()), C(2 * (size_t)M);
  expect(orc_assemble(m.coords.data(), tp64.data(), m.gk, m.sources.data(), m.N, pt64.data(),
                      m.pk, m.E, 0, 2.0, 0.0, m.sources.data(), M, 3, 2, 7, F.data(), G.data(),
                      C.data(), 1) == 0,
         "oracle dense assembly");
}

}  // namespace

// The polar rules of discontinuous physics tile the reference triangle (weights sum to 1/2)
// for every node of the three orders.
void check_polar() {
  for (int kind : {BIE_DTRI1, BIE_DTRI3, BIE_DTRI6})
    for (int c = 0; c < (kind & 0xff); ++c) {
      double pu, pv;
      bie::discontinuous_node(kind, 0.2, c, &pu, &pv);
      std::vector<double> u, v, w;
      bie::polar_rule(pu, pv, 39, u, v, w);  // 40 radii and angles
      double sum = 0.0;
      for (double x : w) sum += x;
      expect(u.size() == 3 * 1600 && std::fabs(sum - 0.5) < 1e-13,
             "polar rule tiles the triangle");
    }
}

int main(int argc, char** argv) {
  check_rules();
  check_subdivision();
  check_polar();
  for (int a = 1; a < argc; ++a) {
    Mesh m;
    if (!read_mesh(argv[a], m)) {
      std::fprintf(stderr, "cannot read %s\n", argv[a]);
      return 2;
    }
    check_mesh(m);
  }
  if (failures) return 1;
  std::printf("sanitized host checks passed\n");
  return 0;
}
