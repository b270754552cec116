// C-ABI entry points of libbie_amd (declared in include/bie_amd.h).
//
// Ownership follows the drop-in contract: callers own F, G, C (host or device); the
// library owns only device mesh copies and rule-dependent tables inside a bie_dmesh,
// released by bie_dmesh_destroy.  Errors become status codes plus a thread-local
// message, never exceptions across the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "bie_common.hpp"
#include "schedule.hpp"

namespace bie {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

namespace {

template <class F>
int guarded(F&& f) {
  try {
    f();
    g_last_error.clear();
    return BIE_OK;
  } catch (const Error& e) {
    g_last_error = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return BIE_ENOMEM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return BIE_EINVAL;
  }
}

template <class T>
T* dalloc(size_t count) {
  if (count == 0) count = 1;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, count * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error{BIE_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e)};
  }
  return static_cast<T*>(p);
}

void dfree(void* p) {
  if (p) (void)hipFree(p);
}

void validate_mesh(const bie_mesh* m) {
  BIE_REQUIRE(m != nullptr, BIE_EINVAL, "mesh is NULL");
  const int gkd = m->geometry_kind, pkd = m->physics_kind;
  BIE_REQUIRE(gkd == BIE_TRI3 || gkd == BIE_TRI6 || gkd == BIE_QUAD4 || gkd == BIE_QUAD4S ||
                  gkd == BIE_QUAD8 || gkd == BIE_QUAD9,
              BIE_EUNSUPPORTED,
              "geometry_kind must be BIE_TRI3, BIE_TRI6, BIE_QUAD4, BIE_QUAD4S, BIE_QUAD8 or "
              "BIE_QUAD9");
  if (kind_is_quad(gkd)) {
    BIE_REQUIRE(pkd == gkd || (gkd == BIE_QUAD9 && pkd == BIE_QUAD4) || pkd == BIE_DQUAD1 ||
                    pkd == BIE_DQUAD4 || pkd == BIE_DQUAD9,
                BIE_EUNSUPPORTED,
                "physics on quadrilateral geometry must be the geometry kind, BIE_QUAD4 on "
                "BIE_QUAD9, or BIE_DQUAD1 / BIE_DQUAD4 / BIE_DQUAD9");
    if (pkd == BIE_DQUAD4 || pkd == BIE_DQUAD9)
      BIE_REQUIRE(std::isfinite(m->physics_beta) && 1.0 - m->physics_beta != 0.0, BIE_EINVAL,
                  "physics_beta must be finite and differ from 1");
  } else {
    BIE_REQUIRE(pkd == BIE_TRI3 || pkd == BIE_TRI6 || pkd == BIE_DTRI1 || pkd == BIE_DTRI3 ||
                    pkd == BIE_DTRI6,
                BIE_EUNSUPPORTED,
                "physics on triangles must be BIE_TRI3, BIE_TRI6, BIE_DTRI1, BIE_DTRI3 or "
                "BIE_DTRI6");
    BIE_REQUIRE(!(gkd == BIE_TRI3 && pkd == BIE_TRI6), BIE_EUNSUPPORTED,
                "quadratic physics on linear geometry is not a reference configuration");
    if (pkd == BIE_DTRI3 || pkd == BIE_DTRI6)
      BIE_REQUIRE(std::isfinite(m->physics_beta) && 1.0 - 3.0 * m->physics_beta != 0.0,
                  BIE_EINVAL, "physics_beta must be finite and differ from 1/3");
  }
  BIE_REQUIRE(m->index_base == 0 || m->index_base == 1, BIE_EINVAL, "index_base must be 0 or 1");
  BIE_REQUIRE((m->flags & ~BIE_MESH_SCALED_HYPOT) == 0, BIE_EINVAL, "unknown mesh flag bits");
  BIE_REQUIRE(m->n_elements >= 0 && m->n_coordinates >= 0 && m->n_sources >= 0, BIE_EINVAL,
              "negative mesh size");
  BIE_REQUIRE(m->n_elements <= (1LL << 28) && m->n_coordinates < (1LL << 31) &&
                  m->n_sources < (1LL << 31),
              BIE_EINVAL, "mesh too large for 32-bit element/node indices");
  if (m->n_elements > 0) {
    BIE_REQUIRE(m->coordinates && m->topology && m->sources && m->physics_topology, BIE_EINVAL,
                "mesh array pointer is NULL");
  }
}

std::vector<int> convert_topology(const int64_t* t, int rows, int64_t E, int64_t base,
                                  int64_t limit, const char* what) {
  std::vector<int> out((size_t)rows * E);
  for (int64_t e = 0; e < E; ++e)
    for (int r = 0; r < rows; ++r) {
      const int64_t v = t[r + e * rows] - base;
      if (v < 0 || v >= limit)
        throw Error{BIE_EINVAL, std::string(what) + " index out of range at element " +
                                    std::to_string(e + base)};
      out[(size_t)e * rows + r] = (int)v;
    }
  return out;
}

void ensure_rules(bie_dmesh* dm, int n, int m, hipStream_t s) {
  if (dm->cn == n && dm->cm == m && dm->d_elem) return;
  RuleSet rs = build_ruleset(dm->gkind, dm->pkind, dm->beta, n, m);
  if (dm->d_elem) (void)hipDeviceSynchronize();  // in-flight kernels may read the old tables
  dfree(dm->d_pts);
  dfree(dm->d_phi);
  dfree(dm->d_elem);
  dm->d_pts = dm->d_phi = dm->d_elem = nullptr;
  dm->cn = dm->cm = -1;
  dm->geom_valid = false;
  dm->d_pts = dalloc<double>(rs.pts.size());
  // basis table, padded to phi_pad doubles, then the trig table of the slot kernels
  const int nphi = rs.R * rs.Qmax * dm->pk;
  BIE_REQUIRE((size_t)nphi == rs.phi.size(), BIE_EINVAL, "internal: basis table size mismatch");
  std::vector<double> phi_trig((size_t)phi_pad(nphi) + kTrigDoubles, 0.0);
  std::copy(rs.phi.begin(), rs.phi.end(), phi_trig.begin());
  for (int j = 0; j < kTrigN; ++j) {  // (cos, sin)(j pi / 64), rounded from long double
    const long double th = (long double)j * (3.14159265358979323846264338327950288L / 64);
    phi_trig[phi_pad(nphi) + 2 * j] = (double)cosl(th);
    phi_trig[phi_pad(nphi) + 2 * j + 1] = (double)sinl(th);
  }
  dm->d_phi = dalloc<double>(phi_trig.size());
  dm->d_elem = dalloc<double>((size_t)std::max<int64_t>(dm->E, 1) * rs.Pe * 8);
  BIE_HIP_CHECK(hipMemcpyAsync(dm->d_pts, rs.pts.data(), rs.pts.size() * sizeof(double),
                               hipMemcpyHostToDevice, s));
  BIE_HIP_CHECK(hipMemcpyAsync(dm->d_phi, phi_trig.data(), phi_trig.size() * sizeof(double),
                               hipMemcpyHostToDevice, s));
  BIE_HIP_CHECK(hipStreamSynchronize(s));
  dm->rs = std::move(rs);
  dm->cn = n;
  dm->cm = m;
}

// Element geometry tables (y, n, J w / 4 pi at every rule point) of the cached rule set.
void refresh_geometry(bie_dmesh* dm, hipStream_t s) {
  launch_geometry(dm->d_coords, dm->d_topo, dm->gk, (int)dm->E, dm->d_pts, dm->rs.Pe,
                  dm->d_elem, s);
  dm->geom_valid = true;
}

// Event record for bie_timing_*: (call start, dense kernel start, dense kernel end,
// mid-edge kernel end) per call.
struct Timing {
  std::mutex mu;
  bool on = false;
  std::vector<hipEvent_t> pool;
  std::vector<std::array<hipEvent_t, 4>> rec;
  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    BIE_HIP_CHECK(hipEventCreate(&e));
    return e;
  }
};
Timing g_timing;

// Which assembly kernel to use: the LDS-slot kernel (default) or the all-atomics kernel
// (BIE_AMD_KERNEL=atomic), kept for small problems and as a cross-check.
bool use_atomic_kernel() {
  const char* v = std::getenv("BIE_AMD_KERNEL");
  return v && std::strcmp(v, "atomic") == 0;
}

// LDS slots per workgroup of the single-k slot kernel: as many as fit beside the basis
// and trig tables, the staged table rows and the mid-edge queue (one workgroup per CU);
// BIE_AMD_SLOTS=<n> asks for fewer.
int slot_occupancy() {
  const char* v = std::getenv("BIE_AMD_OCC");
  return (v && std::atoi(v) == 2) ? 2 : 1;
}

// Elements each wave takes per step of the single-k slot kernel (1 or 2).
int slot_elems_per_wave() {
  const char* v = std::getenv("BIE_AMD_V");
  return (v && std::atoi(v) == 2) ? 2 : 1;
}

int slot_split() {
  const char* v = std::getenv("BIE_AMD_SPLIT");
  return (v && std::atoi(v) == 2) ? 2 : 1;
}

// Waves per workgroup of the single-k slot kernel on the real-k unrolled instances:
// kWideWaves (three per SIMD) for three-node physics, kSlotWaves for six-node physics,
// whose wide instance needs more registers than three waves per SIMD leave (measured
// slower); BIE_AMD_WAVES=8 or 12 forces either.
int slot_waves(int NP) {
  if (const char* v = std::getenv("BIE_AMD_WAVES")) {
    const int x = std::atoi(v);
    if (x == kWideWaves || x == kSlotWaves) return x;
  }
  return NP == 3 ? kWideWaves : kSlotWaves;
}

int plain_slot_count(int nphi, int Qc, int occ = 1, int split = 1, int waves = kSlotWaves) {
  // LDS of one workgroup when occ of them share a CU (allocation granularity left spare)
  const long budget = occ == 2 ? kLdsBytes / 2 - 1024 : kLdsBytes;
  const int stage = split == 2 ? split_stage_doubles(Qc, 2) : stage_doubles(Qc, waves);
  int S = fitting_slots(phi_pad(nphi) + kTrigDoubles + stage, 1, budget);
  if (const char* v = std::getenv("BIE_AMD_SLOTS")) {
    const int x = std::atoi(v);
    if (x >= 8 && x < S) S = x;
  }
  return S;
}

template <class T>
T* upload_vec(const std::vector<T>& v, hipStream_t s) {
  T* d = dalloc<T>(v.size());
  if (!v.empty())
    BIE_HIP_CHECK(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return d;
}

void free_schedule(bie_dmesh::Sched& sc) {
  // kernels still in flight may read the schedule being replaced
  if (sc.d_step_elems) (void)hipDeviceSynchronize();
  for (int** p : {&sc.d_seg_step_begin, &sc.d_step_elems, &sc.d_step_slot, &sc.d_flush_begin,
                  &sc.d_flush_node, &sc.d_flush_slot, &sc.d_zero_cols}) {
    dfree(*p);
    *p = nullptr;
  }
  dfree(sc.d_step_sel);
  sc.d_step_sel = nullptr;
  sc.key = -1;
  sc.S = -1;
}

void free_schedules(bie_dmesh* dm) {
  for (auto& sc : dm->sched) free_schedule(sc);
}

// Segment count of the single-k slot kernel's schedule: at least 8 tiles (row block x
// segment) per CU, and among counts up to twice that, the one whose tile count fills whole
// rounds of resident workgroups (one per CU) best, so that the last round leaves few CUs
// idle.  (The schedule may split further where a segment's front exceeds the slots.)
int choose_segment_count(int device, int64_t rowblocks, int64_t E, int per_cu = 1) {
  static int cus_of[16] = {0};
  int& cus = cus_of[device & 15];
  if (cus == 0) {
    BIE_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    if (cus <= 0) cus = 1;
  }
  const int64_t round = (int64_t)cus * per_cu;  // workgroups resident at once
  const int64_t hi_e = std::max<int64_t>(1, E);
  const int64_t base = std::max<int64_t>(1, std::min<int64_t>(hi_e, (8LL * round + rowblocks - 1) / rowblocks));
  int64_t best = base;
  double best_waste = 1e300;
  for (int64_t p = base; p <= std::min<int64_t>(hi_e, 2 * base); ++p) {
    const int64_t tiles = rowblocks * p;
    const double waste = (double)((tiles + round - 1) / round * round) / (double)tiles;
    if (waste < best_waste - 1e-12) {
      best_waste = waste;
      best = p;
    }
  }
  return (int)best;
}

// Schedule `which` (0: single-k kernel, 1: sweep kernel, 2: batch kernel) for S live
// slots and at least min_segments segments.
bie_dmesh::Sched& ensure_schedule(bie_dmesh* dm, int which, int S, int min_segments,
                                  hipStream_t s, int wv = kSlotWaves) {
  bie_dmesh::Sched& out = dm->sched[which];
  if (out.key == min_segments && out.S == S && out.wv == wv && out.d_step_elems) return out;
  free_schedule(out);
  std::vector<double> cent((size_t)dm->E * 3);
  for (int64_t e = 0; e < dm->E; ++e)
    for (int d = 0; d < 3; ++d) {
      double acc = 0;
      for (int k = 0; k < 3; ++k) acc += dm->h_coords[3 * (size_t)dm->h_topo[e * dm->gk + k] + d];
      cent[3 * e + d] = acc / 3.0;
    }
  SlotSchedule sc = build_slot_schedule(dm->h_ptopo.data(), dm->pk, (int)dm->E, (int)dm->N,
                                        cent.data(), wv, S, min_segments);
  out.d_seg_step_begin = upload_vec(sc.seg_step_begin, s);
  out.h_seg_step_begin = sc.seg_step_begin;
  out.d_step_elems = upload_vec(sc.step_elems, s);
  // per (step, wave) copies of the selection nodes and slots, in the order the kernel
  // reads them
  const size_t nsw = (size_t)sc.n_steps * wv;
  std::vector<int> step_slot(nsw * dm->pk, 0);
  std::vector<double> step_sel(nsw * dm->pk * 4, 0.0);
  for (size_t t = 0; t < nsw; ++t) {
    const int e = sc.step_elems[t];
    if (e < 0) continue;
    for (int k = 0; k < dm->pk; ++k) {
      step_slot[t * dm->pk + k] = sc.slot_of[(size_t)e * dm->pk + k];
      const int node = dm->h_ptopo[(size_t)e * dm->pk + k];
      for (int d = 0; d < 3; ++d)
        step_sel[(t * dm->pk + k) * 4 + d] = dm->h_sources[3 * (size_t)node + d];
    }
  }
  out.d_step_slot = upload_vec(step_slot, s);
  out.d_step_sel = upload_vec(step_sel, s);
  out.d_flush_begin = upload_vec(sc.flush_begin, s);
  out.d_flush_node = upload_vec(sc.flush_node, s);
  out.d_flush_slot = upload_vec(sc.flush_slot, s);
  out.d_zero_cols = upload_vec(sc.zero_cols, s);
  BIE_HIP_CHECK(hipStreamSynchronize(s));
  out.segments = sc.n_segments;
  out.zero_cols = (int)sc.zero_cols.size();
  out.max_live = sc.max_live;
  out.key = min_segments;
  out.S = S;
  out.wv = wv;
  return out;
}

SlotArgs slot_args(bie_dmesh* dm, const bie_dmesh::Sched& sc, int64_t rowblocks) {
  SlotArgs sa{};
  sa.W = sc.wv;
  sa.S = sc.S;
  sa.n_segments = sc.segments;
  sa.n_rowblocks = (int)rowblocks;
  sa.n_tiles = rowblocks * sc.segments;
  const char* ord = std::getenv("BIE_AMD_TILE_ORDER");
  sa.seg_major = (ord && std::strcmp(ord, "rowblock") == 0) ? 0 : 1;
  sa.phi_in_lds = 1;
  sa.seg_step_begin = sc.d_seg_step_begin;
  sa.step_elems = sc.d_step_elems;
  sa.step_slot = sc.d_step_slot;
  sa.step_sel = sc.d_step_sel;
  sa.flush_begin = sc.d_flush_begin;
  sa.flush_node = sc.d_flush_node;
  sa.flush_slot = sc.d_flush_slot;
  sa.mid_count = dm->d_mid_count;
  sa.mid_queue = dm->d_mid_queue;
  sa.mid_cap = dm->mid_cap;
  sa.push_mid = 1;
  return sa;
}

void ensure_mid_queue(bie_dmesh* dm, int64_t cap) {
  if (!dm->d_mid_count) dm->d_mid_count = dalloc<unsigned>(1);
  if (dm->mid_cap >= cap && dm->d_mid_queue) return;
  dfree(dm->d_mid_queue);
  dm->d_mid_queue = nullptr;
  dm->mid_cap = 0;
  dm->d_mid_queue = dalloc<int2>((size_t)cap);
  dm->mid_cap = cap;
}

// Records the collocation mid-edge pairs that the slot kernel queued: the pairs sorted by
// (element, target), and the distinct (column, row) entries of F and G they add to, sorted
// by (column, row), each with the list of (pair, node) contributions.  Releases the queue.
// Host work once per mesh; synchronises the stream.
void record_mid_pairs(bie_dmesh* dm, hipStream_t s) {
  unsigned count = 0;
  BIE_HIP_CHECK(hipMemcpyAsync(&count, dm->d_mid_count, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  BIE_HIP_CHECK(hipStreamSynchronize(s));
  const int64_t n = std::min<int64_t>((int64_t)count, dm->mid_cap);
  BIE_REQUIRE(n < (int64_t)1 << 28, BIE_EINVAL, "too many mid-edge pairs to record");
  std::vector<int2> q((size_t)n);
  if (n)
    BIE_HIP_CHECK(hipMemcpy(q.data(), dm->d_mid_queue, sizeof(int2) * n, hipMemcpyDeviceToHost));
  std::vector<unsigned long long> keys((size_t)n);
  for (int64_t p = 0; p < n; ++p)
    keys[p] = ((unsigned long long)(q[p].y & 0x0fffffff) << 34) |
              ((unsigned long long)(unsigned)q[p].x << 2) | (unsigned long long)(q[p].y >> 28);
  std::sort(keys.begin(), keys.end());
  std::vector<std::pair<unsigned long long, int>> con((size_t)n * 6);
  for (int64_t p = 0; p < n; ++p) {
    const int e = (int)(keys[p] >> 34);
    const unsigned row = (unsigned)((keys[p] >> 2) & 0xffffffffull);
    for (int k = 0; k < 6; ++k) {
      const unsigned col = (unsigned)dm->h_ptopo[(size_t)e * 6 + k];
      con[(size_t)p * 6 + k] = {((unsigned long long)col << 32) | row, (int)(p * 6 + k)};
    }
  }
  std::sort(con.begin(), con.end());
  std::vector<unsigned long long> entries;
  std::vector<int> off, refs(con.size());
  for (size_t t = 0; t < con.size(); ++t) {
    if (t == 0 || con[t].first != con[t - 1].first) {
      entries.push_back(con[t].first);
      off.push_back((int)t);
    }
    refs[t] = con[t].second;
  }
  off.push_back((int)con.size());
  auto up = [&](auto& v) { return upload_vec(v, s); };
  unsigned long long* d_keys = up(keys);
  unsigned long long* d_entries = up(entries);
  int* d_off = up(off);
  int* d_refs = up(refs);
  double* d_vals = dalloc<double>((size_t)std::max<int64_t>(n, 1) * 24);
  BIE_HIP_CHECK(hipStreamSynchronize(s));
  for (void* p : {(void*)dm->d_mid_keys, (void*)dm->d_mid_entries, (void*)dm->d_mid_off,
                  (void*)dm->d_mid_refs, (void*)dm->d_mid_vals})
    dfree(p);
  dm->d_mid_keys = d_keys;
  dm->d_mid_entries = d_entries;
  dm->d_mid_off = d_off;
  dm->d_mid_refs = d_refs;
  dm->d_mid_vals = d_vals;
  dm->n_mid_keys = n;
  dm->n_mid_entries = (int64_t)entries.size();
  dfree(dm->d_mid_queue);  // the M x E queue is no longer needed for collocation
  dm->d_mid_queue = nullptr;
  dm->mid_cap = 0;
}

AssemblyArgs make_args(bie_dmesh* dm, const double* d_targets, int64_t M, double kr,
                       double ki, uint32_t flags, double* dF, double* dG, double* dC) {
  AssemblyArgs a{};
  a.targets = d_targets;
  a.M = M;
  a.N = dm->N;
  a.E = (int)dm->E;
  a.NP = dm->pk;
  a.elem = dm->d_elem;
  a.Pe = dm->rs.Pe;
  a.selnodes = dm->d_selnodes;
  a.ptopo = dm->d_ptopo;
  a.phi = dm->d_phi;
  a.Qmax = dm->rs.Qmax;
  for (int c = 0; c < 6; ++c) {
    a.off[c] = dm->rs.off[c];
    a.len[c] = dm->rs.len[c];
  }
  a.R = dm->rs.R;
  a.hyp = dm->hyp;
  a.kr = kr;
  a.ki = ki;
  a.flags = flags;
  a.F = dF;
  a.G = dG;
  a.C = dC;
  return a;
}

// Whether stream s is capturing a graph (work enqueued there is ordered by the graph's
// later launches, not now).
bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  BIE_HIP_CHECK(hipStreamIsCapturing(s, &st));
  return st != hipStreamCaptureStatusNone;
}

// Calls on one handle share its device scratch (element tables, schedules, the mid-edge
// queue, counter and values), so a call's kernels must not overlap the previous call's:
// order_after_previous makes stream s wait for the last call on the handle when that call
// ran on another stream, and mark_done records this call's end.  Called under dm->mu.
void order_after_previous(bie_dmesh* dm, hipStream_t s) {
  if (!dm->done_valid || dm->done_stream == s || capturing(s)) return;
  BIE_HIP_CHECK(hipStreamWaitEvent(s, dm->done, 0));
}

void mark_done(bie_dmesh* dm, hipStream_t s) {
  if (capturing(s)) return;
  BIE_HIP_CHECK(hipEventRecord(dm->done, s));
  dm->done_stream = s;
  dm->done_valid = true;
}

// Largest distance from the mesh centre to any of n points (3 x n, column-major).
double radius_about(const double* c, const double* x, int64_t n) {
  double r2 = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const double a = x[3 * i] - c[0], b = x[3 * i + 1] - c[1], d = x[3 * i + 2] - c[2];
    r2 = std::max(r2, a * a + b * b + d * d);
  }
  return std::sqrt(r2);
}

// Whether every argument k r of the slot kernel's sincos stays below kTabFastMaxArg: r is
// at most |x - c| + |y - c| for a target x and an element point y, and y = sum N_k X_k with
// sum |N_k| <= 2 on linear and quadratic triangles, so |y - c| <= 2 coord_radius (taken
// twice again for margin).  target_radius < 0: unknown (device targets), never fast.
bool fast_trig_ok(const bie_dmesh* dm, double kr, double ki, double target_radius) {
  if (ki != 0.0 || !(target_radius >= 0.0)) return false;
  const double reach = target_radius + 4.0 * dm->coord_radius;
  return std::isfinite(reach) && std::fabs(kr) * reach < 0.5 * kTabFastMaxArg;
}

void assemble_device_impl(bie_dmesh* dm, double kr, double ki, const double* d_targets,
                          int64_t M, int n, int m, uint32_t flags, double* dF, double* dG,
                          double* dC, hipStream_t s, double target_radius = -1.0) {
  BIE_REQUIRE(dm != nullptr, BIE_EINVAL, "dmesh is NULL");
  BIE_REQUIRE(M >= 0, BIE_EINVAL, "negative number of targets");
  BIE_REQUIRE((flags & ~BIE_FGC) == 0, BIE_EINVAL, "unknown flag bits");
  BIE_REQUIRE(dF && dG && dC, BIE_EINVAL, "output pointer is NULL");
  BIE_REQUIRE(std::isfinite(kr) && std::isfinite(ki), BIE_EINVAL, "k must be finite");
  if (d_targets == nullptr) {
    d_targets = dm->d_sources;
    BIE_REQUIRE(M == dm->N, BIE_EINVAL, "collocation (NULL targets) needs n_targets == n_sources");
    target_radius = dm->source_radius;
  }
  BIE_HIP_CHECK(hipSetDevice(dm->device));
  std::lock_guard<std::mutex> lk(dm->mu);
  order_after_previous(dm, s);
  ensure_rules(dm, n, m, s);
  std::array<hipEvent_t, 4> ev{};
  bool timed = false;
  {
    std::lock_guard<std::mutex> tl(g_timing.mu);
    if (g_timing.on) {
      timed = true;
      for (auto& e : ev) e = g_timing.get();
    }
  }
  if (timed) BIE_HIP_CHECK(hipEventRecord(ev[0], s));
  const bool atomic_path = use_atomic_kernel() || M == 0 || dm->E == 0 ||
                           sizeof(double) * (size_t)dm->rs.R * dm->rs.Qmax * dm->pk > 24 * 1024;
  if (atomic_path || !(flags & BIE_F)) launch_fill_zero(dF, 2 * M * dm->N, s);
  if (atomic_path || !(flags & BIE_G)) launch_fill_zero(dG, 2 * M * dm->N, s);
  launch_fill_zero(dC, 2 * M, s);
  if (M > 0 && dm->E > 0) {
    refresh_geometry(dm, s);
    AssemblyArgs a = make_args(dm, d_targets, M, kr, ki, flags, dF, dG, dC);
    const bool phi_fits = sizeof(double) * (size_t)a.R * a.Qmax * a.NP <= 24 * 1024;
    if (use_atomic_kernel() || !phi_fits) {
      if (timed) BIE_HIP_CHECK(hipEventRecord(ev[1], s));
      launch_assemble(a, s);
      if (timed) BIE_HIP_CHECK(hipEventRecord(ev[2], s));
    } else {
      const int64_t rowblocks = (M + 63) / 64;
      const bool fast = fast_trig_ok(dm, kr, ki, target_radius);
      const bool qc_inst = fast && (a.len[0] == 4 || a.len[0] == 6 || a.len[0] == 16);
      // the occupancy, split and two-element variants exist for three- and six-node physics
      const bool tuned = qc_inst && (dm->pk == 3 || dm->pk == 6);
      const int occ = tuned ? slot_occupancy() : 1;
      const int split = tuned && occ == 1 ? slot_split() : 1;
      const int waves = tuned && occ == 1 && split == 1 ? slot_waves(dm->pk) : kSlotWaves;
      const int vpw =
          tuned && occ == 1 && split == 1 && waves == kSlotWaves ? slot_elems_per_wave() : 1;
      const int min_segments = choose_segment_count(dm->device, rowblocks, dm->E, occ);
      const bie_dmesh::Sched& sc = ensure_schedule(
          dm, 0, plain_slot_count(a.R * a.Qmax * a.NP, a.len[0], occ, split, waves),
          min_segments, s, waves * vpw);
      // collocation on quadratic physics: after the first call the mid-edge pairs come
      // from the recorded sorted list instead of the slot kernel's queue
      const bool colloc = d_targets == dm->d_sources && M == dm->N;
      const bool mids = uses_mid_rules(dm->gk, dm->pk);
      const bool recorded = mids && colloc && dm->n_mid_keys >= 0;
      if (mids && !recorded) {
        ensure_mid_queue(dm, M * dm->E);
        launch_fill_zero_u32(dm->d_mid_count, s);
      } else if (!dm->d_mid_count) {
        dm->d_mid_count = dalloc<unsigned>(1);
      }
      SlotArgs sa = slot_args(dm, sc, rowblocks);
      sa.push_mid = recorded ? 0 : 1;
      sa.occupancy = occ;
      sa.split = split;
      // The slot kernel stores owned columns once; only shared/untouched columns need
      // zeros (a disabled output is zeroed whole above and never written).
      if (flags & BIE_F) launch_zero_columns(dF, M, sc.d_zero_cols, sc.zero_cols, s);
      if (flags & BIE_G) launch_zero_columns(dG, M, sc.d_zero_cols, sc.zero_cols, s);
      // recorded mid-edge pairs: their values do not depend on the slot kernel, so they are
      // computed on the handle's side stream while it runs (in a captured graph: in order)
      const bool side = recorded && dm->n_mid_keys > 0 && !capturing(s);
      if (side) {
        if (!dm->side) {
          BIE_HIP_CHECK(hipStreamCreateWithFlags(&dm->side, hipStreamNonBlocking));
          BIE_HIP_CHECK(hipEventCreateWithFlags(&dm->side_fork, hipEventDisableTiming));
          BIE_HIP_CHECK(hipEventCreateWithFlags(&dm->side_join, hipEventDisableTiming));
        }
        BIE_HIP_CHECK(hipEventRecord(dm->side_fork, s));  // after the zero fills and geometry
        BIE_HIP_CHECK(hipStreamWaitEvent(dm->side, dm->side_fork, 0));
        launch_mid_compute(a, dm->d_mid_keys, dm->n_mid_keys, dm->d_mid_vals, dm->side);
        BIE_HIP_CHECK(hipEventRecord(dm->side_join, dm->side));
      }
      if (timed) BIE_HIP_CHECK(hipEventRecord(ev[1], s));
      launch_assemble_slots(a, sa, fast, s);
      if (timed) BIE_HIP_CHECK(hipEventRecord(ev[2], s));
      if (side) {
        BIE_HIP_CHECK(hipStreamWaitEvent(s, dm->side_join, 0));
        launch_mid_scatter(a, dm->n_mid_keys, dm->d_mid_vals, dm->d_mid_entries, dm->d_mid_off,
                           dm->d_mid_refs, dm->n_mid_entries, s);
      } else if (recorded) {
        launch_mid_recorded(a, dm->d_mid_keys, dm->n_mid_keys, dm->d_mid_vals, dm->d_mid_entries,
                            dm->d_mid_off, dm->d_mid_refs, dm->n_mid_entries, s);
      } else {
        launch_assemble_mid(a, sa, s);
        if (mids && colloc) record_mid_pairs(dm, s);
      }
    }
  } else if (timed) {
    BIE_HIP_CHECK(hipEventRecord(ev[1], s));
    BIE_HIP_CHECK(hipEventRecord(ev[2], s));
  }
  if (timed) {
    BIE_HIP_CHECK(hipEventRecord(ev[3], s));
    std::lock_guard<std::mutex> tl(g_timing.mu);
    g_timing.rec.push_back(ev);
  }
  mark_done(dm, s);
}

// Batched sweep over nk real wavenumbers (d_k on the device), or with taylor the k
// derivatives of orders 0..nk-1 at k_re + i k_im (d_k unused).  prepare_only builds every
// cache (rules, schedule, queues) without launching, so that a later call can be
// captured into a graph.
void sweep_device_impl(bie_dmesh* dm, const double* d_k, int nk, const double* d_targets,
                       int64_t M, int n, int m, uint32_t flags, double* dF, double* dG,
                       int64_t stride, double* dC, hipStream_t s, bool prepare_only,
                       bool taylor = false, double k_re = 0.0, double k_im = 0.0) {
  BIE_REQUIRE(dm != nullptr, BIE_EINVAL, "dmesh is NULL");
  BIE_REQUIRE(!kind_is_quad(dm->gkind), BIE_EUNSUPPORTED,
              "sweeps and k-derivatives: triangular meshes only");
  BIE_REQUIRE(nk >= 0 && M >= 0, BIE_EINVAL, "negative size");
  BIE_REQUIRE((flags & ~BIE_FGC) == 0, BIE_EINVAL, "unknown flag bits");
  BIE_REQUIRE(taylor || nk == 0 || d_k, BIE_EINVAL, "k list is NULL");
  BIE_REQUIRE(!taylor || (std::isfinite(k_re) && std::isfinite(k_im)), BIE_EINVAL,
              "wavenumber is not finite");
  // a disabled F or G may be NULL (then nothing is written there)
  BIE_REQUIRE(((flags & BIE_F) == 0 || dF) && ((flags & BIE_G) == 0 || dG) && dC, BIE_EINVAL,
              "output pointer is NULL");
  BIE_REQUIRE(stride >= M * dm->N, BIE_EINVAL, "stride smaller than M * N");
  if (d_targets == nullptr) {
    d_targets = dm->d_sources;
    BIE_REQUIRE(M == dm->N, BIE_EINVAL, "collocation (NULL targets) needs n_targets == n_sources");
  }
  BIE_HIP_CHECK(hipSetDevice(dm->device));
  ensure_rules(dm, n, m, s);
  const int KB = sweep_kb(dm->pk);
  BIE_REQUIRE(sizeof(double) * (size_t)dm->rs.R * dm->rs.Qmax * dm->pk <= 24 * 1024,
              BIE_EUNSUPPORTED, "sweep supports quadrature orders up to n*m' = 36");
  const int fit = fitting_slots(sweep_fixed_doubles(dm->rs.R * dm->rs.Qmax * dm->pk), KB);
  const int S = dm->sweep_slots ? std::min(dm->sweep_slots, fit) : fit;
  const int min_seg = dm->sweep_elems_per_tile
                          ? (int)std::max<int64_t>(1, (dm->E + dm->sweep_elems_per_tile - 1) /
                                                          dm->sweep_elems_per_tile)
                          : 1;
  const bie_dmesh::Sched& sc = ensure_schedule(dm, 1, S, min_seg, s);
  if (!dm->d_mid_count) dm->d_mid_count = dalloc<unsigned>(1);
  if (uses_mid_rules(dm->gk, dm->pk)) ensure_mid_queue(dm, std::max<int64_t>(1, M * dm->E));
  const int KBk = sweep_block_size(dm->pk, (flags & BIE_G) != 0);
  const int nkb = (nk + KBk - 1) / KBk;
  if (dm->block_uniform_cap < std::max(nkb, 1)) {
    dfree(dm->d_block_uniform);
    dm->d_block_uniform = nullptr;
    dm->block_uniform_cap = 0;
    dm->d_block_uniform = dalloc<int>((size_t)std::max(nkb, 1));
    dm->block_uniform_cap = std::max(nkb, 1);
  }
  if (prepare_only) return;
  order_after_previous(dm, s);
  launch_fill_zero(dC, 2 * M, s);
  launch_fill_zero_u32(dm->d_mid_count, s);
  for (int which = 0; which < 2; ++which) {
    double* X = which == 0 ? dF : dG;
    const bool on = (flags & (which == 0 ? BIE_F : BIE_G)) != 0;
    if (on)
      launch_zero_columns_multi(X, stride, nk, M, sc.d_zero_cols, sc.zero_cols, s);
    else if (X)
      for (int q = 0; q < nk; ++q) launch_fill_zero(X + 2 * (size_t)q * stride, 2 * M * dm->N, s);
  }
  if (M == 0 || dm->E == 0 || nk == 0) {
    mark_done(dm, s);
    return;
  }
  refresh_geometry(dm, s);
  if (!taylor) launch_block_uniform(d_k, nk, KBk, dm->d_block_uniform, s);
  const int64_t rowblocks = (M + 63) / 64;
  AssemblyArgs a = make_args(dm, d_targets, M, 0.0, 0.0, flags, dF, dG, dC);
  SlotArgs sa = slot_args(dm, sc, rowblocks);
  SweepArgs sw{};
  sw.k = taylor ? nullptr : d_k;
  sw.nk = nk;
  sw.block_uniform = taylor ? nullptr : dm->d_block_uniform;
  sw.stride = stride;
  sw.n_tiles = (int64_t)nkb * sc.segments * rowblocks;
  sw.taylor = taylor ? 1 : 0;
  sw.k_re = k_re;
  sw.k_im = k_im;
  launch_sweep(a, sa, sw, s);
  mark_done(dm, s);
}


// ---- uneven batch -------------------------------------------------------------------

// Tile list of the batch for its current configuration: every (instance, row block,
// segment), costed by the segment's step count, ordered longest-first (order 0; ties in
// instance, segment, row block order) or interleaved across instances (order 1).
void batch_configure(bie_batch* b, hipStream_t s) {
  const int key = b->elems_per_tile * 1000 + b->slots * 10 + b->order;
  bie_dmesh* d0 = b->meshes[0];
  // slots = 32 or 64 asks for that many, capped at what fits beside the tables and staging
  const int fit = fitting_slots(
      phi_pad(d0->rs.R * d0->rs.Qmax * d0->pk) + kTrigDoubles + stage_doubles(d0->rs.len[0]), 1);
  const int S = b->slots ? std::min(b->slots, fit) : fit;
  bool current = b->configured_key == key && b->d_tiles != nullptr && b->slots_used == S;
  // a mesh's batch schedule may have been rebuilt by another batch that shares it
  for (size_t ii = 0; current && ii < b->meshes.size(); ++ii) {
    const bie_dmesh::Sched& sc = b->meshes[ii]->sched[2];
    current = sc.d_step_elems != nullptr && sc.S == S && sc.key == b->min_segments[ii];
  }
  if (current) return;
  struct Tile {
    int steps, inst, rb, seg;
    int64_t rank;
  };
  std::vector<Tile> tiles;
  const int n_inst = (int)b->meshes.size();
  b->configured_key = -1;
  b->rowblocks.assign(n_inst, 0);
  b->min_segments.assign(n_inst, 0);
  for (int ii = 0; ii < n_inst; ++ii) {
    bie_dmesh* dm = b->meshes[ii];
    const int64_t rbs = (dm->N + 63) / 64;
    b->rowblocks[ii] = rbs;
    const int min_seg = (int)std::max<int64_t>(1, (dm->E + b->elems_per_tile - 1) / b->elems_per_tile);
    b->min_segments[ii] = min_seg;
    const bie_dmesh::Sched& sc = ensure_schedule(dm, 2, S, min_seg, s);
    if (dm->N == 0) continue;
    int64_t rank = 0;
    for (int g = 0; g < sc.segments; ++g) {
      const int steps = sc.h_seg_step_begin[g + 1] - sc.h_seg_step_begin[g];
      for (int64_t rb = 0; rb < rbs; ++rb) tiles.push_back(Tile{steps, ii, (int)rb, g, rank++});
    }
  }
  BIE_REQUIRE(tiles.size() < (size_t)INT32_MAX, BIE_EINVAL, "batch has too many tiles");
  if (b->order == 0)
    std::stable_sort(tiles.begin(), tiles.end(),
                     [](const Tile& x, const Tile& y) { return x.steps > y.steps; });
  else
    std::stable_sort(tiles.begin(), tiles.end(),
                     [](const Tile& x, const Tile& y) { return x.rank < y.rank; });
  std::vector<int4> h(tiles.size());
  for (size_t t = 0; t < tiles.size(); ++t)
    h[t] = make_int4(tiles[t].inst, tiles[t].rb, tiles[t].seg, 0);
  if (b->d_tiles) (void)hipDeviceSynchronize();  // a previous launch may still read it
  dfree(b->d_tiles);
  b->d_tiles = nullptr;
  b->n_tiles = 0;
  b->d_tiles = upload_vec(h, s);
  BIE_HIP_CHECK(hipStreamSynchronize(s));
  b->n_tiles = (int)h.size();
  b->slots_used = S;
  b->configured_key = key;
}

struct MeshLocks {
  std::vector<std::unique_lock<std::mutex>> held;
  explicit MeshLocks(const std::vector<bie_dmesh*>& ms) {
    std::vector<bie_dmesh*> order(ms);
    std::sort(order.begin(), order.end());
    order.erase(std::unique(order.begin(), order.end()), order.end());
    for (bie_dmesh* dm : order) held.emplace_back(dm->mu);
  }
};

void batch_validate(bie_batch* b, const double* k_re, const double* k_im, uint32_t flags,
                    double* const* dF, double* const* dG, double* const* dC) {
  BIE_REQUIRE(b != nullptr, BIE_EINVAL, "batch is NULL");
  BIE_REQUIRE((flags & ~BIE_FGC) == 0, BIE_EINVAL, "unknown flag bits");
  BIE_REQUIRE(k_re && dF && dG && dC, BIE_EINVAL, "NULL argument array");
  const int n_inst = (int)b->meshes.size();
  for (int ii = 0; ii < n_inst; ++ii) {
    const double ki = k_im ? k_im[ii] : 0.0;
    BIE_REQUIRE(std::isfinite(k_re[ii]) && std::isfinite(ki), BIE_EINVAL, "k must be finite");
    BIE_REQUIRE(dC[ii], BIE_EINVAL, "C output pointer is NULL");
    BIE_REQUIRE(!(flags & BIE_F) || dF[ii], BIE_EINVAL, "F output pointer is NULL");
    BIE_REQUIRE(!(flags & BIE_G) || dG[ii], BIE_EINVAL, "G output pointer is NULL");
  }
}

// What the batch's kernels need besides the instance table.
struct BatchPlan {
  bool cplx = false;
  bool fast = true;  // every instance: real k with |k| r below the fast sincos bound
  bool gon = true;
  int64_t max_work = 1;
  // compile-time rule length of the fast kernel instance (0: the generic instance)
  int qc(int len0) const {
    return fast && !cplx && (len0 == 4 || len0 == 6 || len0 == 16) ? len0 : 0;
  }
};

// Host half of a batched assembly (under the batch and mesh locks): every per-mesh cache,
// the tile list, and the instance table in b->h_inst (after the previous upload has left
// it).
BatchPlan batch_fill(bie_batch* b, const double* k_re, const double* k_im, uint32_t flags,
                     double* const* dF, double* const* dG, double* const* dC, hipStream_t s) {
  const int n_inst = (int)b->meshes.size();
  for (bie_dmesh* dm : b->meshes) {
    ensure_rules(dm, b->n, b->m, s);
    if (!dm->geom_valid) refresh_geometry(dm, s);
    if (uses_mid_rules(dm->gk, dm->pk))
      ensure_mid_queue(dm, std::max<int64_t>(1, dm->N * dm->E));
    else if (!dm->d_mid_count)
      dm->d_mid_count = dalloc<unsigned>(1);
  }
  batch_configure(b, s);
  BIE_HIP_CHECK(hipEventSynchronize(b->staged));  // the previous upload has left h_inst
  BatchPlan plan;
  plan.gon = (flags & BIE_G) != 0;
  for (int ii = 0; ii < n_inst; ++ii) {
    bie_dmesh* dm = b->meshes[ii];
    const double ki = k_im ? k_im[ii] : 0.0;
    plan.cplx = plan.cplx || ki != 0.0;
    plan.fast = plan.fast && fast_trig_ok(dm, k_re[ii], ki, dm->source_radius);
    const bie_dmesh::Sched& sc = dm->sched[2];
    BatchInst& bi = b->h_inst[ii];
    bi.a = make_args(dm, dm->d_sources, dm->N, k_re[ii], ki, flags, dF[ii], dG[ii], dC[ii]);
    bi.sa = slot_args(dm, sc, b->rowblocks[ii]);
    bi.zero_cols = sc.d_zero_cols;
    bi.n_zero = sc.zero_cols;
    bi.pad = 0;
    plan.max_work = std::max<int64_t>(plan.max_work, sc.zero_cols);  // columns to zero per output
    if ((!(flags & BIE_F) && dF[ii]) || (!(flags & BIE_G) && dG[ii]))
      plan.max_work = std::max<int64_t>(plan.max_work, dm->N);
  }
  return plan;
}

// Device half: zero fills, the persistent assembly over the tile list, the mid-edge pairs.
// Kernels only (no memset, no host synchronisation), so it can be captured.
void batch_kernels(bie_batch* b, const BatchInst* d_inst, const int4* d_tiles,
                   unsigned* d_counter, const BatchPlan& plan, hipStream_t s) {
  const int n_inst = (int)b->meshes.size();
  launch_batch_prep(d_inst, n_inst, plan.max_work, s);
  if (b->n_tiles > 0) {
    bie_dmesh* d0 = b->meshes[0];
    const int nphi = d0->rs.R * d0->rs.Qmax * d0->pk;
    int cus = 0;
    BIE_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
    const int nstage = stage_doubles(d0->rs.len[0]);
    const int qc = plan.qc(d0->rs.len[0]);
    const int per_cu =
        batch_blocks_per_cu(b->pk, plan.cplx, plan.gon, qc, b->slots_used, nphi, nstage);
    const int grid = (int)std::min<int64_t>((int64_t)std::max(cus, 1) * per_cu, b->n_tiles);
    launch_assemble_batch(d_inst, d_tiles, b->n_tiles, d_counter, d0->d_phi, nphi, nstage,
                          b->slots_used, b->pk, plan.cplx, plan.gon, qc, grid, s);
  }
  if (uses_mid_rules(b->gk, b->pk)) launch_assemble_mid_batch(d_inst, n_inst, plan.cplx, s);
}

void batch_assemble_impl(bie_batch* b, const double* k_re, const double* k_im, uint32_t flags,
                         double* const* dF, double* const* dG, double* const* dC,
                         hipStream_t s) {
  batch_validate(b, k_re, k_im, flags, dF, dG, dC);
  const int n_inst = (int)b->meshes.size();
  BIE_HIP_CHECK(hipSetDevice(b->device));
  std::lock_guard<std::mutex> bl(b->mu);
  MeshLocks locks(b->meshes);
  for (bie_dmesh* dm : b->meshes) order_after_previous(dm, s);
  const BatchPlan plan = batch_fill(b, k_re, k_im, flags, dF, dG, dC, s);
  BIE_HIP_CHECK(hipMemcpyAsync(b->d_inst, b->h_inst, sizeof(BatchInst) * n_inst,
                               hipMemcpyHostToDevice, s));
  BIE_HIP_CHECK(hipEventRecord(b->staged, s));
  batch_kernels(b, b->d_inst, b->d_tiles, b->d_counter, plan, s);
  for (bie_dmesh* dm : b->meshes) mark_done(dm, s);
}

// w_i = alpha_i^T (dF_i/dk) p_i for every instance of the batch, in one launch.
void batch_dfdk_impl(bie_batch* b, const double* k_re, const double* k_im,
                     const double* const* d_alpha, const double* const* d_p, double* d_w,
                     hipStream_t s) {
  BIE_REQUIRE(b != nullptr, BIE_EINVAL, "batch is NULL");
  BIE_REQUIRE(k_re && d_alpha && d_p && d_w, BIE_EINVAL, "NULL argument array");
  const int n_inst = (int)b->meshes.size();
  for (int ii = 0; ii < n_inst; ++ii) {
    const double ki = k_im ? k_im[ii] : 0.0;
    BIE_REQUIRE(std::isfinite(k_re[ii]) && std::isfinite(ki), BIE_EINVAL, "k must be finite");
    BIE_REQUIRE(d_alpha[ii] && d_p[ii], BIE_EINVAL, "alpha or p pointer is NULL");
  }
  BIE_HIP_CHECK(hipSetDevice(b->device));
  std::lock_guard<std::mutex> bl(b->mu);
  MeshLocks locks(b->meshes);
  std::vector<DfdkInst> h(n_inst);
  std::vector<int4> tiles;
  constexpr int ech = 64;  // elements per tile
  for (int ii = 0; ii < n_inst; ++ii) {
    bie_dmesh* dm = b->meshes[ii];
    order_after_previous(dm, s);
    ensure_rules(dm, b->n, b->m, s);
    if (!dm->geom_valid) refresh_geometry(dm, s);
    DfdkInst& a = h[ii];
    a = DfdkInst{};
    a.targets = dm->d_sources;
    a.M = dm->N;
    a.E = (int)dm->E;
    a.Pe = dm->rs.Pe;
    a.elem = dm->d_elem;
    a.selnodes = dm->d_selnodes;
    a.ptopo = dm->d_ptopo;
    a.phi = dm->d_phi;
    a.Qmax = dm->rs.Qmax;
    a.R = dm->rs.R;
    a.hyp = dm->hyp;
    for (int c = 0; c < 6; ++c) {
      a.off[c] = dm->rs.off[c];
      a.len[c] = dm->rs.len[c];
    }
    a.kr = k_re[ii];
    a.ki = k_im ? k_im[ii] : 0.0;
    a.alpha = reinterpret_cast<const double2*>(d_alpha[ii]);
    a.p = reinterpret_cast<const double2*>(d_p[ii]);
    a.w = d_w + 2 * (size_t)ii;
    const int64_t rbs = (dm->N + 255) / 256, chunks = (dm->E + ech - 1) / ech;
    for (int64_t rb = 0; rb < rbs; ++rb)
      for (int64_t ch = 0; ch < chunks; ++ch) tiles.push_back(make_int4(ii, (int)rb, (int)ch, 0));
  }
  BIE_REQUIRE(tiles.size() < (size_t)INT32_MAX, BIE_EINVAL, "batch has too many tiles");
  DfdkInst* d_inst = dalloc<DfdkInst>(h.size());
  int4* d_tiles = nullptr;
  try {
    d_tiles = dalloc<int4>(tiles.size());
    BIE_HIP_CHECK(hipMemcpyAsync(d_inst, h.data(), sizeof(DfdkInst) * h.size(),
                                 hipMemcpyHostToDevice, s));
    if (!tiles.empty())
      BIE_HIP_CHECK(hipMemcpyAsync(d_tiles, tiles.data(), sizeof(int4) * tiles.size(),
                                   hipMemcpyHostToDevice, s));
    launch_fill_zero(d_w, 2 * (int64_t)n_inst, s);
    launch_dfdk_batch(d_inst, d_tiles, (int)tiles.size(), ech, b->pk, s);
    // the per-call tables are freed once the launch has read them
    BIE_HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    dfree(d_inst);
    dfree(d_tiles);
    throw;
  }
  dfree(d_inst);
  dfree(d_tiles);
  for (bie_dmesh* dm : b->meshes) mark_done(dm, s);
}

void batch_destroy_impl(bie_batch* b) {
  (void)hipSetDevice(b->device);
  if (b->staged) {
    (void)hipEventSynchronize(b->staged);
    (void)hipEventDestroy(b->staged);
  }
  if (b->h_inst) (void)hipHostFree(b->h_inst);
  dfree(b->d_inst);
  dfree(b->d_tiles);
  dfree(b->d_counter);
  delete b;
}

}  // namespace
}  // namespace bie

using namespace bie;

extern "C" {

const char* bie_last_error(void) { return g_last_error.c_str(); }

int bie_version(void) { return 3; }

int bie_dmesh_create(const bie_mesh* mesh, int device, bie_dmesh** out) {
  return guarded([&] {
    BIE_REQUIRE(out != nullptr, BIE_EINVAL, "out is NULL");
    *out = nullptr;
    validate_mesh(mesh);
    // Host-side validation first: a malformed mesh is rejected before any HIP call.
    const int64_t base = mesh->index_base;
    std::vector<int> topo = convert_topology(mesh->topology,
                                             geometry_nodes(mesh->geometry_kind),
                                             mesh->n_elements, base, mesh->n_coordinates,
                                             "topology");
    std::vector<int> ptopo = convert_topology(mesh->physics_topology,
                                              physics_nodes(mesh->physics_kind),
                                              mesh->n_elements, base, mesh->n_sources,
                                              "physics_topology");
    BIE_REQUIRE(device >= 0, BIE_EINVAL, "device ordinal out of range");
    int ndev = 0;
    BIE_HIP_CHECK(hipGetDeviceCount(&ndev));
    BIE_REQUIRE(device < ndev, BIE_EINVAL, "device ordinal out of range");
    BIE_HIP_CHECK(hipSetDevice(device));
    std::unique_ptr<bie_dmesh> dm(new bie_dmesh);
    dm->h_topo = topo;
    dm->h_ptopo = ptopo;
    dm->h_coords.assign(mesh->coordinates, mesh->coordinates + 3 * mesh->n_coordinates);
    dm->h_sources.assign(mesh->sources, mesh->sources + 3 * mesh->n_sources);
    dm->device = device;
    dm->hyp = (mesh->flags & BIE_MESH_SCALED_HYPOT) ? 1 : 0;
    dm->gk = geometry_nodes(mesh->geometry_kind);
    dm->gkind = mesh->geometry_kind;
    dm->pk = physics_nodes(mesh->physics_kind);
    dm->pkind = mesh->physics_kind;
    dm->beta = kind_has_beta(mesh->physics_kind) ? mesh->physics_beta : 0.0;
    dm->Nc = mesh->n_coordinates;
    dm->E = mesh->n_elements;
    dm->N = mesh->n_sources;
    if (dm->Nc > 0) {  // bounding-box centre of the coordinates
      for (int d = 0; d < 3; ++d) {
        double lo = INFINITY, hi = -INFINITY;
        for (int64_t v = 0; v < dm->Nc; ++v) {
          lo = std::min(lo, dm->h_coords[3 * v + d]);
          hi = std::max(hi, dm->h_coords[3 * v + d]);
        }
        dm->center[d] = 0.5 * (lo + hi);
      }
    }
    dm->coord_radius = radius_about(dm->center, dm->h_coords.data(), dm->Nc);
    dm->source_radius = radius_about(dm->center, dm->h_sources.data(), dm->N);
    try {
      dm->d_coords = dalloc<double>((size_t)3 * dm->Nc);
      dm->d_sources = dalloc<double>((size_t)3 * dm->N);
      dm->d_topo = dalloc<int>(topo.size());
      dm->d_ptopo = dalloc<int>(ptopo.size());
      dm->d_selnodes = dalloc<double>((size_t)dm->E * dm->pk * 4);
      if (dm->Nc)
        BIE_HIP_CHECK(hipMemcpy(dm->d_coords, mesh->coordinates, 24 * dm->Nc, hipMemcpyHostToDevice));
      if (dm->N)
        BIE_HIP_CHECK(hipMemcpy(dm->d_sources, mesh->sources, 24 * dm->N, hipMemcpyHostToDevice));
      if (!topo.empty())
        BIE_HIP_CHECK(hipMemcpy(dm->d_topo, topo.data(), topo.size() * sizeof(int), hipMemcpyHostToDevice));
      if (!ptopo.empty())
        BIE_HIP_CHECK(hipMemcpy(dm->d_ptopo, ptopo.data(), ptopo.size() * sizeof(int), hipMemcpyHostToDevice));
      launch_gather_selnodes(dm->d_sources, dm->d_ptopo, dm->pk, (int)dm->E, dm->d_selnodes, nullptr);
      BIE_HIP_CHECK(hipEventCreateWithFlags(&dm->done, hipEventDisableTiming));
      BIE_HIP_CHECK(hipDeviceSynchronize());
    } catch (...) {
      bie_dmesh_destroy(dm.release());
      throw;
    }
    *out = dm.release();
  });
}

int bie_dmesh_destroy(bie_dmesh* dm) {
  if (!dm) return BIE_OK;
  (void)hipSetDevice(dm->device);
  if (dm->done) {  // the handle's last call may still be running
    (void)hipEventSynchronize(dm->done);
    (void)hipEventDestroy(dm->done);
  }
  if (dm->side) {
    (void)hipStreamSynchronize(dm->side);
    (void)hipStreamDestroy(dm->side);
    (void)hipEventDestroy(dm->side_fork);
    (void)hipEventDestroy(dm->side_join);
  }
  dfree(dm->d_coords);
  dfree(dm->d_topo);
  dfree(dm->d_sources);
  dfree(dm->d_ptopo);
  dfree(dm->d_selnodes);
  dfree(dm->d_pts);
  dfree(dm->d_phi);
  dfree(dm->d_elem);
  free_schedules(dm);
  dfree(dm->d_block_uniform);
  dfree(dm->d_mid_count);
  dfree(dm->d_mid_queue);
  dfree(dm->d_mid_keys);
  dfree(dm->d_mid_vals);
  dfree(dm->d_mid_entries);
  dfree(dm->d_mid_off);
  dfree(dm->d_mid_refs);
  delete dm;
  return BIE_OK;
}

int64_t bie_dmesh_n_sources(const bie_dmesh* dm) { return dm ? dm->N : -1; }
int64_t bie_dmesh_n_elements(const bie_dmesh* dm) { return dm ? dm->E : -1; }
const double* bie_dmesh_sources(const bie_dmesh* dm) { return dm ? dm->d_sources : nullptr; }

int bie_assemble_device(bie_dmesh* dm, double k_re, double k_im, const double* d_targets,
                        int64_t n_targets, int32_t n, int32_t m, uint32_t flags, double* dF,
                        double* dG, double* dC, void* stream) {
  return guarded([&] {
    assemble_device_impl(dm, k_re, k_im, d_targets, n_targets, n, m, flags, dF, dG, dC,
                         static_cast<hipStream_t>(stream));
  });
}

int bie_select_rules_device(bie_dmesh* dm, const double* d_targets, int64_t n_targets,
                            int8_t* d_rules, void* stream) {
  return guarded([&] {
    BIE_REQUIRE(dm && d_rules, BIE_EINVAL, "NULL argument");
    BIE_REQUIRE(n_targets >= 0, BIE_EINVAL, "negative number of targets");
    if (!d_targets) {
      d_targets = dm->d_sources;
      BIE_REQUIRE(n_targets == dm->N, BIE_EINVAL, "collocation needs n_targets == n_sources");
    }
    BIE_HIP_CHECK(hipSetDevice(dm->device));
    const int R = kind_is_quad(dm->gkind) ? 1 : (dm->gk == 6 ? 6 : 3);
    launch_select(d_targets, n_targets, dm->d_selnodes, dm->pk, (int)dm->E, R, dm->hyp,
                  d_rules, static_cast<hipStream_t>(stream));
  });
}

int bie_element_geometry(bie_dmesh* dm, int32_t rule, int32_t n, int32_t m, double* y,
                         double* normals, double* jw) {
  return guarded([&] {
    BIE_REQUIRE(dm && y && normals && jw, BIE_EINVAL, "NULL argument");
    BIE_HIP_CHECK(hipSetDevice(dm->device));
    std::lock_guard<std::mutex> lk(dm->mu);
    order_after_previous(dm, nullptr);
    ensure_rules(dm, n, m, nullptr);
    BIE_REQUIRE(rule >= 1 && rule <= dm->rs.R, BIE_EINVAL, "rule index out of range");
    const RuleSet& rs = dm->rs;
    refresh_geometry(dm, nullptr);
    std::vector<double> h((size_t)dm->E * rs.Pe * 8);
    BIE_HIP_CHECK(hipMemcpy(h.data(), dm->d_elem, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    mark_done(dm, nullptr);
    const int Q = rs.len[rule - 1], off = rs.off[rule - 1];
    for (int64_t e = 0; e < dm->E; ++e)
      for (int q = 0; q < Q; ++q) {
        const double* p = &h[((size_t)e * rs.Pe + off + q) * 8];
        const size_t o = (size_t)e * Q + q;
        for (int d = 0; d < 3; ++d) {
          y[3 * o + d] = p[d];
          normals[3 * o + d] = p[3 + d];
        }
        jw[o] = p[6] * 4.0 * kPi;
      }
  });
}

namespace {

// The host-array entry point keeps the device form of the last mesh it was given (upload,
// rule tables, slot schedule, recorded mid-edge pairs), so a caller that assembles one
// mesh at many wavenumbers - a frequency loop calling assemble_parallel! once per k - pays
// that setup once.  A mesh is matched byte for byte against a copy of its fields and
// arrays.  The cache is never destroyed (device memory is released with the process, not
// by a static destructor that could run after the HIP runtime has gone).
struct HostMeshCache {
  std::mutex mu;
  std::vector<unsigned char> key;
  std::shared_ptr<bie_dmesh> dm;
  int64_t hits = 0, builds = 0;
};

HostMeshCache& host_mesh_cache() {
  static HostMeshCache* c = new HostMeshCache;
  return *c;
}

void append_bytes(std::vector<unsigned char>& key, const void* p, size_t n) {
  const unsigned char flag = p ? 1 : 0;
  key.push_back(flag);
  if (p && n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    key.insert(key.end(), b, b + n);
  }
}

std::vector<unsigned char> mesh_key(const bie_mesh* m) {
  std::vector<unsigned char> key;
  const int64_t scal[7] = {m->n_coordinates, m->n_elements, m->n_sources, m->geometry_kind,
                           m->physics_kind, m->index_base, m->flags};
  append_bytes(key, scal, sizeof(scal));
  const double beta = kind_has_beta(m->physics_kind) ? m->physics_beta : 0.0;
  append_bytes(key, &beta, sizeof(beta));
  append_bytes(key, m->coordinates, sizeof(double) * 3 * (size_t)m->n_coordinates);
  append_bytes(key, m->topology,
               sizeof(int64_t) * (size_t)geometry_nodes(m->geometry_kind) * m->n_elements);
  append_bytes(key, m->sources, sizeof(double) * 3 * (size_t)m->n_sources);
  append_bytes(key, m->physics_topology,
               sizeof(int64_t) * (size_t)physics_nodes(m->physics_kind) * m->n_elements);
  return key;
}

int host_mesh_handle(const bie_mesh* mesh, std::shared_ptr<bie_dmesh>* out) {
  std::vector<unsigned char> key;
  int rc = guarded([&] {
    validate_mesh(mesh);
    key = mesh_key(mesh);
  });
  if (rc != BIE_OK) return rc;
  HostMeshCache& c = host_mesh_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  if (c.dm && c.key == key) {
    *out = c.dm;
    ++c.hits;
    return BIE_OK;
  }
  bie_dmesh* raw = nullptr;
  rc = bie_dmesh_create(mesh, 0, &raw);
  if (rc != BIE_OK) return rc;
  ++c.builds;
  // a replaced handle is destroyed when its last caller releases it
  std::shared_ptr<bie_dmesh> sp(raw, [](bie_dmesh* d) { bie_dmesh_destroy(d); });
  c.dm = sp;
  c.key = std::move(key);
  *out = std::move(sp);
  return BIE_OK;
}

}  // namespace

namespace {

// Device-to-host copy of large outputs into caller (pageable) memory: DMA into one of two
// pinned staging buffers while host threads copy the other one out, instead of the
// runtime's single-threaded pageable path.  The buffers are allocated once per process and
// shared by all callers under a mutex; without them (allocation failed) the plain copy
// is used.
struct HostStaging {
  std::mutex mu;
  bool tried = false;
  char* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipStream_t s = nullptr;
};
constexpr size_t kStageBytes = size_t(128) << 20;

HostStaging& host_staging() {
  static HostStaging* h = new HostStaging;  // never destroyed (see host_mesh_cache)
  return *h;
}

void parallel_copy(char* dst, const char* src, size_t n) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>(std::min(8u, hw), std::max<size_t>(1, n >> 23));
  if (nt <= 1) {
    std::memcpy(dst, src, n);
    return;
  }
  const size_t part = (n + nt - 1) / nt;
  std::vector<std::thread> th;
  for (size_t t = 1; t < nt; ++t) {
    const size_t b = t * part, e = std::min(n, b + part);
    if (b < e) th.emplace_back([=] { std::memcpy(dst + b, src + b, e - b); });
  }
  std::memcpy(dst, src, std::min(n, part));
  for (auto& x : th) x.join();
}

void copy_to_host(void* dst, const void* src, size_t n) {
  if (n == 0) return;
  HostStaging& h = host_staging();
  std::lock_guard<std::mutex> lk(h.mu);
  if (!h.tried) {
    h.tried = true;
    bool ok = hipStreamCreateWithFlags(&h.s, hipStreamNonBlocking) == hipSuccess;
    for (int i = 0; ok && i < 2; ++i) {
      ok = hipHostMalloc(reinterpret_cast<void**>(&h.buf[i]), kStageBytes,
                         hipHostMallocDefault) == hipSuccess &&
           hipEventCreateWithFlags(&h.ev[i], hipEventDisableTiming) == hipSuccess;
    }
    if (!ok) {
      for (int i = 0; i < 2; ++i) {
        if (h.buf[i]) (void)hipHostFree(h.buf[i]);
        if (h.ev[i]) (void)hipEventDestroy(h.ev[i]);
        h.buf[i] = nullptr;
        h.ev[i] = nullptr;
      }
      if (h.s) (void)hipStreamDestroy(h.s);
      h.s = nullptr;
      (void)hipGetLastError();
    }
  }
  if (!h.s || n < 2 * kStageBytes) {
    BIE_HIP_CHECK(hipMemcpy(dst, src, n, hipMemcpyDeviceToHost));
    return;
  }
  const char* s = static_cast<const char*>(src);
  char* d = static_cast<char*>(dst);
  const size_t nchunk = (n + kStageBytes - 1) / kStageBytes;
  for (size_t c = 0; c <= nchunk; ++c) {
    if (c < nchunk) {  // chunk c into buffer c % 2 (its previous contents were copied out)
      const size_t off = c * kStageBytes, len = std::min(kStageBytes, n - off);
      BIE_HIP_CHECK(hipMemcpyAsync(h.buf[c & 1], s + off, len, hipMemcpyDeviceToHost, h.s));
      BIE_HIP_CHECK(hipEventRecord(h.ev[c & 1], h.s));
    }
    if (c > 0) {  // chunk c - 1 out of its buffer while chunk c is in flight
      const size_t off = (c - 1) * kStageBytes, len = std::min(kStageBytes, n - off);
      BIE_HIP_CHECK(hipEventSynchronize(h.ev[(c - 1) & 1]));
      parallel_copy(d + off, h.buf[(c - 1) & 1], len);
    }
  }
}

}  // namespace

int bie_assemble(const bie_mesh* mesh, double k_re, double k_im, const double* targets,
                 int64_t n_targets, int32_t n, int32_t m, uint32_t flags, double* F,
                 double* G, double* C) {
  std::shared_ptr<bie_dmesh> handle;
  int rc = host_mesh_handle(mesh, &handle);
  if (rc != BIE_OK) return rc;
  bie_dmesh* dm = handle.get();
  // targets equal to the mesh sources value for value are collocation (the reference's own
  // call passes mesh.sources as in_sources): the collocation path then applies (recorded
  // mid-edge pairs, no M x E pair queue)
  if (targets && n_targets == mesh->n_sources &&
      (n_targets == 0 || std::memcmp(targets, mesh->sources, 24 * (size_t)n_targets) == 0))
    targets = nullptr;
  double* dT = nullptr;
  double* dF = nullptr;
  double* dG = nullptr;
  double* dC = nullptr;
  rc = guarded([&] {
    BIE_REQUIRE(F && G && C, BIE_EINVAL, "output pointer is NULL");
    BIE_REQUIRE(n_targets >= 0, BIE_EINVAL, "negative number of targets");
    const int64_t M = targets ? n_targets : dm->N;
    BIE_REQUIRE(targets != nullptr || n_targets == dm->N || n_targets == 0, BIE_EINVAL,
                "collocation (NULL targets) needs n_targets == n_sources");
    const size_t fg = (size_t)2 * M * dm->N;
    dF = dalloc<double>(fg);
    dG = dalloc<double>(fg);
    dC = dalloc<double>((size_t)2 * M);
    if (targets) {
      dT = dalloc<double>((size_t)3 * M);
      if (M) BIE_HIP_CHECK(hipMemcpy(dT, targets, 24 * M, hipMemcpyHostToDevice));
    }
    assemble_device_impl(dm, k_re, k_im, targets ? dT : nullptr, M, n, m, flags, dF, dG, dC,
                         nullptr, targets ? radius_about(dm->center, targets, M) : -1.0);
    BIE_HIP_CHECK(hipDeviceSynchronize());
    if (fg) {
      copy_to_host(F, dF, fg * sizeof(double));
      copy_to_host(G, dG, fg * sizeof(double));
    }
    if (M) BIE_HIP_CHECK(hipMemcpy(C, dC, 16 * M, hipMemcpyDeviceToHost));
    if (targets) {  // the M x E pair queue of explicit targets is not kept in the cache
      std::lock_guard<std::mutex> lk(dm->mu);
      dfree(dm->d_mid_queue);
      dm->d_mid_queue = nullptr;
      dm->mid_cap = 0;
    }
  });
  dfree(dT);
  dfree(dF);
  dfree(dG);
  dfree(dC);
  return rc;
}

}  // extern "C"

// ---- sparse near field ------------------------------------------------------------------

struct bie_sparse {
  bie_dmesh* dm = nullptr;       // not owned
  int depth = 1;
  int64_t nnz = 0, n_pairs = 0;
  std::vector<int64_t> row_ptr;  // host pattern
  std::vector<int> cols;
  int2* d_pairs = nullptr;
  int* d_pos = nullptr;
  double* d_gsel = nullptr;
  double* d_rules = nullptr;
  int off[6] = {0, 0, 0, 0, 0, 0};
  int len[6] = {0, 0, 0, 0, 0, 0};
  // batches of pairs sharing an element (k_sparse_batch; 8 ints each), none when too few
  // pairs would share
  int* d_batches = nullptr;
  int64_t n_batches = 0;
  double* d_trig = nullptr;  // the sincos table of the kernels (kTrigDoubles)
  // the one rule of the partial assembly (gauss_points_triangle(partial_n)), built on demand
  double* d_partial = nullptr;
  int partial_n = -1, partial_len = 0;
};

namespace bie {
namespace {

void sparse_destroy_impl(bie_sparse* sp) {
  if (!sp) return;
  if (sp->dm) (void)hipSetDevice(sp->dm->device);
  (void)hipDeviceSynchronize();  // a launch may still read the tables
  dfree(sp->d_pairs);
  dfree(sp->d_pos);
  dfree(sp->d_gsel);
  dfree(sp->d_rules);
  dfree(sp->d_batches);
  dfree(sp->d_trig);
  dfree(sp->d_partial);
  delete sp;
}

// Pattern, pairs, selection table and subdivision rules of a sparse assembly on dm.
// Point records of a rule for the sparse kernel: (u, v, w, 0), sparse_record_doubles() each;
// the kernel evaluates the geometry and physics bases at (u, v).
void append_rule_records(std::vector<double>& rec, const std::vector<double>& u,
                         const std::vector<double>& v, const std::vector<double>& w) {
  BIE_REQUIRE(sparse_record_doubles() == 4, BIE_EINVAL, "internal: sparse record layout");
  for (size_t q = 0; q < u.size(); ++q) {
    rec.push_back(u[q]);
    rec.push_back(v[q]);
    rec.push_back(w[q]);
    rec.push_back(0.0);
  }
}

// connected_sources of either physics (continuous: shared nodes; discontinuous: the
// element's own rows, shared corners at depth 2).
void near_field_pattern(const int* topo, int gkind, const int* ptopo, int pkind, int64_t E,
                        int64_t N, int depth, std::vector<int64_t>& row_ptr,
                        std::vector<int>& cols, std::vector<int64_t>& pair_ptr,
                        std::vector<int>& pair_elem) {
  BIE_REQUIRE(!kind_is_quad(gkind), BIE_EUNSUPPORTED,
              "sparse near field: triangular elements only");
  if (physics_discontinuous(pkind))
    sparse_pattern_disc(topo, geometry_nodes(gkind), ptopo, physics_nodes(pkind), E, N, depth,
                        row_ptr, cols, pair_ptr, pair_elem);
  else
    sparse_pattern(ptopo, physics_nodes(pkind), E, N, depth, row_ptr, cols, pair_ptr,
                   pair_elem);
}

bie_sparse* sparse_create_impl(bie_dmesh* dm, int depth, double offset, int n_gauss) {
  BIE_REQUIRE(dm != nullptr, BIE_EINVAL, "dmesh is NULL");
  BIE_REQUIRE(depth == 1 || depth == 2, BIE_EINVAL, "depth must be 1 or 2");
  BIE_REQUIRE(!kind_is_quad(dm->gkind), BIE_EUNSUPPORTED,
              "sparse near field: triangular elements only");
  BIE_REQUIRE(!std::isinf(offset), BIE_EINVAL, "offset must be finite (NaN: the default 0)");
  const double offr = std::isnan(offset) ? 0.0 : offset;
  const bool disc = physics_discontinuous(dm->pkind);
  // the reference forces Ngauss = 40 for quadratic discontinuous physics
  if (dm->pkind == BIE_DTRI6) n_gauss = 40;
  BIE_REQUIRE(!disc || (n_gauss >= 1 && n_gauss <= 400), BIE_EINVAL,
              "Ngauss must be between 1 and 400");
  BIE_HIP_CHECK(hipSetDevice(dm->device));
  std::unique_ptr<bie_sparse> sp(new bie_sparse);
  sp->dm = dm;
  sp->depth = depth;
  const int NP = dm->pk, NG = dm->gk;
  std::vector<int64_t> pair_ptr;
  std::vector<int> pair_elem;
  near_field_pattern(dm->h_topo.data(), dm->gkind, dm->h_ptopo.data(), dm->pkind, dm->E, dm->N,
                     depth, sp->row_ptr, sp->cols, pair_ptr, pair_elem);
  sp->nnz = (int64_t)sp->cols.size();
  sp->n_pairs = (int64_t)pair_elem.size();
  BIE_REQUIRE(sp->nnz < (int64_t)INT32_MAX && sp->n_pairs * NP < (int64_t)INT32_MAX, BIE_EINVAL,
              "sparse pattern too large");
  std::vector<int2> pairs((size_t)sp->n_pairs);
  std::vector<int> pos((size_t)sp->n_pairs * NP);
  for (int64_t i = 0; i < dm->N; ++i)
    for (int64_t t = pair_ptr[i]; t < pair_ptr[i + 1]; ++t) {
      const int e = pair_elem[t];
      pairs[t] = make_int2((int)i, e);
      const int* c0 = sp->cols.data() + sp->row_ptr[i];
      const int* c1 = sp->cols.data() + sp->row_ptr[i + 1];
      for (int s = 0; s < NP; ++s) {
        const int col = dm->h_ptopo[(size_t)e * NP + s];
        const int* it = std::lower_bound(c0, c1, col);
        BIE_REQUIRE(it != c1 && *it == col, BIE_EINVAL, "internal: sparse column missing");
        pos[(size_t)t * NP + s] = (int)(sp->row_ptr[i] + (it - c0));
      }
    }
  // geometry node coordinates per element, 4 doubles each (the selection's layout)
  std::vector<double> gsel((size_t)dm->E * NG * 4, 0.0);
  for (int64_t e = 0; e < dm->E; ++e)
    for (int g = 0; g < NG; ++g)
      for (int d = 0; d < 3; ++d)
        gsel[((size_t)e * NG + g) * 4 + d] = dm->h_coords[3 * (size_t)dm->h_topo[e * NG + g] + d];
  // the six rules (sources at the vertices, then the edge midpoints 1-2, 2-3, 3-1;
  // assembly_collocation_losses.jl:227-232), geometry and physics bases at their points;
  // all six for linear elements too, as the reference builds them (the selection's
  // fallthrough uses the sixth)
  const double srcs[6][3] = {{1.0, offr, offr}, {offr, 1.0, offr}, {offr, offr, 1.0},
                             {0.5, 0.5, offr},  {offr, 0.5, 0.5},  {0.5, offr, 0.5}};
  // Discontinuous physics (losses.jl:233-257): the polar rules around the physics nodes;
  // the rule of the closest geometry node c is the polar rule of physics node c, and the
  // reference fills the slots beyond the node count with the rules of nodes 1, 2, 3
  // (constant physics: node 1 for all six).
  const int RS = sparse_record_doubles();
  std::vector<double> rec;
  for (int c = 0; c < 6; ++c) {
    std::vector<double> u, v, w;
    if (disc) {
      const int node = dm->pkind == BIE_DTRI1 ? 0 : (dm->pkind == BIE_DTRI3 ? c % 3 : c);
      if (node != c) {  // a filler: the same records as the slot of that node
        sp->off[c] = sp->off[node];
        sp->len[c] = sp->len[node];
        continue;
      }
      double pu, pv;
      discontinuous_node(dm->pkind, dm->beta, node, &pu, &pv);
      polar_rule(pu, pv, n_gauss, u, v, w);
    } else {
      subdivision_rule(srcs[c], u, v, w);
    }
    sp->off[c] = (int)(rec.size() / RS);
    sp->len[c] = (int)u.size();
    append_rule_records(rec, u, v, w);
  }
  // Batches for k_sparse_batch: each element's pairs ordered by the rule the mesh sources
  // select (nearest geometry node in plain squared distance: only a grouping, the kernel
  // selects every row's rule exactly), cut into runs of at most B pairs of one rule.
  std::vector<int> brec;
  int64_t nb = 0;
  {
    const int B = sparse_batch_rows(NP);
    std::vector<int64_t> first((size_t)dm->E + 1, 0);
    for (int64_t t = 0; t < sp->n_pairs; ++t) ++first[(size_t)pairs[t].y + 1];
    for (int64_t e = 0; e < dm->E; ++e) first[e + 1] += first[e];
    std::vector<int64_t> order((size_t)sp->n_pairs), fill(first.begin(), first.end() - 1);
    for (int64_t t = 0; t < sp->n_pairs; ++t) order[fill[pairs[t].y]++] = t;
    std::vector<std::pair<int, int64_t>> key;
    for (int64_t e = 0; e < dm->E; ++e) {
      key.clear();
      for (int64_t u = first[e]; u < first[e + 1]; ++u) {
        const int64_t t = order[u];
        const double* x = dm->h_sources.data() + 3 * (size_t)pairs[t].x;
        int best = 0;
        double bd = INFINITY;
        for (int g = 0; g < NG; ++g) {
          const double* y = gsel.data() + ((size_t)e * NG + g) * 4;
          const double d = (y[0] - x[0]) * (y[0] - x[0]) + (y[1] - x[1]) * (y[1] - x[1]) +
                           (y[2] - x[2]) * (y[2] - x[2]);
          if (d < bd) {
            bd = d;
            best = g;
          }
        }
        key.push_back({sp->off[best], t});  // slots sharing records share a rule
      }
      std::stable_sort(key.begin(), key.end(),
                       [](const std::pair<int, int64_t>& l, const std::pair<int, int64_t>& r) {
                         return l.first < r.first;
                       });
      size_t k = 0;
      while (k < key.size()) {
        int r8[8] = {(int)e, 0, -1, -1, -1, -1, 0, 0};
        const int rule = key[k].first;
        while (k < key.size() && r8[1] < B && key[k].first == rule)
          r8[2 + r8[1]++] = (int)key[k++].second;
        brec.insert(brec.end(), r8, r8 + 8);
        ++nb;
      }
    }
  }
  sp->d_pairs = upload_vec(pairs, nullptr);
  sp->d_pos = upload_vec(pos, nullptr);
  sp->d_gsel = upload_vec(gsel, nullptr);
  sp->d_rules = upload_vec(rec, nullptr);
  {
    std::vector<double> trig(kTrigDoubles);
    for (int j = 0; j < kTrigN; ++j) {  // (cos, sin)(j pi / 64), rounded from long double
      const long double th = (long double)j * (3.14159265358979323846264338327950288L / 64);
      trig[2 * j] = (double)cosl(th);
      trig[2 * j + 1] = (double)sinl(th);
    }
    sp->d_trig = upload_vec(trig, nullptr);
  }
  // share only where it pays: on average at least 1.5 pairs per batch (the batched kernel
  // holds B rows' sums and runs at lower occupancy)
  if (nb > 0 && (double)sp->n_pairs >= 1.5 * (double)nb) {
    sp->d_batches = upload_vec(brec, nullptr);
    sp->n_batches = nb;
  }
  BIE_HIP_CHECK(hipDeviceSynchronize());
  return sp.release();
}

void sparse_assemble_impl(bie_sparse* sp, double kr, double ki, const double* d_sources,
                          uint32_t flags, double* dF, double* dG, hipStream_t s) {
  BIE_REQUIRE(sp != nullptr, BIE_EINVAL, "sparse handle is NULL");
  BIE_REQUIRE((flags & ~(BIE_F | BIE_G)) == 0, BIE_EINVAL, "flags: BIE_F and BIE_G only");
  BIE_REQUIRE(std::isfinite(kr) && std::isfinite(ki), BIE_EINVAL, "k must be finite");
  BIE_REQUIRE(((flags & BIE_F) == 0 || dF) && ((flags & BIE_G) == 0 || dG), BIE_EINVAL,
              "output pointer is NULL");
  bie_dmesh* dm = sp->dm;
  BIE_HIP_CHECK(hipSetDevice(dm->device));
  std::lock_guard<std::mutex> lk(dm->mu);
  order_after_previous(dm, s);
  if (dF) launch_fill_zero(dF, 2 * sp->nnz, s);
  if (dG) launch_fill_zero(dG, 2 * sp->nnz, s);
  SparseLaunch L{};
  L.src = d_sources ? d_sources : dm->d_sources;
  L.pairs = sp->d_pairs;
  L.pos = sp->d_pos;
  L.n_pairs = (flags & (BIE_F | BIE_G)) ? sp->n_pairs : 0;
  L.coords = dm->d_coords;
  L.topo = dm->d_topo;
  L.gsel = sp->d_gsel;
  L.rules = sp->d_rules;
  L.trig = sp->d_trig;
  L.batches = sp->d_batches;
  L.n_batches = sp->n_batches;
  for (int c = 0; c < 6; ++c) {
    L.off[c] = sp->off[c];
    L.len[c] = sp->len[c];
  }
  L.NG = dm->gk;
  L.NP = dm->pk;
  L.hyp = dm->hyp;
  L.disc = physics_discontinuous(dm->pkind);
  L.beta = dm->beta;
  L.kr = kr;
  L.ki = ki;
  L.flags = flags;
  L.F = dF;
  L.G = dG;
  launch_sparse_pairs(L, s);
  mark_done(dm, s);
}

// The partial assembly (partial_assemble_parallel!, fast_multipole_method.jl:11-81): the
// same pattern and pairs, every pair integrated with the one symmetric rule
// gauss_points_triangle(n) - the quadrature a fast operator applies to the near field,
// which its correction subtracts.
void sparse_partial_impl(bie_sparse* sp, double kr, double ki, int n_gauss,
                         const double* d_sources, uint32_t flags, double* dF, double* dG,
                         hipStream_t s) {
  BIE_REQUIRE(sp != nullptr, BIE_EINVAL, "sparse handle is NULL");
  BIE_REQUIRE((flags & ~(BIE_F | BIE_G)) == 0, BIE_EINVAL, "flags: BIE_F and BIE_G only");
  BIE_REQUIRE(std::isfinite(kr) && std::isfinite(ki), BIE_EINVAL, "k must be finite");
  BIE_REQUIRE(((flags & BIE_F) == 0 || dF) && ((flags & BIE_G) == 0 || dG), BIE_EINVAL,
              "output pointer is NULL");
  bie_dmesh* dm = sp->dm;
  BIE_HIP_CHECK(hipSetDevice(dm->device));
  std::lock_guard<std::mutex> lk(dm->mu);
  if (sp->partial_n != n_gauss) {
    std::vector<double> u, v, w;
    BIE_REQUIRE(gauss_points_triangle(n_gauss, u, v, w), BIE_EINVAL,
                "Triangular quadrature not implemented for n=" + std::to_string(n_gauss) +
                    " (gauss_points_triangle has n = 1, 3, 4, 7, 12, 13)");
    std::vector<double> rec;
    append_rule_records(rec, u, v, w);
    if (sp->d_partial) {
      BIE_HIP_CHECK(hipDeviceSynchronize());  // a launch may still read the old rule
      dfree(sp->d_partial);
      sp->d_partial = nullptr;
      sp->partial_n = -1;
    }
    sp->d_partial = upload_vec(rec, nullptr);
    BIE_HIP_CHECK(hipDeviceSynchronize());
    sp->partial_n = n_gauss;
    sp->partial_len = (int)u.size();
  }
  order_after_previous(dm, s);
  if (dF) launch_fill_zero(dF, 2 * sp->nnz, s);
  if (dG) launch_fill_zero(dG, 2 * sp->nnz, s);
  SparseLaunch L{};
  L.src = d_sources ? d_sources : dm->d_sources;
  L.pairs = sp->d_pairs;
  L.pos = sp->d_pos;
  L.n_pairs = (flags & (BIE_F | BIE_G)) ? sp->n_pairs : 0;
  L.coords = dm->d_coords;
  L.topo = dm->d_topo;
  L.gsel = sp->d_gsel;
  L.rules = sp->d_partial;
  L.trig = sp->d_trig;
  for (int c = 0; c < 6; ++c) {  // whatever the selection, the one rule
    L.off[c] = 0;
    L.len[c] = sp->partial_len;
  }
  L.NG = dm->gk;
  L.NP = dm->pk;
  L.hyp = dm->hyp;
  L.disc = physics_discontinuous(dm->pkind);
  L.beta = dm->beta;
  L.kr = kr;
  L.ki = ki;
  L.flags = flags;
  L.F = dF;
  L.G = dG;
  launch_sparse_pairs(L, s);
  mark_done(dm, s);
}

}  // namespace
}  // namespace bie

extern "C" {

int bie_sparse_create_ex(bie_dmesh* dm, int32_t depth, double offset, int32_t n_gauss,
                         bie_sparse** out) {
  return guarded([&] {
    BIE_REQUIRE(out != nullptr, BIE_EINVAL, "out is NULL");
    *out = nullptr;
    *out = sparse_create_impl(dm, depth, offset, n_gauss);
  });
}

int bie_sparse_create(bie_dmesh* dm, int32_t depth, double offset, bie_sparse** out) {
  return bie_sparse_create_ex(dm, depth, offset, 40, out);
}

int bie_sparse_destroy(bie_sparse* sp) {
  sparse_destroy_impl(sp);
  return BIE_OK;
}

int64_t bie_sparse_nnz(const bie_sparse* sp) { return sp ? sp->nnz : -1; }

int bie_sparse_pattern(const bie_sparse* sp, int32_t index_base, int64_t* I, int64_t* J) {
  return guarded([&] {
    BIE_REQUIRE(sp && I && J, BIE_EINVAL, "NULL argument");
    BIE_REQUIRE(index_base == 0 || index_base == 1, BIE_EINVAL, "index_base must be 0 or 1");
    for (size_t i = 0; i + 1 < sp->row_ptr.size(); ++i)
      for (int64_t t = sp->row_ptr[i]; t < sp->row_ptr[i + 1]; ++t) {
        I[t] = (int64_t)i + index_base;
        J[t] = (int64_t)sp->cols[t] + index_base;
      }
  });
}

int bie_sparse_assemble_device(bie_sparse* sp, double k_re, double k_im, const double* d_sources,
                               uint32_t flags, double* dF, double* dG, void* stream) {
  return guarded([&] {
    sparse_assemble_impl(sp, k_re, k_im, d_sources, flags, dF, dG,
                         static_cast<hipStream_t>(stream));
  });
}

int bie_sparse_partial_device(bie_sparse* sp, double k_re, double k_im, int32_t n_gauss,
                              const double* d_sources, uint32_t flags, double* dF, double* dG,
                              void* stream) {
  return guarded([&] {
    sparse_partial_impl(sp, k_re, k_im, n_gauss, d_sources, flags, dF, dG,
                        static_cast<hipStream_t>(stream));
  });
}

int bie_sparse_count(const bie_mesh* mesh, int32_t depth, int64_t* nnz) {
  return guarded([&] {
    validate_mesh(mesh);
    BIE_REQUIRE(nnz != nullptr, BIE_EINVAL, "nnz is NULL");
    BIE_REQUIRE(!kind_is_quad(mesh->geometry_kind), BIE_EUNSUPPORTED,
                "sparse near field: triangular elements only");
    std::vector<int> topo = convert_topology(mesh->topology, geometry_nodes(mesh->geometry_kind),
                                             mesh->n_elements, mesh->index_base,
                                             mesh->n_coordinates, "topology");
    std::vector<int> ptopo = convert_topology(mesh->physics_topology,
                                              physics_nodes(mesh->physics_kind),
                                              mesh->n_elements, mesh->index_base,
                                              mesh->n_sources, "physics_topology");
    std::vector<int64_t> row_ptr, pair_ptr;
    std::vector<int> cols, pair_elem;
    near_field_pattern(topo.data(), mesh->geometry_kind, ptopo.data(), mesh->physics_kind,
                       mesh->n_elements, mesh->n_sources, depth, row_ptr, cols, pair_ptr,
                       pair_elem);
    *nnz = (int64_t)cols.size();
  });
}

int bie_sparse_assemble(const bie_mesh* mesh, double k_re, double k_im, const double* sources,
                        int64_t n_sources, int32_t depth, double offset, uint32_t flags,
                        int64_t nnz, int64_t* I, int64_t* J, double* F, double* G) {
  return bie_sparse_assemble_ex(mesh, k_re, k_im, sources, n_sources, depth, offset, 40, flags,
                                nnz, I, J, F, G);
}

int bie_sparse_assemble_ex(const bie_mesh* mesh, double k_re, double k_im,
                           const double* sources, int64_t n_sources, int32_t depth,
                           double offset, int32_t n_gauss, uint32_t flags, int64_t nnz,
                           int64_t* I, int64_t* J, double* F, double* G) {
  std::shared_ptr<bie_dmesh> handle;
  int rc = host_mesh_handle(mesh, &handle);
  if (rc != BIE_OK) return rc;
  bie_dmesh* dm = handle.get();
  bie_sparse* sp = nullptr;
  double* dT = nullptr;
  double* dF = nullptr;
  double* dG = nullptr;
  rc = guarded([&] {
    BIE_REQUIRE(I && J && F && G, BIE_EINVAL, "output pointer is NULL");
    BIE_REQUIRE(sources == nullptr || n_sources == dm->N, BIE_EINVAL,
                "the sparse assembly needs one source per mesh source (n_sources)");
    sp = sparse_create_impl(dm, depth, offset, n_gauss);
    BIE_REQUIRE(nnz == sp->nnz, BIE_EINVAL,
                "nnz does not match the pattern (bie_sparse_count gives " +
                    std::to_string(sp->nnz) + ")");
    dF = dalloc<double>((size_t)2 * nnz);
    dG = dalloc<double>((size_t)2 * nnz);
    if (sources) {
      dT = dalloc<double>((size_t)3 * n_sources);
      if (n_sources)
        BIE_HIP_CHECK(hipMemcpy(dT, sources, 24 * n_sources, hipMemcpyHostToDevice));
    }
    sparse_assemble_impl(sp, k_re, k_im, dT, flags, dF, dG, nullptr);
    BIE_HIP_CHECK(hipDeviceSynchronize());
    if (nnz) {
      BIE_HIP_CHECK(hipMemcpy(F, dF, 16 * nnz, hipMemcpyDeviceToHost));
      BIE_HIP_CHECK(hipMemcpy(G, dG, 16 * nnz, hipMemcpyDeviceToHost));
    }
    const int base = mesh->index_base;
    for (size_t i = 0; i + 1 < sp->row_ptr.size(); ++i)
      for (int64_t t = sp->row_ptr[i]; t < sp->row_ptr[i + 1]; ++t) {
        I[t] = (int64_t)i + base;
        J[t] = (int64_t)sp->cols[t] + base;
      }
  });
  sparse_destroy_impl(sp);
  dfree(dT);
  dfree(dF);
  dfree(dG);
  return rc;
}

int bie_partial_assemble(const bie_mesh* mesh, double k_re, double k_im, const double* sources,
                         int64_t n_sources, int32_t depth, int32_t n_gauss, uint32_t flags,
                         int64_t nnz, int64_t* I, int64_t* J, double* F, double* G) {
  std::shared_ptr<bie_dmesh> handle;
  int rc = host_mesh_handle(mesh, &handle);
  if (rc != BIE_OK) return rc;
  bie_dmesh* dm = handle.get();
  bie_sparse* sp = nullptr;
  double* dT = nullptr;
  double* dF = nullptr;
  double* dG = nullptr;
  rc = guarded([&] {
    BIE_REQUIRE(I && J && F && G, BIE_EINVAL, "output pointer is NULL");
    BIE_REQUIRE(sources == nullptr || n_sources == dm->N, BIE_EINVAL,
                "the partial assembly needs one source per mesh source (n_sources)");
    sp = sparse_create_impl(dm, depth, NAN, 40);
    BIE_REQUIRE(nnz == sp->nnz, BIE_EINVAL,
                "nnz does not match the pattern (bie_sparse_count gives " +
                    std::to_string(sp->nnz) + ")");
    dF = dalloc<double>((size_t)2 * nnz);
    dG = dalloc<double>((size_t)2 * nnz);
    if (sources) {
      dT = dalloc<double>((size_t)3 * n_sources);
      if (n_sources)
        BIE_HIP_CHECK(hipMemcpy(dT, sources, 24 * n_sources, hipMemcpyHostToDevice));
    }
    sparse_partial_impl(sp, k_re, k_im, n_gauss, dT, flags, dF, dG, nullptr);
    BIE_HIP_CHECK(hipDeviceSynchronize());
    if (nnz) {
      BIE_HIP_CHECK(hipMemcpy(F, dF, 16 * nnz, hipMemcpyDeviceToHost));
      BIE_HIP_CHECK(hipMemcpy(G, dG, 16 * nnz, hipMemcpyDeviceToHost));
    }
    const int base = mesh->index_base;
    for (size_t i = 0; i + 1 < sp->row_ptr.size(); ++i)
      for (int64_t t = sp->row_ptr[i]; t < sp->row_ptr[i + 1]; ++t) {
        I[t] = (int64_t)i + base;
        J[t] = (int64_t)sp->cols[t] + base;
      }
  });
  sparse_destroy_impl(sp);
  dfree(dT);
  dfree(dF);
  dfree(dG);
  return rc;
}

int bie_host_cache_info(int64_t* hits, int64_t* builds, int64_t* n_sources,
                        int64_t* queue_entries) {
  return guarded([&] {
    HostMeshCache& c = host_mesh_cache();
    std::lock_guard<std::mutex> lk(c.mu);
    if (hits) *hits = c.hits;
    if (builds) *builds = c.builds;
    if (n_sources) *n_sources = c.dm ? c.dm->N : -1;
    if (queue_entries) {
      *queue_entries = 0;
      if (c.dm) {
        std::lock_guard<std::mutex> dl(c.dm->mu);
        *queue_entries = c.dm->d_mid_queue ? c.dm->mid_cap : 0;
      }
    }
  });
}

int bie_host_cache_release(void) {
  return guarded([&] {
    {
      HostMeshCache& c = host_mesh_cache();
      std::lock_guard<std::mutex> lk(c.mu);
      c.dm.reset();  // destroyed now, or when a call still using it returns
      c.key.clear();
    }
    HostStaging& h = host_staging();
    std::lock_guard<std::mutex> lk(h.mu);
    if (h.s) BIE_HIP_CHECK(hipStreamSynchronize(h.s));
    for (int i = 0; i < 2; ++i) {
      if (h.buf[i]) (void)hipHostFree(h.buf[i]);
      if (h.ev[i]) (void)hipEventDestroy(h.ev[i]);
      h.buf[i] = nullptr;
      h.ev[i] = nullptr;
    }
    if (h.s) (void)hipStreamDestroy(h.s);
    h.s = nullptr;
    h.tried = false;  // the next large copy allocates them again
  });
}

int bie_schedule_check(const bie_mesh* mesh, int32_t min_segments, int64_t* stats) {
  return guarded([&] {
    validate_mesh(mesh);
    BIE_REQUIRE(min_segments >= 1, BIE_EINVAL, "min_segments must be >= 1");
    const int64_t base = mesh->index_base;
    const int ng = geometry_nodes(mesh->geometry_kind);
    std::vector<int> topo = convert_topology(mesh->topology, ng, mesh->n_elements, base,
                                             mesh->n_coordinates, "topology");
    const int np = physics_nodes(mesh->physics_kind);
    std::vector<int> ptopo = convert_topology(mesh->physics_topology, np,
                                              mesh->n_elements, base, mesh->n_sources,
                                              "physics_topology");
    const int64_t E = mesh->n_elements;
    std::vector<double> cent((size_t)E * 3);
    for (int64_t e = 0; e < E; ++e)
      for (int d = 0; d < 3; ++d) {
        double acc = 0;
        for (int k = 0; k < 3; ++k)
          acc += mesh->coordinates[3 * (size_t)topo[e * ng + k] + d];
        cent[3 * e + d] = acc / 3.0;
      }
    SlotSchedule sc = build_slot_schedule(ptopo.data(), np, (int)E,
                                          (int)mesh->n_sources, cent.data(), kSlotWaves,
                                          kSlotCount, min_segments);
    validate_slot_schedule(sc, ptopo.data(), np, (int)E, (int)mesh->n_sources);
    // the schedule of two elements per wave and step as well
    SlotSchedule sc2 = build_slot_schedule(ptopo.data(), np, (int)E,
                                           (int)mesh->n_sources, cent.data(), 2 * kSlotWaves,
                                           kSlotCount, min_segments);
    validate_slot_schedule(sc2, ptopo.data(), np, (int)E, (int)mesh->n_sources);
    if (stats) {
      stats[0] = sc.n_segments;
      stats[1] = sc.n_steps;
      stats[2] = sc.max_live;
      stats[3] = sc.n_shared;
      stats[4] = (int64_t)sc.zero_cols.size();
      stats[5] = (int64_t)sc.flush_node.size();
    }
  });
}

namespace {

void assembly_vjp_impl(bie_dmesh* dm, double k, double ki, bool cplx, const double* d_targets,
                       int64_t n_targets, const int32_t* d_target_node, const double* d_alpha,
                       const double* d_beta, const double* d_p, int32_t n, int32_t m,
                       double* d_grad_coords, double* d_grad_k, hipStream_t s) {
  BIE_REQUIRE(dm != nullptr, BIE_EINVAL, "dmesh is NULL");
  BIE_REQUIRE(n_targets >= 0, BIE_EINVAL, "negative number of targets");
  BIE_REQUIRE(d_alpha && d_p && d_grad_coords, BIE_EINVAL, "NULL argument");
  BIE_REQUIRE(std::isfinite(k) && std::isfinite(ki), BIE_EINVAL, "k must be finite");
  BIE_REQUIRE(!physics_discontinuous(dm->pkind) && !kind_is_quad(dm->gkind), BIE_EUNSUPPORTED,
              "assembly adjoint: continuous physics on triangles only");
  if (!d_targets) {
    d_targets = dm->d_sources;
    BIE_REQUIRE(n_targets == dm->N, BIE_EINVAL, "collocation needs n_targets == n_sources");
  }
  std::lock_guard<std::mutex> lk(dm->mu);
  BIE_HIP_CHECK(hipSetDevice(dm->device));
  order_after_previous(dm, s);
  ensure_rules(dm, n, m, s);
  AdjointArgs a{};
  a.targets = d_targets;
  a.M = n_targets;
  a.tvert = reinterpret_cast<const int*>(d_target_node);
  a.alpha = reinterpret_cast<const double2*>(d_alpha);
  a.beta = reinterpret_cast<const double2*>(d_beta);
  a.p = reinterpret_cast<const double2*>(d_p);
  a.coords = dm->d_coords;
  a.topo = dm->d_topo;
  a.ptopo = dm->d_ptopo;
  a.selnodes = dm->d_selnodes;
  a.pts = dm->d_pts;
  a.phi = dm->d_phi;
  a.Qmax = dm->rs.Qmax;
  for (int c = 0; c < 6; ++c) {
    a.off[c] = dm->rs.off[c];
    a.len[c] = dm->rs.len[c];
  }
  a.R = dm->rs.R;
  a.E = (int)dm->E;
  a.hyp = dm->hyp;
  a.k = k;
  a.ki = ki;
  a.gradX = d_grad_coords;
  a.gradk = d_grad_k;
  launch_adjoint(a, dm->gk, dm->pk, cplx, s);
  mark_done(dm, s);
}

}  // namespace

int bie_assembly_vjp_device(bie_dmesh* dm, double k, const double* d_targets, int64_t n_targets,
                            const int32_t* d_target_node, const double* d_alpha,
                            const double* d_beta, const double* d_p, int32_t n, int32_t m,
                            double* d_grad_coords, double* d_grad_k, void* stream) {
  return guarded([&] {
    assembly_vjp_impl(dm, k, 0.0, false, d_targets, n_targets, d_target_node, d_alpha, d_beta,
                      d_p, n, m, d_grad_coords, d_grad_k, static_cast<hipStream_t>(stream));
  });
}

int bie_assembly_vjp_complex_device(bie_dmesh* dm, double k_re, double k_im,
                                    const double* d_targets, int64_t n_targets,
                                    const int32_t* d_target_node, const double* d_alpha,
                                    const double* d_beta, const double* d_p, int32_t n,
                                    int32_t m, double* d_grad_coords, double* d_grad_k,
                                    void* stream) {
  return guarded([&] {
    assembly_vjp_impl(dm, k_re, k_im, true, d_targets, n_targets, d_target_node, d_alpha,
                      d_beta, d_p, n, m, d_grad_coords, d_grad_k,
                      static_cast<hipStream_t>(stream));
  });
}

int bie_zgemm_device(int64_t M, int64_t nrhs, int64_t K, const double* dA, int64_t lda,
                     const double* dX, int64_t ldx, double* dY, int64_t ldy, void* stream) {
  return guarded([&] {
    zgemm_device(M, nrhs, K, dA, lda, dX, ldx, dY, ldy, static_cast<hipStream_t>(stream));
  });
}

int bie_lu_factor_device(int64_t n, double* dA, int64_t lda, const double* dC, int32_t* d_ipiv,
                         void* stream) {
  return guarded([&] { lu_factor_device(n, dA, lda, dC, d_ipiv, static_cast<hipStream_t>(stream)); });
}

int bie_lu_solve_device(int64_t n, const double* dA, int64_t lda, const int32_t* d_ipiv,
                        int32_t trans, double* dB, int64_t ldb, int64_t nrhs, void* stream) {
  return guarded([&] {
    lu_solve_device(n, dA, lda, d_ipiv, trans, dB, ldb, nrhs, static_cast<hipStream_t>(stream));
  });
}

int bie_zgemv_device(int32_t trans, int64_t m, int64_t n, const double* dA, int64_t lda,
                     const double* dx, double* dy, void* stream) {
  return guarded([&] { zgemv_device(trans, m, n, dA, lda, dx, dy, static_cast<hipStream_t>(stream)); });
}

int bie_gmres_batched_device(int64_t n, int32_t n_systems, const double* dF, int64_t strideF,
                             int64_t ldf, const double* dC, const double* dB, int64_t ldb,
                             double* dX, int64_t ldx, double tol, int32_t restart,
                             int32_t maxiter, int32_t* iterations, double* residuals,
                             void* stream) {
  return guarded([&] {
    gmres_batched_device(n, n_systems, dF, strideF, ldf, dC, dB, ldb, dX, ldx, tol, restart,
                         maxiter, iterations, residuals, static_cast<hipStream_t>(stream));
  });
}

int bie_zgemm_device_noalloc(int64_t M, int64_t nrhs, int64_t K, const double* dA,
                             int64_t lda, const double* dX, int64_t ldx, double* dY,
                             int64_t ldy, void* stream) {
  return guarded([&] {
    zgemm_device_noalloc(M, nrhs, K, dA, lda, dX, ldx, dY, ldy, static_cast<hipStream_t>(stream));
  });
}

int bie_zgemv_batched_device(int64_t M, int64_t K, const double* dA, int64_t strideA,
                             int64_t lda, const double* dX, int64_t ldx, double* dY, int64_t ldy,
                             const int32_t* d_idx, int32_t n_idx, void* stream) {
  return guarded([&] {
    zgemv_batched_device(M, K, dA, strideA, lda, dX, ldx, dY, ldy, d_idx, n_idx,
                         static_cast<hipStream_t>(stream));
  });
}

int bie_assemble_sweep_device(bie_dmesh* dm, const double* d_k, int32_t nk,
                              const double* d_targets, int64_t n_targets, int32_t n, int32_t m,
                              uint32_t flags, double* dF, double* dG, int64_t stride,
                              double* dC, void* stream) {
  return guarded([&] {
    BIE_REQUIRE(dm != nullptr, BIE_EINVAL, "dmesh is NULL");
    std::lock_guard<std::mutex> lk(dm->mu);
    sweep_device_impl(dm, d_k, nk, d_targets, n_targets, n, m, flags, dF, dG, stride, dC,
                      static_cast<hipStream_t>(stream), false);
  });
}

int bie_sweep_tune(bie_dmesh* dm, const double* d_k, int32_t nk, const double* d_targets,
                   int64_t n_targets, int32_t n, int32_t m, uint32_t flags, double* dF,
                   double* dG, int64_t stride, double* dC, void* stream,
                   int32_t* elems_per_tile, int32_t* slots, double* best_ms) {
  return guarded([&] {
    BIE_REQUIRE(dm != nullptr, BIE_EINVAL, "dmesh is NULL");
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lk(dm->mu);
    BIE_HIP_CHECK(hipSetDevice(dm->device));
    hipEvent_t e0, e1;
    BIE_HIP_CHECK(hipEventCreate(&e0));
    BIE_HIP_CHECK(hipEventCreate(&e1));
    // tile = 64 targets (one per lane) x elems_per_tile elements (0: the whole surface, one
    // segment per row block), with as many LDS slots as fit or 32
    const int cand_elems[] = {16, 32, 64, 128, 0};
    const int cand_slots[] = {0, 32};
    const int old_e = dm->sweep_elems_per_tile, old_s = dm->sweep_slots;
    double best = 1e300;
    int be = old_e, bs = old_s;
    try {
      for (int ce : cand_elems)
        for (int cs : cand_slots) {
          dm->sweep_elems_per_tile = ce;
          dm->sweep_slots = cs;
          // builds the schedule and warms up, then one timed sweep
          sweep_device_impl(dm, d_k, nk, d_targets, n_targets, n, m, flags, dF, dG, stride, dC,
                            s, false);
          BIE_HIP_CHECK(hipEventRecord(e0, s));
          sweep_device_impl(dm, d_k, nk, d_targets, n_targets, n, m, flags, dF, dG, stride, dC,
                            s, false);
          BIE_HIP_CHECK(hipEventRecord(e1, s));
          BIE_HIP_CHECK(hipEventSynchronize(e1));
          float ms = 0.f;
          BIE_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
          if (ms < best) {
            best = ms;
            be = ce;
            bs = cs;
          }
        }
    } catch (...) {
      dm->sweep_elems_per_tile = old_e;
      dm->sweep_slots = old_s;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    dm->sweep_elems_per_tile = be;
    dm->sweep_slots = bs;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (elems_per_tile) *elems_per_tile = be;
    if (slots) *slots = bs;
    if (best_ms) *best_ms = best;
  });
}

int bie_assemble_taylor_device(bie_dmesh* dm, double k_re, double k_im, int32_t n_orders,
                               const double* d_targets, int64_t n_targets, int32_t n, int32_t m,
                               uint32_t flags, double* dF, double* dG, int64_t stride,
                               double* dC, void* stream) {
  return guarded([&] {
    BIE_REQUIRE(dm != nullptr, BIE_EINVAL, "dmesh is NULL");
    BIE_REQUIRE(n_orders >= 0 && n_orders <= kTaylorMaxOrders, BIE_EINVAL,
                "n_orders must be in 0..64");
    std::lock_guard<std::mutex> lk(dm->mu);
    sweep_device_impl(dm, nullptr, n_orders, d_targets, n_targets, n, m, flags, dF, dG, stride,
                      dC, static_cast<hipStream_t>(stream), false, true, k_re, k_im);
  });
}

int bie_taylor_expand_device(int32_t n_terms, int64_t n, const double* dA, int64_t strideA,
                             const double* coef, double* dY, void* stream) {
  return guarded([&] {
    taylor_expand_device(n_terms, n, dA, strideA, coef, dY, static_cast<hipStream_t>(stream));
  });
}

struct bie_graph {
  int device = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<void*> owned;  // device buffers the captured kernels read (freed with the graph)
};

int bie_sweep_graph_create(bie_dmesh* dm, const double* d_k, int32_t nk,
                           const double* d_targets, int64_t n_targets, int32_t n, int32_t m,
                           uint32_t flags, double* dF, double* dG, int64_t stride, double* dC,
                           bie_graph** out) {
  return guarded([&] {
    BIE_REQUIRE(dm != nullptr && out != nullptr, BIE_EINVAL, "NULL argument");
    *out = nullptr;
    std::lock_guard<std::mutex> lk(dm->mu);
    BIE_HIP_CHECK(hipSetDevice(dm->device));
    hipStream_t cs = nullptr;
    BIE_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    std::unique_ptr<bie_graph> g(new bie_graph);
    g->device = dm->device;
    try {
      // build every cache and set kernel attributes outside the capture
      sweep_device_impl(dm, d_k, nk, d_targets, n_targets, n, m, flags, dF, dG, stride, dC, cs,
                        true);
      sweep_prepare_attributes(dm->pk, dm->rs.R * dm->rs.Qmax * dm->pk, false,
                               (flags & BIE_G) != 0);
      BIE_HIP_CHECK(hipStreamSynchronize(cs));
      BIE_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      try {
        sweep_device_impl(dm, d_k, nk, d_targets, n_targets, n, m, flags, dF, dG, stride, dC,
                          cs, false);
      } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(cs, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
      }
      BIE_HIP_CHECK(hipStreamEndCapture(cs, &g->graph));
      BIE_HIP_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    } catch (...) {
      (void)hipStreamDestroy(cs);
      if (g->exec) (void)hipGraphExecDestroy(g->exec);
      if (g->graph) (void)hipGraphDestroy(g->graph);
      throw;
    }
    (void)hipStreamDestroy(cs);
    *out = g.release();
  });
}

int bie_graph_launch(bie_graph* g, void* stream) {
  return guarded([&] {
    BIE_REQUIRE(g != nullptr, BIE_EINVAL, "graph is NULL");
    BIE_HIP_CHECK(hipSetDevice(g->device));
    BIE_HIP_CHECK(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
  });
}

int bie_graph_destroy(bie_graph* g) {
  if (!g) return BIE_OK;
  (void)hipSetDevice(g->device);
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (!g->owned.empty()) (void)hipDeviceSynchronize();  // a replay may still read them
  for (void* p : g->owned) (void)hipFree(p);
  delete g;
  return BIE_OK;
}

int bie_batch_graph_create(bie_batch* b, const double* k_re, const double* k_im,
                           uint32_t flags, double* const* dF, double* const* dG,
                           double* const* dC, bie_graph** out) {
  return guarded([&] {
    BIE_REQUIRE(out != nullptr, BIE_EINVAL, "out is NULL");
    *out = nullptr;
    batch_validate(b, k_re, k_im, flags, dF, dG, dC);
    const int n_inst = (int)b->meshes.size();
    BIE_HIP_CHECK(hipSetDevice(b->device));
    std::lock_guard<std::mutex> bl(b->mu);
    MeshLocks locks(b->meshes);
    hipStream_t cs = nullptr;
    BIE_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    std::unique_ptr<bie_graph> g(new bie_graph);
    g->device = b->device;
    try {
      // every cache, the tile list and the instance table outside the capture; the graph
      // gets its own copies of the table, the tile list and the tile counter
      for (bie_dmesh* dm : b->meshes) order_after_previous(dm, cs);
      const BatchPlan plan = batch_fill(b, k_re, k_im, flags, dF, dG, dC, cs);
      // quadratic instances queue their mid-edge pairs: into queues of the graph's own (a
      // mesh's queue is released once its collocation pairs are recorded, and regrown by
      // calls with more targets)
      for (int ii = 0; ii < n_inst; ++ii) {
        bie_dmesh* dm = b->meshes[ii];
        if (!uses_mid_rules(dm->gk, dm->pk)) continue;
        const int64_t cap = std::max<int64_t>(1, dm->N * dm->E);
        int2* q = dalloc<int2>((size_t)cap);
        g->owned.push_back(q);
        unsigned* cnt = dalloc<unsigned>(1);
        g->owned.push_back(cnt);
        b->h_inst[ii].sa.mid_queue = q;
        b->h_inst[ii].sa.mid_count = cnt;
        b->h_inst[ii].sa.mid_cap = cap;
      }
      BatchInst* g_inst = dalloc<BatchInst>(std::max(n_inst, 1));
      g->owned.push_back(g_inst);
      int4* g_tiles = dalloc<int4>(std::max(b->n_tiles, 1));
      g->owned.push_back(g_tiles);
      unsigned* g_counter = dalloc<unsigned>(kXcds);
      g->owned.push_back(g_counter);
      BIE_HIP_CHECK(hipMemcpyAsync(g_inst, b->h_inst, sizeof(BatchInst) * n_inst,
                                   hipMemcpyHostToDevice, cs));
      if (b->n_tiles > 0)
        BIE_HIP_CHECK(hipMemcpyAsync(g_tiles, b->d_tiles, sizeof(int4) * b->n_tiles,
                                     hipMemcpyDeviceToDevice, cs));
      // one eager run sets the kernels' attributes (LDS sizes) outside the capture
      batch_kernels(b, g_inst, g_tiles, g_counter, plan, cs);
      BIE_HIP_CHECK(hipStreamSynchronize(cs));
      BIE_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      try {
        batch_kernels(b, g_inst, g_tiles, g_counter, plan, cs);
      } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(cs, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
      }
      BIE_HIP_CHECK(hipStreamEndCapture(cs, &g->graph));
      BIE_HIP_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    } catch (...) {
      (void)hipStreamDestroy(cs);
      if (g->exec) (void)hipGraphExecDestroy(g->exec);
      if (g->graph) (void)hipGraphDestroy(g->graph);
      for (void* p : g->owned) (void)hipFree(p);
      g->owned.clear();
      throw;
    }
    (void)hipStreamDestroy(cs);
    *out = g.release();
  });
}

int bie_batch_create(bie_dmesh* const* meshes, int32_t n_inst, int32_t n, int32_t m,
                     bie_batch** out) {
  return guarded([&] {
    BIE_REQUIRE(out != nullptr, BIE_EINVAL, "out is NULL");
    *out = nullptr;
    BIE_REQUIRE(meshes != nullptr && n_inst >= 1, BIE_EINVAL, "empty batch");
    BIE_REQUIRE(n_inst <= 65535, BIE_EINVAL, "too many batch instances");
    BIE_REQUIRE(n >= 1 && m >= 1, BIE_EINVAL, "quadrature orders must be positive");
    for (int ii = 0; ii < n_inst; ++ii) {
      BIE_REQUIRE(meshes[ii] != nullptr, BIE_EINVAL, "batch mesh is NULL");
      BIE_REQUIRE(meshes[ii]->device == meshes[0]->device, BIE_EINVAL,
                  "batch meshes live on different devices");
      BIE_REQUIRE(!kind_is_quad(meshes[ii]->gkind), BIE_EUNSUPPORTED,
                  "batches: triangular meshes only");
      BIE_REQUIRE(meshes[ii]->gk == meshes[0]->gk && meshes[ii]->pkind == meshes[0]->pkind &&
                      meshes[ii]->beta == meshes[0]->beta,
                  BIE_EUNSUPPORTED, "batch meshes must share geometry and physics kinds");
    }
    // one mid-edge queue per mesh: each may appear once
    if (uses_mid_rules(meshes[0]->gk, meshes[0]->pk)) {
      std::vector<bie_dmesh*> ms(meshes, meshes + n_inst);
      std::sort(ms.begin(), ms.end());
      BIE_REQUIRE(std::adjacent_find(ms.begin(), ms.end()) == ms.end(), BIE_EINVAL,
                  "a quadratic-physics mesh may appear only once in a batch");
    }
    {
      RuleSet rs = build_ruleset(meshes[0]->gkind, meshes[0]->pkind, meshes[0]->beta, n, m);
      BIE_REQUIRE(sizeof(double) * (size_t)rs.R * rs.Qmax * rs.pk <= 24 * 1024, BIE_EUNSUPPORTED,
                  "batch supports quadrature orders up to n*m' = 36");
    }
    BIE_HIP_CHECK(hipSetDevice(meshes[0]->device));
    std::unique_ptr<bie_batch> b(new bie_batch);
    b->device = meshes[0]->device;
    b->gk = meshes[0]->gk;
    b->pk = meshes[0]->pk;
    b->pkind = meshes[0]->pkind;
    b->beta = meshes[0]->beta;
    b->n = n;
    b->m = m;
    b->meshes.assign(meshes, meshes + n_inst);
    try {
      void* h = nullptr;
      BIE_HIP_CHECK(hipHostMalloc(&h, sizeof(BatchInst) * n_inst, hipHostMallocDefault));
      b->h_inst = static_cast<BatchInst*>(h);
      b->d_inst = dalloc<BatchInst>(n_inst);
      b->d_counter = dalloc<unsigned>(kXcds);
      BIE_HIP_CHECK(hipEventCreateWithFlags(&b->staged, hipEventDisableTiming));
      BIE_HIP_CHECK(hipEventRecord(b->staged, nullptr));
    } catch (...) {
      batch_destroy_impl(b.release());
      throw;
    }
    *out = b.release();
  });
}

int bie_batch_destroy(bie_batch* b) {
  if (!b) return BIE_OK;
  batch_destroy_impl(b);
  return BIE_OK;
}

int bie_batch_configure(bie_batch* b, int32_t elems_per_tile, int32_t slots, int32_t order) {
  return guarded([&] {
    BIE_REQUIRE(b != nullptr, BIE_EINVAL, "batch is NULL");
    BIE_REQUIRE(elems_per_tile >= kSlotWaves && elems_per_tile <= (1 << 20), BIE_EINVAL,
                "elements per tile out of range");
    BIE_REQUIRE(slots == 0 || slots == 32 || slots == 64, BIE_EINVAL,
                "slots must be 32, 64 or 0 (as many as fit)");
    BIE_REQUIRE(order == 0 || order == 1, BIE_EINVAL, "order must be 0 or 1");
    std::lock_guard<std::mutex> bl(b->mu);
    b->elems_per_tile = elems_per_tile;
    b->slots = slots;
    b->order = order;
  });
}

int bie_batch_info(bie_batch* b, int32_t* elems_per_tile, int32_t* slots, int32_t* order,
                   int64_t* n_tiles) {
  return guarded([&] {
    BIE_REQUIRE(b != nullptr, BIE_EINVAL, "batch is NULL");
    std::lock_guard<std::mutex> bl(b->mu);
    if (elems_per_tile) *elems_per_tile = b->elems_per_tile;
    if (slots) *slots = b->slots;
    if (order) *order = b->order;
    if (n_tiles) *n_tiles = b->configured_key >= 0 ? b->n_tiles : -1;
  });
}

int bie_batch_assemble(bie_batch* b, const double* k_re, const double* k_im, uint32_t flags,
                       double* const* dF, double* const* dG, double* const* dC, void* stream) {
  return guarded([&] {
    batch_assemble_impl(b, k_re, k_im, flags, dF, dG, dC, static_cast<hipStream_t>(stream));
  });
}

int bie_batch_dfdk_device(bie_batch* b, const double* k_re, const double* k_im,
                          const double* const* d_alpha, const double* const* d_p, double* d_w,
                          void* stream) {
  return guarded([&] {
    batch_dfdk_impl(b, k_re, k_im, d_alpha, d_p, d_w, static_cast<hipStream_t>(stream));
  });
}

int bie_batch_tune(bie_batch* b, const double* k_re, const double* k_im, uint32_t flags,
                   double* const* dF, double* const* dG, double* const* dC, void* stream,
                   double* best_ms) {
  return guarded([&] {
    BIE_REQUIRE(b != nullptr, BIE_EINVAL, "batch is NULL");
    hipStream_t s = static_cast<hipStream_t>(stream);
    BIE_HIP_CHECK(hipSetDevice(b->device));
    hipEvent_t e0, e1;
    BIE_HIP_CHECK(hipEventCreate(&e0));
    BIE_HIP_CHECK(hipEventCreate(&e1));
    struct Cand {
      int T, S, order;
    };
    std::vector<Cand> cands;
    for (int T : {32, 64, 128})
      for (int S : {32, 0})
        for (int o : {0, 1}) cands.push_back(Cand{T, S, o});
    double best = INFINITY;
    Cand pick = cands[0];
    try {
      for (const Cand& c : cands) {
        {
          std::lock_guard<std::mutex> bl(b->mu);
          b->elems_per_tile = c.T;
          b->slots = c.S;
          b->order = c.order;
        }
        batch_assemble_impl(b, k_re, k_im, flags, dF, dG, dC, s);  // builds, warms up
        BIE_HIP_CHECK(hipEventRecord(e0, s));
        for (int r = 0; r < 2; ++r) batch_assemble_impl(b, k_re, k_im, flags, dF, dG, dC, s);
        BIE_HIP_CHECK(hipEventRecord(e1, s));
        BIE_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        BIE_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms / 2.0 < best) {
          best = ms / 2.0;
          pick = c;
        }
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    {
      std::lock_guard<std::mutex> bl(b->mu);
      b->elems_per_tile = pick.T;
      b->slots = pick.S;
      b->order = pick.order;
    }
    batch_assemble_impl(b, k_re, k_im, flags, dF, dG, dC, s);  // outputs of the pick
    if (best_ms) *best_ms = best;
  });
}

int bie_timing_enable(int on) {
  return guarded([&] {
    std::lock_guard<std::mutex> tl(g_timing.mu);
    g_timing.on = on != 0;
  });
}

int bie_timing_collect(double* kernel_ms, double* mid_ms, double* call_ms, int64_t* n_calls) {
  return guarded([&] {
    std::vector<std::array<hipEvent_t, 4>> rec;
    {
      std::lock_guard<std::mutex> tl(g_timing.mu);
      rec.swap(g_timing.rec);
    }
    double ks = 0.0, ms = 0.0, cs = 0.0;
    for (auto& ev : rec) {
      BIE_HIP_CHECK(hipEventSynchronize(ev[3]));
      float a = 0.f, b = 0.f, c = 0.f;
      BIE_HIP_CHECK(hipEventElapsedTime(&a, ev[1], ev[2]));
      BIE_HIP_CHECK(hipEventElapsedTime(&b, ev[2], ev[3]));
      BIE_HIP_CHECK(hipEventElapsedTime(&c, ev[0], ev[3]));
      ks += a;
      ms += b;
      cs += c;
    }
    {
      std::lock_guard<std::mutex> tl(g_timing.mu);
      for (auto& ev : rec)
        for (auto e : ev) g_timing.pool.push_back(e);
    }
    if (kernel_ms) *kernel_ms = ks;
    if (mid_ms) *mid_ms = ms;
    if (call_ms) *call_ms = cs;
    if (n_calls) *n_calls = (int64_t)rec.size();
  });
}

int bie_device_malloc(void** ptr, int64_t bytes) {
  return guarded([&] {
    BIE_REQUIRE(ptr && bytes >= 0, BIE_EINVAL, "bad argument");
    *ptr = dalloc<char>((size_t)bytes);
  });
}

int bie_device_free(void* ptr) {
  return guarded([&] { dfree(ptr); });
}

int bie_memcpy_h2d(void* dst, const void* src, int64_t bytes) {
  return guarded([&] {
    if (bytes > 0) BIE_HIP_CHECK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
  });
}

int bie_memcpy_d2h(void* dst, const void* src, int64_t bytes) {
  return guarded([&] {
    if (bytes > 0) BIE_HIP_CHECK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
  });
}

int bie_device_synchronize(void) {
  return guarded([&] { BIE_HIP_CHECK(hipDeviceSynchronize()); });
}

}  // extern "C"
