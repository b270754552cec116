// Host scheduler for the LDS-accumulator assembly kernel.
//
// The dense assembly writes F[i, j] = sum over elements e incident to node j of that
// element's contribution for target i.  A workgroup owns 64 target rows and one
// *segment* of elements; it sweeps the segment in steps of W elements (one per wave) and
// accumulates node contributions in LDS *slots*.  A node occupies a slot from the step
// of its first incident element in the segment to the step of its last; the kernel
// flushes it to F and G during the following step (overlapped with that step's
// quadrature, one barrier per step), so the slot becomes free for a new node one step
// after that: a node whose last step is t holds its slot through step t + 1 and the slot
// is reused at step t + 2 at the earliest.  Nodes whose elements all lie in one segment
// ("owned") are written once with a plain store; nodes touched by several segments
// ("shared") are added atomically into columns zeroed beforehand.
//
// Segments come from recursive coordinate splitting of the element centroids into the
// requested number of parts; each is swept along its longest axis, which keeps the number
// of occupied slots (the sweep front) near the segment's width.  Segments are split
// further until their front fits the slot budget.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "schedule.hpp"

namespace bie {
namespace {

struct SegOrder {
  std::vector<int> elems;  // sweep order
};

int longest_axis(const std::vector<int>& idx, const double* cent) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int e : idx)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], cent[3 * (size_t)e + d]);
      hi[d] = std::max(hi[d], cent[3 * (size_t)e + d]);
    }
  int ax = 0;
  for (int d = 1; d < 3; ++d)
    if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
  return ax;
}

void sort_along(std::vector<int>& idx, const double* cent, int ax) {
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    return cent[3 * (size_t)a + ax] < cent[3 * (size_t)b + ax];
  });
}

// Largest number of slots occupied at any step when idx (already in sweep order) is
// swept in steps of W elements: a node occupies its slot from its first step through
// the step after its last.
int front_width(const std::vector<int>& idx, const int* ptopo, int NP, int W,
                std::vector<int>& first, std::vector<int>& last, std::vector<int>& touched) {
  touched.clear();
  const int nsteps = ((int)idx.size() + W - 1) / W;
  for (size_t t = 0; t < idx.size(); ++t) {
    const int step = (int)t / W;
    for (int s = 0; s < NP; ++s) {
      const int node = ptopo[(size_t)idx[t] * NP + s];
      if (first[node] < 0) {
        first[node] = step;
        touched.push_back(node);
      }
      last[node] = step;
    }
  }
  std::vector<int> delta(nsteps + 2, 0);
  for (int node : touched) {
    delta[first[node]] += 1;
    delta[last[node] + 2] -= 1;
  }
  int live = 0, best = 0;
  for (int t = 0; t < nsteps; ++t) {
    live += delta[t];
    best = std::max(best, live);
  }
  for (int node : touched) first[node] = last[node] = -1;
  return best;
}

// Splits idx into `parts` segments of nearly equal size along longest axes (halves of
// the part count at every level, so any count is reachable), then splits further any
// segment whose front exceeds S slots.
void split(std::vector<int> idx, int parts, const double* cent, const int* ptopo, int NP, int W,
           int S, std::vector<int>& first, std::vector<int>& last, std::vector<int>& touched,
           std::vector<SegOrder>& out) {
  const int ax = longest_axis(idx, cent);
  sort_along(idx, cent, ax);
  if ((int)idx.size() < parts) parts = std::max(1, (int)idx.size());
  if (parts <= 1) {
    if (idx.size() <= 1 || front_width(idx, ptopo, NP, W, first, last, touched) <= S) {
      out.push_back(SegOrder{std::move(idx)});
      return;
    }
    parts = 2;
  }
  const int pa = parts / 2, pb = parts - pa;
  const size_t h = idx.size() * (size_t)pa / (size_t)parts;
  std::vector<int> a(idx.begin(), idx.begin() + h), b(idx.begin() + h, idx.end());
  split(std::move(a), pa, cent, ptopo, NP, W, S, first, last, touched, out);
  split(std::move(b), pb, cent, ptopo, NP, W, S, first, last, touched, out);
}

}  // namespace

SlotSchedule build_slot_schedule(const int* ptopo, int NP, int E, int N, const double* cent,
                                 int W, int S, int min_segments) {
  SlotSchedule sc;
  sc.W = W;
  sc.S = S;
  sc.NP = NP;
  std::vector<int> first(N, -1), last(N, -1), touched;
  std::vector<SegOrder> segs;
  if (E > 0) {
    std::vector<int> all(E);
    std::iota(all.begin(), all.end(), 0);
    split(std::move(all), std::max(1, min_segments), cent, ptopo, NP, W, S, first, last, touched,
          segs);
  }
  sc.n_segments = (int)segs.size();
  // how many segments touch each node
  std::vector<int> nseg(N, 0), mark(N, -1);
  for (int g = 0; g < sc.n_segments; ++g)
    for (int e : segs[g].elems)
      for (int s = 0; s < NP; ++s) {
        const int node = ptopo[(size_t)e * NP + s];
        if (mark[node] != g) {
          mark[node] = g;
          nseg[node] += 1;
        }
      }
  for (int j = 0; j < N; ++j)
    if (nseg[j] != 1) sc.zero_cols.push_back(j);  // shared or untouched
  sc.n_shared = 0;
  for (int j = 0; j < N; ++j) sc.n_shared += nseg[j] > 1;

  sc.slot_of.assign((size_t)E * NP, -1);
  std::vector<int> node_slot(N, -1);
  sc.seg_step_begin.push_back(0);
  sc.flush_begin.push_back(0);
  sc.max_live = 0;
  for (int g = 0; g < sc.n_segments; ++g) {
    const std::vector<int>& el = segs[g].elems;
    const int nsteps = ((int)el.size() + W - 1) / W;
    // first / last step per node within this segment
    touched.clear();
    for (size_t t = 0; t < el.size(); ++t) {
      const int step = (int)t / W;
      for (int s = 0; s < NP; ++s) {
        const int node = ptopo[(size_t)el[t] * NP + s];
        if (first[node] < 0) {
          first[node] = step;
          touched.push_back(node);
        }
        last[node] = step;
      }
    }
    std::vector<std::vector<int>> starts(nsteps), ends(nsteps);
    for (int node : touched) {
      starts[first[node]].push_back(node);
      ends[last[node]].push_back(node);
    }
    // free slots, lowest index first; slots of nodes that ended at the previous step are
    // released after this step (they are flushed during it)
    std::vector<int> freelist, pending;
    for (int q = S - 1; q >= 0; --q) freelist.push_back(q);
    int live = 0;
    for (int t = 0; t < nsteps; ++t) {
      std::sort(starts[t].begin(), starts[t].end());
      for (int node : starts[t]) {
        if (freelist.empty()) throw Error{BIE_EINVAL, "slot schedule overflow"};
        node_slot[node] = freelist.back();
        freelist.pop_back();
        ++live;
      }
      sc.max_live = std::max(sc.max_live, live);
      for (int w = 0; w < W; ++w) {
        const size_t p = (size_t)t * W + w;
        const int e = p < el.size() ? el[p] : -1;
        sc.step_elems.push_back(e);
        if (e >= 0)
          for (int s = 0; s < NP; ++s)
            sc.slot_of[(size_t)e * NP + s] = node_slot[ptopo[(size_t)e * NP + s]];
      }
      std::sort(ends[t].begin(), ends[t].end());
      for (int node : ends[t]) {
        sc.flush_node.push_back(node);
        sc.flush_slot.push_back(node_slot[node] | (nseg[node] > 1 ? kSharedBit : 0));
      }
      for (int sl : pending) {
        freelist.push_back(sl);
        --live;
      }
      pending.clear();
      for (int node : ends[t]) pending.push_back(node_slot[node]);
      sc.flush_begin.push_back((int)sc.flush_node.size());
    }
    for (int node : touched) first[node] = last[node] = -1;
    sc.seg_step_begin.push_back(sc.seg_step_begin.back() + nsteps);
  }
  sc.n_steps = sc.seg_step_begin.back();
  return sc;
}

// Replays a schedule and checks every invariant the slot kernel relies on: each element
// is swept exactly once; a slot holds one node from its first to its last step; every
// touched node is flushed exactly once per segment, from the slot it occupies; owned and
// shared flags agree with the segment count; zero_cols = shared + untouched columns.
void validate_slot_schedule(const SlotSchedule& sc, const int* ptopo, int NP, int E, int N) {
  auto fail = [](const std::string& m) { throw Error{BIE_EINVAL, "schedule: " + m}; };
  if ((int)sc.step_elems.size() != sc.n_steps * sc.W) fail("step_elems size");
  if ((int)sc.flush_begin.size() != sc.n_steps + 1) fail("flush_begin size");
  std::vector<int> seen(E, 0), nseg(N, 0), seg_mark(N, -1);
  std::vector<int> slot_node(sc.S, -1), node_slot(N, -1), flushed(N, -1);
  for (int g = 0; g < sc.n_segments; ++g) {
    std::vector<int> pending;  // slots flushed during the current step, free after it
    for (int t = sc.seg_step_begin[g]; t < sc.seg_step_begin[g + 1]; ++t) {
      for (int w = 0; w < sc.W; ++w) {
        const int e = sc.step_elems[(size_t)t * sc.W + w];
        if (e < 0) continue;
        if (e >= E) fail("element id out of range");
        if (seen[e]++) fail("element swept twice");
        for (int s = 0; s < NP; ++s) {
          const int node = ptopo[(size_t)e * NP + s];
          const int sl = sc.slot_of[(size_t)e * NP + s];
          if (sl < 0 || sl >= sc.S) fail("slot out of range");
          if (seg_mark[node] != g) {
            seg_mark[node] = g;
            nseg[node]++;
          }
          if (node_slot[node] < 0) {  // first touch in this segment
            if (flushed[node] == g) fail("node touched after its flush");
            if (slot_node[sl] >= 0) fail("slot reused while occupied");
            slot_node[sl] = node;
            node_slot[node] = sl;
          } else if (node_slot[node] != sl) {
            fail("node moved between slots");
          }
        }
      }
      // the flush of step t - 1 runs during step t: its slots are free from step t + 1
      for (int sl : pending) slot_node[sl] = -1;
      pending.clear();
      for (int f = sc.flush_begin[t]; f < sc.flush_begin[t + 1]; ++f) {
        const int node = sc.flush_node[f];
        const int sl = sc.flush_slot[f] & 0xffff;
        if (node < 0 || node >= N) fail("flush node out of range");
        if (node_slot[node] != sl || slot_node[sl] != node) fail("flush of a slot not held");
        if (flushed[node] == g) fail("node flushed twice in a segment");
        flushed[node] = g;
        node_slot[node] = -1;
        pending.push_back(sl);
      }
    }
    for (int sl : pending) slot_node[sl] = -1;
    for (int q = 0; q < sc.S; ++q)
      if (slot_node[q] >= 0) fail("slot still occupied at segment end");
  }
  for (int e = 0; e < E; ++e)
    if (seen[e] != 1) fail("element not swept");
  std::vector<char> zero(N, 0);
  for (int j : sc.zero_cols) zero[j] = 1;
  for (int j = 0; j < N; ++j)
    if ((nseg[j] != 1) != (zero[j] != 0)) fail("zero_cols disagrees with segment counts");
  for (size_t f = 0; f < sc.flush_node.size(); ++f) {
    const bool shared = (sc.flush_slot[f] & kSharedBit) != 0;
    if (shared != (nseg[sc.flush_node[f]] > 1)) fail("shared flag wrong");
  }
}

}  // namespace bie
