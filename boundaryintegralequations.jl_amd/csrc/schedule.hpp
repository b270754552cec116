// Slot schedule of the LDS-accumulator assembly kernel (built by schedule.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "bie_common.hpp"

namespace bie {

constexpr int kSharedBit = 1 << 30;

struct SlotSchedule {
  int W = 0;            // elements per step (waves per workgroup)
  int S = 0;            // LDS slots available
  int NP = 0;
  int n_segments = 0;
  int n_steps = 0;
  int n_shared = 0;     // nodes touched by more than one segment
  int max_live = 0;     // largest number of live slots over all steps
  std::vector<int> seg_step_begin;  // n_segments + 1
  std::vector<int> step_elems;      // n_steps * W (element id or -1)
  std::vector<int> slot_of;         // E * NP slot of each element-local node
  std::vector<int> flush_begin;     // n_steps + 1
  std::vector<int> flush_node;      // node column flushed after the step
  std::vector<int> flush_slot;      // slot | kSharedBit when the node is shared
  std::vector<int> zero_cols;       // columns to zero before the kernel
};

// min_segments: the number of segments to aim for (exactly that many unless a segment's
// front exceeds S slots, which splits it further).
SlotSchedule build_slot_schedule(const int* ptopo, int NP, int E, int N, const double* cent,
                                 int W, int S, int min_segments);

void validate_slot_schedule(const SlotSchedule& sc, const int* ptopo, int NP, int E, int N);

}  // namespace bie
