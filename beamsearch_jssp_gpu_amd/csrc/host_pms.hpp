// host_pms.hpp -- BS-PMS (cpu/v1_pms/solver.cpp:8-70) on the host.
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/jssp.h"
#include "rand_stream.hpp"

namespace jssp {

struct HostPmsResult {
  int makespan = JSSP_NO_MAKESPAN;  // best complete cmax, whatever the bound (as the reference)
  bool found = false;
  int64_t expanded_states = 0;      // children evaluated
  std::vector<int32_t> sequence;    // 1-based op ids of the first child reaching makespan
};

// One pass of beam width K pruning children with cmax >= best_known; draws
// the reference's rand() values from rs.  Host only; thread-safe for
// distinct rs.
HostPmsResult host_pms(const jssp_instance& in, int K, int best_known, RandStream& rs, bool want_sequence);

}  // namespace jssp
