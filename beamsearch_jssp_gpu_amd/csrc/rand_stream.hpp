// rand_stream.hpp -- the values the reference's rand() calls return, from
// either the process-wide generator or a private one.
//
// glibc's rand() is random() over a 128-byte TYPE_3 state (31-word additive
// feedback, x[i] = x[i-3] + x[i-31], output >> 1), and srand(s) is
// srandom(s).  initstate_r/random_r over a private 128-byte buffer select the
// same TYPE_3 generator and seeding, so RandStream(seed) yields exactly the
// sequence srand(seed); rand(); rand(); ... would, without touching the
// process-wide state.  That lets one process replay several MPI ranks of the
// reference drivers (each rank seeds its own process with srand(time + rank),
// gpu/v1/main.cu:60, cpu/v1_pms/main.cpp:56).
#pragma once
#include <cstdint>
#include <cstdlib>
#include <cstring>

namespace jssp {

class RandStream {
 public:
  // process_wide = true draws from rand() itself (the reference solvers'
  // convention: the caller has seeded it with srand); seed is then unused.
  RandStream(bool process_wide, uint32_t seed) : global_(process_wide) {
    if (!global_) {
      std::memset(&rd_, 0, sizeof rd_);
      std::memset(state_, 0, sizeof state_);
      initstate_r(seed, state_, sizeof state_, &rd_);
    }
  }
  // random_data points into state_: the object must stay where it is.
  RandStream(const RandStream&) = delete;
  RandStream& operator=(const RandStream&) = delete;

  int next() {
    if (global_) return rand();
    int32_t v = 0;
    random_r(&rd_, &v);
    return (int)v;
  }

 private:
  bool global_;
  random_data rd_;
  alignas(8) char state_[128];
};

}  // namespace jssp
