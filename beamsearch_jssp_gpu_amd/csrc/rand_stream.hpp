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
//
// The process-wide form continues the caller's rand() stream (the reference
// solvers' convention: the caller seeds it with srand, cpu/v0_ich/main.cpp:39).
// It does not draw with rand() itself: other code in the same process may
// draw from or reseed the process-wide generator while a solve runs (ROCm's
// code-object manager, libamd_comgr, calls srand() once and rand() tens of
// times when the device is first used), which would change both the solve's
// values and what the caller sees afterwards.  Instead the generator's state and
// position are copied into a private generator when the stream is made, the
// solve draws from that copy, and the copy is written back into the
// process-wide generator when the stream is destroyed: the caller's next
// rand() is then the one that follows exactly the draws the reference solver
// would have made.  The copy uses glibc's documented state exchange
// (initstate/setstate return the previous state block, whose first word
// holds the generator type and position).
#pragma once
#include <cstdint>
#include <cstdlib>
#include <cstring>

namespace jssp {

namespace rand_detail {

constexpr size_t kTypeBytes = 128;  // TYPE_3: the state srand()/rand() use by default
constexpr size_t kMaxBytes = 256;   // TYPE_4, the largest glibc type

// Bytes of a state block: one header word (type + 5 * position) and the
// type's words (TYPE_0 keeps one LCG word; TYPE_1..4: 7, 15, 31, 63).
inline size_t state_bytes(const char* block) {
  int32_t head;
  std::memcpy(&head, block, 4);
  static const int kDeg[5] = {1, 7, 15, 31, 63};
  const int type = head % 5;
  return 4u * (1u + (unsigned)kDeg[type < 0 ? 0 : type]);
}

// Copies the process-wide generator's state block, position included, into
// out (kMaxBytes): switch the generator to a scratch block (which records
// the current block's position in that block's first word), copy the block,
// switch back.  Returns its size; *home = the block itself.
inline size_t capture_process_rand(char* out, char** home) {
  alignas(8) char scratch[kMaxBytes];
  char* old = initstate(1, scratch, kTypeBytes);
  const size_t bytes = state_bytes(old);
  std::memcpy(out, old, bytes);
  setstate(old);
  *home = old;
  return bytes;
}

// Loads a block captured as above back into the generator's block `home` and
// makes the process-wide generator resume from it, at the captured position.
inline void load_process_rand(const char* block, size_t bytes, char* home) {
  alignas(8) char scratch[kMaxBytes];
  initstate(1, scratch, kTypeBytes);  // away from home, so its header may be rewritten
  std::memcpy(home, block, bytes);
  setstate(home);
}

}  // namespace rand_detail

// Keeps the process-wide generator where it was for the guard's lifetime:
// whatever draws from or reseeds it meanwhile (the GPU runtime's code-object
// loader does, once per process, when the device is first used), the caller
// sees the same next rand() afterwards.
class ProcessRandGuard {
 public:
  ProcessRandGuard() { bytes_ = rand_detail::capture_process_rand(block_, &home_); }
  ~ProcessRandGuard() { rand_detail::load_process_rand(block_, bytes_, home_); }
  ProcessRandGuard(const ProcessRandGuard&) = delete;
  ProcessRandGuard& operator=(const ProcessRandGuard&) = delete;

 private:
  size_t bytes_;
  char* home_;
  alignas(8) char block_[rand_detail::kMaxBytes];
};

class RandStream {
 public:
  // process_wide = true continues the process-wide stream (seed unused).
  RandStream(bool process_wide, uint32_t seed) : global_(process_wide) {
    std::memset(&rd_, 0, sizeof rd_);
    std::memset(spare_, 0, sizeof spare_);
    std::memset(state_, 0, sizeof state_);
    if (!global_) {
      initstate_r(seed, state_, rand_detail::kTypeBytes, &rd_);
      return;
    }
    bytes_ = rand_detail::capture_process_rand(state_, &home_);
    // A private generator over the copy: initialised on the spare block
    // first, since setstate_r records the current block's position there.
    initstate_r(1, spare_, rand_detail::kTypeBytes, &rd_);
    setstate_r(state_, &rd_);
  }
  ~RandStream() {
    if (!global_) return;
    // Record the private position in the copy's first word (by switching the
    // private generator away from it), then load the copy into the
    // process-wide generator.
    setstate_r(spare_, &rd_);
    rand_detail::load_process_rand(state_, bytes_, home_);
  }
  // random_data points into the state blocks: the object must stay where it is.
  RandStream(const RandStream&) = delete;
  RandStream& operator=(const RandStream&) = delete;

  int next() {
    int32_t v = 0;
    random_r(&rd_, &v);
    return (int)v;
  }

 private:
  bool global_;
  size_t bytes_ = rand_detail::kTypeBytes;
  char* home_ = nullptr;
  random_data rd_;
  alignas(8) char spare_[rand_detail::kMaxBytes];
  alignas(8) char state_[rand_detail::kMaxBytes];
};

}  // namespace jssp
