// engine.hpp -- internal state of a jssp_engine, shared by the solve, batch
// and serving translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/jssp.h"
#include "level_kernels.hpp"
#include "rand_stream.hpp"

// Kernel kinds timed by the optional profiler.
enum { kKindScore = 0, kKindFilter = 1, kKindRank = 2, kNumKinds = 3 };

struct KernelTime {
  double ms = 0.0;
  int64_t launches = 0;
  double bytes = 0.0;
};

struct jssp_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  char* dbuf = nullptr;   // device arena (grows, reused across solves)
  size_t dcap = 0;
  char* hbuf = nullptr;   // pinned host staging (grows)
  size_t hcap = 0;
  char* xbuf = nullptr;   // exact modes: device-visible coherent host memory the kernels
  size_t xcap = 0;        // read and write directly (per-level children, heap order)
  std::string err;
  bool profiling = false;
  std::vector<hipEvent_t> pev;  // event pool for per-kernel timing
  KernelTime kt[kNumKinds];
  // Per launch of each kind in the last profiled single-beam device solve,
  // in level order: duration (ms) and algorithmic bytes.
  std::vector<float> kl_ms[kNumKinds];
  std::vector<double> kl_bytes[kNumKinds];
  std::vector<jssp::SelState> level_log;  // per-level selection state of the last profiled solve
  // Batched solves: their own arena (kept apart from the single-solve one)
  // and the captured step graph of the last batch layout.
  char* bbuf = nullptr;
  size_t bcap = 0;
  // Captured step graphs of batch segments, keyed by their launch layout.
  std::vector<std::pair<std::vector<uint64_t>, hipGraphExec_t>> step_graphs;
  int autotune = 1;       // consult/extend the lanes-per-parent tuning table
  std::string tune_path;  // JSON file of the tuning table ("" = in memory only)
  std::vector<std::pair<uint64_t, int>> tune;  // (n, m, K bucket) -> lanes per parent
  std::recursive_mutex mu;
};

namespace jssp {

// Bump allocator used twice: a dry run (base == nullptr) to size a plan,
// then the real carve over the arena.
struct Bump {
  char* base = nullptr;
  size_t off = 0;
  template <class T>
  T* take(size_t count) {
    off = (off + 255) & ~(size_t)255;
    T* p = reinterpret_cast<T*>(base + off);
    off += count * sizeof(T) + 0;
    return p;
  }
};

int set_error(jssp_engine* e, int code, const char* where, const std::string& msg);
int ensure_device_arena(jssp_engine* e, size_t bytes);
int ensure_host_staging(jssp_engine* e, size_t bytes);
int check_instance(jssp_engine* e, const jssp_instance* in);
int check_params(jssp_engine* e, const jssp_instance* in, const jssp_params* p);
void release_batch_state(jssp_engine* e);

// ICH_EXACT / PMS_EXACT solve drawing the reference's rand() values from rs.
int solve_exact(jssp_engine* e, const jssp_instance* in, const jssp_params* p, RandStream& rs,
                jssp_result* res, int32_t* trail_out, int32_t* pcount_out);
// Takes the device address of the no-allocation entry point's scratch array
// on `device` (current), once per device (csrc/noalloc.hip).
hipError_t noalloc_setup(int device);
// n_seeds beams per pass, n_passes passes; pass e prunes with the best
// makespan found before it.  Beam i of pass e uses seeds[e * n_seeds + i]
// (seed + e * n_seeds + i when seeds is null); pass_best[e], when given,
// receives the best makespan after pass e.
int solve_multiseed(jssp_engine* e, const jssp_instance* in, const jssp_params* p, jssp_result* res,
                    const uint32_t* seeds, int32_t* pass_best);

}  // namespace jssp

#define JSSP_HIP(e, call)                                                              \
  do {                                                                                 \
    hipError_t _r = (call);                                                            \
    if (_r != hipSuccess)                                                              \
      return ::jssp::set_error((e), JSSP_ERR_HIP, #call, hipGetErrorString(_r));       \
  } while (0)
