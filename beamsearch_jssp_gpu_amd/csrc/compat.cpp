// compat.cpp -- reference-signature shims over the C ABI (see
// include/jssp_compat.hpp for the behaviour contract).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/jssp.h"
#include "../../include/jssp_compat.hpp"
#include "rand_stream.hpp"

namespace {

// One process-wide engine on device 0, created on first use; the reference's
// solvers are process-global too (c_ops, default stream, rand()).  Creating
// it leaves the caller's rand() stream where it was (jssp_engine_create).
jssp_engine* default_engine() {
  static std::once_flag once;
  static jssp_engine* eng = nullptr;
  std::call_once(once, [] {
    if (jssp_engine_create(0, &eng) != JSSP_OK) {
      std::fprintf(stderr, "jssp: cannot create an engine on device 0\n");
      std::exit(EXIT_FAILURE);
    }
  });
  return eng;
}

struct Flat {
  std::vector<int32_t> mach, dur;
  jssp_instance in;
};

// JSSPInstance (1-based ops 1..nm, 1-based machines) -> flat 0-based arrays.
Flat flatten(const JSSPInstance& inst) {
  Flat f;
  const int n = inst.nbjob, m = inst.nbmachine;
  f.mach.resize((size_t)n * m);
  f.dur.resize((size_t)n * m);
  for (int op = 1; op <= n * m && op < (int)inst.operations.size(); ++op) {
    f.mach[op - 1] = inst.operations[op].num_mc - 1;
    f.dur[op - 1] = inst.operations[op].dure;
  }
  f.in.n_jobs = n;
  f.in.n_machines = m;
  f.in.machine = f.mach.data();
  f.in.duration = f.dur.data();
  return f;
}

int run(const JSSPInstance& inst, int K, int mode, int phases, int best_known, int seed, unsigned flags,
        int* seq) {
  jssp_engine* e = default_engine();
  Flat f = flatten(inst);
  jssp_params p = {};
  p.beam_width = K;
  p.mode = mode;
  p.phases = phases;
  p.best_known = best_known;
  p.seed = (uint32_t)seed;
  p.flags = flags;
  jssp_result r = {};
  r.sequence = seq;
  int rc = jssp_solve(e, &f.in, &p, &r);
  if (rc != JSSP_OK) {
    std::fprintf(stderr, "jssp error %d: %s\n", rc, jssp_last_error(e));
    std::exit(EXIT_FAILURE);
  }
  return r.makespan;
}

}  // namespace

bool resolve_instance_file(const char* input_arg, char* out_path, int path_size) {
  return jssp_resolve_instance_file(input_arg, out_path, path_size) != 0;
}

// Builds the reference's JSSPInstance layout (cpu/common/instance_parser.cpp:
// 51-92): dummy op 0, ops 1..nm job-major with 1-based machines, dummy end
// op nm+1, opr_mach[machine] = op ids on that machine.
bool parse_instance(const char* filepath, JSSPInstance& inst) {
  int32_t n = 0, m = 0;
  if (jssp_parse_instance_file(filepath, &n, &m, nullptr, nullptr, 0) != JSSP_OK) {
    std::fprintf(stderr, "Error: cannot read instance file: %s\n", filepath);
    return false;
  }
  std::vector<int32_t> mach((size_t)n * m), dur((size_t)n * m);
  if (jssp_parse_instance_file(filepath, &n, &m, mach.data(), dur.data(), n * m) != JSSP_OK) {
    std::fprintf(stderr, "Error: malformed instance file: %s\n", filepath);
    return false;
  }
  inst.nbjob = n;
  inst.nbmachine = m;
  inst.nboperation = n * m + 2;
  inst.operations.clear();
  inst.opr_mach.assign(m + 1, std::vector<int>());
  inst.operations.push_back(Operation{0, 0, 0, 0});
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < m; ++k) {
      const int op = j * m + k + 1;
      inst.operations.push_back(Operation{op, mach[op - 1] + 1, dur[op - 1], j + 1});
      if (mach[op - 1] + 1 >= 0 && mach[op - 1] + 1 <= m) inst.opr_mach[mach[op - 1] + 1].push_back(op);
    }
  inst.operations.push_back(Operation{n * m + 1, 0, 0, 0});
  return true;
}

int beam_search_ich(const JSSPInstance& inst, int beam_width) {
  return run(inst, beam_width, JSSP_MODE_ICH_EXACT, 2, JSSP_NO_MAKESPAN, 0, JSSP_FLAG_GLOBAL_RAND, nullptr);
}

int beam_search_pms(const JSSPInstance& inst, int beam_width, int best_known) {
  return run(inst, beam_width, JSSP_MODE_PMS_EXACT, 2, best_known, 0, JSSP_FLAG_GLOBAL_RAND, nullptr);
}

// The GPU solvers draw no rand() values (gpu/v1/solver.cu:178-298): the
// caller's stream is left as it was, whatever the device runtime does.
int run_gpu_beam_search_v1(const JSSPInstance& inst, int beam_width, int best_known, int random_seed,
                           int* out_best_sequence) {
  jssp::ProcessRandGuard keep_rand;
  return run(inst, beam_width, JSSP_MODE_PMS_GPU, 1, best_known, random_seed, 0, out_best_sequence);
}

int run_gpu_beam_search_v2(const JSSPInstance& inst, int beam_width, int best_known, int random_seed,
                           int* out_best_sequence) {
  jssp::ProcessRandGuard keep_rand;
  return run(inst, beam_width, JSSP_MODE_PMS_GPU, 2, best_known, random_seed, 0, out_best_sequence);
}

namespace {

int hybrid_threads() {
  const char* env = std::getenv("OMP_NUM_THREADS");
  const int t = env ? std::atoi(env) : (int)std::thread::hardware_concurrency();
  return std::max(1, std::min(t, 1024));
}

int hybrid(const JSSPInstance& inst, int gpu_beam_width, int cpu_beam_width, int best_known, int rank,
           int iteration, int phases) {
  jssp_engine* e = default_engine();
  Flat f = flatten(inst);
  const uint32_t gpu_seed = (uint32_t)std::rand();  // hybrid/v1/solver.cu:28, thread 0
  jssp::ProcessRandGuard keep_rand;                  // the caller's stream: that one draw only
  const uint32_t cpu_base = (uint32_t)std::time(nullptr) + (uint32_t)(rank * 1000 + iteration * 100);
  jssp_result r = {};
  int rc = jssp_hybrid_iteration(e, &f.in, gpu_beam_width, cpu_beam_width, best_known, gpu_seed, phases,
                                 hybrid_threads(), cpu_base, &r);
  if (rc != JSSP_OK) {
    std::fprintf(stderr, "jssp error %d: %s\n", rc, jssp_last_error(e));
    std::exit(EXIT_FAILURE);
  }
  return r.makespan;
}

}  // namespace

int run_hybrid_iteration_v1(const JSSPInstance& inst, int gpu_beam_width, int cpu_beam_width, int best_known,
                            int rank, int iteration) {
  return hybrid(inst, gpu_beam_width, cpu_beam_width, best_known, rank, iteration, 1);
}

int run_hybrid_iteration_v2(const JSSPInstance& inst, int gpu_beam_width, int cpu_beam_width, int best_known,
                            int rank, int iteration) {
  return hybrid(inst, gpu_beam_width, cpu_beam_width, best_known, rank, iteration, 2);
}
