// jssp_compat.hpp -- the reference's C++ solver API, implemented on the
// MI355X engine, so reference-style drivers compile and link unchanged
// against libjssp_hip.so.
//
// The types mirror cpu/common/jssp_types.h:8-23 field for field (an interface
// that has to be alike for source compatibility); the functions have the
// signatures of:
//   resolve_instance_file   cpu/common/instance_parser.h:6
//   parse_instance          cpu/common/instance_parser.h:9
//   beam_search_ich         cpu/v0_ich/solver.h:6      (ICH_EXACT: same rand() calls, same result)
//   beam_search_pms         cpu/v1_pms/solver.h:6      (PMS_EXACT: same rand() calls, same result)
//   run_gpu_beam_search_v1  gpu/v1/solver.cuh:6-12     (PMS_GPU, single-phase)
//   run_gpu_beam_search_v2  gpu/v2_twophases/solver.cuh:6-12 (PMS_GPU, two-phase)
//   run_hybrid_iteration_v1 hybrid/v1/solver.cuh:6-13 (device pass + host BS-PMS threads)
//   run_hybrid_iteration_v2 hybrid/v2_twophases/solver.cuh:6-13
// Behaviour notes:
//   * the GPU functions return min(best_known, best found) and fill
//     out_best_sequence (n*m 1-based op ids) only when they found a
//     schedule better than best_known (the reference copies its best buffer
//     whenever the result is < 500000, stale or not);
//   * the reference's GPU solvers keep invalid branch slots in the beam
//     (gpu/v1/solver.cu:39,261); here only valid children survive;
//   * the hybrid iterations run the device pass with seed = rand() (the
//     caller's process-wide stream, as the reference's thread 0) and
//     threads - 1 host passes, thread tid drawing from its own stream seeded
//     time(NULL) + rank*1000 + iteration*100 + tid (hybrid/v1/solver.cu:35);
//     threads = OMP_NUM_THREADS if set, else the hardware thread count.  In
//     the reference the workers' srand() calls and thread 0's rand() share
//     one generator and race; here each thread's values are the ones it
//     would draw alone;
//   * on an engine error these functions print the message to stderr and
//     exit(EXIT_FAILURE), which is what the reference's CUDA_CHECK does
//     (gpu/common/gpu_common.cuh:9-17).  The C ABI (jssp.h) never exits.
#pragma once

#include <vector>

struct Operation {
  int num_op;   // 1-indexed operation id
  int num_mc;   // 1-indexed machine id
  int dure;     // processing duration
  int num_j;    // 1-indexed job id
};

struct JSSPInstance {
  int nbjob = 0;
  int nbmachine = 0;
  int nboperation = 0;
  std::vector<Operation> operations;
  std::vector<std::vector<int>> opr_mach;
};

bool resolve_instance_file(const char* input_arg, char* out_path, int path_size);
bool parse_instance(const char* filepath, JSSPInstance& inst);

int beam_search_ich(const JSSPInstance& inst, int beam_width);
int beam_search_pms(const JSSPInstance& inst, int beam_width, int best_known);
int run_gpu_beam_search_v1(const JSSPInstance& inst, int beam_width, int best_known, int random_seed,
                           int* out_best_sequence = nullptr);
int run_gpu_beam_search_v2(const JSSPInstance& inst, int beam_width, int best_known, int random_seed,
                           int* out_best_sequence = nullptr);
int run_hybrid_iteration_v1(const JSSPInstance& inst, int gpu_beam_width, int cpu_beam_width, int best_known,
                            int rank, int iteration);
int run_hybrid_iteration_v2(const JSSPInstance& inst, int gpu_beam_width, int cpu_beam_width, int best_known,
                            int rank, int iteration);
