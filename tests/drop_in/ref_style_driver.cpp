// ref_style_driver.cpp -- a driver written against the reference's solver API
// only (include/jssp_compat.hpp, linked to libjssp_hip.so), calling it the way
// the reference's own drivers do:
//   srand(seed); beam_search_ich(inst, K)          cpu/v0_ich/main.cpp:39-42
//   beam_search_pms(inst, K, best_known)            cpu/v1_pms/main.cpp:62-63
//   run_gpu_beam_search_v1/_v2(inst, K, best_known, seed[, out_best_sequence])
//                                                   gpu/v1/main.cu:62-64,
//                                                   gpu/v2_twophases/main.cu:62-64
// TEST INFRASTRUCTURE: tests/test_gpu_drop_in.py runs it and checks what it
// prints against the oracle, the reference's golden makespans and the C
// library's rand() stream.
//
//   ref_style_driver <instance> <K> <srand seed> <gpu seed> <best_known>
// prints one JSON object.  The CPU-solver shims consume the process-wide
// rand() stream like the reference's solvers, so the value of the next
// rand() after each call is printed as well.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "jssp_compat.hpp"

static void print_seq(const char* name, const std::vector<int>& s) {
  std::printf(", \"%s\": [", name);
  for (size_t i = 0; i < s.size(); ++i) std::printf("%s%d", i ? ", " : "", s[i]);
  std::printf("]");
}

int main(int argc, char** argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s <instance> <K> <srand seed> <gpu seed> <best_known>\n", argv[0]);
    return 2;
  }
  const int K = std::atoi(argv[2]);
  const unsigned s = (unsigned)std::strtoul(argv[3], nullptr, 10);
  const int gpu_seed = std::atoi(argv[4]);
  const int best_known = std::atoi(argv[5]);
  char path[4096];
  if (!resolve_instance_file(argv[1], path, sizeof path)) {
    std::fprintf(stderr, "instance not found: %s\n", argv[1]);
    return 1;
  }
  JSSPInstance inst;
  if (!parse_instance(path, inst)) return 1;
  const int nm = inst.nbjob * inst.nbmachine;

  std::srand(s);
  const int ich = beam_search_ich(inst, K);
  const int ich_next = std::rand();
  std::srand(s);
  const int pms = beam_search_pms(inst, K, best_known);
  const int pms_next = std::rand();
  std::vector<int> seq1(nm, -7), seq2(nm, -7);
  const int v1 = run_gpu_beam_search_v1(inst, K, best_known, gpu_seed, seq1.data());
  const int v2 = run_gpu_beam_search_v2(inst, K, best_known, gpu_seed, seq2.data());
  const int v1_plain = run_gpu_beam_search_v1(inst, K, best_known, gpu_seed);  // default argument

  std::printf("{\"n\": %d, \"m\": %d, \"nboperation\": %d, \"ich\": %d, \"ich_next_rand\": %d, \"pms\": %d, "
              "\"pms_next_rand\": %d, \"v1\": %d, \"v2\": %d, \"v1_plain\": %d",
              inst.nbjob, inst.nbmachine, inst.nboperation, ich, ich_next, pms, pms_next, v1, v2, v1_plain);
  print_seq("v1_seq", seq1);
  print_seq("v2_seq", seq2);
  std::printf("}\n");
  return 0;
}
