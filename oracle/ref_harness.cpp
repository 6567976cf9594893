// ref_harness.cpp -- C entry points over the reference's own CPU solvers.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference sources where they lie (cpu/common/instance_parser.cpp,
// cpu/v0_ich/solver.cpp, cpu/v1_pms/solver.cpp) into oracle/_ref/.  It is
// used to write golden fixtures, to validate oracle.cpp, and as the CPU
// baseline timed by bench.py.  No reference source is copied here: the
// reference headers are included from the reference tree at build time.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unistd.h>

#include "common/instance_parser.h"   // cpu/common/instance_parser.h
#include "v0_ich/solver.h"            // cpu/v0_ich/solver.h:6
int beam_search_pms(const JSSPInstance& inst, int beam_width, int best_known);  // cpu/v1_pms/solver.h:6

namespace {

// Writes the instance in the reference file format (header "n m", then one
// row of (machine, duration) pairs per job) and parses it with the
// reference's own parse_instance (cpu/common/instance_parser.cpp:37-93).
bool load(int n, int m, const int* mach, const int* dur, JSSPInstance& inst) {
  char path[] = "/tmp/jssp_ref_XXXXXX";
  int fd = mkstemp(path);
  if (fd < 0) return false;
  FILE* f = fdopen(fd, "w");
  fprintf(f, "%d %d\n", n, m);
  for (int j = 0; j < n; ++j) {
    for (int k = 0; k < m; ++k) fprintf(f, " %d %d", mach[j * m + k], dur[j * m + k]);
    fprintf(f, "\n");
  }
  fclose(f);
  bool ok = parse_instance(path, inst);
  unlink(path);
  return ok;
}

}  // namespace

extern "C" {

// srand(seed); beam_search_ich(inst, K)  -- the call sequence of
// cpu/v0_ich/main.cpp:93-96.
int ref_ich(int n, int m, const int* mach, const int* dur, int beam_width, unsigned seed) {
  JSSPInstance inst;
  if (!load(n, m, mach, dur, inst)) return -1;
  srand(seed);
  return beam_search_ich(inst, beam_width);
}

// srand(seed); beam_search_pms(inst, K, best_known) -- one pass of
// cpu/v1_pms/main.cpp:62-63 with a fixed seed.
int ref_pms(int n, int m, const int* mach, const int* dur, int beam_width, int best_known,
            unsigned seed) {
  JSSPInstance inst;
  if (!load(n, m, mach, dur, inst)) return -1;
  srand(seed);
  return beam_search_pms(inst, beam_width, best_known);
}

// The loop of cpu/v1_pms/main.cpp:55-73 with `ranks` MPI ranks replayed in
// one process: each rank's rand() stream is a glibc state buffer seeded
// initstate(time_seed + r) and made current with setstate() around that
// rank's beam_search_pms call, so every rank sees exactly the values its own
// process would (srand(time + rank) then rand()...).  pass_best[it] = the
// all-reduced minimum after iteration it.  The process-wide state is
// restored afterwards.
int ref_pms_driver(int n, int m, const int* mach, const int* dur, int beam_width, int ranks,
                   int iterations, unsigned time_seed, int* pass_best) {
  JSSPInstance inst;
  if (!load(n, m, mach, dur, inst) || ranks < 1 || ranks > 64) return -1;
  static char states[64][128];
  char* saved = initstate(time_seed, states[0], 128);
  for (int r = 1; r < ranks; ++r) initstate(time_seed + (unsigned)r, states[r], 128);
  int local[64];
  for (int r = 0; r < ranks; ++r) local[r] = 500000;
  int global_best = 500000;
  for (int it = 0; it < iterations; ++it) {
    for (int r = 0; r < ranks; ++r) {
      setstate(states[r]);
      int makespan = beam_search_pms(inst, beam_width, global_best);
      if (makespan < local[r]) local[r] = makespan;
    }
    int red = 500000;
    for (int r = 0; r < ranks; ++r) red = red < local[r] ? red : local[r];
    global_best = red;
    if (pass_best) pass_best[it] = global_best;
  }
  setstate(saved);
  return global_best;
}

// Parses a file with the reference parser and returns it 0-based
// (machine = num_mc - 1), job-major.  Returns n*m, or -1 on failure / when
// cap is too small.
int ref_parse(const char* path, int* n, int* m, int* mach, int* dur, int cap) {
  JSSPInstance inst;
  if (!parse_instance(path, inst)) return -1;
  *n = inst.nbjob;
  *m = inst.nbmachine;
  int total = inst.nbjob * inst.nbmachine;
  if (total > cap) return -1;
  for (int op = 1; op <= total; ++op) {
    mach[op - 1] = inst.operations[op].num_mc - 1;
    dur[op - 1] = inst.operations[op].dure;
  }
  return total;
}

// resolve_instance_file (cpu/common/instance_parser.cpp:8-34).
int ref_resolve(const char* arg, char* out, int size) {
  return resolve_instance_file(arg, out, size) ? 1 : 0;
}

}  // extern "C"
