// host_io.cpp -- instance files, synthetic instances and schedule checks.
// Host only; no device is touched.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

#include "../../include/jssp.h"

namespace {

bool file_ok(const char* p) {
  std::ifstream f(p);
  return f.good();
}

uint64_t splitmix64(uint64_t& s) {
  s += 0x9E3779B97F4A7C15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

extern "C" {

// Same search order as resolve_instance_file (cpu/common/instance_parser.cpp:8-34).
int jssp_resolve_instance_file(const char* arg, char* out, int32_t size) {
  if (!arg || !out || size <= 1) return 0;
  std::snprintf(out, (size_t)size, "%s", arg);
  if (file_ok(out)) return 1;
  static const char* const kFormats[] = {
      "./la%s.txt",
      "./ta%s.txt",
      "./lawrence_instances/la%s.txt",
      "../lawrence_instances/la%s.txt",
      "../../lawrence_instances/la%s.txt",
      "./taillard_instances/ta%s",
      "../taillard_instances/ta%s",
      "../../taillard_instances/ta%s",
  };
  for (const char* fmt : kFormats) {
    std::snprintf(out, (size_t)size, fmt, arg);
    if (file_ok(out)) return 1;
  }
  return 0;
}

// File format of parse_instance (cpu/common/instance_parser.cpp:37-93):
// "n m" then n rows of m (machine, duration) pairs with 0-based machines.
int jssp_parse_instance_file(const char* path, int32_t* n_jobs, int32_t* n_machines,
                             int32_t* machine, int32_t* duration, int32_t cap) {
  if (!path || !n_jobs || !n_machines) return JSSP_ERR_INVALID;
  FILE* f = std::fopen(path, "r");
  if (!f) return JSSP_ERR_IO;
  int n = 0, m = 0;
  if (std::fscanf(f, "%d %d", &n, &m) != 2 || n <= 0 || m <= 0) {
    std::fclose(f);
    return JSSP_ERR_INVALID;
  }
  *n_jobs = n;
  *n_machines = m;
  const long long total = (long long)n * m;
  if (!machine || !duration) {  // size query
    std::fclose(f);
    return JSSP_OK;
  }
  if (total > cap) {
    std::fclose(f);
    return JSSP_ERR_LIMIT;
  }
  for (long long i = 0; i < total; ++i) {
    int mc = 0, d = 0;
    if (std::fscanf(f, "%d %d", &mc, &d) != 2) {
      std::fclose(f);
      return JSSP_ERR_INVALID;  // the reference silently keeps machine 0 / duration 0 here
    }
    machine[i] = mc;
    duration[i] = d;
  }
  std::fclose(f);
  return JSSP_OK;
}

int jssp_generate_instance(uint64_t seed, int32_t n, int32_t m, int32_t dmin, int32_t dmax,
                           int32_t* machine, int32_t* duration) {
  if (n <= 0 || m <= 0 || dmin > dmax || dmin < 0 || !machine || !duration) return JSSP_ERR_INVALID;
  uint64_t s = seed;
  const uint64_t span = (uint64_t)((int64_t)dmax - dmin + 1);
  std::vector<int32_t> perm(m);
  for (int j = 0; j < n; ++j) {
    for (int k = 0; k < m; ++k) perm[k] = k;
    for (int k = m - 1; k > 0; --k) {  // Fisher-Yates
      int r = (int)(splitmix64(s) % (uint64_t)(k + 1));
      int t = perm[k]; perm[k] = perm[r]; perm[r] = t;
    }
    for (int k = 0; k < m; ++k) {
      machine[j * m + k] = perm[k];
      duration[j * m + k] = (int32_t)(dmin + (int64_t)(splitmix64(s) % span));
    }
  }
  return JSSP_OK;
}

// Replays a sequence with the reference's rule (cpu/common/jssp_types.h:75-97,
// including its M == 1 predecessor quirk).
int jssp_validate_sequence(const jssp_instance* inst, const int32_t* seq, int32_t* makespan) {
  if (!inst || !seq || !makespan) return JSSP_ERR_INVALID;
  const int n = inst->n_jobs, m = inst->n_machines;
  std::vector<int> mr(m, 0), jr(n, 0), nx(n, 0);
  int cmax = 0;
  *makespan = -1;
  for (int i = 0; i < n * m; ++i) {
    const int op = seq[i] - 1;
    if (op < 0 || op >= n * m) return JSSP_OK;
    const int j = op / m, k = op % m;
    if (nx[j] != k) return JSSP_OK;
    const int mc = inst->machine[op];
    if (mc < 0 || mc >= m) return JSSP_ERR_INVALID;
    const int pred = (m == 1) ? (j > 0 ? jr[j - 1] : 0) : jr[j];
    const int f = (mr[mc] > pred ? mr[mc] : pred) + inst->duration[op];
    mr[mc] = f; jr[j] = f; nx[j]++;
    if (f > cmax) cmax = f;
  }
  *makespan = cmax;
  return JSSP_OK;
}

int32_t jssp_lower_bound(const jssp_instance* inst) {
  if (!inst) return -1;
  const int n = inst->n_jobs, m = inst->n_machines;
  std::vector<long long> load(m, 0);
  long long best = 0;
  for (int j = 0; j < n; ++j) {
    long long len = 0;
    for (int k = 0; k < m; ++k) {
      len += inst->duration[j * m + k];
      int mc = inst->machine[j * m + k];
      if (mc >= 0 && mc < m) load[mc] += inst->duration[j * m + k];
    }
    if (len > best) best = len;
  }
  for (long long l : load)
    if (l > best) best = l;
  return (int32_t)best;
}

const char* jssp_version(void) { return "beamsearch_jssp_gpu_amd 0.1 (gfx950)"; }

}  // extern "C"
