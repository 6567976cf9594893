// driver.hip -- the reference's multi-rank pass loops (plain and hybrid
// CPU+GPU) on one device.
//
// The reference runs world_size MPI processes; each seeds its own rand()
// with srand(time(NULL) + rank), then for every iteration solves once with
// best_known = global_best and all-reduces the minimum
// (gpu/v1/main.cu:55-71, gpu/v2_twophases/main.cu:55-71,
// cpu/v1_pms/main.cpp:55-73; the benchmark loops gpu/v1/benchmark.cu:52-66,
// cpu/v1_pms/benchmark.cpp:48-66 are the same per instance).  Here every
// rank is a private RandStream; PMS_GPU ranks of one pass are the beams of
// one batched solve, PMS_EXACT ranks run one after the other, each consuming
// its own stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "host_pms.hpp"

namespace jssp {
namespace {

constexpr int kMaxRanks = 4096;

int run_passes(jssp_engine* e, const jssp_instance* in, const jssp_driver_params* d, jssp_result* res,
               int32_t* pass_best) {
  const int R = d->ranks, E = d->iterations;
  std::vector<std::unique_ptr<RandStream>> streams;
  streams.reserve(R);
  for (int r = 0; r < R; ++r) streams.emplace_back(new RandStream(false, d->time_seed + (uint32_t)r));

  jssp_params p = {};
  p.beam_width = d->beam_width;
  p.mode = d->mode;
  p.phases = d->phases;
  p.best_known = JSSP_NO_MAKESPAN;
  int rc = check_params(e, in, &p);
  if (rc) return rc;

  if (d->mode == JSSP_MODE_PMS_GPU) {
    // seed = rand() per iteration on each rank (gpu/v1/main.cu:63); the
    // streams are independent, so draw them pass-major up front.
    std::vector<uint32_t> seeds((size_t)R * E);
    for (int ep = 0; ep < E; ++ep)
      for (int r = 0; r < R; ++r) seeds[(size_t)ep * R + r] = (uint32_t)streams[r]->next();
    p.n_seeds = R;
    p.n_passes = E;
    return solve_multiseed(e, in, &p, res, seeds.data(), pass_best);
  }

  // PMS_EXACT: beam_search_pms(inst, K, global_best) per rank and pass.
  const int L = in->n_jobs * in->n_machines;
  std::vector<int32_t> seq(res->sequence ? (size_t)L : 0);
  int global = JSSP_NO_MAKESPAN, best_idx = -1;
  int64_t expanded = 0;
  double ms = 0.0;
  for (int ep = 0; ep < E; ++ep) {
    int pass_min = global;
    for (int r = 0; r < R; ++r) {
      p.best_known = global;
      jssp_result one = {};
      one.sequence = res->sequence ? seq.data() : nullptr;
      rc = solve_exact(e, in, &p, *streams[r], &one, nullptr, nullptr);
      if (rc) return rc;
      expanded += one.expanded_states;
      ms += one.device_ms;
      if (one.found && one.makespan < pass_min) {
        pass_min = one.makespan;
        best_idx = ep * R + r;
        if (res->sequence) std::memcpy(res->sequence, seq.data(), (size_t)L * sizeof(int32_t));
      }
    }
    global = pass_min;
    if (pass_best) pass_best[ep] = global;
  }
  res->makespan = global;
  res->found = best_idx >= 0 ? 1 : 0;
  res->best_seed_index = std::max(0, best_idx);
  res->expanded_states = expanded;
  res->levels = L;
  res->device_ms = ms;
  return JSSP_OK;
}

// One hybrid pass for R ranks (hybrid/v1/solver.cu:11-44 per rank): the R
// device passes (PMS_GPU, seeds gpu_seeds[r]) as one batch on this thread,
// and R * (T - 1) host BS-PMS passes (rank r, thread tid: private stream
// srand(cpu_seed(r) + tid)) on a pool of host threads, all pruning with
// best_known.  res->makespan = min(best_known, every pass); the first pass
// to reach it, in the order device passes (rank order) then host passes
// (rank, tid), gives best_seed_index = r * T + tid and the sequence.
int hybrid_pass(jssp_engine* e, const jssp_instance* in, int Kg, int Kc, int best_known, const uint32_t* gpu_seeds,
                const uint32_t* cpu_seed_base, int R, int T, int phases, jssp_result* res) {
  const int L = in->n_jobs * in->n_machines;
  const int tasks = R * (T - 1);
  std::vector<HostPmsResult> host(tasks);
  std::atomic<int> next{0};
  auto worker = [&]() {
    for (int t = next.fetch_add(1); t < tasks; t = next.fetch_add(1)) {
      const int r = t / (T - 1), tid = 1 + t % (T - 1);
      RandStream rs(false, cpu_seed_base[r] + (uint32_t)tid);
      host[t] = host_pms(*in, Kc, best_known, rs, res->sequence != nullptr);
    }
  };
  const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
  const int nthreads = std::min(tasks, std::max(1, hw - 1));
  std::vector<std::thread> pool;
  pool.reserve(nthreads);
  for (int i = 0; i < nthreads; ++i) pool.emplace_back(worker);

  jssp_params p = {};
  p.beam_width = Kg;
  p.mode = JSSP_MODE_PMS_GPU;
  p.phases = phases;
  p.best_known = best_known;
  p.n_seeds = R;
  p.n_passes = 1;
  jssp_result g = {};
  std::vector<int32_t> gseq(res->sequence ? (size_t)L : 0);
  g.sequence = res->sequence ? gseq.data() : nullptr;
  const int rc = solve_multiseed(e, in, &p, &g, gpu_seeds, nullptr);
  for (std::thread& th : pool) th.join();
  if (rc) return rc;

  int best = best_known, idx = -1;
  int64_t expanded = g.expanded_states;
  if (g.found && g.makespan < best) {
    best = g.makespan;
    idx = g.best_seed_index * T;
    if (res->sequence) std::memcpy(res->sequence, gseq.data(), (size_t)L * sizeof(int32_t));
  }
  for (int t = 0; t < tasks; ++t) {
    const HostPmsResult& h = host[t];
    expanded += h.expanded_states;
    if (h.found && h.makespan < best) {
      best = h.makespan;
      idx = (t / (T - 1)) * T + 1 + t % (T - 1);
      if (res->sequence) std::memcpy(res->sequence, h.sequence.data(), (size_t)L * sizeof(int32_t));
    }
  }
  res->makespan = best;
  res->found = idx >= 0 ? 1 : 0;
  res->best_seed_index = std::max(0, idx);
  res->expanded_states = expanded;
  res->levels = L;
  res->device_ms = g.device_ms;
  return JSSP_OK;
}

int check_hybrid(jssp_engine* e, const jssp_instance* in, int Kg, int Kc, int phases, int threads) {
  jssp_params p = {};
  p.beam_width = Kg;
  p.mode = JSSP_MODE_PMS_GPU;
  p.phases = phases;
  int rc = check_params(e, in, &p);
  if (rc) return rc;
  if (phases != 1 && phases != 2) return set_error(e, JSSP_ERR_INVALID, "hybrid", "phases must be 1 or 2");
  if (Kc < 1 || Kc > JSSP_MAX_BEAM) return set_error(e, JSSP_ERR_INVALID, "hybrid", "cpu beam width out of range");
  if (threads < 1 || threads > 1024) return set_error(e, JSSP_ERR_INVALID, "hybrid", "threads must be in [1, 1024]");
  return JSSP_OK;
}

}  // namespace
}  // namespace jssp

extern "C" {

int jssp_hybrid_iteration(jssp_engine* e, const jssp_instance* in, int32_t gpu_beam_width, int32_t cpu_beam_width,
                          int32_t best_known, uint32_t gpu_seed, int32_t phases, int32_t threads,
                          uint32_t cpu_seed_base, jssp_result* res) {
  if (!e || !in || !res) return JSSP_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  int rc = jssp::check_instance(e, in);
  if (rc) return rc;
  rc = jssp::check_hybrid(e, in, gpu_beam_width, cpu_beam_width, phases, threads);
  if (rc) return rc;
  JSSP_HIP(e, hipSetDevice(e->device));
  return jssp::hybrid_pass(e, in, gpu_beam_width, cpu_beam_width, best_known, &gpu_seed, &cpu_seed_base, 1,
                           threads, phases, res);
}

int jssp_run_hybrid(jssp_engine* e, const jssp_instance* in, const jssp_hybrid_params* h, jssp_result* res,
                    int32_t* pass_best) {
  if (!e || !in || !h || !res) return JSSP_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  int rc = jssp::check_instance(e, in);
  if (rc) return rc;
  if (h->accel_factor < 1) return jssp::set_error(e, JSSP_ERR_INVALID, "jssp_run_hybrid", "accel_factor must be >= 1");
  const int Kc = std::max(1, h->gpu_beam_width / h->accel_factor);  // hybrid/v1/main.cu:45
  rc = jssp::check_hybrid(e, in, h->gpu_beam_width, Kc, h->phases, h->threads);
  if (rc) return rc;
  if (h->ranks < 1 || h->ranks > jssp::kMaxRanks)
    return jssp::set_error(e, JSSP_ERR_INVALID, "jssp_run_hybrid", "ranks must be in [1, 4096]");
  if (h->iterations < 1) return jssp::set_error(e, JSSP_ERR_INVALID, "jssp_run_hybrid", "iterations must be >= 1");
  JSSP_HIP(e, hipSetDevice(e->device));
  const int R = h->ranks, T = h->threads, L = in->n_jobs * in->n_machines;
  std::vector<std::unique_ptr<jssp::RandStream>> streams;
  for (int r = 0; r < R; ++r) streams.emplace_back(new jssp::RandStream(false, h->time_seed + (uint32_t)r));
  std::vector<uint32_t> gseeds(R), cbase(R);
  std::vector<int32_t> seq(res->sequence ? (size_t)L : 0);
  int global = JSSP_NO_MAKESPAN, best_idx = -1;
  int64_t expanded = 0;
  double ms = 0.0;
  for (int it = 0; it < h->iterations; ++it) {
    for (int r = 0; r < R; ++r) {
      gseeds[r] = (uint32_t)streams[r]->next();  // thread 0: seed = rand() (hybrid/v1/solver.cu:28)
      // workers: srand(time + rank*1000 + iteration*100 + tid) (hybrid/v1/solver.cu:35)
      cbase[r] = h->time_seed + (uint32_t)(r * 1000 + it * 100);
    }
    jssp_result one = {};
    one.sequence = res->sequence ? seq.data() : nullptr;
    rc = jssp::hybrid_pass(e, in, h->gpu_beam_width, Kc, global, gseeds.data(), cbase.data(), R, T, h->phases,
                           &one);
    if (rc) return rc;
    expanded += one.expanded_states;
    ms += one.device_ms;
    if (one.found && one.makespan < global) {
      global = one.makespan;
      best_idx = it * R * T + one.best_seed_index;
      if (res->sequence) std::memcpy(res->sequence, seq.data(), (size_t)L * sizeof(int32_t));
    }
    if (pass_best) pass_best[it] = global;
  }
  res->makespan = global;
  res->found = best_idx >= 0 ? 1 : 0;
  res->best_seed_index = std::max(0, best_idx);
  res->expanded_states = expanded;
  res->levels = L;
  res->device_ms = ms;
  return JSSP_OK;
}

int jssp_run_passes(jssp_engine* e, const jssp_instance* in, const jssp_driver_params* d, jssp_result* res,
                    int32_t* pass_best) {
  if (!e || !in || !d || !res) return JSSP_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  int rc = jssp::check_instance(e, in);
  if (rc) return rc;
  if (d->mode != JSSP_MODE_PMS_GPU && d->mode != JSSP_MODE_PMS_EXACT)
    return jssp::set_error(e, JSSP_ERR_INVALID, "jssp_run_passes", "the pass drivers run PMS_GPU or PMS_EXACT");
  if (d->ranks < 1 || d->ranks > jssp::kMaxRanks)
    return jssp::set_error(e, JSSP_ERR_INVALID, "jssp_run_passes", "ranks must be in [1, 4096]");
  if (d->iterations < 1) return jssp::set_error(e, JSSP_ERR_INVALID, "jssp_run_passes", "iterations must be >= 1");
  JSSP_HIP(e, hipSetDevice(e->device));
  res->best_seed_index = 0;
  return jssp::run_passes(e, in, d, res, pass_best);
}

void jssp_rand_stream(uint32_t seed, int32_t count, int32_t* out) {
  if (!out || count <= 0) return;
  jssp::RandStream rs(false, seed);
  for (int32_t i = 0; i < count; ++i) out[i] = rs.next();
}

}  // extern "C"
