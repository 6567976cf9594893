// driver.hip -- the reference's multi-rank pass loop on one device.
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
#include <cstring>
#include <memory>
#include <vector>

#include "engine.hpp"

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

}  // namespace
}  // namespace jssp

extern "C" {

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
