// host_pms.cpp -- BS-PMS on the host, for the hybrid CPU+GPU driver.
//
// The reference's hybrid iteration (hybrid/v1/solver.cu:11-44) runs
// beam_search_pms (cpu/v1_pms/solver.cpp:8-70) on every OpenMP worker thread
// next to the GPU pass.  This is the same search on the engine's compact
// state: machine ready times, job ready times, each job's next operation and
// the ready list in the reference's order (erase, then append), so a child
// costs O(n + m) words instead of the reference's O(nm) vectors.  Selection
// replays the reference exactly: rand() for the pivot and the noise in the
// same order (cpu/v1_pms/solver.cpp:32,52), the FP64 score
// cmax / ops_scheduled + (rand() % 1000) / 1e6, the prune cmax < best_known,
// and the bounded libstdc++ heap of insert_into_beam
// (cpu/common/jssp_types.h:47-58) on (score, state handle) pairs, which
// compares only scores and so permutes exactly like the reference's vector
// of states.
#include "host_pms.hpp"

#include <algorithm>
#include <cstring>

namespace jssp {
namespace {

struct Handle {
  double score;
  int idx;  // state slot in the level's pool
};
struct HandleLess {
  bool operator()(const Handle& a, const Handle& b) const { return a.score < b.score; }
};

// One level's states: fixed-size records [mr(m) | jr(n) | nx(n) | ready(n)]
// plus cmax, ready count and the (parent, job) that made it.
struct Pool {
  int m = 0, n = 0, words = 0;
  std::vector<int32_t> w;
  std::vector<int32_t> cmax, nready, parent, job;
  void init(int m_, int n_, int cap) {
    m = m_; n = n_; words = m + 3 * n;
    w.assign((size_t)cap * words, 0);
    cmax.assign(cap, 0);
    nready.assign(cap, 0);
    parent.assign(cap, -1);
    job.assign(cap, -1);
  }
  int32_t* mr(int i) { return w.data() + (size_t)i * words; }
  int32_t* jr(int i) { return mr(i) + m; }
  int32_t* nx(int i) { return mr(i) + m + n; }
  int32_t* ready(int i) { return mr(i) + m + 2 * n; }
};

}  // namespace

HostPmsResult host_pms(const jssp_instance& in, int K, int best_known, RandStream& rs, bool want_sequence) {
  const int n = in.n_jobs, m = in.n_machines, L = n * m;
  HostPmsResult res;
  res.makespan = JSSP_NO_MAKESPAN;
  // Pools of candidates (at most K) and children of the current level (at
  // most K kept; a child replaced in the heap frees its slot for the next).
  Pool cur, nxt;
  cur.init(m, n, 1);
  nxt.init(m, n, std::max(1, K) + 1);
  for (int j = 0; j < n; ++j) {
    cur.ready(0)[j] = j;  // create_initial_state: op j*m+1 of every job
    cur.nx(0)[j] = 0;
  }
  cur.nready[0] = n;
  int ncur = 1;
  std::vector<int> order(1, 0);      // candidates in heap-array order
  std::vector<Handle> heap;
  heap.reserve(K + 1);
  std::vector<int> free_slots;
  // Trail for the sequence: per level, (parent position in the previous
  // level's order) << 8 | job of each kept child.
  std::vector<std::vector<uint32_t>> trail;
  int best_pos = -1, best_job = -1;
  for (int step = 1; step <= L + 10 && ncur > 0; ++step) {
    heap.clear();
    free_slots.clear();
    for (int s = K; s >= 0; --s) free_slots.push_back(s);
    for (int ic = 0; ic < ncur; ++ic) {
      const int ci = order[ic];
      const int nb = cur.nready[ci];
      if (nb <= 0) continue;
      const int32_t* rdy = cur.ready(ci);
      const int pivot = rs.next() % nb;
      const int tmc = in.machine[rdy[pivot] * m + cur.nx(ci)[rdy[pivot]]];
      for (int r = 0; r < nb; ++r) {
        const int j = rdy[r];
        const int k = cur.nx(ci)[j];
        const int mc = in.machine[j * m + k];
        if (mc != tmc) continue;
        // schedule_operation (cpu/common/jssp_types.h:75-97); with M == 1 the
        // predecessor is the previous job's operation.
        const int pred = (m == 1) ? (j > 0 ? cur.jr(ci)[j - 1] : 0) : cur.jr(ci)[j];
        const int f = std::max(cur.mr(ci)[mc], pred) + in.duration[j * m + k];
        const int cm = std::max(cur.cmax[ci], f);
        const double noise = (double)(rs.next() % 1000) / 1000000.0;
        const double score = (double)cm / (double)step + noise;
        ++res.expanded_states;
        if (step == L && cm < res.makespan) {
          res.makespan = cm;
          best_pos = ic;
          best_job = j;
        }
        if (!(cm < best_known)) continue;
        int slot;
        if ((int)heap.size() < K) {
          slot = free_slots.back();
          free_slots.pop_back();
          heap.push_back({score, slot});
          std::push_heap(heap.begin(), heap.end(), HandleLess{});
        } else if (score < heap.front().score) {
          std::pop_heap(heap.begin(), heap.end(), HandleLess{});
          slot = heap.back().idx;  // the evicted child's slot is reused
          heap.back() = {score, slot};
          std::push_heap(heap.begin(), heap.end(), HandleLess{});
        } else {
          continue;
        }
        std::memcpy(nxt.mr(slot), cur.mr(ci), sizeof(int32_t) * nxt.words);
        nxt.mr(slot)[mc] = f;
        nxt.jr(slot)[j] = f;
        nxt.nx(slot)[j] = k + 1;
        // ready list: erase j (order kept), append j if it has ops left
        int32_t* nr = nxt.ready(slot);
        int cnt = 0;
        for (int t = 0; t < nb; ++t)
          if (rdy[t] != j) nr[cnt++] = rdy[t];
        if (k + 1 < m) nr[cnt++] = j;
        nxt.nready[slot] = cnt;
        nxt.cmax[slot] = cm;
        nxt.parent[slot] = ic;
        nxt.job[slot] = j;
      }
    }
    // The next candidates, in heap-array order.
    const int cnt = (int)heap.size();
    if (want_sequence) {
      trail.emplace_back(cnt);
      for (int i = 0; i < cnt; ++i)
        trail.back()[i] = ((uint32_t)nxt.parent[heap[i].idx] << 8) | (uint32_t)nxt.job[heap[i].idx];
    }
    order.resize(cnt);
    for (int i = 0; i < cnt; ++i) order[i] = heap[i].idx;
    std::swap(cur, nxt);
    if ((int)nxt.cmax.size() < K + 1) nxt.init(m, n, K + 1);
    ncur = cnt;
  }
  res.found = res.makespan < JSSP_NO_MAKESPAN;
  if (res.found && want_sequence) {
    // Jobs scheduled at steps L, L-1, ..., 1: the best complete child's job,
    // then each parent's (trail[s-1] holds the children kept at step s).
    std::vector<int> jobs(L);
    jobs[L - 1] = best_job;
    int pos = best_pos;
    for (int st = L - 1; st >= 1; --st) {
      const uint32_t t = trail[st - 1][pos];
      jobs[st - 1] = (int)(t & 0xFFu);
      pos = (int)(t >> 8);
    }
    std::vector<int> next_op(n, 0);
    res.sequence.resize(L);
    for (int st = 0; st < L; ++st) res.sequence[st] = jobs[st] * m + (next_op[jobs[st]]++) + 1;
  }
  return res;
}

}  // namespace jssp

extern "C" int jssp_host_solve_pms(const jssp_instance* in, int32_t beam_width, int32_t best_known, uint32_t seed,
                                   jssp_result* result) {
  if (!in || !in->machine || !in->duration || !result) return JSSP_ERR_INVALID;
  const int n = in->n_jobs, m = in->n_machines;
  if (n <= 0 || m <= 0 || beam_width < 1) return JSSP_ERR_INVALID;
  if (n > JSSP_MAX_JOBS || m > JSSP_MAX_MACHINES || beam_width > JSSP_MAX_BEAM) return JSSP_ERR_LIMIT;
  for (int i = 0; i < n * m; ++i)
    if (in->machine[i] < 0 || in->machine[i] >= m || in->duration[i] < 0) return JSSP_ERR_INVALID;
  jssp::RandStream rs(false, seed);
  jssp::HostPmsResult r = jssp::host_pms(*in, beam_width, best_known, rs, result->sequence != nullptr);
  result->makespan = r.makespan;
  result->found = r.found ? 1 : 0;
  result->expanded_states = r.expanded_states;
  result->levels = n * m;
  result->best_seed_index = 0;
  result->device_ms = 0.0;
  if (r.found && result->sequence)
    for (int i = 0; i < n * m; ++i) result->sequence[i] = r.sequence[i];
  return JSSP_OK;
}
