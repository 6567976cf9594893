// oracle.cpp -- CPU restatement of the reference's beam-search semantics.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
// calls this file; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it, and only as the checker.
//
// Functions and what pins each one (all citations are to the reference tree):
//   orc_ich_exact   cpu/v0_ich/solver.cpp:8-54 + cpu/common/jssp_types.h:40-97.
//                   Pinned by golden makespans of the compiled reference
//                   (tests/golden/golden_ref.json) and by brute force on tiny
//                   instances (K wide enough to keep every state).
//   orc_pms_exact   cpu/v1_pms/solver.cpp:8-70.  Pinned by golden makespans of
//                   the compiled reference and by brute force.
//   orc_gpu_solve   gpu/v1/solver.cu:9-176 + gpu/v2_twophases/solver.cu:9-254,
//                   with the stale-slot behaviour replaced by clean semantics
//                   (only valid, unpruned children survive; the survivor count
//                   is min(V, K)).  There is no executable GPU reference here,
//                   so it is pinned by: brute force (K >= all states), the
//                   closed-form hash vectors in tests, re-simulation of every
//                   returned sequence, and an independent replay of its
//                   trail (tests/test_oracle.py) which recomputes every
//                   child's (score bits, child index) key from the replayed
//                   parents and requires each level's survivors to be the
//                   min(V, K) smallest keys in ascending order (the prefix of
//                   CUB's stable SortPairs, gpu/v1/solver.cu:255-261), the
//                   expanded count, and the best child's makespan and
//                   sequence.
//   orc_simulate    cpu/common/jssp_types.h:75-97 applied to a sequence.
//   orc_pivot_hash  gpu/v1/solver.cu:51-53; pinned by the published inverse
//                   of lowbias32.
//   orc_noise_hash  gpu/v1/solver.cu:122-125; pinned by its inverse built from
//                   the multiplicative inverses of its two odd constants (the
//                   three xor-shifts by 16 are involutions), which must give
//                   back seed ^ cid*c1 ^ bid*c2.
//   orc_gpu_score   gpu/v1/solver.cu:121-127; pinned by exact FP32 values
//                   computed from the definition in numpy.
//   orc_bruteforce  exhaustive search over all operation sequences.
//   orc_glibc_rand  glibc's TYPE_3 random(), the generator behind rand()
//                   (stdlib/random_r.c, glibc 2.35).  Pinned against the C
//                   library's own srand()/rand() in tests.
//   orc_pms_driver  cpu/v1_pms/main.cpp:55-73 over orc_glibc_rand streams.
//                   Pinned: one rank and one pass reduce to orc_pms_exact after
//                   srand(time_seed) (itself pinned by the goldens); the
//                   reference's own beam_search_pms run under the same
//                   per-rank streams (oracle/_ref) gives the same values.
//   orc_hybrid_driver  hybrid/v1/main.cu:45-85 + hybrid/v1/solver.cu:11-44
//                   over orc_glibc_rand streams.  Pinned: one thread reduces
//                   to orc_gpu_driver; each host pass is orc_pms_exact after
//                   srand(its seed) (pinned by the goldens).
//   orc_gpu_driver  gpu/v1/main.cu:55-71 over orc_glibc_rand streams and
//                   orc_gpu_solve.  Pinned: one rank and one pass reduce to
//                   orc_gpu_solve with seed = the first rand() after
//                   srand(time_seed); later passes prune with the earlier
//                   minimum, checked against independent orc_gpu_solve calls.
//
// Float rules: compile without -ffast-math and with -ffp-contract=off; the
// FP32 score is evaluated as (float)cmax / (float)ns + (float)(h % 1000) / 1e9f
// in that order, as gpu/v1/solver.cu:121-127 does.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

constexpr int kNoMakespan = 500000;  // cpu/common/jssp_types.h:33

// ---------------------------------------------------------------------------
// Reference-exact modes.  The state mirrors JobShopState
// (cpu/common/jssp_types.h:26-37): 1-based machine ready times, the
// scheduled sequence, the ready list of 1-based op ids and the finish time of
// every op (nm + 2 entries).
// ---------------------------------------------------------------------------
struct RefInst {
  int n, m, nops;           // nops = n*m + 2 (instance_parser.cpp:51)
  std::vector<int> mc;      // 1-based machine per op id (0 and nm+1 are dummies)
  std::vector<int> dur;     // duration per op id
};

RefInst make_ref_inst(int n, int m, const int* mach, const int* dur) {
  RefInst r;
  r.n = n; r.m = m; r.nops = n * m + 2;
  r.mc.assign(r.nops, 0);
  r.dur.assign(r.nops, 0);
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < m; ++k) {
      int op = j * m + k + 1;               // instance_parser.cpp:58-82
      r.mc[op] = mach[j * m + k] + 1;
      r.dur[op] = dur[j * m + k];
    }
  return r;
}

struct RefState {
  std::vector<int> tr, seq, c, tf;
  int cmax = kNoMakespan;
  double score = 0.0;
};

struct RefCmp {  // jssp_types.h:40-44: max-heap on score
  bool operator()(const RefState& a, const RefState& b) const { return a.score < b.score; }
};

void ref_insert(std::vector<RefState>& beam, const RefState& s, int k) {  // jssp_types.h:47-58
  if ((int)beam.size() < k) {
    beam.push_back(s);
    std::push_heap(beam.begin(), beam.end(), RefCmp{});
  } else if (s.score < beam.front().score) {
    std::pop_heap(beam.begin(), beam.end(), RefCmp{});
    beam.back() = s;
    std::push_heap(beam.begin(), beam.end(), RefCmp{});
  }
}

RefState ref_root(const RefInst& in) {  // jssp_types.h:61-72
  RefState s;
  s.tr.assign(in.m + 1, 0);
  s.tf.assign(in.nops, 0);
  s.cmax = 0;
  for (int j = 0; j < in.n; ++j) s.c.push_back(j * in.m + 1);
  return s;
}

void ref_schedule(int op, const RefInst& in, RefState& s) {  // jssp_types.h:75-97
  const int mc = in.mc[op], M = in.m;
  s.seq.push_back(op);
  int pred = (op % M == 1) ? 0 : s.tf[op - 1];
  int start = std::max(s.tr[mc], pred);
  s.tr[mc] = start + in.dur[op];
  s.tf[op] = s.tr[mc];
  if (op % M != 0) s.c.push_back(op + 1);
  s.cmax = 0;
  for (int k = 1; k <= M; ++k) s.cmax = std::max(s.cmax, s.tr[k]);
}

}  // namespace

extern "C" {

// BS-ICH, cpu/v0_ich/solver.cpp:8-54.  Consumes the process-global rand()
// stream (srand(srand_seed) first when use_srand != 0).  Returns the best
// makespan of the final beam; best_seq_out (n*m, 1-based op ids, may be NULL)
// receives the sequence of the first state in final-beam array order that
// attains it.  *expanded = number of children generated.
int orc_ich_exact(int n, int m, const int* mach, const int* dur, int beam_width,
                  unsigned srand_seed, int use_srand, int* best_seq_out,
                  long long* expanded) {
  if (use_srand) srand(srand_seed);
  RefInst in = make_ref_inst(n, m, mach, dur);
  const int total = n * m;
  std::vector<RefState> beam, next;
  beam.push_back(ref_root(in));
  long long ex = 0;
  for (int step = 0; step < total; ++step) {
    next.clear();
    for (const RefState& cur : beam) {
      for (size_t idx = 0; idx < cur.c.size(); ++idx) {
        RefState child = cur;
        int op = cur.c[idx];
        child.c.erase(child.c.begin() + idx);
        ref_schedule(op, in, child);
        child.score = (double)child.cmax + (rand() % 100) * 0.01;
        ++ex;
        ref_insert(next, child, beam_width);
      }
    }
    beam = next;
    if (beam.empty()) break;
  }
  int best = kNoMakespan;
  const RefState* arg = nullptr;
  for (const RefState& s : beam)
    if (s.cmax < best) { best = s.cmax; arg = &s; }
  if (best_seq_out && arg)
    for (int i = 0; i < total && i < (int)arg->seq.size(); ++i) best_seq_out[i] = arg->seq[i];
  if (expanded) *expanded = ex;
  return best;
}

}  // extern "C"

namespace {

// glibc random() over the TYPE_3 state that rand()/srand() use (glibc
// stdlib/random_r.c; glibc 2.35 in this image): 31 words, state[0] = seed
// (0 -> 1), state[i] = 16807 * state[i-1] mod (2^31 - 1) by Schrage's method,
// then the additive feedback state[f] += state[r] with f three ahead of r,
// output state[f] >> 1, and the first 310 outputs discarded.
struct GlibcRand {
  uint32_t st[31];
  int f = 3, r = 0;
  explicit GlibcRand(unsigned seed) {
    if (seed == 0) seed = 1;
    st[0] = seed;
    int32_t word = (int32_t)seed;
    for (int i = 1; i < 31; ++i) {
      long hi = word / 127773, lo = word % 127773;
      word = (int32_t)(16807 * lo - 2836 * hi);
      if (word < 0) word += 2147483647;
      st[i] = (uint32_t)word;
    }
    for (int i = 0; i < 310; ++i) next();
  }
  int next() {
    st[f] += st[r];
    const int out = (int)(st[f] >> 1);
    f = (f + 1) % 31;
    r = (r + 1) % 31;
    return out;
  }
};

// BS-PMS, cpu/v1_pms/solver.cpp:8-70 (one pass); rnd() stands for rand().
template <class Rand>
int pms_pass(int n, int m, const int* mach, const int* dur, int beam_width, int best_known, Rand&& rnd,
             long long* expanded) {
  RefInst in = make_ref_inst(n, m, mach, dur);
  const int total = n * m;
  std::vector<RefState> cand, next;
  cand.push_back(ref_root(in));
  int best = kNoMakespan;
  long long ex = 0;
  for (int step = 1; step <= total + 10 && !cand.empty(); ++step) {
    next.clear();
    for (int ic = 0; ic < (int)cand.size(); ++ic) {
      const std::vector<int>& ready = cand[ic].c;
      int nb = (int)ready.size();
      if (nb <= 0) continue;
      int pivot = rnd() % nb;
      int tmc = in.mc[ready[pivot]];
      for (int r = 0; r < nb; ++r) {
        if (in.mc[ready[r]] != tmc) continue;
        RefState child = cand[ic];
        int op = ready[r];
        auto it = std::find(child.c.begin(), child.c.end(), op);
        if (it != child.c.end()) child.c.erase(it);
        ref_schedule(op, in, child);
        double noise = (double)(rnd() % 1000) / 1000000.0;
        child.score = ((double)child.cmax / child.seq.size()) + noise;
        ++ex;
        if (child.cmax < best_known) ref_insert(next, child, beam_width);
        if ((int)child.seq.size() == total && child.cmax < best) best = child.cmax;
      }
    }
    cand.swap(next);
  }
  if (expanded) *expanded = ex;
  return best;
}

}  // namespace

extern "C" {

// BS-PMS, cpu/v1_pms/solver.cpp:8-70 (one pass).  Consumes the global rand().
int orc_pms_exact(int n, int m, const int* mach, const int* dur, int beam_width,
                  int best_known, unsigned srand_seed, int use_srand,
                  long long* expanded) {
  if (use_srand) srand(srand_seed);
  return pms_pass(n, m, mach, dur, beam_width, best_known, [] { return rand(); }, expanded);
}

// out[i] = the i-th rand() after srand(seed), from the restated generator.
void orc_glibc_rand(unsigned seed, int count, int* out) {
  GlibcRand g(seed);
  for (int i = 0; i < count; ++i) out[i] = g.next();
}

// The PMS driver loop, cpu/v1_pms/main.cpp:55-73: rank r seeds its stream
// with srand(time_seed + r); each iteration every rank runs
// beam_search_pms(inst, K, global_best), local_best = min(local_best,
// result), then global_best = MPI_Allreduce(local_best, MIN).
// pass_best[e] = global_best after iteration e; *best_index = e * ranks + r
// of the first solve reaching the final value (-1 if none);
// *expanded = children summed over all solves.  Returns global_best.
int orc_pms_driver(int n, int m, const int* mach, const int* dur, int beam_width, int ranks,
                   int iterations, unsigned time_seed, int* pass_best, int* best_index,
                   long long* expanded) {
  std::vector<GlibcRand> streams;
  for (int r = 0; r < ranks; ++r) streams.emplace_back(time_seed + (unsigned)r);
  std::vector<int> local(ranks, kNoMakespan);
  int global = kNoMakespan, bi = -1;
  long long ex = 0;
  for (int it = 0; it < iterations; ++it) {
    for (int r = 0; r < ranks; ++r) {
      long long e1 = 0;
      GlibcRand& g = streams[r];
      const int ms = pms_pass(n, m, mach, dur, beam_width, global, [&g] { return g.next(); }, &e1);
      ex += e1;
      if (ms < local[r]) local[r] = ms;
    }
    int red = kNoMakespan;
    for (int r = 0; r < ranks; ++r) red = std::min(red, local[r]);
    if (red < global) {
      for (int r = 0; r < ranks && bi < it * ranks; ++r)
        if (local[r] == red) bi = it * ranks + r;
    }
    global = red;
    if (pass_best) pass_best[it] = global;
  }
  if (best_index) *best_index = bi;
  if (expanded) *expanded = ex;
  return global;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Clean GPU semantics (gpu/v1/solver.cu, gpu/v2_twophases/solver.cu).
// ---------------------------------------------------------------------------
namespace {

uint32_t pivot_hash(uint32_t h) {  // gpu/v1/solver.cu:51-53
  h ^= h >> 16; h *= 0x7feb352dU;
  h ^= h >> 15; h *= 0x846ca68bU;
  h ^= h >> 16;
  return h;
}

uint32_t noise_hash(uint32_t seed, uint32_t cid, uint32_t bid) {  // gpu/v1/solver.cu:122-125
  uint32_t h = seed ^ (cid * 0x85ebca6bU) ^ (bid * 0xc2b2ae35U);
  h ^= h >> 16; h *= 0x85ebca6bU;
  h ^= h >> 16; h *= 0xc2b2ae35U;
  h ^= h >> 16;
  return h;
}

float gpu_score(int cmax, int ns, uint32_t h) {  // gpu/v1/solver.cu:121-127
  float base = (ns == 0) ? (float)cmax : ((float)cmax / (float)ns);
  float noise = (float)(h % 1000u) / 1000000000.0f;
  return base + noise;
}

struct GState {
  std::vector<int> mr, jr, nx, ready;  // 0-based machines / jobs; ready = job ids
  int cmax = 0;
};

struct Child {
  uint64_t key;
  int p, j;
};

// Schedules the next op of job j on state s (gpu_snapshot.cuh:47-70 with
// 0-based ids).  The reference's predecessor rule (op % M == 1 ? 0 :
// tf[op-1]) reads the previous job's op when M == 1; that is reproduced.
int sched_finish(const GState& s, int n, int m, const int* mach, const int* dur, int j, int* mc_out) {
  (void)n;
  int k = s.nx[j];
  int op = j * m + k;
  int mc = mach[op];
  int pred = (m == 1) ? (j > 0 ? s.jr[j - 1] : 0) : s.jr[j];
  int start = std::max(s.mr[mc], pred);
  if (mc_out) *mc_out = mc;
  return start + dur[op];
}

}  // namespace

extern "C" {

// Mode numbers match include/jssp.h: 1 = ICH_DET (operation-level branching,
// ready list in erase-then-append order), 2 = PMS_GPU (hash-chosen pivot
// machine, ready list in swap-remove order).
//
// Returns min(best_known, best complete makespan).  *found = 1 when a
// complete child with cmax < best_known exists; seq_out (n*m 1-based op ids)
// then holds the sequence of the complete child minimising (cmax, child
// index).  trail_out, if non-NULL, receives ((n*m - 1) x beam_width) entries:
// trail[l*K + q] = (parent rank << 8) | job of survivor q after level l.
// pcount_out (n*m + 1 entries) receives the parent count per level.
int orc_gpu_solve(int n, int m, const int* mach, const int* dur, int beam_width,
                  int mode, int best_known, unsigned seed, int* seq_out,
                  long long* expanded, int* found, int* trail_out, int* pcount_out) {
  const int L = n * m, B = n, K = beam_width;
  std::vector<GState> beam(1), next;
  GState& root = beam[0];
  root.mr.assign(m, 0); root.jr.assign(n, 0); root.nx.assign(n, 0);
  for (int j = 0; j < n; ++j) root.ready.push_back(j);
  root.cmax = 0;
  long long ex = 0;
  uint64_t best = ~0ull;      // (cmax << 32) | child index at the last level
  int best_parent = -1, best_job = -1;
  std::vector<int> trail;     // (L-1) x K
  if (L > 1) trail.assign((size_t)(L - 1) * K, -1);
  if (pcount_out) {
    pcount_out[0] = 1;
    pcount_out[L] = 0;  // the last level keeps no beam
  }
  std::vector<Child> kids;
  for (int step = 0; step < L; ++step) {
    const uint32_t lseed = seed + (uint32_t)step;
    const int ns = step + 1;
    const bool last = (step == L - 1);
    kids.clear();
    for (int p = 0; p < (int)beam.size(); ++p) {
      const GState& st = beam[p];
      const int r = (int)st.ready.size();
      if (r == 0) continue;
      std::vector<int> slots;
      if (mode == 1) {
        slots = st.ready;
      } else {
        uint32_t h = pivot_hash(lseed + (uint32_t)p);
        int pj = st.ready[h % (uint32_t)r];
        int tmc = mach[pj * m + st.nx[pj]];
        for (int q = 0; q < r; ++q) {
          int jj = st.ready[q];
          if (mach[jj * m + st.nx[jj]] == tmc) slots.push_back(jj);
        }
      }
      for (int s = 0; s < (int)slots.size(); ++s) {
        int j = slots[s];
        int f = sched_finish(st, n, m, mach, dur, j, nullptr);
        int cmax = std::max(st.cmax, f);
        ++ex;
        if (!(cmax < best_known)) continue;
        uint32_t x = (uint32_t)(p * B + s);
        if (last) {
          uint64_t cand = ((uint64_t)(uint32_t)cmax << 32) | x;
          if (cand < best) { best = cand; best_parent = p; best_job = j; }
          continue;
        }
        float sc = gpu_score(cmax, ns, noise_hash(lseed, (uint32_t)p, (uint32_t)s));
        uint32_t bits;
        std::memcpy(&bits, &sc, 4);
        kids.push_back({((uint64_t)bits << 32) | x, p, j});
      }
    }
    if (last) break;
    std::sort(kids.begin(), kids.end(), [](const Child& a, const Child& b) { return a.key < b.key; });
    int keep = std::min((int)kids.size(), K);
    next.assign(keep, GState());
    for (int q = 0; q < keep; ++q) {
      const Child& c = kids[q];
      const GState& st = beam[c.p];
      GState& ch = next[q];
      ch = st;
      int mc;
      int f = sched_finish(st, n, m, mach, dur, c.j, &mc);
      ch.mr[mc] = f; ch.jr[c.j] = f; ch.nx[c.j] += 1;
      ch.cmax = std::max(st.cmax, f);
      auto it = std::find(ch.ready.begin(), ch.ready.end(), c.j);
      if (mode == 1) {
        ch.ready.erase(it);
      } else {
        *it = ch.ready.back();
        ch.ready.pop_back();
      }
      if (ch.nx[c.j] < m) ch.ready.push_back(c.j);
      trail[(size_t)step * K + q] = (c.p << 8) | c.j;
    }
    if (pcount_out) pcount_out[step + 1] = keep;
    beam.swap(next);
  }
  if (expanded) *expanded = ex;
  if (trail_out && L > 1) std::memcpy(trail_out, trail.data(), trail.size() * sizeof(int));
  int ok = (best != ~0ull);
  if (found) *found = ok;
  if (ok && seq_out) {
    std::vector<int> jobs(L);
    jobs[L - 1] = best_job;
    int r = best_parent;
    for (int l = L - 2; l >= 0; --l) {
      int t = trail[(size_t)l * K + r];
      jobs[l] = t & 0xFF;
      r = t >> 8;
    }
    std::vector<int> cnt(n, 0);
    for (int l = 0; l < L; ++l) {
      int j = jobs[l];
      seq_out[l] = j * m + cnt[j] + 1;
      cnt[j]++;
    }
  }
  int bc = ok ? (int)(best >> 32) : kNoMakespan;
  return ok ? std::min(bc, best_known) : best_known;
}

// Re-simulates a sequence of 1-based op ids with the reference's scheduling
// rule (cpu/common/jssp_types.h:75-97).  Returns the makespan, or -1 if the
// sequence is not a permutation of all n*m ops in job order.
int orc_simulate(int n, int m, const int* mach, const int* dur, const int* seq) {
  std::vector<int> mr(m, 0), jr(n, 0), nx(n, 0);
  int cmax = 0;
  for (int i = 0; i < n * m; ++i) {
    int op = seq[i] - 1;
    if (op < 0 || op >= n * m) return -1;
    int j = op / m, k = op % m;
    if (nx[j] != k) return -1;
    int mc = mach[op];
    int pred = (m == 1) ? (j > 0 ? jr[j - 1] : 0) : jr[j];
    int f = std::max(mr[mc], pred) + dur[op];
    mr[mc] = f; jr[j] = f; nx[j]++;
    cmax = std::max(cmax, f);
  }
  return cmax;
}

static int bf_rec(int n, int m, const int* mach, const int* dur, std::vector<int>& mr,
                  std::vector<int>& jr, std::vector<int>& nx, int left, int cmax, int bestsofar) {
  if (left == 0) return cmax;
  int best = bestsofar;
  for (int j = 0; j < n; ++j) {
    if (nx[j] >= m) continue;
    int op = j * m + nx[j];
    int mc = mach[op];
    int pred = (m == 1) ? (j > 0 ? jr[j - 1] : 0) : jr[j];
    int f = std::max(mr[mc], pred) + dur[op];
    int c2 = std::max(cmax, f);
    if (c2 >= best) continue;  // cmax never decreases
    int smr = mr[mc], sjr = jr[j];
    mr[mc] = f; jr[j] = f; nx[j]++;
    int v = bf_rec(n, m, mach, dur, mr, jr, nx, left - 1, c2, best);
    if (v < best) best = v;
    mr[mc] = smr; jr[j] = sjr; nx[j]--;
  }
  return best;
}

// The GPU driver loop, gpu/v1/main.cu:55-71 and gpu/v2_twophases/main.cu:
// 55-71: rank r seeds its stream with srand(time_seed + r); each iteration
// every rank draws seed = rand() and runs the PMS_GPU beam search with
// best_known = global_best; local_best = min(local_best, result), then
// global_best = MPI_Allreduce(local_best, MIN).  Outputs as orc_pms_driver;
// seq_out (n*m) receives the sequence of the first solve reaching the final
// value.
int orc_gpu_driver(int n, int m, const int* mach, const int* dur, int beam_width, int ranks,
                   int iterations, unsigned time_seed, int* pass_best, int* best_index,
                   long long* expanded, int* seq_out) {
  std::vector<GlibcRand> streams;
  for (int r = 0; r < ranks; ++r) streams.emplace_back(time_seed + (unsigned)r);
  std::vector<int> local(ranks, kNoMakespan), seq((size_t)n * m);
  int global = kNoMakespan, bi = -1;
  long long ex = 0;
  for (int it = 0; it < iterations; ++it) {
    int running = global;
    for (int r = 0; r < ranks; ++r) {
      const unsigned seed = (unsigned)streams[r].next();
      long long e1 = 0;
      int found = 0;
      const int ms = orc_gpu_solve(n, m, mach, dur, beam_width, 2, global, seed, seq.data(), &e1, &found,
                                   nullptr, nullptr);
      ex += e1;
      if (ms < local[r]) local[r] = ms;
      if (found && ms < running) {
        running = ms;
        bi = it * ranks + r;
        if (seq_out) std::memcpy(seq_out, seq.data(), seq.size() * sizeof(int));
      }
    }
    int red = kNoMakespan;
    for (int r = 0; r < ranks; ++r) red = std::min(red, local[r]);
    global = red;
    if (pass_best) pass_best[it] = global;
  }
  if (best_index) *best_index = bi;
  if (expanded) *expanded = ex;
  return global;
}

// The hybrid driver loop, hybrid/v1/main.cu:45-85 with the iteration of
// hybrid/v1/solver.cu:11-44: rank r seeds srand(time_seed + r); each
// iteration, thread 0 draws seed = rand() and runs the PMS_GPU search with
// beam gpu_k, threads tid = 1 .. threads-1 run beam_search_pms with beam
// cpu_k from srand(time_seed + r*1000 + it*100 + tid), all with best_known =
// global_best; iteration best = min(best_known, every result); local_best =
// min(local_best, iteration best); global_best = MPI_Allreduce(MIN).  Every
// thread draws from its own generator (in the reference the threads share
// one, and race).  pass_best[it] = global_best after iteration it; returns
// global_best.
int orc_hybrid_driver(int n, int m, const int* mach, const int* dur, int gpu_k, int cpu_k, int threads, int ranks,
                      int iterations, unsigned time_seed, int* pass_best, long long* expanded) {
  std::vector<GlibcRand> streams;
  for (int r = 0; r < ranks; ++r) streams.emplace_back(time_seed + (unsigned)r);
  std::vector<int> local(ranks, kNoMakespan), seq((size_t)n * m);
  int global = kNoMakespan;
  long long ex = 0;
  for (int it = 0; it < iterations; ++it) {
    for (int r = 0; r < ranks; ++r) {
      int best_iter = global;
      const unsigned seed = (unsigned)streams[r].next();
      long long e1 = 0;
      int found = 0;
      int v = orc_gpu_solve(n, m, mach, dur, gpu_k, 2, global, seed, seq.data(), &e1, &found, nullptr, nullptr);
      ex += e1;
      best_iter = std::min(best_iter, v);
      for (int tid = 1; tid < threads; ++tid) {
        GlibcRand g(time_seed + (unsigned)(r * 1000 + it * 100 + tid));
        long long e2 = 0;
        v = pms_pass(n, m, mach, dur, cpu_k, global, [&g] { return g.next(); }, &e2);
        ex += e2;
        best_iter = std::min(best_iter, v);
      }
      if (best_iter < local[r]) local[r] = best_iter;
    }
    int red = kNoMakespan;
    for (int r = 0; r < ranks; ++r) red = std::min(red, local[r]);
    global = red;
    if (pass_best) pass_best[it] = global;
  }
  if (expanded) *expanded = ex;
  return global;
}

// Exhaustive minimum makespan over every operation sequence (tiny instances).
int orc_bruteforce(int n, int m, const int* mach, const int* dur) {
  std::vector<int> mr(m, 0), jr(n, 0), nx(n, 0);
  return bf_rec(n, m, mach, dur, mr, jr, nx, n * m, 0, 0x7fffffff);
}

// The two hashes, exported for their known-answer / inverse tests.
uint32_t orc_pivot_hash(uint32_t h) { return pivot_hash(h); }
uint32_t orc_noise_hash(uint32_t seed, uint32_t cid, uint32_t bid) { return noise_hash(seed, cid, bid); }
float orc_gpu_score(int cmax, int ns, uint32_t h) { return gpu_score(cmax, ns, h); }

}  // extern "C"
