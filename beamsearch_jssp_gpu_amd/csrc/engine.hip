// engine.hip -- single-instance solves: workspace planning, the level loop,
// the exact (host-replay) modes and the C ABI entry points around them.
//
// Reference counterparts: run_gpu_beam_search_v1 (gpu/v1/solver.cu:178-298),
// run_gpu_beam_search_v2 (gpu/v2_twophases/solver.cu:256-386),
// beam_search_ich (cpu/v0_ich/solver.cpp:8-54), beam_search_pms
// (cpu/v1_pms/solver.cpp:8-70).  Differences by design: the device arena and
// stream live in the engine and are reused across solves (the reference
// mallocs and frees nine buffers per call), the level loop enqueues the same
// three launches at every level with every count and the level itself kept
// on the device (the reference sizes grids from min(P*n, K) on the host), and
// errors are returned, never exit().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "engine.hpp"
#include "rand_stream.hpp"
#include "level_kernels.hpp"
#include "snapshot.hpp"
#include "workspace.hpp"

namespace jssp {

int set_error(jssp_engine* e, int code, const char* where, const std::string& msg) {
  if (e) e->err = std::string(where) + ": " + msg;
  return code;
}

int ensure_device_arena(jssp_engine* e, size_t bytes) {
  if (bytes <= e->dcap) return JSSP_OK;
  if (e->dbuf) {
    JSSP_HIP(e, hipStreamSynchronize(e->stream));
    JSSP_HIP(e, hipFree(e->dbuf));
    e->dbuf = nullptr;
    e->dcap = 0;
  }
  size_t cap = bytes + bytes / 4;
  if (hipMalloc(&e->dbuf, cap) != hipSuccess) {
    e->dbuf = nullptr;
    (void)hipGetLastError();
    return set_error(e, JSSP_ERR_NOMEM, "hipMalloc", "device arena of " + std::to_string(cap) + " bytes");
  }
  e->dcap = cap;
  return JSSP_OK;
}

int ensure_exact_host(jssp_engine* e, size_t bytes) {
  if (bytes <= e->xcap) return JSSP_OK;
  if (e->xbuf) {
    JSSP_HIP(e, hipStreamSynchronize(e->stream));
    JSSP_HIP(e, hipHostFree(e->xbuf));
    e->xbuf = nullptr;
    e->xcap = 0;
  }
  const size_t cap = bytes + bytes / 4 + 4096;
  if (hipHostMalloc(&e->xbuf, cap, hipHostMallocCoherent) != hipSuccess) {
    e->xbuf = nullptr;
    (void)hipGetLastError();
    return set_error(e, JSSP_ERR_NOMEM, "hipHostMalloc", "exact-mode host arrays of " + std::to_string(cap) + " bytes");
  }
  e->xcap = cap;
  return JSSP_OK;
}

int ensure_host_staging(jssp_engine* e, size_t bytes) {
  if (bytes <= e->hcap) return JSSP_OK;
  if (e->hbuf) {
    JSSP_HIP(e, hipStreamSynchronize(e->stream));
    JSSP_HIP(e, hipHostFree(e->hbuf));
    e->hbuf = nullptr;
    e->hcap = 0;
  }
  size_t cap = bytes + bytes / 4 + 4096;
  if (hipHostMalloc(&e->hbuf, cap, hipHostMallocDefault) != hipSuccess) {
    e->hbuf = nullptr;
    (void)hipGetLastError();
    return set_error(e, JSSP_ERR_NOMEM, "hipHostMalloc", "pinned staging of " + std::to_string(cap) + " bytes");
  }
  e->hcap = cap;
  return JSSP_OK;
}

int check_instance(jssp_engine* e, const jssp_instance* in) {
  if (!in || !in->machine || !in->duration) return set_error(e, JSSP_ERR_INVALID, "instance", "null instance");
  const int n = in->n_jobs, m = in->n_machines;
  if (n <= 0 || m <= 0) return set_error(e, JSSP_ERR_INVALID, "instance", "n_jobs and n_machines must be >= 1");
  if (n > JSSP_MAX_JOBS || m > JSSP_MAX_MACHINES)
    return set_error(e, JSSP_ERR_LIMIT, "instance", "n_jobs <= 128 and n_machines <= 255 required");
  long long total = 0;
  for (int i = 0; i < n * m; ++i) {
    if (in->machine[i] < 0 || in->machine[i] >= m)
      return set_error(e, JSSP_ERR_INVALID, "instance", "machine id out of range at op " + std::to_string(i));
    if (in->duration[i] < 0)
      return set_error(e, JSSP_ERR_INVALID, "instance", "negative duration at op " + std::to_string(i));
    total += in->duration[i];
  }
  if (total >= (1ll << 30)) return set_error(e, JSSP_ERR_LIMIT, "instance", "sum of durations must be < 2^30");
  return JSSP_OK;
}

int check_params(jssp_engine* e, const jssp_instance* in, const jssp_params* p) {
  if (!p) return set_error(e, JSSP_ERR_INVALID, "params", "null params");
  if (p->beam_width < 1) return set_error(e, JSSP_ERR_INVALID, "params", "beam_width must be >= 1");
  if (p->beam_width > JSSP_MAX_BEAM) return set_error(e, JSSP_ERR_LIMIT, "params", "beam_width > 65536");
  if (p->mode < JSSP_MODE_ICH_EXACT || p->mode > JSSP_MODE_PMS_EXACT)
    return set_error(e, JSSP_ERR_INVALID, "params", "unknown mode");
  if (p->phases != 0 && p->phases != 1 && p->phases != 2)
    return set_error(e, JSSP_ERR_INVALID, "params", "phases must be 1 or 2");
  if ((long long)p->beam_width * in->n_jobs >= (1ll << 31))
    return set_error(e, JSSP_ERR_LIMIT, "params", "beam_width * n_jobs must be < 2^31");
  return JSSP_OK;
}

// ---------------------------------------------------------------------------
// Workspace planning.
int default_lanes(int n) { return n <= 16 ? 16 : (n <= 32 ? 32 : 64); }

// Dynamic LDS of the rank kernel (rank_lds_base, plus the op table where the
// kernel stages it).
static size_t rank_lds_for(int K, int n, int m, bool fused, int waves) {
  return rank_lds_base(K, n, m, fused, waves) + (rank_ops_in_lds(K, n, m, fused, waves) ? (size_t)n * m * 8 : 0);
}

bool fused_eligible(const jssp_instance* in, const jssp_params* p) {
  const bool exact = (p->mode == JSSP_MODE_ICH_EXACT || p->mode == JSSP_MODE_PMS_EXACT);
  return !exact && p->phases == 2 && !(p->flags & (JSSP_FLAG_GENERIC_DIGIT | JSSP_FLAG_UNFUSED)) &&
         p->beam_width <= kRankLdsMax &&
         rank_lds_for(p->beam_width, in->n_jobs, in->n_machines, true, kRankWaves) <= kRankFusedLdsMax &&
         exact_digit_applies(in);
}

Shape make_shape(const jssp_instance* in, const jssp_params* p, int lanes, int rank_waves, int rank_batch,
                 int filter_cap) {
  Shape s;
  s.n = in->n_jobs; s.m = in->n_machines;
  s.S = 2 * s.n + s.m + 1;
  s.B = s.n;
  s.K = p->beam_width;
  s.L = s.n * s.m;
  s.exact = (p->mode == JSSP_MODE_ICH_EXACT || p->mode == JSSP_MODE_PMS_EXACT);
  s.single = !s.exact && p->phases == 1;
  s.fused = (lanes == 0 || lanes == kLanesFused) && fused_eligible(in, p);
  int G = lanes;
  if (G != 16 && G != 32 && G != 64) G = default_lanes(s.n);
  if (G < s.n || s.single || s.fused) G = 64;
  s.G = G;
  const long long N = (long long)s.K * s.B;
  const int parents_per_block = kScoreWaves * (64 / s.G);
  s.nblk_score = (int)std::max(1, std::min((s.K + parents_per_block - 1) / parents_per_block, 512));
  // Throughput shapes (batches, the server): at most 64 filter blocks per
  // instance, each taking several 1024-child tiles, kFilterTilesTp per pass,
  // so fewer blocks re-read the instance's histogram copies.
  const long long cap = filter_cap > 0 ? filter_cap : (rank_waves == kRankWavesThroughput ? 64 : 512);
  const long long tiles = (N + kFilterThreads - 1) / kFilterThreads;
  const long long per = rank_waves == kRankWavesThroughput ? kFilterTilesTp : 1;
  s.nblk_filter = (int)std::max(1ll, std::min((tiles + per - 1) / per, cap));
  s.rank_waves = rank_waves == kRankWavesThroughput ? kRankWavesThroughput : kRankWaves;
  s.rank_batch = s.rank_waves == kRankWavesThroughput ? std::max(1, std::min(rank_batch, kRankBatch)) : 1;
  s.nblk_rank = rank_blocks(s.K, s.rank_waves, s.rank_batch);
  return s;
}

size_t carve(const Shape& s, char* base, Work& w) {
  Bump b;
  b.base = base;
  const size_t N = (size_t)s.K * s.B;
  w.ops = b.take<int2>((size_t)s.n * s.m);
  w.beam[0] = b.take<int>((size_t)s.K * s.S);
  w.beam[1] = b.take<int>((size_t)s.K * s.S);
  w.cbuf = s.single ? b.take<int>(N * child_stride(s.S)) : nullptr;
  w.keys = b.take<unsigned long long>(std::max(N, (size_t)2 * s.K));
  w.crec[0] = b.take<unsigned long long>(N);
  w.crec[1] = b.take<unsigned long long>(N);
  w.cand = b.take<unsigned long long>(N);
  w.ref_buf = b.take<unsigned long long>(N + s.K);
  w.pvalid = b.take<uint8_t>(s.K);
  w.tcnt = b.take<uint32_t>((N + kFilterThreads - 1) / kFilterThreads + 2);
  w.hist[0] = b.take<uint32_t>((size_t)kNB * kHistCopies);
  w.hist[1] = b.take<uint32_t>((size_t)kNB * kHistCopies);
  w.bstat = b.take<BlockStat>(std::max(s.nblk_score, s.nblk_rank));
  w.rstat = b.take<RankStat>(s.nblk_rank);
  w.sel[0] = b.take<SelState>(1);
  w.sel[1] = b.take<SelState>(1);
  w.sctr[0] = b.take<uint32_t>((size_t)kSureSubs * kCtrStride);
  w.sctr[1] = b.take<uint32_t>((size_t)kSureSubs * kCtrStride);
  w.pcount = b.take<int>(s.L + 1);
  w.trail = b.take<int>((size_t)std::max(1, s.L - 1) * s.K);
  w.jump = s.exact ? nullptr : b.take<uint16_t>((size_t)((std::max(1, s.L - 1) + 31) / 32) * s.K);
  w.expanded = b.take<unsigned long long>(1);
  w.lvl_exp = b.take<uint32_t>(s.L);
  w.sel_log = b.take<SelState>(s.L);
  w.seq = b.take<int>(s.L);
  w.result = b.take<int>(4);
  if (s.exact) {
    w.ex_cmax = b.take<int>(N);
    w.ex_job = b.take<uint8_t>(N);
    w.ex_mc = b.take<uint8_t>(N);
    w.ex_rcount = b.take<int>(s.K);
    w.ex_order = b.take<int>(s.K);
  } else {
    w.ex_cmax = nullptr; w.ex_job = nullptr; w.ex_mc = nullptr; w.ex_rcount = nullptr; w.ex_order = nullptr;
  }
  return b.off + 256;
}

// The exact selection digit and its key word (cmax << 11 | the score's ulps
// above fl(cmax / ns)) are valid when
// (1) at every level a lower makespan always gives a lower FP32 score: for
//     cmax c1 < c2 at ns scheduled ops, score(c1) <= c1/ns + 1e-6 + ulp and
//     score(c2) >= c2/ns - ulp/2, so it suffices that 1/L > 1e-6 +
//     2 ulp(2 dmax + 1), since every score is below dmax + 1e-6 (a
//     semi-active schedule grows cmax by at most one duration per op, so
//     cmax/ns <= dmax); the key word then orders children as the score does;
// (2) cmax fits 21 bits below the all-ones word: cmax <= the sum of all
//     durations (every start is a finish time of an earlier op);
// (3) the noise offset fits 11 bits: the busiest machine has at least the
//     average load, so cmax >= ns dmin / m and fl(cmax / ns) >= 2^-7 when
//     128 dmin >= m; an ulp there is at least 2^-30 and the noise (below
//     1e-6) adds at most 1075 ulps.
bool exact_digit_applies(const jssp_instance* in) {
  int dmax = 0, dmin = INT32_MAX;
  long long dsum = 0;
  for (int i = 0; i < in->n_jobs * in->n_machines; ++i) {
    dmax = std::max(dmax, (int)in->duration[i]);
    dmin = std::min(dmin, (int)in->duration[i]);
    dsum += in->duration[i];
  }
  if (dmax <= 0 || dmin <= 0) return false;
  if (dsum >= (1ll << (32 - 11)) - 1) return false;
  if (128ll * dmin < in->n_machines) return false;
  int e = 0;
  std::frexp(2.0 * dmax + 1.0, &e);  // 2 dmax + 1 in [2^(e-1), 2^e)
  const double ulp = std::ldexp(1.0, (e - 1) - 23);
  const double L = (double)in->n_jobs * in->n_machines;
  return 1.0 / L > 1e-6 + 2.0 * ulp;
}

LevelArgs base_args(const Shape& s, const Work& w, const jssp_params* p) {
  LevelArgs a;
  std::memset(&a, 0, sizeof(a));
  a.ops = w.ops;
  a.beam_cur = w.beam[0];
  a.beam_next = w.beam[1];
  a.cbuf = w.cbuf;
  a.keys = w.keys;
  a.crec = w.crec[0];
  a.cand = w.cand;
  a.ref_buf = w.ref_buf;
  a.pvalid = w.pvalid;
  a.tcnt = w.tcnt;
  a.hist = w.hist[0];
  a.hist_next = w.hist[1];
  a.bstat = w.bstat;
  a.rstat = w.rstat;
  a.sel = w.sel[0];
  a.sctr = w.sctr[0];
  a.sctr_next = w.sctr[1];
  a.sel_prev = w.sel[1];
  a.sel_next = w.sel[1];
  a.sel_log = nullptr;
  a.pcount = w.pcount;
  a.trail = w.trail;
  a.expanded = w.expanded;
  a.lvl_exp = w.lvl_exp;
  a.seq_out = w.seq;
  a.result = w.result;
  a.ex_cmax = w.ex_cmax; a.ex_job = w.ex_job; a.ex_mc = w.ex_mc;
  a.ex_rcount = w.ex_rcount; a.ex_order = w.ex_order;
  a.n = s.n; a.m = s.m; a.S = s.S; a.B = s.B; a.K = s.K; a.L = s.L;
  a.level = 0;
  a.best_known = p->best_known;
  a.nblk_score = s.nblk_score;
  a.nblk_rank = s.nblk_rank;
  a.bm_words = (p->flags & JSSP_FLAG_RADIX_CUT) ? 0u
               : (uint32_t)std::min<size_t>(((size_t)s.K * s.B + 31) / 32, kBitmapWordsMax);
  a.lseed = p->seed;
  a.bt_parent = -1; a.bt_job = -1;
  a.mode = (p->mode == JSSP_MODE_PMS_GPU) ? 2 : 1;
  a.G = s.G;
  a.hist_copies = hist_copies_for(s.K);
  return a;
}

InstDesc make_desc(const Shape& s, const Work& w, const jssp_params* p, const jssp_instance* in) {
  InstDesc d;
  std::memset(&d, 0, sizeof(d));
  d.base = base_args(s, w, p);
  d.base.exact_digit = (!(p->flags & JSSP_FLAG_GENERIC_DIGIT) && exact_digit_applies(in)) ? 1 : 0;
  d.beam[0] = w.beam[0]; d.beam[1] = w.beam[1];
  d.hist[0] = w.hist[0]; d.hist[1] = w.hist[1];
  d.sel[0] = w.sel[0]; d.sel[1] = w.sel[1];
  d.sctr[0] = w.sctr[0]; d.sctr[1] = w.sctr[1];
  d.crec[0] = w.crec[0]; d.crec[1] = w.crec[1];
  d.nblk_score = s.nblk_score;
  d.nblk_filter = s.nblk_filter;
  d.nblk_rank = s.nblk_rank;
  d.fused = s.fused ? 1 : 0;
  // Score partials per level: one per score block, or one per rank block
  // when the rank kernel scores the next level.
  d.base.nblk_score = s.fused ? s.nblk_rank : s.nblk_score;
  d.lvl_score = 0;
  d.lvl_filter = 0;
  d.lvl_rank = -1;
  return d;
}

void append_maps(const Shape& s, int inst, std::vector<BlockRef>& score, std::vector<BlockRef>& filter,
                 std::vector<BlockRef>& rank) {
  if (!s.fused)
    for (int b = 0; b < s.nblk_score; ++b) score.push_back(make_int2(inst, b));
  for (int b = 0; b < s.nblk_filter; ++b) filter.push_back(make_int2(inst, b));
  for (int b = 0; b < s.nblk_rank; ++b) rank.push_back(make_int2(inst, b));
}

size_t filter_lds_bytes(const Shape& s, const jssp_params* p) {
  if (p->flags & JSSP_FLAG_RADIX_CUT) return 0;
  return 4 * std::min<size_t>(((size_t)s.K * s.B + 31) / 32, kBitmapWordsMax);
}

size_t rank_lds_bytes(const Shape& s) { return rank_lds_for(s.K, s.n, s.m, s.fused, s.rank_waves); }

int rank_variant(const Shape& s, const jssp_params* p) {
  if (!s.fused) return s.K > kRankLdsMax ? kRankUnfusedWide : kRankUnfused;
  const int sz = rank_size_class(s.S, s.n);
  if (p->mode == JSSP_MODE_PMS_GPU) return sz == 0 ? kRankFusedPmsSmall : (sz == 1 ? kRankFusedPmsMedium : kRankFusedPms);
  return sz == 0 ? kRankFusedIchSmall : (sz == 1 ? kRankFusedIchMedium : kRankFusedIch);
}

int upload_ops(jssp_engine* e, const jssp_instance* in, int2* d_ops) {
  const int total = in->n_jobs * in->n_machines;
  int rc = ensure_host_staging(e, (size_t)total * sizeof(int2));
  if (rc) return rc;
  int2* h = reinterpret_cast<int2*>(e->hbuf);
  for (int i = 0; i < total; ++i) { h[i].x = in->machine[i]; h[i].y = in->duration[i]; }
  JSSP_HIP(e, hipMemcpyAsync(d_ops, h, (size_t)total * sizeof(int2), hipMemcpyHostToDevice, e->stream));
  // The staging buffer is reused right after; finish the copy first.
  JSSP_HIP(e, hipStreamSynchronize(e->stream));
  return JSSP_OK;
}

hipEvent_t Prof::get() {
  if (next >= e->pev.size()) {
    hipEvent_t ev;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    e->pev.push_back(ev);
  }
  return e->pev[next++];
}
const EventPair* Prof::pair(int kind, EventPair* ep, bool on) {
  if (!e->profiling || !on) return nullptr;
  ep->start = get();
  ep->stop = get();
  if (!ep->start || !ep->stop) return nullptr;
  marks.push_back({kind, {ep->start, ep->stop}});
  return ep;
}
void Prof::collect() {
  for (auto& v : e->kl_ms) v.clear();
  for (auto& mk : marks) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, mk.second.first, mk.second.second) == hipSuccess) {
      e->kt[mk.first].ms += ms;
      e->kt[mk.first].launches += 1;
      e->kl_ms[mk.first].push_back(ms);
    } else {
      e->kl_ms[mk.first].push_back(-1.f);
    }
  }
}

int enqueue_single_solve(jssp_engine* e, const Shape& s, const jssp_params* q, InstDesc* d_desc, Prof* prof) {
  const size_t flds = filter_lds_bytes(s, q), rlds = rank_lds_bytes(s);
  const int variant = rank_variant(s, q);
  JSSP_HIP(e, hipEventRecord(e->ev_start, e->stream));
  JSSP_HIP(e, launch_step_init(d_desc, nullptr, 1, e->stream));
  for (int l = 0; l < s.L; ++l) {
    EventPair ep;
    if (!s.fused)
      JSSP_HIP(e, launch_step_score(d_desc, nullptr, s.nblk_score, s.single, e->stream,
                                    prof ? prof->pair(kKindScore, &ep) : nullptr));
    JSSP_HIP(e, launch_step_filter(d_desc, nullptr, s.nblk_filter, flds, false, e->stream,
                                   prof ? prof->pair(kKindFilter, &ep, l < s.L - 1) : nullptr));
    JSSP_HIP(e, launch_step_rank(d_desc, nullptr, s.nblk_rank, rlds, variant, s.rank_waves, e->stream,
                                 prof ? prof->pair(kKindRank, &ep, l < s.L - 1) : nullptr));
  }
  JSSP_HIP(e, launch_step_jump(d_desc, s.K, s.L, e->stream));
  JSSP_HIP(e, launch_step_backtrack(d_desc, nullptr, 1, e->stream));
  JSSP_HIP(e, hipEventRecord(e->ev_stop, e->stream));
  return JSSP_OK;
}

namespace {

void add_algorithmic_bytes(jssp_engine* e, const Shape& s, const std::vector<int>& pc,
                           const std::vector<uint32_t>& lexp) {
  // Algorithmic bytes per launch: a score pass reads the parent snapshots
  // (4S each) and writes, per child, its 4-byte key word and its 8-byte
  // record (and, single-phase, its 4S-byte snapshot); the filter reads every
  // key word once and writes the survivors' 8-byte (key word, child) keys;
  // materialisation reads each survivor's key, its parent and its record and
  // writes its snapshot.  C counts the children evaluated (every valid slot;
  // the pruned ones get a key word and no record, which the count includes).
  const double S4 = 4.0 * s.S;
  for (auto& v : e->kl_bytes) v.clear();
  auto add = [&](int kind, double b) {
    e->kt[kind].bytes += b;
    e->kl_bytes[kind].push_back(b);
  };
  // Fused: the rank kernel of level l also scores level l + 1 from the
  // survivors' words in registers (12 bytes per child, no parent read).
  for (int l = 0; l < s.L; ++l) {
    const double P = pc[l], C = lexp[l];
    const double Kp = (double)pc[l + 1];
    if (s.single) add(kKindScore, S4 * P + (S4 + 12.0) * C);
    else if (!s.fused) add(kKindScore, S4 * P + 12.0 * C);
    if (l == s.L - 1) break;
    add(kKindFilter, 4.0 * C + 8.0 * Kp);
    add(kKindRank, 2.0 * S4 * Kp + 8.0 * Kp + (s.fused ? 12.0 * lexp[l + 1] : 0.0));
  }
}

// PMS_GPU / ICH_DET, one beam: a batch of one on the step kernels.
int solve_device(jssp_engine* e, const jssp_instance* in, const jssp_params* p, uint32_t seed, int bound,
                 jssp_result* res, int32_t* trail_out, int32_t* pcount_out) {
  jssp_params q = *p;
  q.seed = seed;
  q.best_known = bound;
  const Shape s = make_shape(in, &q, 0);
  Work w;
  Bump tail;
  const size_t need = carve(s, nullptr, w);
  int rc = ensure_device_arena(e, need + sizeof(InstDesc) + 1024);
  if (rc) return rc;
  carve(s, e->dbuf, w);
  tail.base = e->dbuf + need;
  InstDesc* d_desc = tail.take<InstDesc>(1);
  rc = upload_ops(e, in, w.ops);
  if (rc) return rc;
  InstDesc desc = make_desc(s, w, &q, in);
  desc.base.jump = w.jump;  // this solve runs k_jump_b before its backtrack
  if (e->profiling) desc.base.sel_log = w.sel_log;
  // The descriptor goes up through the pinned staging buffer; the step
  // kernels run with no block map (one instance, local block = grid block).
  rc = ensure_host_staging(e, std::max(sizeof(InstDesc), (size_t)64 + (size_t)s.L * 4 + 64));
  if (rc) return rc;
  std::memcpy(e->hbuf, &desc, sizeof(InstDesc));
  JSSP_HIP(e, hipMemcpyAsync(d_desc, e->hbuf, sizeof(InstDesc), hipMemcpyHostToDevice, e->stream));
  JSSP_HIP(e, hipStreamSynchronize(e->stream));
  Prof prof(e);
  rc = enqueue_single_solve(e, s, &q, d_desc, &prof);
  if (rc) return rc;
  int* h_res = reinterpret_cast<int*>(e->hbuf);
  unsigned long long* h_exp = reinterpret_cast<unsigned long long*>(e->hbuf + 32);
  int* h_seq = reinterpret_cast<int*>(e->hbuf + 64);
  JSSP_HIP(e, hipMemcpyAsync(h_res, w.result, 3 * sizeof(int), hipMemcpyDeviceToHost, e->stream));
  JSSP_HIP(e, hipMemcpyAsync(h_exp, w.expanded, sizeof(unsigned long long), hipMemcpyDeviceToHost, e->stream));
  JSSP_HIP(e, hipMemcpyAsync(h_seq, w.seq, (size_t)s.L * sizeof(int), hipMemcpyDeviceToHost, e->stream));
  JSSP_HIP(e, hipStreamSynchronize(e->stream));
  float msec = 0.f;
  JSSP_HIP(e, hipEventElapsedTime(&msec, e->ev_start, e->ev_stop));
  res->makespan = h_res[0];
  res->found = h_res[1];
  res->expanded_states = (int64_t)*h_exp;
  res->levels = s.L;
  res->device_ms = msec;
  if (res->found && res->sequence) std::memcpy(res->sequence, h_seq, (size_t)s.L * sizeof(int));
  if (e->profiling) {
    prof.collect();
    std::vector<int> pc(s.L + 1);
    std::vector<uint32_t> le(s.L);
    JSSP_HIP(e, hipMemcpy(pc.data(), w.pcount, (s.L + 1) * sizeof(int), hipMemcpyDeviceToHost));
    JSSP_HIP(e, hipMemcpy(le.data(), w.lvl_exp, s.L * sizeof(uint32_t), hipMemcpyDeviceToHost));
    add_algorithmic_bytes(e, s, pc, le);
    e->level_log.resize((size_t)s.L);
    JSSP_HIP(e, hipMemcpy(e->level_log.data(), w.sel_log, (size_t)s.L * sizeof(SelState), hipMemcpyDeviceToHost));
  }
  if (trail_out && s.L > 1)
    JSSP_HIP(e, hipMemcpy(trail_out, w.trail, (size_t)(s.L - 1) * s.K * sizeof(int), hipMemcpyDeviceToHost));
  if (pcount_out) JSSP_HIP(e, hipMemcpy(pcount_out, w.pcount, (s.L + 1) * sizeof(int), hipMemcpyDeviceToHost));
  return JSSP_OK;
}

// ICH_EXACT / PMS_EXACT: the device expands every child of every parent
// (cmax, job, machine per erase-order ready position); the host consumes the
// rand() stream (process-wide, or private and seeded) in the reference's
// order and replays the libstdc++ bounded heap (cpu/common/jssp_types.h:47-58) on order-isomorphic
// keys; the device then materialises the heap's array in order, which is the
// parent order of the next level (cpu/v0_ich/solver.cpp:22,42).
struct HeapEntry {
  long long ikey;   // ICH: 100*cmax + rand()%100 (orders exactly like the double score)
  double dkey;      // PMS: cmax/ns + (rand()%1000)/1e6, cpu/v1_pms/solver.cpp:166-167
  int x;
};
struct IchLess {
  bool operator()(const HeapEntry& a, const HeapEntry& b) const { return a.ikey < b.ikey; }
};
struct PmsLess {
  bool operator()(const HeapEntry& a, const HeapEntry& b) const { return a.dkey < b.dkey; }
};

template <class Less>
void heap_insert(std::vector<HeapEntry>& beam, const HeapEntry& h, int k) {
  if ((int)beam.size() < k) {
    beam.push_back(h);
    std::push_heap(beam.begin(), beam.end(), Less{});
  } else if (Less{}(h, beam.front())) {
    std::pop_heap(beam.begin(), beam.end(), Less{});
    beam.back() = h;
    std::push_heap(beam.begin(), beam.end(), Less{});
  }
}

void set_exact_level(LevelArgs& a, const Work& w, int level) {
  a.level = level;
  a.beam_cur = w.beam[level & 1];
  a.beam_next = w.beam[(level + 1) & 1];
}

}  // namespace

int solve_exact(jssp_engine* e, const jssp_instance* in, const jssp_params* p, RandStream& rs,
                jssp_result* res, int32_t* trail_out, int32_t* pcount_out) {
  const bool ich = (p->mode == JSSP_MODE_ICH_EXACT);
  const Shape s = make_shape(in, p, 64);
  Work w;
  const size_t need = carve(s, nullptr, w);
  int rc = ensure_device_arena(e, need);
  if (rc) return rc;
  carve(s, e->dbuf, w);
  rc = upload_ops(e, in, w.ops);
  if (rc) return rc;
  // The level's children (count, cmax, job, machine per slot) and the heap
  // order live in coherent pinned host memory that the exact kernels write
  // and read directly: no copy command per level, one wait for the expand.
  const size_t N = (size_t)s.K * s.B;
  rc = ensure_exact_host(e, N * (sizeof(int) + 2) + (size_t)s.K * 2 * sizeof(int) + 1024);
  if (rc) return rc;
  int* h_cmax = reinterpret_cast<int*>(e->xbuf);
  uint8_t* h_job = reinterpret_cast<uint8_t*>(h_cmax + N);
  uint8_t* h_mc = h_job + N;
  int* h_r = reinterpret_cast<int*>(((uintptr_t)(h_mc + N) + 15) & ~(uintptr_t)15);
  int* h_order = h_r + s.K;
  char* xdev = nullptr;
  JSSP_HIP(e, hipHostGetDevicePointer(reinterpret_cast<void**>(&xdev), e->xbuf, 0));
  auto dev = [&](const void* hp) { return xdev + (reinterpret_cast<const char*>(hp) - e->xbuf); };
  w.ex_cmax = reinterpret_cast<int*>(dev(h_cmax));
  w.ex_job = reinterpret_cast<uint8_t*>(dev(h_job));
  w.ex_mc = reinterpret_cast<uint8_t*>(dev(h_mc));
  w.ex_rcount = reinterpret_cast<int*>(dev(h_r));
  w.ex_order = reinterpret_cast<int*>(dev(h_order));

  LevelArgs a = base_args(s, w, p);
  JSSP_HIP(e, hipEventRecord(e->ev_start, e->stream));
  JSSP_HIP(e, launch_init(a, e->stream));
  std::vector<HeapEntry> heap;
  heap.reserve(s.K);
  int P = 1;
  long long expanded = 0;
  int best = JSSP_NO_MAKESPAN, best_p = -1, best_j = -1;
  int levels = 0;
  for (int l = 0; l < s.L && P > 0; ++l) {
    set_exact_level(a, w, l);
    JSSP_HIP(e, launch_exact_expand(a, P, e->stream));
    JSSP_HIP(e, hipStreamSynchronize(e->stream));
    ++levels;
    heap.clear();
    const int ns = l + 1;
    for (int pp = 0; pp < P; ++pp) {
      const int r = h_r[pp];
      const size_t base = (size_t)pp * s.B;
      if (ich) {
        for (int qq = 0; qq < r; ++qq) {
          HeapEntry h;
          h.ikey = 100ll * h_cmax[base + qq] + (rs.next() % 100);  // cpu/v0_ich/solver.cpp:35
          h.dkey = 0.0;
          h.x = (int)(base + qq);
          ++expanded;
          heap_insert<IchLess>(heap, h, s.K);
        }
      } else {
        if (r <= 0) continue;
        const int pivot = rs.next() % r;                       // cpu/v1_pms/solver.cpp:32
        const int tmc = h_mc[base + pivot];
        for (int qq = 0; qq < r; ++qq) {
          if (h_mc[base + qq] != tmc) continue;
          const int cm = h_cmax[base + qq];
          const double noise = (double)(rs.next() % 1000) / 1000000.0;
          HeapEntry h;
          h.ikey = 0;
          h.dkey = ((double)cm / (double)ns) + noise;
          h.x = (int)(base + qq);
          ++expanded;
          if (cm < p->best_known) heap_insert<PmsLess>(heap, h, s.K);
          if (ns == s.L && cm < best) { best = cm; best_p = pp; best_j = h_job[base + qq]; }
        }
      }
    }
    if (l == s.L - 1) {
      if (ich) {
        for (const HeapEntry& h : heap) {
          const int cm = h_cmax[h.x];
          if (cm < best) { best = cm; best_p = h.x / s.B; best_j = h_job[h.x]; }
        }
      }
      break;
    }
    const int count = (int)heap.size();
    for (int i = 0; i < count; ++i) h_order[i] = heap[i].x;
    if (count > 0) JSSP_HIP(e, launch_exact_materialize(a, count, e->stream));
    P = count;
  }
  const bool found = best < JSSP_NO_MAKESPAN && best_p >= 0;
  if (found) {
    a.bt_parent = best_p;
    a.bt_job = best_j;
    JSSP_HIP(e, launch_backtrack(a, e->stream));
  }
  JSSP_HIP(e, hipEventRecord(e->ev_stop, e->stream));
  if (found && res->sequence)
    JSSP_HIP(e, hipMemcpyAsync(res->sequence, w.seq, (size_t)s.L * sizeof(int), hipMemcpyDeviceToHost, e->stream));
  JSSP_HIP(e, hipStreamSynchronize(e->stream));
  float ms = 0.f;
  JSSP_HIP(e, hipEventElapsedTime(&ms, e->ev_start, e->ev_stop));
  res->makespan = best;
  res->found = found ? 1 : 0;
  res->expanded_states = expanded;
  res->levels = levels;
  res->device_ms = ms;
  if (trail_out && s.L > 1)
    JSSP_HIP(e, hipMemcpy(trail_out, w.trail, (size_t)(s.L - 1) * s.K * sizeof(int), hipMemcpyDeviceToHost));
  if (pcount_out) JSSP_HIP(e, hipMemcpy(pcount_out, w.pcount, (s.L + 1) * sizeof(int), hipMemcpyDeviceToHost));
  return JSSP_OK;
}


int solve_any(jssp_engine* e, const jssp_instance* in, const jssp_params* p, jssp_result* res,
              int32_t* trail_out, int32_t* pcount_out) {
  int rc = check_instance(e, in);
  if (rc) return rc;
  rc = check_params(e, in, p);
  if (rc) return rc;
  if (!res) return set_error(e, JSSP_ERR_INVALID, "jssp_solve", "null result");
  JSSP_HIP(e, hipSetDevice(e->device));
  res->best_seed_index = 0;
  if (p->mode == JSSP_MODE_ICH_EXACT || p->mode == JSSP_MODE_PMS_EXACT) {
    RandStream rs((p->flags & JSSP_FLAG_GLOBAL_RAND) != 0, p->seed);
    return solve_exact(e, in, p, rs, res, trail_out, pcount_out);
  }
  if (std::max(1, p->n_seeds) > 1 || std::max(1, p->n_passes) > 1)
    return solve_multiseed(e, in, p, res, nullptr, nullptr);
  return solve_device(e, in, p, p->seed, p->best_known, res, trail_out, pcount_out);
}

}  // namespace jssp

extern "C" {

int jssp_engine_create(int device, jssp_engine** out) {
  if (!out) return JSSP_ERR_INVALID;
  *out = nullptr;
  // The first use of the device loads the code objects, and the loader
  // draws from and reseeds the process-wide rand(): keep the caller's stream
  // where it was (the exact modes continue it, as the reference's CPU
  // solvers do).
  jssp::ProcessRandGuard keep_rand;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    (void)hipGetLastError();
    return JSSP_ERR_HIP;
  }
  if (device < 0 || device >= count) return JSSP_ERR_INVALID;
  jssp_engine* e = new jssp_engine();
  e->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e->ev_start) != hipSuccess || hipEventCreate(&e->ev_stop) != hipSuccess) {
    (void)hipGetLastError();
    delete e;
    return JSSP_ERR_HIP;
  }
  // Device address of the no-allocation entry point's scratch array, taken
  // once per device with the device current (a failure only disables that
  // entry point, which then reports it).
  (void)jssp::noalloc_setup(device);
  *out = e;
  return JSSP_OK;
}

void jssp_engine_destroy(jssp_engine* e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  jssp::release_batch_state(e);
  for (hipEvent_t ev : e->pev) (void)hipEventDestroy(ev);
  if (e->ev_start) (void)hipEventDestroy(e->ev_start);
  if (e->ev_stop) (void)hipEventDestroy(e->ev_stop);
  if (e->dbuf) (void)hipFree(e->dbuf);
  if (e->hbuf) (void)hipHostFree(e->hbuf);
  if (e->xbuf) (void)hipHostFree(e->xbuf);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

const char* jssp_last_error(const jssp_engine* e) { return e ? e->err.c_str() : "null engine"; }

int jssp_solve(jssp_engine* e, const jssp_instance* in, const jssp_params* p, jssp_result* res) {
  if (!e) return JSSP_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  return jssp::solve_any(e, in, p, res, nullptr, nullptr);
}

int jssp_solve_trace(jssp_engine* e, const jssp_instance* in, const jssp_params* p, jssp_result* res,
                     int32_t* trail, int32_t* pcount) {
  if (!e) return JSSP_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  if (p && (p->n_seeds > 1 || p->n_passes > 1))
    return jssp::set_error(e, JSSP_ERR_INVALID, "jssp_solve_trace", "trace needs a single beam");
  return jssp::solve_any(e, in, p, res, trail, pcount);
}

int jssp_engine_set_profiling(jssp_engine* e, int32_t on) {
  if (!e) return JSSP_ERR_INVALID;
  e->profiling = on != 0;
  for (auto& k : e->kt) k = KernelTime();
  return JSSP_OK;
}

int jssp_engine_level_stats(const jssp_engine* e, int32_t level, uint32_t* out, int32_t count) {
  if (!e || !out || level < 0 || (size_t)level >= e->level_log.size()) return JSSP_ERR_INVALID;
  const jssp::SelState& st = e->level_log[(size_t)level];
  const uint32_t v[10] = {st.V, st.Kp, st.lo, st.s1, st.T1, st.cum_lt, st.t1, st.need, st.sure_ctr, st.tie_ctr};
  for (int i = 0; i < count && i < 10; ++i) out[i] = v[i];
  return JSSP_OK;
}

int jssp_engine_kernel_launches(const jssp_engine* e, int32_t kind, float* ms, double* bytes, int32_t capacity) {
  if (!e || kind < 0 || kind >= kNumKinds || capacity < 0) return -1;
  const auto& t = e->kl_ms[kind];
  const auto& b = e->kl_bytes[kind];
  const size_t n = std::min(t.size(), b.size());
  for (size_t i = 0; i < n && i < (size_t)capacity; ++i) {
    if (ms) ms[i] = t[i];
    if (bytes) bytes[i] = b[i];
  }
  return (int)n;
}

int jssp_engine_kernel_time(const jssp_engine* e, int32_t kind, double* total_ms, int64_t* launches,
                            double* bytes) {
  if (!e || kind < 0 || kind >= kNumKinds) return JSSP_ERR_INVALID;
  if (total_ms) *total_ms = e->kt[kind].ms;
  if (launches) *launches = e->kt[kind].launches;
  if (bytes) *bytes = e->kt[kind].bytes;
  return JSSP_OK;
}

}  // extern "C"
