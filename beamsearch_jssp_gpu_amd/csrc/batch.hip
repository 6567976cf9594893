// batch.hip -- many solves advanced together, one level per step.
//
// Every instance of a batch has its own workspace and descriptor; a step is
// one launch per kernel kind (score when some instance scores separately,
// filter, and the rank kernel of each variant present) whose grid is the
// concatenation of its instances' blocks (block map entry = instance, local
// block).  The level lives on the device, so a step is the same launch
// sequence whatever the instances' levels.  An instance of L levels is done
// after L steps, which the host knows in advance: the batch runs in segments
// that end where some instance ends, and each segment's maps hold only the
// instances still running, so no launch dispatches blocks of finished
// instances.  A segment's steps are captured in a hipGraph (cached by launch
// layout) and replayed.  The reference solves one
// instance per call and one call per MPI rank (gpu/v1/main.cu:60-71); the
// multi-seed driver below runs its seeds as one batch instead, sharing the
// best bound through a device atomic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <vector>

#include "engine.hpp"
#include "level_kernels.hpp"
#include "workspace.hpp"

namespace jssp {

void release_step_graphs(jssp_engine* e) {
  for (auto& g : e->step_graphs) (void)hipGraphExecDestroy(g.second);
  e->step_graphs.clear();
}

void release_batch_state(jssp_engine* e) {
  release_step_graphs(e);
  if (e->bbuf) (void)hipFree(e->bbuf);
  e->bbuf = nullptr;
  e->bcap = 0;
}

namespace {

constexpr int kStepsPerGraph = 16;
constexpr size_t kGraphCacheMax = 64;

// Steps [first, first + steps) of a batch and the blocks of the instances
// still running during them, per kernel kind; rank blocks per kernel variant
// and rank block shape (index v + kRankVariants * shape).
constexpr int kRankGroups = kRankVariants * kRankShapes;
struct Segment {
  int first = 0, steps = 0;
  int nms = 0, nmf = 0;
  BlockRef *d_ms = nullptr, *d_mf = nullptr;
  int nmr[kRankGroups] = {};
  BlockRef* d_mr[kRankGroups] = {};
  size_t flds = 0;
  size_t rlds[kRankShapes] = {};
};

struct BatchPlan {
  int count = 0;
  int Lmax = 0;
  std::vector<Shape> shapes;
  std::vector<Work> works;
  std::vector<InstDesc> descs;
  InstDesc* d_desc = nullptr;
  BlockRef* d_inst = nullptr;
  std::vector<Segment> segs;
  int* d_res = nullptr;                 // count x 4
  unsigned long long* d_exp = nullptr;  // count
  int* d_seq = nullptr;                 // sum of L
  int* d_bound = nullptr;               // shared bound (multi-seed)
  std::vector<size_t> seq_off;
};

int grow_batch_arena(jssp_engine* e, size_t bytes) {
  if (bytes <= e->bcap) return JSSP_OK;
  JSSP_HIP(e, hipStreamSynchronize(e->stream));
  release_step_graphs(e);  // they name buffers of the old arena
  if (e->bbuf) JSSP_HIP(e, hipFree(e->bbuf));
  e->bbuf = nullptr;
  e->bcap = 0;
  const size_t cap = bytes + bytes / 4;
  if (hipMalloc(&e->bbuf, cap) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(e, JSSP_ERR_NOMEM, "hipMalloc", "batch arena of " + std::to_string(cap) + " bytes");
  }
  e->bcap = cap;
  return JSSP_OK;
}

// Lays out and uploads a batch: packed op tables, per-instance workspaces,
// descriptors, block maps and a shared result region.  codes[i] (or null:
// all 0) is instance i's tuned shape code (lanes, rank shape, filter cap).
int plan_batch(jssp_engine* e, int count, const jssp_instance* insts, const jssp_params* params,
               const int* codes, bool shared_bound, BatchPlan& bp) {
  bp = BatchPlan();
  bp.count = count;
  std::vector<jssp_params> ps(params, params + count);
  size_t ops_total = 0, seq_total = 0, work_total = 0;
  std::vector<size_t> work_need(count);
  std::vector<BlockRef> mi;
  std::vector<int> ends;
  for (int i = 0; i < count; ++i) {
    ps[i].phases = 2;  // batches materialise survivors only; beams are identical either way
    Shape s = make_batch_shape(&insts[i], &ps[i], codes ? codes[i] : 0);
    bp.shapes.push_back(s);
    Work w;
    work_need[i] = carve(s, nullptr, w);
    work_total += work_need[i];
    ops_total += (size_t)s.n * s.m;
    bp.seq_off.push_back(seq_total);
    seq_total += (size_t)s.L;
    bp.Lmax = std::max(bp.Lmax, s.L);
    mi.push_back(make_int2(i, 0));
    ends.push_back(s.L);
  }
  // Segments end where some instance ends; all maps go in one array, in
  // segment order (score, filter, then the rank blocks of each variant).
  std::sort(ends.begin(), ends.end());
  ends.erase(std::unique(ends.begin(), ends.end()), ends.end());
  std::vector<BlockRef> maps;
  std::vector<std::vector<size_t>> offs;  // per segment: score, filter, rank per variant
  for (size_t j = 0; j < ends.size(); ++j) {
    Segment sg;
    sg.first = j ? ends[j - 1] : 0;
    sg.steps = ends[j] - sg.first;
    std::vector<BlockRef> ms, mf, mr[kRankGroups];
    for (int i = 0; i < count; ++i) {
      const Shape& s = bp.shapes[i];
      if (s.L <= sg.first) continue;  // finished before this segment
      const int rs = rank_shape_of(s);
      append_maps(s, i, ms, mf, mr[rank_variant(s, &ps[i]) + kRankVariants * rs]);
      sg.flds = std::max(sg.flds, filter_lds_bytes(s, &ps[i]));
      sg.rlds[rs] = std::max(sg.rlds[rs], rank_lds_bytes(s));
    }
    std::vector<size_t> o;
    auto add = [&](const std::vector<BlockRef>& v) {
      o.push_back(maps.size());
      maps.insert(maps.end(), v.begin(), v.end());
    };
    add(ms);
    add(mf);
    for (int v = 0; v < kRankGroups; ++v) add(mr[v]);
    sg.nms = (int)ms.size();
    sg.nmf = (int)mf.size();
    for (int v = 0; v < kRankGroups; ++v) sg.nmr[v] = (int)mr[v].size();
    offs.push_back(o);
    bp.segs.push_back(sg);
  }
  Bump sz;
  sz.take<int2>(ops_total);
  sz.take<InstDesc>(count);
  sz.take<BlockRef>(maps.size());
  sz.take<BlockRef>(mi.size());
  sz.take<int>((size_t)count * 4);
  sz.take<unsigned long long>(count);
  sz.take<int>(seq_total);
  sz.take<int>(4);
  const size_t head = sz.off + 256;
  int rc = grow_batch_arena(e, head + work_total + 256 * (size_t)count);
  if (rc) return rc;
  Bump b;
  b.base = e->bbuf;
  int2* d_ops = b.take<int2>(ops_total);
  bp.d_desc = b.take<InstDesc>(count);
  BlockRef* d_maps = b.take<BlockRef>(maps.size());
  bp.d_inst = b.take<BlockRef>(mi.size());
  for (size_t j = 0; j < bp.segs.size(); ++j) {
    Segment& sg = bp.segs[j];
    sg.d_ms = d_maps + offs[j][0];
    sg.d_mf = d_maps + offs[j][1];
    for (int v = 0; v < kRankGroups; ++v) sg.d_mr[v] = d_maps + offs[j][2 + v];
  }
  bp.d_res = b.take<int>((size_t)count * 4);
  bp.d_exp = b.take<unsigned long long>(count);
  bp.d_seq = b.take<int>(seq_total);
  bp.d_bound = b.take<int>(4);
  size_t off = head;
  size_t ops_off = 0;
  for (int i = 0; i < count; ++i) {
    Work w;
    carve(bp.shapes[i], e->bbuf + off, w);
    off += work_need[i];
    off = (off + 255) & ~(size_t)255;
    w.ops = d_ops + ops_off;
    ops_off += (size_t)bp.shapes[i].n * bp.shapes[i].m;
    w.result = bp.d_res + 4 * i;
    w.expanded = bp.d_exp + i;
    w.seq = bp.d_seq + bp.seq_off[i];
    bp.works.push_back(w);
    InstDesc d = make_desc(bp.shapes[i], w, &ps[i], &insts[i]);
    d.shared_bound = shared_bound ? bp.d_bound : nullptr;
    bp.descs.push_back(d);
  }
  // One staging pass and one copy per region.
  const size_t hbytes = ops_total * sizeof(int2) + (size_t)count * sizeof(InstDesc) +
                        (maps.size() + mi.size()) * sizeof(BlockRef) + 64;
  rc = ensure_host_staging(e, hbytes);
  if (rc) return rc;
  char* h = e->hbuf;
  size_t hoff = 0;
  auto put = [&](void* dst, const void* src, size_t bytes) -> int {
    if (!bytes) return JSSP_OK;
    std::memcpy(h + hoff, src, bytes);
    JSSP_HIP(e, hipMemcpyAsync(dst, h + hoff, bytes, hipMemcpyHostToDevice, e->stream));
    hoff += bytes;
    return JSSP_OK;
  };
  {
    int2* hops = reinterpret_cast<int2*>(h);
    size_t k = 0;
    for (int i = 0; i < count; ++i)
      for (int q = 0; q < insts[i].n_jobs * insts[i].n_machines; ++q, ++k) {
        hops[k].x = insts[i].machine[q];
        hops[k].y = insts[i].duration[q];
      }
    JSSP_HIP(e, hipMemcpyAsync(d_ops, hops, ops_total * sizeof(int2), hipMemcpyHostToDevice, e->stream));
    hoff = ops_total * sizeof(int2);
  }
  if ((rc = put(bp.d_desc, bp.descs.data(), (size_t)count * sizeof(InstDesc)))) return rc;
  if ((rc = put(d_maps, maps.data(), maps.size() * sizeof(BlockRef)))) return rc;
  if ((rc = put(bp.d_inst, mi.data(), mi.size() * sizeof(BlockRef)))) return rc;
  JSSP_HIP(e, hipStreamSynchronize(e->stream));
  return JSSP_OK;
}

int enqueue_step(jssp_engine* e, const BatchPlan& bp, const Segment& sg) {
  if (sg.nms > 0) JSSP_HIP(e, launch_step_score(bp.d_desc, sg.d_ms, sg.nms, false, e->stream));
  JSSP_HIP(e, launch_step_filter(bp.d_desc, sg.d_mf, sg.nmf, sg.flds, true, e->stream));
  for (int g = 0; g < kRankGroups; ++g) {
    if (sg.nmr[g] == 0) continue;
    const int v = g % kRankVariants, rs = g / kRankVariants;
    JSSP_HIP(e, launch_step_rank(bp.d_desc, sg.d_mr[g], sg.nmr[g], sg.rlds[rs], v, rank_shape_waves(rs), e->stream,
                                  nullptr, rank_shape_batch(rs)));
  }
  return JSSP_OK;
}

// n steps of segment sg: whole groups of kStepsPerGraph steps replay the
// segment's captured graph (captured once per launch layout), the rest are
// launched directly.
int enqueue_segment(jssp_engine* e, const BatchPlan& bp, const Segment& sg, int n, bool use_graph) {
  int done = 0;
  if (use_graph && n >= kStepsPerGraph) {
    std::vector<uint64_t> key = {(uint64_t)(uintptr_t)bp.d_desc, (uint64_t)sg.nms, (uint64_t)sg.nmf,
                                 (uint64_t)(uintptr_t)sg.d_ms, (uint64_t)(uintptr_t)sg.d_mf, (uint64_t)sg.flds};
    for (int rs = 0; rs < kRankShapes; ++rs) key.push_back((uint64_t)sg.rlds[rs]);
    for (int v = 0; v < kRankGroups; ++v) {
      key.push_back((uint64_t)sg.nmr[v]);
      key.push_back((uint64_t)(uintptr_t)sg.d_mr[v]);
    }
    hipGraphExec_t ge = nullptr;
    for (const auto& g : e->step_graphs)
      if (g.first == key) { ge = g.second; break; }
    if (!ge) {
      hipGraph_t g = nullptr;
      JSSP_HIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
      int rc = JSSP_OK;
      for (int t = 0; t < kStepsPerGraph && rc == JSSP_OK; ++t) rc = enqueue_step(e, bp, sg);
      hipError_t ce = hipStreamEndCapture(e->stream, &g);
      if (rc) return rc;
      if (ce != hipSuccess) return set_error(e, JSSP_ERR_HIP, "hipStreamEndCapture", hipGetErrorString(ce));
      hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) return set_error(e, JSSP_ERR_HIP, "hipGraphInstantiate", hipGetErrorString(ie));
      if (e->step_graphs.size() >= kGraphCacheMax) {
        // Full cache: let the enqueued launches finish, then drop it.
        JSSP_HIP(e, hipStreamSynchronize(e->stream));
        release_step_graphs(e);
      }
      e->step_graphs.push_back({key, ge});
    }
    for (; done + kStepsPerGraph <= n; done += kStepsPerGraph) JSSP_HIP(e, hipGraphLaunch(ge, e->stream));
  }
  for (; done < n; ++done) {
    int rc = enqueue_step(e, bp, sg);
    if (rc) return rc;
  }
  return JSSP_OK;
}

// Steps [from, from + n) of the batch (steps past every instance's last
// level are no-ops and are skipped).
int enqueue_steps(jssp_engine* e, const BatchPlan& bp, int from, int n, bool use_graph) {
  const int end = from + n;
  for (const Segment& sg : bp.segs) {
    const int lo = std::max(from, sg.first), hi = std::min(end, sg.first + sg.steps);
    if (lo >= hi) continue;
    int rc = enqueue_segment(e, bp, sg, hi - lo, use_graph);
    if (rc) return rc;
  }
  return JSSP_OK;
}

int run_plan(jssp_engine* e, const BatchPlan& bp, bool use_graph) {
  JSSP_HIP(e, launch_step_init(bp.d_desc, bp.d_inst, bp.count, e->stream));
  int rc = enqueue_steps(e, bp, 0, bp.Lmax, use_graph);
  if (rc) return rc;
  JSSP_HIP(e, launch_step_backtrack(bp.d_desc, bp.d_inst, bp.count, e->stream));
  return JSSP_OK;
}

struct HostResults {
  std::vector<int> res;
  std::vector<unsigned long long> exp;
  std::vector<int> seq;
};

int fetch_results(jssp_engine* e, const BatchPlan& bp, HostResults& hr) {
  hr.res.resize((size_t)bp.count * 4);
  hr.exp.resize(bp.count);
  size_t seq_total = bp.seq_off.empty() ? 0 : bp.seq_off.back() + bp.shapes.back().L;
  hr.seq.resize(seq_total);
  JSSP_HIP(e, hipMemcpyAsync(hr.res.data(), bp.d_res, hr.res.size() * sizeof(int), hipMemcpyDeviceToHost, e->stream));
  JSSP_HIP(e, hipMemcpyAsync(hr.exp.data(), bp.d_exp, hr.exp.size() * 8, hipMemcpyDeviceToHost, e->stream));
  JSSP_HIP(e, hipMemcpyAsync(hr.seq.data(), bp.d_seq, seq_total * sizeof(int), hipMemcpyDeviceToHost, e->stream));
  JSSP_HIP(e, hipStreamSynchronize(e->stream));
  return JSSP_OK;
}

// ---- lanes-per-parent tuning ------------------------------------------------
uint64_t tune_key(int n, int m, int K) {
  int kb = 0;
  while ((1 << kb) < K) ++kb;  // K rounded up to a power of two
  return ((uint64_t)n << 40) | ((uint64_t)m << 32) | (uint64_t)kb;
}

int lookup_tune(const jssp_engine* e, uint64_t key) {
  for (const auto& kv : e->tune)
    if (kv.first == key) return kv.second;
  return 0;
}

void load_tune_file(jssp_engine* e) {
  if (e->tune_path.empty()) return;
  std::ifstream f(e->tune_path);
  if (!f.good()) return;
  // Format: one "n m kbits lanes [rank_shape [filter_cap_index]]" record per line.
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream in(line);
    int n, m, kb, g, rs = kRankShapeBatched, fc = 0;
    if (!(in >> n >> m >> kb >> g)) continue;
    in >> rs >> fc;
    const uint64_t key = ((uint64_t)n << 40) | ((uint64_t)m << 32) | (uint64_t)kb;
    if (lookup_tune(e, key) || !(g == kLanesFused || g == 16 || g == 32 || g == 64)) continue;
    if (rs < 0 || rs >= kRankShapes || fc < 0 || fc > 3) continue;
    e->tune.push_back({key, make_code(g, rs, fc)});
  }
}

void save_tune_file(const jssp_engine* e) {
  if (e->tune_path.empty()) return;
  std::ofstream f(e->tune_path);
  for (const auto& kv : e->tune)
    f << (kv.first >> 40) << " " << ((kv.first >> 32) & 0xFF) << " " << (kv.first & 0xFFFFFFFF) << " "
      << code_lanes(kv.second) << " " << code_rank_shape(kv.second) << " " << ((kv.second >> 12) & 0x3) << "\n";
}

// Device time of `steps` steps of a batch of `sub` with every member on
// shape `code`, starting at step w0 (the steps before run untimed).
int time_shape(jssp_engine* e, const std::vector<jssp_instance>& sub, const std::vector<jssp_params>& subp,
               int code, int w0, int steps, double& ms_out) {
  std::vector<int> codes(sub.size(), code);
  BatchPlan bp;
  int rc = plan_batch(e, (int)sub.size(), sub.data(), subp.data(), codes.data(), false, bp);
  if (rc) return rc;
  JSSP_HIP(e, launch_step_init(bp.d_desc, bp.d_inst, bp.count, e->stream));
  rc = enqueue_steps(e, bp, 0, w0, false);
  if (rc) return rc;
  JSSP_HIP(e, hipEventRecord(e->ev_start, e->stream));
  rc = enqueue_steps(e, bp, w0, steps, false);
  if (rc) return rc;
  JSSP_HIP(e, hipEventRecord(e->ev_stop, e->stream));
  JSSP_HIP(e, hipStreamSynchronize(e->stream));
  float ms = 0.f;
  JSSP_HIP(e, hipEventElapsedTime(&ms, e->ev_start, e->ev_stop));
  ms_out = ms;
  return JSSP_OK;
}

// Children per level of the batch a class is timed on: the class's members
// repeated until the batch fills the device the way a large mixed batch does
// (a few members alone leave most CUs idle and would favour the shapes with
// the most blocks, which lose once the device is full).
constexpr long long kTuneChildren = 1ll << 22;
constexpr int kTuneMaxMembers = 1024;

// For each (n, m, K bucket) class of the batch not yet in the table, times a
// batch of its instances (repeated up to kTuneChildren children per level)
// over up to 16 full-width levels (after the untimed first levels, which
// grow the beam to full width), first with every valid lanes-per-parent
// shape (the fused form, or a separate score kernel with 16 / 32 / 64 lanes),
// then with the winner and each rank block shape (8-wave blocks ranking four
// survivors per wave, or one; 16-wave blocks), then each filter block cap,
// keeping the fastest at each stage.  Results do not depend on the shape.
int autotune_classes(jssp_engine* e, int count, const jssp_instance* insts, const jssp_params* params) {
  if (!e->autotune) return JSSP_OK;
  if (e->tune.empty()) load_tune_file(e);
  bool added = false;
  std::vector<uint64_t> seen;
  for (int i = 0; i < count; ++i) {
    const int n = insts[i].n_jobs, m = insts[i].n_machines, K = params[i].beam_width;
    const uint64_t key = tune_key(n, m, K);
    if (lookup_tune(e, key) || std::find(seen.begin(), seen.end(), key) != seen.end()) continue;
    seen.push_back(key);
    std::vector<jssp_instance> sub;
    std::vector<jssp_params> subp;
    for (int k = i; k < count && (int)sub.size() < 8; ++k)
      if (tune_key(insts[k].n_jobs, insts[k].n_machines, params[k].beam_width) == key) {
        sub.push_back(insts[k]);
        subp.push_back(params[k]);
      }
    long long children = 0;
    for (const auto& q : subp) children += (long long)q.beam_width * n;
    for (size_t k = 0, s0 = sub.size(); children < kTuneChildren && (int)sub.size() < kTuneMaxMembers; ++k) {
      sub.push_back(sub[k % s0]);
      subp.push_back(subp[k % s0]);
      children += (long long)subp.back().beam_width * n;
    }
    const int L = n * m;
    // The beam reaches full width after about log_n(K) levels.
    int grow = 1;
    for (long long w = n; w < K && grow < L; w *= std::max(2, n)) ++grow;
    const int steps = std::max(1, std::min(16, L / 2));
    const int w0 = std::max(0, std::min(grow + 2, L - 1 - steps));
    double t = 0.0, best_t = 1e30;
    int best = make_code(64, kRankShapeBatched, 0);
    const bool can_fuse = fused_eligible(&insts[i], &params[i]);
    for (int g : {kLanesFused, 16, 32, 64}) {
      if (g == kLanesFused && !can_fuse) continue;
      if (g != kLanesFused && g < 64 && g < n) continue;
      const int code = make_code(g, kRankShapeBatched, 0);
      int rc = time_shape(e, sub, subp, code, w0, steps, t);
      if (rc) return rc;
      if (t < best_t) { best_t = t; best = code; }
    }
    for (int rs : {kRankShapeThroughput, kRankShapeSingle}) {
      const int code = make_code(code_lanes(best), rs, 0);
      int rc = time_shape(e, sub, subp, code, w0, steps, t);
      if (rc) return rc;
      if (t < best_t) { best_t = t; best = code; }
    }
    const long long tiles = ((long long)K * n + kFilterThreads - 1) / kFilterThreads;
    for (int fc : {1, 3}) {  // 16 or 256 filter blocks (the default is 64)
      if (fc == 3 && tiles <= 64) continue;
      const int code = make_code(code_lanes(best), code_rank_shape(best), fc);
      int rc = time_shape(e, sub, subp, code, w0, steps, t);
      if (rc) return rc;
      if (t < best_t) { best_t = t; best = code; }
    }
    e->tune.push_back({key, best});
    added = true;
  }
  if (added) save_tune_file(e);
  return JSSP_OK;
}

int check_batch(jssp_engine* e, int count, const jssp_instance* insts, const jssp_params* params) {
  for (int i = 0; i < count; ++i) {
    int rc = check_instance(e, &insts[i]);
    if (rc) return rc;
    rc = check_params(e, &insts[i], &params[i]);
    if (rc) return rc;
    if (params[i].mode != JSSP_MODE_ICH_DET && params[i].mode != JSSP_MODE_PMS_GPU)
      return set_error(e, JSSP_ERR_INVALID, "batch", "batched solves take modes ICH_DET and PMS_GPU");
    if (params[i].n_seeds > 1 || params[i].n_passes > 1)
      return set_error(e, JSSP_ERR_INVALID, "batch", "n_seeds / n_passes apply to jssp_solve");
  }
  return JSSP_OK;
}

}  // namespace

// Batch of independent solves; results[i] equals jssp_solve(insts[i], params[i]).
int solve_batch(jssp_engine* e, int count, const jssp_instance* insts, const jssp_params* params,
                jssp_result* results, bool use_graph) {
  int rc = check_batch(e, count, insts, params);
  if (rc) return rc;
  if (count == 0) return JSSP_OK;
  JSSP_HIP(e, hipSetDevice(e->device));
  rc = autotune_classes(e, count, insts, params);
  if (rc) return rc;
  std::vector<int> codes(count);
  for (int i = 0; i < count; ++i)
    codes[i] = lookup_tune(e, tune_key(insts[i].n_jobs, insts[i].n_machines, params[i].beam_width));
  BatchPlan bp;
  rc = plan_batch(e, count, insts, params, codes.data(), false, bp);
  if (rc) return rc;
  JSSP_HIP(e, hipEventRecord(e->ev_start, e->stream));
  rc = run_plan(e, bp, use_graph);
  if (rc) return rc;
  JSSP_HIP(e, hipEventRecord(e->ev_stop, e->stream));
  HostResults hr;
  rc = fetch_results(e, bp, hr);
  if (rc) return rc;
  float ms = 0.f;
  JSSP_HIP(e, hipEventElapsedTime(&ms, e->ev_start, e->ev_stop));
  for (int i = 0; i < count; ++i) {
    jssp_result& r = results[i];
    r.makespan = hr.res[4 * i];
    r.found = hr.res[4 * i + 1];
    r.expanded_states = (int64_t)hr.exp[i];
    r.levels = bp.shapes[i].L;
    r.best_seed_index = 0;
    r.device_ms = ms;
    if (r.found && r.sequence)
      std::memcpy(r.sequence, hr.seq.data() + bp.seq_off[i], (size_t)bp.shapes[i].L * sizeof(int));
  }
  return JSSP_OK;
}

// Multi-seed / multi-pass search (gpu/v1/main.cu:60-71, cpu/v1_pms/main.cpp:
// 56-73 as one process on one GPU): each pass runs n_seeds independent beams
// as one batch with seeds seed + pass * n_seeds + i; every beam of a pass
// prunes with the device bound as it stood when the pass started, and each
// beam that completes lowers it with atomicMin.  The result is the lowest
// makespan found (ties: earliest pass, lowest seed index) -- the same as
// running those solves one by one with the bound carried between passes.
int solve_multiseed(jssp_engine* e, const jssp_instance* in, const jssp_params* p, jssp_result* res,
                    const uint32_t* seeds, int32_t* pass_best) {
  const int S = std::max(1, p->n_seeds), E = std::max(1, p->n_passes);
  std::vector<jssp_instance> insts(S, *in);
  std::vector<jssp_params> ps(S, *p);
  for (auto& q : ps) { q.n_seeds = 1; q.n_passes = 1; }
  BatchPlan bp;
  int rc = plan_batch(e, S, insts.data(), ps.data(), nullptr, true, bp);
  if (rc) return rc;
  const int L = in->n_jobs * in->n_machines;
  int bound = p->best_known;
  JSSP_HIP(e, hipMemcpyAsync(bp.d_bound, &bound, sizeof(int), hipMemcpyHostToDevice, e->stream));
  std::vector<HostResults> passes(E);
  int64_t expanded = 0;
  JSSP_HIP(e, hipEventRecord(e->ev_start, e->stream));
  for (int ep = 0; ep < E; ++ep) {
    for (int i = 0; i < S; ++i)
      bp.descs[i].base.lseed = seeds ? seeds[(size_t)ep * S + i] : p->seed + (uint32_t)(ep * S + i);
    JSSP_HIP(e, hipStreamSynchronize(e->stream));
    std::memcpy(e->hbuf, bp.descs.data(), (size_t)S * sizeof(InstDesc));
    JSSP_HIP(e, hipMemcpyAsync(bp.d_desc, e->hbuf, (size_t)S * sizeof(InstDesc), hipMemcpyHostToDevice,
                               e->stream));
    rc = run_plan(e, bp, true);
    if (rc) return rc;
    rc = fetch_results(e, bp, passes[ep]);
    if (rc) return rc;
  }
  JSSP_HIP(e, hipEventRecord(e->ev_stop, e->stream));
  JSSP_HIP(e, hipStreamSynchronize(e->stream));
  float ms = 0.f;
  JSSP_HIP(e, hipEventElapsedTime(&ms, e->ev_start, e->ev_stop));
  int best = p->best_known, best_idx = -1;
  const int* best_seq = nullptr;
  for (int ep = 0; ep < E; ++ep) {
    for (int i = 0; i < S; ++i) {
      const HostResults& hr = passes[ep];
      expanded += (int64_t)hr.exp[i];
      if (hr.res[4 * i + 1] && hr.res[4 * i + 2] < best) {
        best = hr.res[4 * i + 2];
        best_idx = ep * S + i;
        best_seq = hr.seq.data() + bp.seq_off[i];
      }
    }
    if (pass_best) pass_best[ep] = best;
  }
  res->makespan = best;
  res->found = best_idx >= 0 ? 1 : 0;
  res->best_seed_index = std::max(0, best_idx);
  res->expanded_states = expanded;
  res->levels = L;
  res->device_ms = ms;
  if (res->found && res->sequence) std::memcpy(res->sequence, best_seq, (size_t)L * sizeof(int));
  return JSSP_OK;
}

}  // namespace jssp

extern "C" {

int jssp_solve_batch(jssp_engine* e, int32_t count, const jssp_instance* insts, const jssp_params* params,
                     jssp_result* results) {
  if (!e || count < 0 || (count > 0 && (!insts || !params || !results))) return JSSP_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  return jssp::solve_batch(e, count, insts, params, results, true);
}

int jssp_solve_batch_ex(jssp_engine* e, int32_t count, const jssp_instance* insts, const jssp_params* params,
                        jssp_result* results, int32_t use_graph) {
  if (!e || count < 0 || (count > 0 && (!insts || !params || !results))) return JSSP_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  return jssp::solve_batch(e, count, insts, params, results, use_graph != 0);
}

int jssp_solve_seeded(jssp_engine* e, const jssp_instance* in, const jssp_params* p, const uint32_t* seeds,
                      jssp_result* res) {
  if (!e || !in || !p || !seeds || !res) return JSSP_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  int rc = jssp::check_instance(e, in);
  if (rc) return rc;
  rc = jssp::check_params(e, in, p);
  if (rc) return rc;
  if (p->mode != JSSP_MODE_ICH_DET && p->mode != JSSP_MODE_PMS_GPU)
    return jssp::set_error(e, JSSP_ERR_INVALID, "jssp_solve_seeded", "seed lists apply to modes ICH_DET and PMS_GPU");
  JSSP_HIP(e, hipSetDevice(e->device));
  res->best_seed_index = 0;
  return jssp::solve_multiseed(e, in, p, res, seeds, nullptr);
}

int jssp_engine_set_autotune(jssp_engine* e, int32_t on, const char* table_path) {
  if (!e) return JSSP_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  e->autotune = on != 0;
  e->tune_path = table_path ? table_path : "";
  e->tune.clear();
  return JSSP_OK;
}

int jssp_engine_tuned_lanes(const jssp_engine* e, int32_t n_jobs, int32_t n_machines, int32_t beam_width) {
  if (!e) return 0;
  return jssp::code_lanes(jssp::lookup_tune(e, jssp::tune_key(n_jobs, n_machines, beam_width)));
}

int jssp_engine_tuned_shape(const jssp_engine* e, int32_t n_jobs, int32_t n_machines, int32_t beam_width,
                            int32_t* lanes, int32_t* rank_waves, int32_t* rank_batch, int32_t* filter_blocks) {
  if (!e) return 0;
  const int code = jssp::lookup_tune(e, jssp::tune_key(n_jobs, n_machines, beam_width));
  if (!code) return 0;
  const int rs = jssp::code_rank_shape(code);
  if (lanes) *lanes = jssp::code_lanes(code);
  if (rank_waves) *rank_waves = jssp::rank_shape_waves(rs);
  if (rank_batch) *rank_batch = jssp::rank_shape_batch(rs);
  if (filter_blocks) *filter_blocks = jssp::code_filter_cap(code) ? jssp::code_filter_cap(code) : 64;
  return 1;
}

}  // extern "C"
