// serve.hip -- continuous-batching solve server.
//
// Requests are admitted into slots while a slot is free and the HBM budget
// has room for the request's workspace (first-fit over one device arena of
// the budget's size, FIFO order).  Every admitted request stays resident on
// the device until it finishes; each serve step advances all resident
// requests by a number of levels with the batched step kernels, so work that
// arrives while others are running overlaps with them on the device.  Results
// are those of jssp_solve on the same request: a request's beams never depend
// on what else is resident.  The reference has no serving layer; its drivers
// solve one instance per process (gpu/v1/main.cu, cpu/v0_ich/main.cpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <unordered_map>
#include <vector>

#include "engine.hpp"
#include "level_kernels.hpp"
#include "workspace.hpp"

namespace {

using jssp::BlockRef;
using jssp::InstDesc;
using jssp::Shape;
using jssp::Work;

enum ReqState { kQueued = 0, kRunning = 1, kDone = 2, kCancelled = 3 };

struct Request {
  std::vector<int32_t> mach, dur;
  jssp_instance inst;
  jssp_params params;
  double t_submit = 0, t_admit = -1, t_done = -1;
  int state = kQueued;
  int slot = -1;
  int makespan = 0, found = 0;
  int64_t expanded = 0;
  std::vector<int32_t> seq;
};

struct Region {
  size_t off, size;
};

}  // namespace

struct jssp_server {
  jssp_engine* e = nullptr;
  jssp_server_config cfg{};
  std::chrono::steady_clock::time_point t0;
  char* arena = nullptr;  // budget-sized resident-beam arena
  size_t arena_bytes = 0;
  std::vector<Region> free_list;  // sorted by offset, coalesced
  char* meta = nullptr;           // descriptors + block maps
  BlockRef* d_maps = nullptr;     // the three block maps, contiguous (d_ms, d_mf, d_mr point into it)
  InstDesc* d_desc = nullptr;
  BlockRef *d_ms = nullptr, *d_mf = nullptr, *d_mr = nullptr;
  // Uploads never wait for the device: the host writes each one into one of
  // two pinned staging buffers, alternately (a buffer is reused once the
  // event recorded after its last copy has passed, long before in practice),
  // and the copies go on the engine's stream behind the step in flight, which
  // they therefore cannot disturb.  Scattered destinations (op tables,
  // descriptors, descriptor fields) are written by one copy of a packed block
  // into d_up and a scatter kernel, instead of one small copy each.
  struct Stage {
    char* h = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
  };
  Stage stage[2];
  int stage_cur = 0;
  char* d_up = nullptr;  // device block of the current upload (grows)
  size_t up_cap = 0;
  // Pinned landing buffer of finished requests' results (grows): copies
  // into pageable memory would wait for the device's other queues.
  char* hv = nullptr;
  size_t hv_cap = 0;
  size_t cap_map = 0;
  int nslots = 0;
  std::vector<uint64_t> slot_ticket;  // 0 = free
  std::vector<Shape> slot_shape;
  std::vector<Work> slot_work;
  std::vector<Region> slot_region;
  std::vector<int> slot_level;  // levels enqueued since the slot's request was admitted
  std::vector<int> slot_rank_cap;  // rank-block partials the slot's workspace holds
  std::deque<uint64_t> queue;
  std::unordered_map<uint64_t, Request> reqs;
  uint64_t next_ticket = 1;
  bool maps_dirty = true;
  int nmr = 0, nact = 0;
  // Resident requests are split into kLanes groups of about equal work, each
  // advanced on its own stream (lane 0 is the engine's stream): requests are
  // independent, so one group's launch gaps and tails overlap the other's
  // kernels.  Group g's score and filter blocks are d_ms[ms_off[g], +
  // ms_cnt[g]) and d_mf[mf_off[g], + mf_cnt[g]); its rank blocks of group q
  // are d_mr[mr_off[g][q], + mr_cnt[g][q]), where the group q of a request of
  // the request's rank block shape rs (jssp::RankShape) and kernel variant v
  // is entry rs * kRankVariants + v.
  static constexpr int kLanes = 2;
  static constexpr int kRankGroups = jssp::kRankShapes * jssp::kRankVariants;
  hipStream_t lane_stream[kLanes] = {};
  hipEvent_t lane_ev[kLanes] = {};
  int ms_off[kLanes] = {}, ms_cnt[kLanes] = {}, mf_off[kLanes] = {}, mf_cnt[kLanes] = {};
  int mr_off[kLanes][kRankGroups] = {}, mr_cnt[kLanes][kRankGroups] = {};
  bool filter_tp = true;  // filter form of the current maps (see rebuild_maps)
  size_t flds = 0, rlds[jssp::kRankShapes] = {};
  std::vector<InstDesc> hdesc;
  // Pipelined steps: the backtrack of the requests that finish in a step
  // writes their finished flags into pinned host memory (status[parity], no
  // copy command) and the step ends with an event; the next serve step reads
  // that status after enqueuing its own kernels, so the device always has
  // the current step queued while the host harvests.  status_ticket records
  // which request held each slot when the step was enqueued (a slot freed and
  // re-admitted since then is skipped).
  int* status[2] = {nullptr, nullptr};
  int* status_dev[2] = {nullptr, nullptr};
  std::vector<uint64_t> status_ticket[2];
  // The requests whose last level runs in the step of each parity, and
  // their backtrack map (pinned host memory the backtrack kernel reads).
  std::vector<int> status_fin[2];
  BlockRef* fin_map[2] = {nullptr, nullptr};
  BlockRef* fin_map_dev[2] = {nullptr, nullptr};
  hipEvent_t status_ev[2] = {nullptr, nullptr};
  int status_pending = -1;  // parity of the step whose status is not read yet
  int parity = 0;
  hipStream_t copy_stream = nullptr;  // result reads of finished requests, beside the running step
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

// Filter blocks that fill the chip: one per CU.
constexpr long long kFillBlocks = 256;
// Rank block shape of the resident requests, from their rank blocks in the
// batched shape (8-wave blocks ranking kRankBatch survivors per wave): at
// kRankFillBatched or more blocks the device is throughput-bound and takes
// that shape; below it, every block counts toward the requests' latency and
// they take 8-wave blocks ranking one survivor per wave, or, at most
// kRankFillSingle blocks in the batched shape (a couple of requests), the
// single-solve shape (16-wave blocks).
constexpr long long kRankFillBatched = 512;
constexpr long long kRankFillSingle = 256;


bool arena_alloc(jssp_server* s, size_t bytes, Region& out) {
  bytes = (bytes + 255) & ~(size_t)255;
  for (size_t i = 0; i < s->free_list.size(); ++i) {
    Region& r = s->free_list[i];
    if (r.size >= bytes) {
      out = {r.off, bytes};
      r.off += bytes;
      r.size -= bytes;
      if (r.size == 0) s->free_list.erase(s->free_list.begin() + (long)i);
      return true;
    }
  }
  return false;
}

void arena_free(jssp_server* s, Region reg) {
  auto it = std::lower_bound(s->free_list.begin(), s->free_list.end(), reg,
                             [](const Region& a, const Region& b) { return a.off < b.off; });
  it = s->free_list.insert(it, reg);
  // coalesce with the next and previous blocks
  if (it + 1 != s->free_list.end() && it->off + it->size == (it + 1)->off) {
    it->size += (it + 1)->size;
    s->free_list.erase(it + 1);
  }
  if (it != s->free_list.begin() && (it - 1)->off + (it - 1)->size == it->off) {
    (it - 1)->size += it->size;
    s->free_list.erase(it);
  }
}

// Pinned staging memory of at least `bytes` that no copy in flight reads
// (the current buffer of the pair).
int stage_take(jssp_server* s, size_t bytes, char** out) {
  jssp_engine* e = s->e;
  jssp_server::Stage& st = s->stage[s->stage_cur];
  if (st.pending) {
    JSSP_HIP(e, hipEventSynchronize(st.ev));
    st.pending = false;
  }
  if (st.cap < bytes) {
    const size_t cap = std::max({bytes, 2 * st.cap, (size_t)64 << 10});
    if (st.h) JSSP_HIP(e, hipHostFree(st.h));
    st.h = nullptr;
    st.cap = 0;
    if (hipHostMalloc((void**)&st.h, cap, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      st.h = nullptr;
      return jssp::set_error(e, JSSP_ERR_NOMEM, "serve", "cannot allocate pinned staging memory");
    }
    st.cap = cap;
  }
  *out = st.h;
  return JSSP_OK;
}

// Marks the current staging buffer as read by the copies just enqueued on
// the engine's stream and switches to the other one.
int stage_done(jssp_server* s) {
  jssp_server::Stage& st = s->stage[s->stage_cur];
  JSSP_HIP(s->e, hipEventRecord(st.ev, s->e->stream));
  st.pending = true;
  s->stage_cur ^= 1;
  return JSSP_OK;
}

// Device upload block of at least `bytes`.  Growing it waits for the
// enqueued work that may still read the old one.
int ensure_up(jssp_server* s, size_t bytes) {
  jssp_engine* e = s->e;
  if (bytes <= s->up_cap) return JSSP_OK;
  JSSP_HIP(e, hipStreamSynchronize(e->stream));
  if (s->d_up) JSSP_HIP(e, hipFree(s->d_up));
  s->d_up = nullptr;
  const size_t cap = std::max({bytes, 2 * s->up_cap, (size_t)256 << 10});
  s->up_cap = 0;
  if (hipMalloc(&s->d_up, cap) != hipSuccess) {
    (void)hipGetLastError();
    s->d_up = nullptr;
    return jssp::set_error(e, JSSP_ERR_NOMEM, "serve", "cannot allocate the upload block");
  }
  s->up_cap = cap;
  return JSSP_OK;
}

inline size_t round16(size_t b) { return (b + 15) & ~(size_t)15; }

// Workspace bytes of a request (two-phase, tuned lanes per parent).  The
// shape has the most rank blocks of the three rank block shapes (8 waves,
// one survivor per wave), so the workspace holds the per-block partials of
// whichever shape the request runs in.
size_t footprint(jssp_server* s, const Request& r, Shape& shape) {
  jssp_params q = r.params;
  q.phases = 2;
  const int lanes = jssp_engine_tuned_lanes(s->e, r.inst.n_jobs, r.inst.n_machines, q.beam_width);
  shape = jssp::make_shape(&r.inst, &q, lanes, jssp::kRankWavesThroughput, 1);
  Work w;
  return jssp::carve(shape, nullptr, w);
}

// Admits queued requests (FIFO) while a slot and budget are free.
int admit(jssp_server* s) {
  jssp_engine* e = s->e;
  std::vector<int> fresh;
  while (!s->queue.empty()) {
    Request& r = s->reqs[s->queue.front()];
    if (r.state != kQueued) {
      s->queue.pop_front();
      continue;
    }
    int slot = -1;
    for (int i = 0; i < s->nslots; ++i)
      if (!s->slot_ticket[i]) { slot = i; break; }
    if (slot < 0) break;
    // Admitted in the shape with the most rank blocks; rebuild_maps picks
    // the shape it runs in.
    Shape shape;
    const size_t bytes = footprint(s, r, shape);
    Region reg;
    if (!arena_alloc(s, bytes, reg)) break;
    Work w;
    jssp::carve(shape, s->arena + reg.off, w);
    jssp_params q = r.params;
    q.phases = 2;
    InstDesc d = jssp::make_desc(shape, w, &q, &r.inst);
    d.base.jump = w.jump;  // the step that ends the request runs k_jump_b before its backtrack
    s->slot_level[slot] = 0;
    s->hdesc[slot] = d;
    s->slot_ticket[slot] = s->queue.front();
    s->slot_shape[slot] = shape;
    s->slot_rank_cap[slot] = shape.nblk_rank;
    s->slot_work[slot] = w;
    s->slot_region[slot] = reg;
    r.slot = slot;
    r.state = kRunning;
    r.t_admit = s->now();
    s->queue.pop_front();
    fresh.push_back(slot);
  }
  if (fresh.empty()) return JSSP_OK;
  // One upload of the admitted requests' op tables and descriptors, the
  // init map and the copy jobs that put the tables and descriptors in place;
  // then the init kernel.  All of it runs behind the step in flight.
  std::vector<jssp::CopyJob> jobs;
  std::vector<BlockRef> init_map;
  size_t off = 0;
  for (int slot : fresh) {
    const size_t tot = (size_t)s->slot_shape[slot].n * s->slot_shape[slot].m;
    jobs.push_back({(unsigned long long)(uintptr_t)s->slot_work[slot].ops, (uint32_t)(off / 4), (uint32_t)(tot * 2)});
    off = round16(off + tot * sizeof(int2));
    jobs.push_back({(unsigned long long)(uintptr_t)(s->d_desc + slot), (uint32_t)(off / 4),
                    (uint32_t)(sizeof(InstDesc) / 4)});
    off = round16(off + sizeof(InstDesc));
    init_map.push_back(make_int2(slot, 0));
  }
  const size_t map_off = off;
  off = round16(off + init_map.size() * sizeof(BlockRef));
  const size_t job_off = off;
  off = round16(off + jobs.size() * sizeof(jssp::CopyJob));
  int rc = ensure_up(s, off);
  if (rc) return rc;
  char* h = nullptr;
  rc = stage_take(s, off, &h);
  if (rc) return rc;
  for (size_t k = 0; k < fresh.size(); ++k) {
    const int slot = fresh[k];
    const Request& r = s->reqs[s->slot_ticket[slot]];
    int2* hops = reinterpret_cast<int2*>(h + 4 * (size_t)jobs[2 * k].src);
    const int tot = r.inst.n_jobs * r.inst.n_machines;
    for (int q = 0; q < tot; ++q) { hops[q].x = r.inst.machine[q]; hops[q].y = r.inst.duration[q]; }
    std::memcpy(h + 4 * (size_t)jobs[2 * k + 1].src, &s->hdesc[slot], sizeof(InstDesc));
  }
  std::memcpy(h + map_off, init_map.data(), init_map.size() * sizeof(BlockRef));
  std::memcpy(h + job_off, jobs.data(), jobs.size() * sizeof(jssp::CopyJob));
  JSSP_HIP(e, hipMemcpyAsync(s->d_up, h, off, hipMemcpyHostToDevice, e->stream));
  rc = stage_done(s);
  if (rc) return rc;
  JSSP_HIP(e, jssp::launch_scatter(reinterpret_cast<const uint32_t*>(s->d_up),
                                   reinterpret_cast<const jssp::CopyJob*>(s->d_up + job_off), (int)jobs.size(),
                                   e->stream));
  JSSP_HIP(e, jssp::launch_step_init(s->d_desc, reinterpret_cast<const BlockRef*>(s->d_up + map_off),
                                     (int)init_map.size(), e->stream));
  s->maps_dirty = true;
  return JSSP_OK;
}

int rebuild_maps(jssp_server* s) {
  jssp_engine* e = s->e;
  constexpr int G = jssp_server::kLanes;
  std::vector<BlockRef> ms, mf, mr;
  int nlive = 0;
  std::vector<BlockRef> msg[G], mfg[G], mrv[G][jssp_server::kRankGroups];
  s->flds = 0;
  for (auto& r : s->rlds) r = 0;
  // Filter form: the throughput one (at most 64 blocks per request, two
  // blocks per CU) while the resident requests' blocks fill the chip; with
  // few requests left, the single-solve one with up to 512 blocks per
  // request.  Only the filter's own tiling depends on it, so it may change
  // between levels; the request's descriptor carries its block count.
  long long tp_blocks = 0;
  for (int i = 0; i < s->nslots; ++i)
    if (s->slot_ticket[i]) tp_blocks += s->slot_shape[i].nblk_filter;
  s->filter_tp = tp_blocks >= kFillBlocks;
  // Rank block shape (the same for every resident request, from their load).
  long long rb_batched = 0;
  for (int i = 0; i < s->nslots; ++i)
    if (s->slot_ticket[i]) rb_batched += jssp::rank_blocks(s->slot_shape[i].K, jssp::kRankWavesThroughput, jssp::kRankBatch);
  const int rshape = rb_batched >= kRankFillBatched ? jssp::kRankShapeBatched
                     : rb_batched > kRankFillSingle ? jssp::kRankShapeThroughput
                                                    : jssp::kRankShapeSingle;
  // Groups: largest requests first, each to the group with the least work
  // so far (work per level ~ survivors x snapshot words).
  std::vector<int> order;
  for (int i = 0; i < s->nslots; ++i)
    if (s->slot_ticket[i]) order.push_back(i);
  auto work = [&](int i) { return (long long)s->slot_shape[i].K * s->slot_shape[i].S; };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
  long long load[G] = {};
  std::vector<int> group(s->nslots, 0);
  for (int i : order) {
    int g = 0;
    for (int h = 1; h < G; ++h) g = load[h] < load[g] ? h : g;
    group[i] = g;
    load[g] += work(i);
  }
  std::vector<std::pair<int, int>> nbf_changed;  // (slot, filter blocks)
  std::vector<int4> folds;                       // requests whose rank block count changes
  for (int i = 0; i < s->nslots; ++i) {
    if (!s->slot_ticket[i]) continue;
    Shape& cur = s->slot_shape[i];
    if (jssp::rank_shape_of(cur) != rshape) {
      // The rank blocks write per-block partials (and, fused, the score
      // partials of the next level) that the next kernels read back over the
      // block count: fold the old blocks' partials into the first entry before
      // the count changes.
      const int old = cur.nblk_rank;
      cur.rank_waves = jssp::rank_shape_waves(rshape);
      cur.rank_batch = jssp::rank_shape_batch(rshape);
      cur.nblk_rank = jssp::rank_blocks(cur.K, cur.rank_waves, cur.rank_batch);
      s->hdesc[i].nblk_rank = cur.nblk_rank;
      s->hdesc[i].base.nblk_rank = cur.nblk_rank;
      if (cur.fused) s->hdesc[i].base.nblk_score = cur.nblk_rank;
      folds.push_back(make_int4(i, old, s->slot_rank_cap[i], cur.fused ? 1 : 0));
    }
    Shape sh = cur;
    if (!s->filter_tp) {
      const long long N = (long long)sh.K * sh.B;
      sh.nblk_filter = (int)std::max(1ll, std::min((N + jssp::kFilterThreads - 1) / jssp::kFilterThreads, 512ll));
    }
    if (s->hdesc[i].nblk_filter != sh.nblk_filter) {
      s->hdesc[i].nblk_filter = sh.nblk_filter;
      nbf_changed.push_back({i, sh.nblk_filter});
    }
    const int g = group[i];
    const int rv = jssp::rank_variant(sh, &s->reqs[s->slot_ticket[i]].params);
    jssp::append_maps(sh, i, msg[g], mfg[g], mrv[g][rshape * jssp::kRankVariants + rv]);
    ++nlive;
    const Request& r = s->reqs[s->slot_ticket[i]];
    s->flds = std::max(s->flds, jssp::filter_lds_bytes(sh, &r.params));
    s->rlds[rshape] = std::max(s->rlds[rshape], jssp::rank_lds_bytes(sh));
  }
  for (int g = 0; g < G; ++g) {
    s->ms_off[g] = (int)ms.size();
    s->ms_cnt[g] = (int)msg[g].size();
    ms.insert(ms.end(), msg[g].begin(), msg[g].end());
    s->mf_off[g] = (int)mf.size();
    s->mf_cnt[g] = (int)mfg[g].size();
    mf.insert(mf.end(), mfg[g].begin(), mfg[g].end());
    for (int v = 0; v < jssp_server::kRankGroups; ++v) {
      s->mr_off[g][v] = (int)mr.size();
      s->mr_cnt[g][v] = (int)mrv[g][v].size();
      mr.insert(mr.end(), mrv[g][v].begin(), mrv[g][v].end());
    }
  }
  s->nmr = (int)mr.size();
  s->nact = nlive;
  const size_t nmaps = ms.size() + mf.size() + mr.size();
  if (nmaps > 3 * s->cap_map)
    return jssp::set_error(e, JSSP_ERR_LIMIT, "serve", "block map capacity exceeded");
  // Staging: the three maps (one copy into d_maps), then the upload block:
  // the changed descriptor fields, the fold jobs and the copy jobs that put
  // the fields in place.
  std::vector<int32_t> vals;
  std::vector<jssp::CopyJob> jobs;
  auto field = [&](int* dst, int v) {
    jobs.push_back({(unsigned long long)(uintptr_t)dst, (uint32_t)vals.size(), 1u});
    vals.push_back(v);
  };
  for (const auto& c : nbf_changed) field(&s->d_desc[c.first].nblk_filter, c.second);
  for (const int4& f : folds) {
    // New rank block counts (only these fields: the device owns the rest).
    const InstDesc& hd = s->hdesc[f.x];
    field(&s->d_desc[f.x].nblk_rank, hd.nblk_rank);
    field(&s->d_desc[f.x].base.nblk_rank, hd.base.nblk_rank);
    field(&s->d_desc[f.x].base.nblk_score, hd.base.nblk_score);
  }
  const size_t maps_bytes = round16(nmaps * sizeof(BlockRef));
  const size_t fold_off = round16(vals.size() * sizeof(int32_t));
  const size_t job_off = round16(fold_off + folds.size() * sizeof(int4));
  const size_t up_bytes = round16(job_off + jobs.size() * sizeof(jssp::CopyJob));
  int rc = ensure_up(s, up_bytes);
  if (rc) return rc;
  char* h = nullptr;
  rc = stage_take(s, maps_bytes + up_bytes, &h);
  if (rc) return rc;
  size_t off = 0;
  BlockRef* const dm = s->d_maps;
  auto put = [&](BlockRef*& dst, const std::vector<BlockRef>& v) {
    dst = dm + off / sizeof(BlockRef);
    if (!v.empty()) std::memcpy(h + off, v.data(), v.size() * sizeof(BlockRef));
    off += v.size() * sizeof(BlockRef);
  };
  put(s->d_ms, ms);
  put(s->d_mf, mf);
  put(s->d_mr, mr);
  char* hu = h + maps_bytes;
  if (!vals.empty()) std::memcpy(hu, vals.data(), vals.size() * sizeof(int32_t));
  if (!folds.empty()) std::memcpy(hu + fold_off, folds.data(), folds.size() * sizeof(int4));
  if (!jobs.empty()) std::memcpy(hu + job_off, jobs.data(), jobs.size() * sizeof(jssp::CopyJob));
  if (nmaps) JSSP_HIP(e, hipMemcpyAsync(dm, h, nmaps * sizeof(BlockRef), hipMemcpyHostToDevice, e->stream));
  if (!jobs.empty()) {
    JSSP_HIP(e, hipMemcpyAsync(s->d_up, hu, up_bytes, hipMemcpyHostToDevice, e->stream));
    JSSP_HIP(e, jssp::launch_scatter(reinterpret_cast<const uint32_t*>(s->d_up),
                                     reinterpret_cast<const jssp::CopyJob*>(s->d_up + job_off), (int)jobs.size(),
                                     e->stream));
    // The rank partials of requests whose block count changes, folded
    // before the next step's kernels read them.
    JSSP_HIP(e, jssp::launch_fold_partials(s->d_desc, reinterpret_cast<const int4*>(s->d_up + fold_off),
                                           (int)folds.size(), e->stream));
  }
  rc = stage_done(s);
  if (rc) return rc;
  s->maps_dirty = false;
  return JSSP_OK;
}

// Collects the requests that the status of step `par` shows finished and
// frees their slots.  Their results are final (every later kernel returns at
// once for a finished request), so they are read on the copy stream while the
// step in flight keeps running.
int harvest(jssp_server* s, int par) {
  jssp_engine* e = s->e;
  JSSP_HIP(e, hipEventSynchronize(s->status_ev[par]));
  const volatile int* done = s->status[par];
  std::vector<int> fin;
  std::vector<size_t> at;
  size_t need = 0;
  for (int i : s->status_fin[par]) {
    const uint64_t t = s->slot_ticket[i];
    if (!t || t != s->status_ticket[par][i] || !done[i]) continue;
    fin.push_back(i);
    at.push_back(need);
    need += round16(32 + (size_t)s->slot_shape[i].L * sizeof(int32_t));
  }
  if (fin.empty()) return JSSP_OK;
  if (s->hv_cap < need) {  // the buffer is idle here: every harvest waits for its copies
    const size_t cap = std::max({need, 2 * s->hv_cap, (size_t)256 << 10});
    if (s->hv) JSSP_HIP(e, hipHostFree(s->hv));
    s->hv = nullptr;
    s->hv_cap = 0;
    if (hipHostMalloc((void**)&s->hv, cap, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      s->hv = nullptr;
      return jssp::set_error(e, JSSP_ERR_NOMEM, "serve", "cannot allocate the result landing buffer");
    }
    s->hv_cap = cap;
  }
  // Each finished request: result (makespan, found, best) at +0, expanded
  // count at +16, sequence at +32.
  for (size_t k = 0; k < fin.size(); ++k) {
    const Work& w = s->slot_work[fin[k]];
    char* d = s->hv + at[k];
    JSSP_HIP(e, hipMemcpyAsync(d, w.result, 3 * sizeof(int), hipMemcpyDeviceToHost, s->copy_stream));
    JSSP_HIP(e, hipMemcpyAsync(d + 16, w.expanded, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               s->copy_stream));
    JSSP_HIP(e, hipMemcpyAsync(d + 32, w.seq, (size_t)s->slot_shape[fin[k]].L * sizeof(int32_t),
                               hipMemcpyDeviceToHost, s->copy_stream));
  }
  JSSP_HIP(e, hipStreamSynchronize(s->copy_stream));
  for (size_t k = 0; k < fin.size(); ++k) {
    const int i = fin[k];
    Request& r = s->reqs[s->slot_ticket[i]];
    const char* d = s->hv + at[k];
    int res[3];
    unsigned long long ex = 0;
    std::memcpy(res, d, sizeof(res));
    std::memcpy(&ex, d + 16, sizeof(ex));
    r.seq.assign(reinterpret_cast<const int32_t*>(d + 32),
                 reinterpret_cast<const int32_t*>(d + 32) + s->slot_shape[i].L);
    r.makespan = res[0];
    r.found = res[1];
    r.expanded = (int64_t)ex;
    r.state = kDone;
    r.t_done = s->now();
    arena_free(s, s->slot_region[i]);
    s->slot_ticket[i] = 0;
    s->maps_dirty = true;
  }
  return JSSP_OK;
}

}  // namespace

extern "C" {

int jssp_server_create(jssp_engine* e, const jssp_server_config* cfg, jssp_server** out) {
  if (!e || !cfg || !out || cfg->hbm_budget_bytes <= 0 || cfg->max_slots <= 0) return JSSP_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  *out = nullptr;
  jssp_server* s = new jssp_server();
  s->e = e;
  s->cfg = *cfg;
  s->t0 = std::chrono::steady_clock::now();
  s->nslots = cfg->max_slots;
  s->arena_bytes = (size_t)cfg->hbm_budget_bytes;
  if (hipSetDevice(e->device) != hipSuccess || hipMalloc(&s->arena, s->arena_bytes) != hipSuccess) {
    (void)hipGetLastError();
    delete s;
    return jssp::set_error(e, JSSP_ERR_NOMEM, "jssp_server_create", "cannot reserve the HBM budget");
  }
  s->free_list.push_back({0, s->arena_bytes});
  s->cap_map = (size_t)s->nslots * 512;
  jssp::Bump b;
  b.take<InstDesc>(s->nslots);
  b.take<BlockRef>(3 * s->cap_map);
  const size_t meta_bytes = b.off + 256;
  if (hipMalloc(&s->meta, meta_bytes) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(s->arena);
    delete s;
    return jssp::set_error(e, JSSP_ERR_NOMEM, "jssp_server_create", "cannot allocate slot descriptors");
  }
  b = jssp::Bump();
  b.base = s->meta;
  s->d_desc = b.take<InstDesc>(s->nslots);
  s->d_maps = b.take<BlockRef>(3 * s->cap_map);
  s->d_ms = s->d_mf = s->d_mr = s->d_maps;
  s->slot_ticket.assign(s->nslots, 0);
  s->slot_rank_cap.assign(s->nslots, 0);
  s->slot_shape.resize(s->nslots);
  s->slot_work.resize(s->nslots);
  s->slot_region.resize(s->nslots);
  s->slot_level.assign(s->nslots, 0);
  s->hdesc.resize(s->nslots);
  for (int k = 0; k < 2; ++k) {
    if (hipEventCreateWithFlags(&s->stage[k].ev, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      s->stage[k].ev = nullptr;
      jssp_server_destroy(s);
      return jssp::set_error(e, JSSP_ERR_NOMEM, "jssp_server_create", "cannot create the staging events");
    }
  }
  for (int k = 0; k < 2; ++k) {
    s->status_ticket[k].assign(s->nslots, 0);
    if (hipHostMalloc((void**)&s->status[k], (size_t)s->nslots * sizeof(int), hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void**)&s->status_dev[k], s->status[k], 0) != hipSuccess ||
        hipHostMalloc((void**)&s->fin_map[k], (size_t)s->nslots * sizeof(BlockRef), hipHostMallocCoherent) !=
            hipSuccess ||
        hipHostGetDevicePointer((void**)&s->fin_map_dev[k], s->fin_map[k], 0) != hipSuccess ||
        hipEventCreateWithFlags(&s->status_ev[k], hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      jssp_server_destroy(s);
      return jssp::set_error(e, JSSP_ERR_NOMEM, "jssp_server_create", "cannot allocate the step status buffers");
    }
  }
  if (hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    jssp_server_destroy(s);
    return jssp::set_error(e, JSSP_ERR_NOMEM, "jssp_server_create", "cannot create the copy stream");
  }
  s->lane_stream[0] = e->stream;
  for (int g = 0; g < jssp_server::kLanes; ++g) {
    if ((g > 0 && hipStreamCreateWithFlags(&s->lane_stream[g], hipStreamNonBlocking) != hipSuccess) ||
        hipEventCreateWithFlags(&s->lane_ev[g], hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      jssp_server_destroy(s);
      return jssp::set_error(e, JSSP_ERR_NOMEM, "jssp_server_create", "cannot create the lane streams");
    }
  }
  *out = s;
  return JSSP_OK;
}

void jssp_server_destroy(jssp_server* s) {
  if (!s) return;
  std::lock_guard<std::recursive_mutex> lk(s->e->mu);
  (void)hipStreamSynchronize(s->e->stream);
  for (int g = 1; g < jssp_server::kLanes; ++g)
    if (s->lane_stream[g]) {
      (void)hipStreamSynchronize(s->lane_stream[g]);
      (void)hipStreamDestroy(s->lane_stream[g]);
    }
  for (int g = 0; g < jssp_server::kLanes; ++g)
    if (s->lane_ev[g]) (void)hipEventDestroy(s->lane_ev[g]);
  if (s->copy_stream) {
    (void)hipStreamSynchronize(s->copy_stream);
    (void)hipStreamDestroy(s->copy_stream);
  }
  for (int k = 0; k < 2; ++k) {
    if (s->status_ev[k]) (void)hipEventDestroy(s->status_ev[k]);
    if (s->stage[k].ev) (void)hipEventDestroy(s->stage[k].ev);
    if (s->stage[k].h) (void)hipHostFree(s->stage[k].h);
    if (s->status[k]) (void)hipHostFree(s->status[k]);
    if (s->fin_map[k]) (void)hipHostFree(s->fin_map[k]);
  }
  if (s->arena) (void)hipFree(s->arena);
  if (s->meta) (void)hipFree(s->meta);
  if (s->d_up) (void)hipFree(s->d_up);
  if (s->hv) (void)hipHostFree(s->hv);
  delete s;
}

int jssp_serve_submit(jssp_server* s, const jssp_instance* inst, const jssp_params* params, uint64_t* ticket) {
  if (!s || !inst || !params || !ticket) return JSSP_ERR_INVALID;
  jssp_engine* e = s->e;
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  int rc = jssp::check_instance(e, inst);
  if (rc) return rc;
  rc = jssp::check_params(e, inst, params);
  if (rc) return rc;
  if (params->mode != JSSP_MODE_ICH_DET && params->mode != JSSP_MODE_PMS_GPU)
    return jssp::set_error(e, JSSP_ERR_INVALID, "serve", "served requests take modes ICH_DET and PMS_GPU");
  if (params->n_seeds > 1 || params->n_passes > 1)
    return jssp::set_error(e, JSSP_ERR_INVALID, "serve", "n_seeds / n_passes apply to jssp_solve");
  Request r;
  const int tot = inst->n_jobs * inst->n_machines;
  r.mach.assign(inst->machine, inst->machine + tot);
  r.dur.assign(inst->duration, inst->duration + tot);
  r.inst = *inst;
  r.inst.machine = r.mach.data();
  r.inst.duration = r.dur.data();
  r.params = *params;
  Shape shape;
  if (footprint(s, r, shape) > s->arena_bytes)
    return jssp::set_error(e, JSSP_ERR_BUDGET, "serve", "request workspace exceeds the HBM budget");
  r.t_submit = s->now();
  const uint64_t t = s->next_ticket++;
  auto it = s->reqs.emplace(t, std::move(r)).first;
  it->second.inst.machine = it->second.mach.data();
  it->second.inst.duration = it->second.dur.data();
  s->queue.push_back(t);
  *ticket = t;
  return JSSP_OK;
}

int jssp_serve_step(jssp_server* s, int32_t levels, int32_t* active) {
  if (!s || levels < 0) return JSSP_ERR_INVALID;
  jssp_engine* e = s->e;
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  JSSP_HIP(e, hipSetDevice(e->device));
  int rc = admit(s);
  if (rc) return rc;
  if (s->maps_dirty) {
    rc = rebuild_maps(s);
    if (rc) return rc;
  }
  const int prev = s->status_pending;
  s->status_pending = -1;
  if (s->nact > 0) {
    // Every lane starts after what the engine's stream holds (admission,
    // the previous step); the lanes are joined again before the backtrack.
    constexpr int G = jssp_server::kLanes;
    JSSP_HIP(e, hipEventRecord(s->lane_ev[0], e->stream));
    for (int g = 1; g < G; ++g)
      if (s->mf_cnt[g] > 0) JSSP_HIP(e, hipStreamWaitEvent(s->lane_stream[g], s->lane_ev[0], 0));
    for (int t = 0; t < levels; ++t) {
      for (int g = 0; g < G; ++g) {
        if (s->mf_cnt[g] == 0) continue;
        hipStream_t st = s->lane_stream[g];
        if (s->ms_cnt[g] > 0)
          JSSP_HIP(e, jssp::launch_step_score(s->d_desc, s->d_ms + s->ms_off[g], s->ms_cnt[g], false, st));
        JSSP_HIP(e, jssp::launch_step_filter(s->d_desc, s->d_mf + s->mf_off[g], s->mf_cnt[g], s->flds,
                                             s->filter_tp, st));
        for (int v = 0; v < jssp_server::kRankGroups; ++v) {
          if (s->mr_cnt[g][v] == 0) continue;
          const int rs = v / jssp::kRankVariants;
          JSSP_HIP(e, jssp::launch_step_rank(s->d_desc, s->d_mr + s->mr_off[g][v], s->mr_cnt[g][v], s->rlds[rs],
                                             v % jssp::kRankVariants, jssp::rank_shape_waves(rs), st, nullptr,
                                             jssp::rank_shape_batch(rs)));
        }
      }
    }
    for (int g = 1; g < G; ++g) {
      if (s->mf_cnt[g] == 0) continue;
      JSSP_HIP(e, hipEventRecord(s->lane_ev[g], s->lane_stream[g]));
      JSSP_HIP(e, hipStreamWaitEvent(e->stream, s->lane_ev[g], 0));
    }
    // This step's status, read by the next serve step (or by this one when
    // it runs no levels).
    const int par = s->parity;
    s->parity ^= 1;
    s->status_ticket[par] = s->slot_ticket;
    // Requests whose last level runs in this step get their jump table (so
    // the backtrack walks their trail in parallel 32-level groups) and their
    // backtrack; a step that ends no request launches neither.  The map
    // buffer of this parity was last read by the step before the previous
    // one, whose status the previous serve step waited for.
    std::vector<int>& fin = s->status_fin[par];
    fin.clear();
    for (int i = 0; i < s->nslots; ++i) {
      if (!s->slot_ticket[i]) continue;
      const Shape& sh = s->slot_shape[i];
      if (s->slot_level[i] < sh.L && s->slot_level[i] + levels >= sh.L) {
        JSSP_HIP(e, jssp::launch_step_jump(s->d_desc + i, sh.K, sh.L, e->stream));
        s->fin_map[par][fin.size()] = make_int2(i, 0);
        fin.push_back(i);
      }
      s->slot_level[i] = std::min(sh.L, s->slot_level[i] + levels);
    }
    if (!fin.empty())
      JSSP_HIP(e, jssp::launch_step_backtrack(s->d_desc, s->fin_map_dev[par], (int)fin.size(), e->stream,
                                              s->status_dev[par]));
    JSSP_HIP(e, hipEventRecord(s->status_ev[par], e->stream));
    s->status_pending = par;
  }
  if (prev >= 0) {
    rc = harvest(s, prev);
    if (rc) return rc;
  }
  if (levels == 0 && s->status_pending >= 0) {
    // A step of no levels reads everything enqueued so far.
    const int cur = s->status_pending;
    s->status_pending = -1;
    rc = harvest(s, cur);
    if (rc) return rc;
  }
  if (active) {
    int a = 0;
    for (int i = 0; i < s->nslots; ++i) a += s->slot_ticket[i] ? 1 : 0;
    *active = a;
  }
  return JSSP_OK;
}

int jssp_serve_poll(jssp_server* s, uint64_t ticket, jssp_result* res) {
  if (!s || !res) return JSSP_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> lk(s->e->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return JSSP_ERR_NOTFOUND;
  const Request& r = it->second;
  if (r.state == kCancelled) return JSSP_ERR_CANCELLED;
  if (r.state != kDone) return JSSP_ERR_PENDING;
  res->makespan = r.makespan;
  res->found = r.found;
  res->expanded_states = r.expanded;
  res->levels = r.inst.n_jobs * r.inst.n_machines;
  res->best_seed_index = 0;
  res->device_ms = (r.t_done - r.t_admit) * 1e3;
  if (r.found && res->sequence) std::memcpy(res->sequence, r.seq.data(), r.seq.size() * sizeof(int32_t));
  return JSSP_OK;
}

int jssp_serve_cancel(jssp_server* s, uint64_t ticket) {
  if (!s) return JSSP_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> lk(s->e->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return JSSP_ERR_NOTFOUND;
  Request& r = it->second;
  if (r.state == kDone || r.state == kCancelled) return JSSP_OK;
  if (r.state == kRunning) {
    // A step may still be running on the slot's workspace: wait for it, then
    // drop the slot from the maps and release its workspace.
    JSSP_HIP(s->e, hipStreamSynchronize(s->e->stream));
    arena_free(s, s->slot_region[r.slot]);
    s->slot_ticket[r.slot] = 0;
    s->maps_dirty = true;
  }
  r.state = kCancelled;
  r.t_done = s->now();
  return JSSP_OK;
}

int jssp_serve_drain(jssp_server* s) {
  if (!s) return JSSP_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> lk(s->e->mu);
  for (;;) {
    int active = 0;
    int rc = jssp_serve_step(s, 32, &active);
    if (rc) return rc;
    bool queued = false;
    for (uint64_t t : s->queue)
      if (s->reqs[t].state == kQueued) { queued = true; break; }
    if (!active && !queued) return JSSP_OK;
    if (!active && queued) {
      // Nothing resident and the head request does not fit: cannot progress.
      return jssp::set_error(s->e, JSSP_ERR_BUDGET, "serve", "queued request does not fit the HBM budget");
    }
  }
}

int jssp_serve_times(jssp_server* s, uint64_t ticket, double* submit_s, double* admit_s, double* done_s) {
  if (!s) return JSSP_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> lk(s->e->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return JSSP_ERR_NOTFOUND;
  if (submit_s) *submit_s = it->second.t_submit;
  if (admit_s) *admit_s = it->second.t_admit;
  if (done_s) *done_s = it->second.t_done;
  return JSSP_OK;
}

}  // extern "C"
