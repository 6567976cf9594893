// workspace.hpp -- device workspace planning and instance descriptors,
// shared by the single-solve, batch and serving paths.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "engine.hpp"
#include "level_kernels.hpp"

namespace jssp {

// Device workspace of one solve (one slot of a batch).
struct Work {
  int2* ops;
  int* beam[2];
  int* cbuf;
  unsigned long long* keys;
  unsigned long long* crec[2];
  unsigned long long* cand;
  unsigned long long* ref_buf;
  uint8_t* pvalid;
  uint32_t* tcnt;
  uint32_t* hist[2];
  BlockStat* bstat;
  RankStat* rstat;
  SelState* sel[2];
  uint32_t* sctr[2];
  int* pcount;
  int* trail;
  uint16_t* jump;
  unsigned long long* expanded;
  uint32_t* lvl_exp;
  SelState* sel_log;
  int* seq;
  int* result;
  int* ex_cmax;
  uint8_t* ex_job;
  uint8_t* ex_mc;
  int* ex_rcount;
  int* ex_order;
};

// Sizes and launch shape of one solve.
struct Shape {
  int n, m, S, B, K, L;
  bool single, exact;
  int G;                       // lanes per parent in the score kernel
  bool fused;                  // the rank kernel scores the next level (no score kernel)
  int rank_waves;              // waves per rank block: kRankWaves, or kRankWavesThroughput
  int rank_batch;              // survivors per rank wave at full width: 1, or kRankBatch
  int nblk_score, nblk_filter, nblk_rank;
};

// Lanes per parent: the smallest group that holds every job (64 above 32).
int default_lanes(int n_jobs);
// Score shape code of the fused form: each level's scoring runs inside the
// previous level's rank kernel, one wave per parent.
constexpr int kLanesFused = 1;
// Whether the fused form applies: two-phase, a device mode, the exact digit.
bool fused_eligible(const jssp_instance* in, const jssp_params* p);
// lanes: 0 = automatic (fused when eligible, else default_lanes), kLanesFused,
// or 16 / 32 / 64 lanes per parent in a separate score kernel (single-phase
// always uses a 64-lane score kernel).
// rank_waves: kRankWaves (single solves) or kRankWavesThroughput (batches,
// the server).
// rank_batch: survivors per rank wave at full width (kRankBatch for batches,
// whose instances fill the chip together; 1 keeps every wave to one survivor).
// filter_cap: most filter blocks per instance in the throughput shapes (0 =
// the default, 64); single solves take up to 512.
Shape make_shape(const jssp_instance* in, const jssp_params* p, int lanes, int rank_waves = kRankWaves,
                 int rank_batch = 1, int filter_cap = 0);

// A batch shape chosen per (n, m, log2 K) class by the autotuner, packed in
// one int: lanes per parent (bits 0-7: kLanesFused, 16, 32 or 64; 0 =
// automatic), the rank block shape (bits 8-11, RankShape) and the filter
// block cap (bits 12-15, an index into kFilterCaps; 0 = the default).
enum RankShape : int {
  kRankShapeBatched = 0,     // 8-wave blocks, kRankBatch survivors per wave (the default)
  kRankShapeThroughput = 1,  // 8-wave blocks, one survivor per wave
  kRankShapeSingle = 2,      // 16-wave blocks, one survivor per wave
  kRankShapes = 3
};
constexpr int kFilterCaps[4] = {0, 16, 64, 256};
inline int code_lanes(int code) { return code & 0xFF; }
inline int code_rank_shape(int code) { return (code >> 8) & 0xF; }
inline int code_filter_cap(int code) { return kFilterCaps[(code >> 12) & 0x3]; }
inline int make_code(int lanes, int rank_shape, int cap_index) { return lanes | (rank_shape << 8) | (cap_index << 12); }
inline int rank_shape_waves(int shape) { return shape == kRankShapeSingle ? kRankWaves : kRankWavesThroughput; }
inline int rank_shape_batch(int shape) { return shape == kRankShapeBatched ? kRankBatch : 1; }
// The batch shape of a tuning code.
inline Shape make_batch_shape(const jssp_instance* in, const jssp_params* p, int code) {
  const int rs = code_rank_shape(code);
  return make_shape(in, p, code_lanes(code), rank_shape_waves(rs), rank_shape_batch(rs), code_filter_cap(code));
}
// Rank blocks of a solve of beam width K in the given block shape: one
// survivor per wave and pass (batch survivors per wave in the batched
// form), at most 4096 waves.
inline int rank_blocks(int K, int waves, int batch) {
  const int per_blk = waves * batch;
  return std::max(1, std::min((K + per_blk - 1) / per_blk, 16 * 256 / waves));
}
// Rank shape of a batch Shape.
inline int rank_shape_of(const Shape& s) {
  return s.rank_waves == kRankWaves ? kRankShapeSingle : (s.rank_batch > 1 ? kRankShapeBatched : kRankShapeThroughput);
}
// Bytes of the workspace (base == nullptr) or carve it at base.
size_t carve(const Shape& s, char* base, Work& w);
LevelArgs base_args(const Shape& s, const Work& w, const jssp_params* p);
bool exact_digit_applies(const jssp_instance* in);
// Descriptor of one solve (host copy; upload it to the device array).
InstDesc make_desc(const Shape& s, const Work& w, const jssp_params* p, const jssp_instance* in);
// Appends this instance's blocks to the three step maps.
void append_maps(const Shape& s, int inst, std::vector<BlockRef>& score, std::vector<BlockRef>& filter,
                 std::vector<BlockRef>& rank);
// Dynamic LDS of the filter (threshold bitmap) and rank (survivor keys) kernels.
size_t filter_lds_bytes(const Shape& s, const jssp_params* p);
size_t rank_lds_bytes(const Shape& s);
// Rank kernel variant of a solve (RankVariant).
int rank_variant(const Shape& s, const jssp_params* p);
// Writes (machine, duration) pairs to the staging buffer and copies them to d_ops.
int upload_ops(jssp_engine* e, const jssp_instance* in, int2* d_ops);

struct Prof;
// Enqueues a whole single-beam device solve (ICH_DET / PMS_GPU) on e->stream
// for a descriptor already in device memory: the init kernel, every level's
// step kernels, the jump table and the backtrack, between the engine's
// start and stop events.  prof (may be null) times the step kernels.
int enqueue_single_solve(jssp_engine* e, const Shape& s, const jssp_params* q, InstDesc* d_desc, Prof* prof);

// Optional per-kernel timing: events recorded by the launches themselves.
struct Prof {
  jssp_engine* e;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;
  size_t next = 0;
  explicit Prof(jssp_engine* eng) : e(eng) {}
  hipEvent_t get();
  // Event pair for one launch of `kind` (stored in *ep), or null when not
  // profiling or when `on` is false.
  const EventPair* pair(int kind, EventPair* ep, bool on = true);
  void collect();
};

}  // namespace jssp
