// level_kernels.hpp -- host-visible interface of the per-level kernels of the
// single-instance path (one solve spread over the whole GPU).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jssp {

// First selection digit, kNB bins.  Exact form (when a lower makespan always
// scores lower): 8 * (cmax - cmin) + min(noise offset in ulps, 7).  Generic
// form: the score bits above a proven lower bound, shifted so the previous
// level's survivor spread covers ~2^kSpreadBins bins.
constexpr int kNB = 2048;
// The score blocks merge their LDS histograms into up to kHistCopies copies
// (block b into copy b % hist_copies_for(K)), which the filter sums:
// same-address atomics serialise, so spreading them cuts the merge time by
// the copy count.
constexpr int kHistCopies = 4;
// Copies a solve uses: one per 64 score/rank blocks at most (a 16-wave rank
// block takes 16 survivors, so K / 16 blocks), which keeps the same-address
// merges per copy at single-solve K = 4096's level while narrower beams read
// fewer copies in every filter block (a batch of 256 instances with K <= 1024:
// +0.9% solves/s with one copy).  A function of K alone, so every block shape
// a server may switch a request to merges into and reads the same copies.
__host__ __device__ inline int hist_copies_for(int K) { return K <= 1024 ? 1 : (K <= 2048 ? 2 : kHistCopies); }
// The shift makes the previous level's survivor spread cover ~2^10 bins.
constexpr int kSpreadBins = 10;
// Shift used at level 0 (no previous level; at most n children anyway).
constexpr uint32_t kLevel0Shift = 20;
// Waves per score block (1024 threads; a wave holds 64 / G parents).
constexpr int kScoreWaves = 16;
// Threads per filter block (one tile of keys per pass) and waves per rank block.
constexpr int kFilterThreads = 1024;
// Tiles a block of the batch / server filter form covers per pass (one
// reservation round trip for all of them).
constexpr int kFilterTilesTp = 4;
constexpr int kRankWaves = 16;
// Rank blocks of batches and the server: half the waves, so two blocks share
// a CU and one block's memory waits overlap another's work (single solves
// keep one 16-wave block per CU: fewer blocks stage the survivor keys).
constexpr int kRankWavesThroughput = 8;
// Survivors a wave of those blocks ranks together (one pass over the LDS keys).
constexpr int kRankBatch = 4;
// Survivors ranked in LDS (8 bytes each) up to this beam width; wider beams
// rank from global memory.
constexpr int kRankLdsMax = 16384;
// Dynamic LDS the fused rank kernel may use (its static LDS is ~9 KB of 160 KB);
// solves that would need more use the separate score kernel.
constexpr size_t kRankFusedLdsMax = 144 * 1024;
// Dynamic LDS of the rank kernel before the op table: survivor keys (K <=
// kRankLdsMax), the tile prefix of tie counts (padded to 16 bytes) and,
// fused, one snapshot view per wave.
__host__ __device__ inline size_t rank_lds_base(int K, int n, int m, bool fused, int waves) {
  const size_t tiles = ((size_t)K * n + kFilterThreads - 1) / kFilterThreads;
  size_t b = (K <= kRankLdsMax ? (size_t)K * 8 : 0) + 4 * ((tiles + 1 + 3) & ~(size_t)3);
  if (fused) b += (size_t)waves * (2 * n + m + 1) * 4;
  return b;
}
// Snapshot size class of the fused rank kernels: 0 (small) = S <= 128 words
// and n <= 64 jobs, 1 (medium) = S <= 256 words, 2 (large) = the rest.
__host__ __device__ inline int rank_size_class(int S, int n) {
  return (S <= 128 && n <= 64) ? 0 : (S <= 256 ? 1 : 2);
}
// Fused throughput rank blocks (batches, the server) with K <= 1024 also
// stage the op table (8 bytes per op, up to kOpsLdsMax ops) after the views:
// their waves then score the next level's children of their survivors from
// LDS, not from memory (batch of 256 instances +2.4% solves/s).  Up to
// K = 1024 the table costs no block per CU (at most ~50 KB per block with
// the ~9 KB of static LDS, against 3 or 4 blocks in 160 KB); wider beams
// would lose one (a serving mix with K = 4096 requests staging it: 0.149 ->
// 0.159 s).  Single-solve blocks do not stage it (measured slower: one
// survivor per wave does not repay the staging).
constexpr int kOpsLdsMax = 2048;
__host__ __device__ inline bool rank_ops_in_lds(int K, int n, int m, bool fused, int waves) {
  return fused && waves == kRankWavesThroughput && K <= 1024 && n * m <= kOpsLdsMax;
}
// The filter reserves the slots of the keys below the threshold bin with
// atomics on kSureSubs counters (filter block b uses counter b % kSureSubs,
// its keys going to cand[r*K, (r+1)*K)), kCtrStride words apart, instead of
// one counter that every block hits in turn.  Used when the threshold bin is
// single-valued, K <= kRankLdsMax and B >= kSureSubs; the rank kernel gathers
// the regions in order.
constexpr int kSureSubs = 8;
constexpr int kCtrStride = 64;
__host__ __device__ inline bool sure_subs(bool single_bin, int K, int B) {
  return single_bin && K <= kRankLdsMax && B >= kSureSubs;
}
// LDS bitmap over child indices for the threshold cut: up to 2^20 children.
constexpr uint32_t kBitmapWordsMax = 32768;

struct BlockStat {
  uint32_t valid;           // kept children
  uint32_t expanded;        // children evaluated
  int min_cmax;             // smallest cmax among the kept children
  uint32_t pad;
  unsigned long long best;  // last level: min (cmax << 32 | child)
};

struct RankStat {
  int min_cmax;     // smallest cmax among the survivors this block materialised
  uint32_t max_hi;  // largest score bits among them
  uint32_t min_hi;  // smallest score bits among them
  uint32_t pad;
};

struct SelState {
  uint32_t lo, s1;          // digit parameters of the level (exact digit: lo written by the
                            // previous level's filter block 0; generic: by score block 0)
  uint32_t V, Kp;           // kept children, survivors = min(V, K)
  uint32_t T1, cum_lt, t1, need;  // threshold bin, keys below it, its size, taken from it
  uint32_t sure_ctr, tie_ctr, done_ctr;  // filter counters, zeroed before the level
  uint32_t exact;           // 1: the digit is the exact (cmax, noise offset) digit
  uint32_t single;          // the threshold bin holds a single score value
  uint32_t sorted_from;     // cand[0, sorted_from) needs ranking; the rest are ties in index order
  uint32_t pad[2];
};

struct LevelArgs {
  const int2* ops;           // (machine, duration) per op j*m + k
  const int* beam_cur;       // P x S
  int* beam_next;            // K x S
  int* cbuf;                 // single-phase children, (K*B) x S
  unsigned long long* keys;  // K*B 8-byte slots: child x's key word as u32 at word x (the filter forms the
                             // (key word, x) key), then the refinement scratch (u64).  Key word: the FP32
                             // score bits, or with the exact digit (cmax << 11 | the score's ulps above
                             // fl(cmax / ns)), which orders the children the same way
  unsigned long long* crec;  // K*B: kept child's record, finish | machine << 32 | job << 40 | next
                             // machine << 48 (two buffers by level parity)
  unsigned long long* cand;  // K*B: candidate survivors (first Kp after the filter)
  unsigned long long* ref_buf;  // K*B + K: scratch of the threshold cut / the ties of each filter tile
  uint8_t* pvalid;           // K: valid child slots of each parent (slots >= it hold no key)
  uint32_t* tcnt;            // ceil(K*B/kFilterThreads): keys of that bin per filter tile
  uint32_t* hist;            // kHistCopies x kNB, this level
  uint32_t* hist_next;       // kHistCopies x kNB, zeroed for the next level
  BlockStat* bstat;          // nblk_score partials
  RankStat* rstat;           // nblk_rank partials of the previous level's survivors
  uint32_t* sctr;            // this level's sure-key counters (kSureSubs, kCtrStride words apart)
  uint32_t* sctr_next;       // the next level's (zeroed by the rank kernel)
  SelState* sel;             // this level
  const SelState* sel_prev;  // previous level
  SelState* sel_next;        // next level (counters zeroed by the rank kernel)
  SelState* sel_log;         // L entries or null: copy of each level's selection state
  int* pcount;               // L+1 parent counts
  int* trail;                // (L-1) x K
  uint16_t* jump;            // null, or ceil((L-1)/32) x K: for survivor q after the last level of
                             // 32-level group g, its ancestor after level 32 g - 1 (k_jump_b)
  unsigned long long* expanded;  // running total
  uint32_t* lvl_exp;         // L: children evaluated per level
  int* seq_out;              // L
  int* result;               // [makespan, found, cmax]
  // exact (host-replay) modes
  int* ex_cmax;              // K*B
  uint8_t* ex_job;           // K*B
  uint8_t* ex_mc;            // K*B
  int* ex_rcount;            // K
  const int* ex_order;       // survivors (child index) in heap-array order
  int n, m, S, B, K, L;
  int level, best_known, nblk_score, nblk_rank;
  uint32_t bm_words;         // LDS bitmap words of the filter's threshold cut (dynamic LDS)
  int exact_digit;           // use the exact (cmax, noise offset) digit (see score_body)
  int mode;                  // 1 ICH_DET, 2 PMS_GPU
  int G;                     // lanes per parent in the score kernel: 16, 32 or 64
  uint32_t lseed;
  int bt_parent, bt_job;     // explicit backtrack start (exact modes), -1 = use bstat best
  int hist_copies;           // histogram copies in use: hist_copies_for(K)
};

// Device-resident state of one solve in flight (a batch is an array of them).
// `base` holds the per-solve constants; the level-dependent views (current /
// next beam, histogram, selection state, level seed) are resolved in each
// kernel from device-side level counters, so one step's launches are
// identical at every level and can be captured once in a hipGraph.
struct InstDesc {
  LevelArgs base;
  int* beam[2];
  uint32_t* hist[2];
  SelState* sel[2];
  unsigned long long* crec[2];
  uint32_t* sctr[2];
  int nblk_score, nblk_filter, nblk_rank;
  // Level each step kernel runs next.  Each counter is advanced by block 0
  // of the kernel that follows its reader in the step (score -> filter ->
  // rank -> score), after the reader has completed, so no block ever sees a
  // counter change while its kernel runs; all saturate at L.
  int lvl_score, lvl_filter, lvl_rank;
  int bt_done;           // sequence rebuilt
  int fused;             // the next level is scored inside the rank kernel (no score launch):
                         // the filter advances lvl_rank, base.nblk_score = nblk_rank partials
  int pad2;
  int* shared_bound;     // multi-seed: device best bound read at init, lowered at the end
};

// Block map entry: (instance index, block index within that instance).
typedef int2 BlockRef;

// Step kernels over an array of descriptors; map[b] names the instance and
// local block of grid block b (map null: instance 0, local block b).
hipError_t launch_step_init(InstDesc* desc, const BlockRef* map, int nblk, hipStream_t s);
// Optional timing events recorded around one launch.
struct EventPair {
  hipEvent_t start, stop;
};

// single: write every kept child's snapshot (single-phase expand).  ev: time
// this launch (null: plain launch).
hipError_t launch_step_score(InstDesc* desc, const BlockRef* map, int nblk, bool single, hipStream_t s,
                             const EventPair* ev = nullptr);
// throughput: the batch / server form of the kernel (two blocks per CU).
hipError_t launch_step_filter(InstDesc* desc, const BlockRef* map, int nblk, size_t lds_bytes, bool throughput,
                              hipStream_t s, const EventPair* ev = nullptr);
// Rank kernel variants: the separate-score form, and the fused form per mode
// (ICH_DET, PMS_GPU) and snapshot size class.  Each returns at once for
// instances of another variant, so a batch mixing them launches each
// variant it holds.
enum RankVariant : int {
  kRankUnfused = 0,
  kRankFusedIch = 1,  // large
  kRankFusedPms = 2,
  kRankFusedIchSmall = 3,
  kRankFusedPmsSmall = 4,
  kRankUnfusedWide = 5,  // separate-score form with K > kRankLdsMax (survivors ranked from global memory)
  kRankFusedIchMedium = 6,
  kRankFusedPmsMedium = 7,
  kRankVariants = 8
};
// waves: kRankWaves or kRankWavesThroughput (the Shape's rank_waves);
// rank_batch > 1 (throughput waves only): the kernels that rank kRankBatch
// survivors of a wave together (the Shape's rank_batch).
hipError_t launch_step_rank(InstDesc* desc, const BlockRef* map, int nblk, size_t lds_bytes, int variant,
                            int waves, hipStream_t s, const EventPair* ev = nullptr, int rank_batch = 1);
// status (optional, device-visible host memory): per instance of the map,
// whether it is finished after this launch.
hipError_t launch_step_backtrack(InstDesc* desc, const BlockRef* map, int nblk, hipStream_t s,
                                 int* status = nullptr);
// Per-block partials of served requests whose rank block count changes
// between levels: job (slot, old rank blocks, partial capacity, fused) folds
// the request's rank partials (and, fused, its score partials, which the
// rank blocks write) [0, old) into entry 0 and clears entries [1, capacity).
// One block per job.
hipError_t launch_fold_partials(InstDesc* desc, const int4* jobs, int count, hipStream_t s);
// Copies of 4-byte words from one uploaded block to scattered device
// destinations (the server's op tables, descriptors and descriptor fields):
// job i copies words [src, src + words) of the block to dst.  One block per job.
struct CopyJob {
  unsigned long long dst;  // device address, 4-byte aligned
  uint32_t src;            // word offset into the uploaded block
  uint32_t words;
};
hipError_t launch_scatter(const uint32_t* block, const CopyJob* jobs, int count, hipStream_t s);
// Single solve: fills base.jump (when set) for the backtrack; K, L of the solve.
hipError_t launch_step_jump(InstDesc* desc, int K, int L, hipStream_t s);

// Exact (host-replay) modes, driven level by level from the host.
hipError_t launch_init(const LevelArgs& a, hipStream_t s);
hipError_t launch_exact_expand(const LevelArgs& a, int P, hipStream_t s);
hipError_t launch_exact_materialize(const LevelArgs& a, int count, hipStream_t s);
hipError_t launch_backtrack(const LevelArgs& a, hipStream_t s);

}  // namespace jssp
