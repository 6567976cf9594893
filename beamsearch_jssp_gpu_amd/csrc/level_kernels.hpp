// level_kernels.hpp -- host-visible interface of the per-level kernels of the
// single-instance path (one solve spread over the whole GPU).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jssp {

// Histogram width of the first selection digit.
constexpr int kNB = 2048;
constexpr int kLog2NB = 11;
// Largest tie bucket sorted in LDS by the tie kernel.
constexpr int kTieLds = 8192;
// Parents per score block (one wave per parent, 1024-thread blocks).
constexpr int kScoreWaves = 16;
// Survivors per rank block (one wave per survivor, 256-thread blocks).
constexpr int kRankWaves = 4;

struct BlockStat {
  uint32_t valid;        // kept children (score phase)
  uint32_t minhi;        // min score bits among them
  uint32_t maxhi;        // max score bits among them
  uint32_t expanded;     // children evaluated
  unsigned long long best;  // last level: min (cmax << 32 | child)
};

struct SelState {
  uint32_t V, lo, hi, shift, Kp;  // valid count, score-bit range, digit shift, survivors
  uint32_t T, cumlt, need, t;      // threshold bin, keys below it, survivors taken from it, its size
  uint32_t pad[7];
};

struct LevelArgs {
  const int2* ops;           // (machine, duration) per op j*m + k
  const int* beam_cur;       // P x S
  int* beam_next;            // K x S
  int* cbuf;                 // single-phase children, (K*B) x S
  unsigned long long* keys;  // K*B (also tie scratch)
  uint8_t* cjob;             // K*B: job of each child slot
  unsigned long long* bucketed;
  uint32_t* hist;            // kNB, this level
  uint32_t* fill;            // kNB, this level
  uint32_t* hist_next;       // kNB, zeroed for the next level
  uint32_t* fill_next;
  uint32_t* excl;            // kNB exclusive prefix of hist
  BlockStat* bstat;          // nblk_score partials
  SelState* sel;
  int* pcount;               // L+1 parent counts
  int* trail;                // (L-1) x K
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
  int level, best_known, nblk_score;
  uint32_t lseed;
  int bt_parent, bt_job;     // explicit backtrack start (exact modes), -1 = use bstat best
};

// mode: 1 ICH_DET, 2 PMS_GPU.  order: kOrderErase / kOrderSwap.
hipError_t launch_init(const LevelArgs& a, hipStream_t s);
hipError_t launch_score(const LevelArgs& a, int mode, bool single, hipStream_t s);
hipError_t launch_hist(const LevelArgs& a, int nblk, hipStream_t s);
hipError_t launch_partition(const LevelArgs& a, int nblk, hipStream_t s);
hipError_t launch_tie(const LevelArgs& a, hipStream_t s);
hipError_t launch_rank(const LevelArgs& a, int order, bool single, int nblk, hipStream_t s);
hipError_t launch_exact_expand(const LevelArgs& a, int P, hipStream_t s);
hipError_t launch_exact_materialize(const LevelArgs& a, int count, hipStream_t s);
hipError_t launch_backtrack(const LevelArgs& a, hipStream_t s);

}  // namespace jssp
