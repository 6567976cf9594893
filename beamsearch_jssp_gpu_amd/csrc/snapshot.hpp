// snapshot.hpp -- compact beam-state layout and the reference's hash/score
// formulas, shared by the device kernels and the host code.
//
// Compact snapshot of one partial schedule: S = 2n + m + 1 int32 words,
//   [0, m)          machine_ready[k]      finish time of the last op on machine k
//   [m, m+n)        job_ready[j]          finish time of the last scheduled op of job j
//   [m+n, m+2n)     job_word[j]           next_op (bits 0-7) | ready position (bits 8-15)
//                                         | machine of next_op (bits 16-23); position 0xFF
//                                         and machine 0 once the job is finished
//   [m+2n]          cmax
// The ready list of the reference (JobShopState::c, cpu/common/jssp_types.h:30;
// the precedence_ready block of gpu/common/gpu_snapshot.cuh:20-31) is kept as
// the position of each unfinished job in that list, so its order survives
// both the erase-then-append update (cpu/v0_ich/solver.cpp:29) and the
// swap-remove update (gpu/v1/solver.cu:101-106) while every other word of a
// child equals its parent's.  The reference snapshot also stores op_finish
// [nm+3] and seq[nm+3]; here the finish time of a job's previous op is
// job_ready[j] and the sequence is rebuilt from a per-level trail.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define JSSP_HD __host__ __device__ __forceinline__
#else
#define JSSP_HD inline
#endif

namespace jssp {

constexpr int kFinishedPos = 0xFF;
constexpr int kNoMakespan = 500000;  // cpu/common/jssp_types.h:33

// Branching order of the ready list.
enum ReadyOrder : int { kOrderErase = 0, kOrderSwap = 1 };

struct Layout {
  int n, m, S;
  int off_mr, off_jr, off_jw, off_cmax;
};

JSSP_HD Layout make_layout(int n, int m) {
  Layout l;
  l.n = n; l.m = m; l.S = 2 * n + m + 1;
  l.off_mr = 0; l.off_jr = m; l.off_jw = m + n; l.off_cmax = m + 2 * n;
  return l;
}

// Row stride (words) of the single-phase child buffer: S rounded up to whole
// 128-byte lines, so every child starts on a line and a writer that also
// fills the padding never leaves a line partly written.
JSSP_HD int child_stride(int S) { return (S + 31) & ~31; }

JSSP_HD int jw_make(int next_op, int pos, int mc) { return next_op | (pos << 8) | (mc << 16); }
JSSP_HD int jw_next(int w) { return w & 0xFF; }
JSSP_HD int jw_pos(int w) { return (w >> 8) & 0xFF; }
JSSP_HD int jw_mc(int w) { return (w >> 16) & 0xFF; }
JSSP_HD int jw_setpos(int w, int pos) { return (w & ~0xFF00) | (pos << 8); }

// Pivot hash, gpu/v1/solver.cu:51-53 (the lowbias32 mixer).
JSSP_HD uint32_t pivot_hash(uint32_t h) {
  h ^= h >> 16; h *= 0x7feb352dU;
  h ^= h >> 15; h *= 0x846ca68bU;
  h ^= h >> 16;
  return h;
}

// Sibling-noise hash, gpu/v1/solver.cu:122-125.
JSSP_HD uint32_t noise_hash(uint32_t seed, uint32_t cid, uint32_t bid) {
  uint32_t h = seed ^ (cid * 0x85ebca6bU) ^ (bid * 0xc2b2ae35U);
  h ^= h >> 16; h *= 0x85ebca6bU;
  h ^= h >> 16; h *= 0xc2b2ae35U;
  h ^= h >> 16;
  return h;
}

// FP32 score, gpu/v1/solver.cu:121-127: base = cmax / ns (ns >= 1 after a
// child is scheduled), plus (h % 1000) / 1e9.  Division is IEEE correctly
// rounded on both sides (no fast-math, no reciprocal).
JSSP_HD float gpu_score_base(int cmax, int ns) {
  return (ns == 0) ? (float)cmax : ((float)cmax / (float)ns);
}
JSSP_HD float gpu_score_add_noise(float base, uint32_t h) {
  float noise = (float)(h % 1000u) / 1000000000.0f;
  return base + noise;
}
JSSP_HD float gpu_score(int cmax, int ns, uint32_t h) {
  return gpu_score_add_noise(gpu_score_base(cmax, ns), h);
}

}  // namespace jssp
