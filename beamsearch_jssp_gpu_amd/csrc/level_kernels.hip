// level_kernels.hip -- per-level kernels of the beam step on CDNA4 (gfx950,
// wave64).
//
// One level = expand/score -> select top-K -> materialise.  The reference does
// this with kernel_expand_beam_layer / kernel_score_candidates (one block per
// parent, one thread per branch, 17 KB of private arrays per thread,
// gpu/v1/solver.cu:9-157, gpu/v2_twophases/solver.cu:9-147), a full CUB radix
// sort of every slot (gpu/v1/solver.cu:255) and a one-thread-per-survivor copy
// (gpu/v1/solver.cu:160-176, gpu/v2_twophases/solver.cu:150-254).  Here a level
// is three launches:
//   k_score   G lanes per parent (G = 16, 32 or 64); lane j holds job j, so the
//   k_expand  ready list, the pivot filter and the child slots come from
//             ballots/popcounts and every child's start/finish/cmax is computed
//             in registers.  Each kept child gets the 64-bit key (score bits <<
//             32 | child index), a 16-bit first selection digit and an 8-byte
//             record (finish, machine, job, next machine), and is counted in an
//             LDS histogram of its digit, merged into one of kHistCopies HBM
//             copies once per block.  k_expand (single-phase) also writes the
//             child's compact snapshot (lanes cover word offsets: coalesced).
//   k_filter  every block sums the histogram copies, finds the bin holding the
//             K-th key (wave prefix scans) and compacts the keys below it with
//             block-aggregated offsets.  When that bin holds one score value
//             (the usual case with the exact digit), its survivors are its
//             smallest child indices: each tile lists its keys of that bin in
//             index order, with a count, and no block waits for another.
//             Otherwise the bin's keys are compacted too and the last block to
//             finish cuts them exactly (score histograms, then index bitmap).
//   k_rank    each block stages the survivor keys in LDS; one wave per
//             survivor counts the smaller keys (its rank = its slot in the next
//             beam, i.e. exact (score bits, child index) order -- the order
//             CUB's stable sort gives the reference), or takes a tie's rank
//             from the tile prefix, and materialises the child from its parent
//             words and record (two-phase) or copies it from the child buffer
//             (single-phase).
// Every count stays on the device and every level counter is advanced by the
// next kernel of the step, so the host enqueues a fixed launch sequence with
// no readback and no device-scope fence, and the sequence is graph-capturable.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>

#include "level_kernels.hpp"
#include "snapshot.hpp"

namespace jssp {
namespace {

typedef unsigned long long u64;

// Global-memory (address space 1) view of a pointer.  The descriptors hold
// generic pointers; accessing through a typed view lets the compiler emit
// global_* instead of flat_* instructions (flat ones also wait on the LDS
// counter).
#define GL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ GL T* gp(T* p) {
  return (GL T*)p;
}

// Whole-struct copies to / from global memory through global pointers (a
// struct's implicit assignment does not take an address-space-qualified
// object), word by word; the compiler merges the words into wide accesses.
template <class T>
__device__ __forceinline__ void store_global(T* dst, const T& v) {
  static_assert(sizeof(T) % 4 == 0, "whole words");
  GL uint32_t* d = reinterpret_cast<GL uint32_t*>(gp(dst));
  const uint32_t* w = reinterpret_cast<const uint32_t*>(&v);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) d[i] = w[i];
}
template <class T>
__device__ __forceinline__ T load_global(const T* src) {
  static_assert(sizeof(T) % 4 == 0, "whole words");
  GL const uint32_t* s = reinterpret_cast<GL const uint32_t*>(gp(src));
  T v;
  uint32_t* w = reinterpret_cast<uint32_t*>(&v);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) w[i] = s[i];
  return v;
}

// Optional phase stamps, compiled in only with -DJSSP_PHASE_STAMPS (a
// separate build, see beamsearch_jssp_gpu_amd/phase_stamps.py): wall-clock
// (100 MHz) stamps at phase boundaries of the filter (thread 0 of each block)
// and the single-solve rank kernel (lane 0 of each wave), for levels
// [kStampLevel0, kStampLevel0 + kStampLevels) of a solve; read back with
// jssp_debug_phase_stamps.
#ifdef JSSP_PHASE_STAMPS
constexpr int kStampLevel0 = 100, kStampLevels = 8, kStampRankBlocks = 256, kStampFilterBlocks = 512;
__device__ unsigned long long g_rank_stamps[kStampLevels][kStampRankBlocks][16][8];
__device__ unsigned long long g_filter_stamps[kStampLevels][kStampFilterBlocks][4];
__device__ __forceinline__ bool stamp_level(int level) {
  return level >= kStampLevel0 && level < kStampLevel0 + kStampLevels;
}
#define RANK_STAMP(k)                                                                                   \
  do {                                                                                                  \
    if (W == 16 && stamp_level(a.level) && lane == 0 && local < kStampRankBlocks)                       \
      g_rank_stamps[a.level - kStampLevel0][local][wave][k] = (unsigned long long)wall_clock64();       \
  } while (0)
#define FILTER_STAMP(k)                                                                                 \
  do {                                                                                                  \
    if (stamp_level(a.level) && tid == 0 && local < kStampFilterBlocks)                                 \
      g_filter_stamps[a.level - kStampLevel0][local][k] = (unsigned long long)wall_clock64();           \
  } while (0)
#else
#define RANK_STAMP(k) ((void)0)
#define FILTER_STAMP(k) ((void)0)
#endif

// Wave-wide reductions and scans on DPP row shifts (a few cycles each) and
// lane reads, instead of LDS-routed lane permutes (an LDS round trip per
// step).  Every lane of the wave must be active: a row shift reads the
// neighbouring lanes' registers whatever their state.
//
// Lane i of row r (lanes 16 r .. 16 r + 15) takes the value of lane i - K of
// its row, or `fill` when that lane is outside the row.
template <int K>
__device__ __forceinline__ uint32_t dpp_row_shr(uint32_t v, uint32_t fill) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, 0x110 + K, 0xF, 0xF, false);
}
// Inclusive scan of op within each row of 16 lanes (id: op's identity):
// after the shifts by 1, 2, 4 and 8, lane i holds op over its row's lanes
// up to i (Hillis-Steele).
template <class Op>
__device__ __forceinline__ uint32_t row_scan(uint32_t v, uint32_t id, Op op) {
  v = op(v, dpp_row_shr<1>(v, id));
  v = op(v, dpp_row_shr<2>(v, id));
  v = op(v, dpp_row_shr<4>(v, id));
  v = op(v, dpp_row_shr<8>(v, id));
  return v;
}
// op over the whole wave from the row scans: the rows' last lanes.
template <class Op>
__device__ __forceinline__ uint32_t wave_reduce(uint32_t v, uint32_t id, Op op) {
  v = row_scan(v, id, op);
  const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 15);
  const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 31);
  const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 47);
  const uint32_t r3 = (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
  return op(op(r0, r1), op(r2, r3));
}
struct OpAdd { __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a + b; } };
struct OpMinU { __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return min(a, b); } };
struct OpMaxU { __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return max(a, b); } };
struct OpOr { __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a | b; } };

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) { return wave_reduce(v, 0u, OpAdd()); }
__device__ __forceinline__ uint32_t wave_min_u(uint32_t v) { return wave_reduce(v, 0xFFFFFFFFu, OpMinU()); }
__device__ __forceinline__ uint32_t wave_max(uint32_t v) { return wave_reduce(v, 0u, OpMaxU()); }
__device__ __forceinline__ uint32_t wave_or(uint32_t v) { return wave_reduce(v, 0u, OpOr()); }
// Signed minimum through the order-preserving map x ^ 0x80000000.
__device__ __forceinline__ int wave_min_i(int v) {
  return (int)(wave_min_u((uint32_t)v ^ 0x80000000u) ^ 0x80000000u);
}
// 64-bit minimum / maximum: the extreme high word, then the extreme low word
// among the lanes holding it.
__device__ __forceinline__ u64 wave_min64(u64 v) {
  const uint32_t hi = wave_min_u((uint32_t)(v >> 32));
  const uint32_t lo = wave_min_u((uint32_t)(v >> 32) == hi ? (uint32_t)v : 0xFFFFFFFFu);
  return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ u64 wave_max64(u64 v) {
  const uint32_t hi = wave_max((uint32_t)(v >> 32));
  const uint32_t lo = wave_max((uint32_t)(v >> 32) == hi ? (uint32_t)v : 0u);
  return ((u64)hi << 32) | lo;
}
// Inclusive prefix sum over the wave's lanes.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
  v = row_scan(v, 0u, OpAdd());
  const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 15);
  const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 31);
  const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 47);
  const int row = lane >> 4;
  return v + (row > 0 ? r0 : 0u) + (row > 1 ? r1 : 0u) + (row > 2 ? r2 : 0u);
}
// Block barrier ordering LDS accesses only.  __syncthreads() is also a
// workgroup-scope release of global memory, so it waits for every load and
// store the wave has in flight; this one waits for LDS operations alone, and
// global loads issued before it keep travelling (their values are waited for
// where they are used).  Only for barriers that protect LDS data.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Device-scope relaxed add on global memory, returning the old value (the
// same as atomicAdd, through a global pointer so no flat instruction holds
// up the wave's later LDS waits).
__device__ __forceinline__ uint32_t dev_add(GL uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Set bits of `mask` below this lane.
__device__ __forceinline__ uint32_t lanes_below(u64 mask, int lane) {
  return (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

// Streams src[0, cnt) through one block with 8 independent loads in flight
// per lane; f(key, valid) is called by every lane of every wave (valid is
// false past the end) so f may use wave-wide ballots.
template <class F>
__device__ __forceinline__ void for_keys(const u64* src, uint32_t cnt, F f) {
  constexpr int U = 8;
  const uint32_t bd = blockDim.x;
  for (uint32_t base = 0; base < cnt; base += bd * U) {
    u64 k[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = base + u * bd + threadIdx.x;
      k[u] = i < cnt ? src[i] : ~0ull;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) f(k[u], base + u * bd + threadIdx.x < cnt);
  }
}

// LDS histogram increment with one atomic per wave when all active lanes hit
// the same bin (the common case: runs of equal scores).
__device__ __forceinline__ void lds_count(uint32_t* hist, uint32_t b, bool in, int lane) {
  const u64 act = __ballot(in);
  if (!act) return;
  const int l0 = __ffsll((long long)act) - 1;
  const uint32_t b0 = (uint32_t)__builtin_amdgcn_readlane((int)b, l0);  // l0 is uniform
  if (__ballot(in && b == b0) == act) {
    if (lane == l0) atomicAdd(&hist[b0], (uint32_t)__popcll(act));
  } else if (in) {
    atomicAdd(&hist[b], 1u);
  }
}

// Wave-aggregated append: returns this lane's slot (valid where `take`).
__device__ __forceinline__ uint32_t wave_append(uint32_t* ctr, bool take, int lane) {
  const u64 m = __ballot(take);
  if (!m) return 0;
  const int l0 = __ffsll((long long)m) - 1;
  uint32_t base = 0;
  if (lane == l0) base = atomicAdd(ctr, (uint32_t)__popcll(m));
  base = (uint32_t)__builtin_amdgcn_readlane((int)base, l0);  // l0 is uniform
  return base + lanes_below(m, lane);
}

// First selection digit.  Monotone non-decreasing in the key whatever lo and
// s1 are, so the selection is exact for any parameters; lo and s1 only decide
// how finely the keys near the threshold are spread over the bins.
__device__ __forceinline__ uint32_t digit1(uint32_t hi, uint32_t lo, uint32_t s1) {
  if (hi < lo) return 0;
  const uint32_t d = (hi - lo) >> s1;
  return d < (uint32_t)kNB ? d : (uint32_t)(kNB - 1);
}

// Exact-digit key word of a child: its cmax above 11 bits of noise offset
// (the score's ulps above fl(cmax / ns)); and its first selection digit,
// 8 * min(cmax - lo, 256) + min(offset, 7), lo a lower bound of the level's
// cmax.  Monotone non-decreasing in the key word, and every bin other than
// the clamp bins (offset 7, or cmax - lo >= 256) holds one key word.
constexpr int kOffBits = 11;
__device__ __forceinline__ uint32_t exact_key(int cmax, uint32_t off) {
  return ((uint32_t)cmax << kOffBits) | off;
}
__device__ __forceinline__ uint32_t exact_digit(uint32_t kw, uint32_t lo) {
  const uint32_t dc = min((kw >> kOffBits) - lo, 256u);
  const uint32_t off = min(kw & ((1u << kOffBits) - 1u), 7u);
  return min(8u * dc + off, (uint32_t)(kNB - 1));
}

// What changes from a parent snapshot to one child.
struct ChildMod {
  int mc, f, j, cm, spos, jlast, newjw;
};

// Record of a kept child, written by the score kernel so that materialising
// a survivor needs only its parent's words: finish time of the scheduled op,
// its machine, the job, the machine of the job's following op.
__device__ __forceinline__ u64 pack_rec(int f, int mc, int j, int next_mc) {
  return (u64)(uint32_t)f | ((u64)(uint32_t)mc << 32) | ((u64)(uint32_t)j << 40) |
         ((u64)(uint32_t)next_mc << 48);
}
__device__ __forceinline__ int rec_finish(u64 r) { return (int)(uint32_t)r; }
__device__ __forceinline__ int rec_machine(u64 r) { return (int)((r >> 32) & 0xFFu); }
__device__ __forceinline__ int rec_job(u64 r) { return (int)((r >> 40) & 0xFFu); }
__device__ __forceinline__ int rec_next_machine(u64 r) { return (int)((r >> 48) & 0xFFu); }

// Machine of op k+1 of job j (0 past the last op).
template <class P>
__device__ __forceinline__ int next_machine(P ops, int j, int m, int k) {
  return k + 1 < m ? ops[j * m + k + 1].x : 0;
}

// Word `o` of the child built from parent word `pw`.  Erase order: the jobs
// behind the removed position move up one (cpu/v0_ich/solver.cpp:29).  Swap
// order: the last job takes the removed position (gpu/v1/solver.cu:101-106).
// Either way the scheduled job, if unfinished, is appended at position r-1
// (encoded in newjw).
template <int ORDER>
__device__ __forceinline__ int child_word(int o, int pw, const ChildMod& c, const Layout& L) {
  int v = pw;
  if (o < L.off_jr) {
    v = (o == c.mc) ? c.f : v;
  } else if (o < L.off_jw) {
    v = (o - L.off_jr == c.j) ? c.f : v;
  } else if (o < L.off_cmax) {
    int jj = o - L.off_jw;
    if (ORDER == kOrderErase) {
      int ps = jw_pos(v);
      v = (ps != kFinishedPos && ps > c.spos) ? v - 256 : v;
    } else {
      v = (jj == c.jlast) ? jw_setpos(v, c.spos) : v;
    }
    v = (jj == c.j) ? c.newjw : v;
  } else {
    v = c.cm;
  }
  return v;
}

// Words per lane of a snapshot (S <= 2*128 + 255 + 1 = 512 -> 8).
constexpr int kMaxWPL = 8;

// ---------------------------------------------------------------------------
__global__ void k_init(LevelArgs a) {
  const Layout Ly = make_layout(a.n, a.m);
  int* root = const_cast<int*>(a.beam_cur);
  for (int i = threadIdx.x; i < Ly.S; i += blockDim.x) {
    int v = 0;
    const int j = i - Ly.off_jw;
    if (i >= Ly.off_jw && i < Ly.off_cmax) v = jw_make(0, j, a.ops[j * a.m].x);  // ready = [0..n-1]
    root[i] = v;
  }
  for (int i = threadIdx.x; i < kNB * kHistCopies; i += blockDim.x) {
    a.hist[i] = 0;
    a.hist_next[i] = 0;
  }
  for (int i = threadIdx.x; i < kSureSubs; i += blockDim.x) {
    a.sctr[i * kCtrStride] = 0u;
    a.sctr_next[i * kCtrStride] = 0u;
  }
  if (threadIdx.x == 0) {
    SelState z = {};
    *a.sel = z;
    *a.sel_next = z;
    a.pcount[0] = 1;
    a.pcount[a.L] = 0;  // no beam after the last level (its filter returns at once)
    *a.expanded = 0ull;
    a.result[0] = a.best_known; a.result[1] = 0; a.result[2] = 0;
  }
}

// ---------------------------------------------------------------------------
// Scoring of one parent's children, shared by the score kernels, the root
// (init) and the rank kernel's fused scoring of the next level.

// Per-wave partials of a score pass.
struct ScoreAcc {
  uint32_t cnt = 0;   // kept children counted in the histogram
  uint32_t exp = 0;   // children evaluated
  int kmin = INT_MAX; // smallest kept cmax
  u64 best = ~0ull;   // last level: min (cmax << 32 | child)
};

// One parent as seen by its lane group: lane j of the group holds job j in
// slot h = 0 and job j + 64 in slot h = 1 (n > 64, G = 64 only).
struct JobView {
  bool rdy[2];
  int pos[2], mc[2], f[2], cm[2], nx[2], nmc[2];
  int r;  // ready jobs of the parent
};

// Where a score pass writes, and the per-level constants it needs.
// keys: the FP32 score bits of each child slot (the child index is the
// slot's position: the filter forms (bits << 32 | x) as it reads them), or
// kNoScore for a pruned child; pvalid[p]: the parent's valid slots (the
// slots past them are not written).
constexpr uint32_t kNoScore = 0xFFFFFFFFu;  // no finite positive float has these bits

struct ScoreDst {
  GL uint32_t* keys;
  GL uint8_t* pvalid;
  GL u64* crec;
  uint32_t lseed;
  int B, mode, exact_digit, best_known;
};
__device__ __forceinline__ ScoreDst score_dst(const LevelArgs& a) {
  ScoreDst d;
  d.keys = reinterpret_cast<GL uint32_t*>(gp(a.keys));
  d.pvalid = gp(a.pvalid);
  d.crec = gp(a.crec);
  d.lseed = a.lseed;
  d.B = a.B;
  d.mode = a.mode;
  d.exact_digit = a.exact_digit;
  d.best_known = a.best_known;
  return d;
}

// Digit and level parameters of a score pass.
struct ScoreLevel {
  uint32_t lo, s1;
  int ns;     // operations scheduled in a child (level + 1)
  bool last;  // last level: only the best complete child is recorded
};

// Word o (per lane) of a snapshot held by the wave as w[i] = word lane + 64 i.
// Every lane of the wave must execute it.
__device__ __forceinline__ int wave_fetch(const int (&w)[kMaxWPL], int NW, int o) {
  int v = 0;
#pragma unroll
  for (int i = 0; i < kMaxWPL; ++i) {
    if (i < NW) {
      const int t = __shfl(w[i], o & 63);
      v = ((o >> 6) == i) ? t : v;
    }
  }
  return v;
}

// The job view of a parent whose words the wave holds (G = 64); cmax = the
// parent's cmax.  Every lane of the wave must execute it.
__device__ __forceinline__ JobView view_from_words(GL const int2* ops, const Layout& Ly, const int (&w)[kMaxWPL],
                                                   int NW, int cmax, int lane) {
  JobView v;
  const int n = Ly.n, m = Ly.m;
  const int NH = (n + 63) >> 6;
  v.r = 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    v.rdy[h] = false; v.pos[h] = kFinishedPos; v.mc[h] = 0; v.f[h] = 0; v.cm[h] = 0; v.nx[h] = 0; v.nmc[h] = 0;
    if (h < NH) {  // wave-uniform
      const int j = h * 64 + lane;
      const bool in = j < n;
      const int jw = wave_fetch(w, NW, in ? Ly.off_jw + j : 0);
      // gpu_snapshot.cuh:104 -- with M == 1 the predecessor is op-1 of the previous job.
      const int pidx = (m == 1) ? (j > 0 ? Ly.off_jr + j - 1 : -1) : Ly.off_jr + j;
      int pred = wave_fetch(w, NW, (in && pidx >= 0) ? pidx : 0);
      pred = (in && pidx >= 0) ? pred : 0;
      const int mc = jw_mc(jw);
      const int mr = wave_fetch(w, NW, in ? Ly.off_mr + mc : 0);
      if (in) {
        v.pos[h] = jw_pos(jw);
        if (v.pos[h] != kFinishedPos) {
          v.rdy[h] = true;
          v.nx[h] = jw_next(jw);
          v.mc[h] = mc;
          const int dur = ops[j * m + v.nx[h]].y;
          v.nmc[h] = next_machine(ops, j, m, v.nx[h]);
          v.f[h] = max(mr, pred) + dur;
          v.cm[h] = max(cmax, v.f[h]);
        }
      }
      v.r += __popcll(__ballot(v.rdy[h]));
    }
  }
  return v;
}

// The job view of a parent whose words the wave has written to LDS at vw
// (G = 64); cmax = the parent's cmax.
// NHM: job slots per lane the caller supports (1 when n <= 64).
// ops: the op table in memory or in LDS.
template <int NHM, class P>
__device__ __forceinline__ JobView view_from_lds(P ops, const Layout& Ly, const int* vw, int cmax, int lane) {
  JobView v;
  const int n = Ly.n, m = Ly.m;
  const int NH = (n + 63) >> 6;
  v.r = 0;
  v.rdy[1] = false; v.pos[1] = kFinishedPos; v.mc[1] = 0; v.f[1] = 0; v.cm[1] = 0; v.nx[1] = 0; v.nmc[1] = 0;
#pragma unroll
  for (int h = 0; h < NHM; ++h) {
    v.rdy[h] = false; v.pos[h] = kFinishedPos; v.mc[h] = 0; v.f[h] = 0; v.cm[h] = 0; v.nx[h] = 0; v.nmc[h] = 0;
    if (h < NH) {  // wave-uniform
      const int j = h * 64 + lane;
      if (j < n) {
        const int jw = vw[Ly.off_jw + j];
        // gpu_snapshot.cuh:104 -- with M == 1 the predecessor is op-1 of the previous job.
        const int pred = (m == 1) ? (j > 0 ? vw[Ly.off_jr + j - 1] : 0) : vw[Ly.off_jr + j];
        v.pos[h] = jw_pos(jw);
        if (v.pos[h] != kFinishedPos) {
          v.rdy[h] = true;
          v.nx[h] = jw_next(jw);
          v.mc[h] = jw_mc(jw);
          const int dur = ops[j * m + v.nx[h]].y;
          v.nmc[h] = next_machine(ops, j, m, v.nx[h]);
          v.f[h] = max(vw[Ly.off_mr + v.mc[h]], pred) + dur;
          v.cm[h] = max(cmax, v.f[h]);
        }
      }
      v.r += __popcll(__ballot(v.rdy[h]));
    }
  }
  return v;
}

// Children of parent p (lane group of G lanes, gl = lane in the group, gmask
// = the group's lanes): branch slots (ICH_DET: ready position; PMS_GPU: rank
// among the hash-chosen pivot machine's ops, gpu/v1/solver.cu:47-70), pruning
// against best_known, the FP32 score and 64-bit key, the first selection
// digit, the child record, the LDS histogram; invalid slots get key ~0.  The
// whole wave executes it; lanes of an inactive parent (pact false) hold no
// ready job.  Fills slot / keep / nvalid for the single-phase writer.
// NHM: job slots per lane the caller supports (1 when n <= 64: the view's
// second slot is then empty).
template <int NHM = 2>
__device__ __forceinline__ void emit_children(const ScoreDst& a, const ScoreLevel& sl, int p, bool pact, int G,
                                              int gl, u64 gmask, int lane, JobView& v, uint32_t* s_h,
                                              ScoreAcc& acc, int (&slot)[2], bool (&keep)[2], int& nvalid) {
  const int B = a.B;
  bool val[2];
  slot[1] = 0;
  val[1] = false;
  keep[1] = false;
  if (a.mode == 1) {
#pragma unroll
    for (int h = 0; h < NHM; ++h) {
      slot[h] = v.pos[h];
      val[h] = v.rdy[h];
    }
    nvalid = v.r;
  } else {
    // Pivot: ready[hash % r] (gpu/v1/solver.cu:51-56); every lane reads its
    // own group's source lane.
    const uint32_t hsh = pivot_hash(a.lseed + (uint32_t)p);
    const int tpos = v.r > 0 ? (int)(hsh % (uint32_t)v.r) : -1;
    const u64 b0 = __ballot(v.rdy[0] && v.pos[0] == tpos) & gmask;
    // A whole-wave group reads the pivot lane's machine with a lane read
    // (the lane is uniform), smaller groups through a permute.
    const int t0 = G == 64 ? __builtin_amdgcn_readlane(v.mc[0], b0 ? __ffsll((long long)b0) - 1 : 0)
                           : __shfl(v.mc[0], b0 ? __ffsll((long long)b0) - 1 : lane);
    int tmc = b0 ? t0 : -1;
    if (NHM > 1) {
      const u64 b1 = __ballot(v.rdy[1] && v.pos[1] == tpos) & gmask;
      const int t1 = G == 64 ? __builtin_amdgcn_readlane(v.mc[1], b1 ? __ffsll((long long)b1) - 1 : 0)
                             : __shfl(v.mc[1], b1 ? __ffsll((long long)b1) - 1 : lane);
      tmc = b0 ? t0 : (b1 ? t1 : -1);
    }
#pragma unroll
    for (int h = 0; h < NHM; ++h) val[h] = v.rdy[h] && v.mc[h] == tmc;
    // Slot = rank of the ready position among the pivot machine's ops:
    // the group's candidate positions as a 128-bit mask.
    u64 plo = 0, phi = 0;
#pragma unroll
    for (int h = 0; h < NHM; ++h) {
      if (val[h]) {
        if (v.pos[h] < 64) plo |= 1ull << v.pos[h];
        else phi |= 1ull << (v.pos[h] - 64);
      }
    }
    if (G == 64) {
      plo = ((u64)wave_or((uint32_t)(plo >> 32)) << 32) | wave_or((uint32_t)plo);
      if (NHM > 1) phi = ((u64)wave_or((uint32_t)(phi >> 32)) << 32) | wave_or((uint32_t)phi);
    } else {
      for (int o = 1; o < G; o <<= 1) {
        plo |= __shfl_xor(plo, o);
        if (NHM > 1) phi |= __shfl_xor(phi, o);
      }
    }
#pragma unroll
    for (int h = 0; h < NHM; ++h) {
      const int ps = v.pos[h];
      slot[h] = ps < 64 ? __popcll(plo & ((1ull << ps) - 1ull))
                        : __popcll(plo) + __popcll(phi & ((1ull << (ps - 64)) - 1ull));
    }
    nvalid = __popcll(plo) + __popcll(phi);
  }
  if (pact && gl == 0) acc.exp += (uint32_t)nvalid;

#pragma unroll
  for (int h = 0; h < NHM; ++h) {
    val[h] = val[h] && slot[h] < B;  // always true for a consistent snapshot; keeps writes in bounds
    keep[h] = val[h] && v.cm[h] < a.best_known;
    uint32_t d = 0;
    bool counted = false;
    if (val[h]) {
      const uint32_t x = (uint32_t)(p * B + slot[h]);
      const int j = h * 64 + gl;
      if (keep[h]) {
        a.crec[x] = pack_rec(v.f[h], v.mc[h], j, v.nmc[h]);
        if (sl.last) {
          const u64 c = ((u64)(uint32_t)v.cm[h] << 32) | x;
          acc.best = c < acc.best ? c : acc.best;
        } else {
          const float base = gpu_score_base(v.cm[h], sl.ns);
          const float sc = gpu_score_add_noise(base, noise_hash(a.lseed, (uint32_t)p, (uint32_t)slot[h]));
          const uint32_t bits = __float_as_uint(sc);
          if (a.exact_digit) {
            // Key word (cmax << 11 | noise offset in ulps): order-isomorphic
            // to the score bits when a lower cmax always scores lower and the
            // offset stays below 2047 (both checked per instance on the host),
            // and it carries what the exact digit needs, so the filter reads
            // no digit array.
            const uint32_t kw = exact_key(v.cm[h], bits - __float_as_uint(base));
            a.keys[x] = kw;
            d = exact_digit(kw, sl.lo);
          } else {
            a.keys[x] = bits;
            d = digit1(bits, sl.lo, sl.s1);
          }
          counted = true;
          acc.cnt++;
          acc.kmin = min(acc.kmin, v.cm[h]);
        }
      } else if (!sl.last) {
        a.keys[x] = kNoScore;
      }
    }
    lds_count(s_h, d, counted, lane);
  }
  // Slots [nvalid, B) of the parent hold no child: the filter reads the
  // count instead of a key for each of them.
  if (!sl.last && pact && gl == 0) a.pvalid[p] = (uint8_t)nvalid;
}

// Block partials of a score pass: bstat[slot] and the LDS histogram merged
// into copy `slot % copies` of hist (copies: 1, 2 or 4).  Every thread of the
// block calls it.
__device__ __forceinline__ void finish_score_block(ScoreAcc acc, const uint32_t* s_h, BlockStat* bstat,
                                                   GL uint32_t* hist, int slot, bool last, int copies) {
  __shared__ uint32_t s_cnt[16], s_exp[16];
  __shared__ int s_kmin[16];
  __shared__ u64 s_best[16];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nwaves = (int)(blockDim.x >> 6);
  acc.cnt = wave_sum(acc.cnt);
  acc.exp = wave_sum(acc.exp);
  acc.kmin = wave_min_i(acc.kmin);
  acc.best = wave_min64(acc.best);
  if (lane == 0) { s_cnt[wave] = acc.cnt; s_exp[wave] = acc.exp; s_kmin[wave] = acc.kmin; s_best[wave] = acc.best; }
  lds_barrier();  // LDS partials and histogram only: the child stores keep draining
  if (threadIdx.x == 0) {
    BlockStat b;
    b.valid = 0; b.expanded = 0; b.min_cmax = INT_MAX; b.pad = 0; b.best = ~0ull;
    for (int w = 0; w < nwaves; ++w) {
      b.valid += s_cnt[w];
      b.expanded += s_exp[w];
      b.min_cmax = min(b.min_cmax, s_kmin[w]);
      b.best = s_best[w] < b.best ? s_best[w] : b.best;
    }
    store_global(bstat + slot, b);
  }
  if (!last) {
    GL uint32_t* hc = hist + (size_t)(slot & (copies - 1)) * kNB;
    for (int i = threadIdx.x; i < kNB; i += blockDim.x)
      if (s_h[i]) __hip_atomic_fetch_add(&hc[i], s_h[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Digit parameters of a level.  Exact digit: lo = a lower bound of every
// parent's cmax, the smallest kept child's cmax of the previous level
// (written by that level's filter; 0 for the root).  Generic digit: every
// child's score is >= fl(cmin / ns), cmin = the smallest parent cmax: cmax
// only grows, correctly rounded division is monotone, and adding the
// non-negative noise cannot round below the base; the shift spreads the
// previous level's survivors' score range over ~2^kSpreadBins bins (from the
// rank kernel's partials).  Every thread of the block calls it.
__device__ __forceinline__ ScoreLevel score_level(const LevelArgs& a, int local) {
  __shared__ int s_rmin[16];
  __shared__ uint32_t s_rmax[16], s_rlo[16];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nwaves = (int)(blockDim.x >> 6);
  ScoreLevel sl;
  sl.ns = a.level + 1;
  sl.last = (a.level == a.L - 1);
  sl.lo = 0;
  sl.s1 = kLevel0Shift;
  GL SelState* sel = gp(a.sel);
  if (a.exact_digit) {
    sl.lo = sel->lo;
  } else if (a.level > 0 && !sl.last) {
    int cmin = INT_MAX;
    uint32_t mxh = 0, mnh = 0xFFFFFFFFu;
    GL const RankStat* rstat = gp(a.rstat);
    for (int i = threadIdx.x; i < a.nblk_rank; i += blockDim.x) {
      cmin = min(cmin, rstat[i].min_cmax);
      mxh = max(mxh, rstat[i].max_hi);
      mnh = min(mnh, rstat[i].min_hi);
    }
    cmin = wave_min_i(cmin);
    mxh = wave_max(mxh);
    mnh = wave_min_u(mnh);
    if (lane == 0) { s_rmin[wave] = cmin; s_rmax[wave] = mxh; s_rlo[wave] = mnh; }
    __syncthreads();
    cmin = INT_MAX;
    mxh = 0;
    mnh = 0xFFFFFFFFu;
    for (int w = 0; w < nwaves; ++w) {
      cmin = min(cmin, s_rmin[w]);
      mxh = max(mxh, s_rmax[w]);
      mnh = min(mnh, s_rlo[w]);
    }
    if (cmin != INT_MAX) sl.lo = __float_as_uint((float)cmin / (float)sl.ns);
    const uint32_t spread = mxh > mnh ? mxh - mnh : 0u;
    const int bl = spread ? 32 - __clz(spread) : 0;
    sl.s1 = (uint32_t)max(0, bl - kSpreadBins);
  }
  if (local == 0 && threadIdx.x == 0) {
    if (!a.exact_digit) sel->lo = sl.lo;
    sel->s1 = sl.s1;
    sel->exact = a.exact_digit ? 1u : 0u;
  }
  return sl;
}

// Single-phase writer: every kept child of parent p (one wave, G = 64) as a
// full snapshot in cbuf at row p*B + slot (row stride child_stride(S): whole
// 128-byte lines; the padding words are written too, so no line is left
// partly written).  Each lane owns the words o = lane + 64 i; what a child
// changes in word o depends only on o's role, so the role is resolved once
// per parent into an id (machine o, 256 + job for job_ready, 512 + job for
// job_word, 1024 for cmax, 2048 for padding) and the ready position of a job
// word, and a child then costs a few compares and selects per word.  The
// kept children's fields are compacted into the wave's LDS rows (finish,
// cmax, new job word, slot | machine << 8 | position << 16 | job << 24) and
// read back kWriteUnroll at a time by every lane (broadcast reads), so
// several children's stores are in flight per wave.  (Loading the parent's
// words with its first loads instead, so that the first child store does
// not wait for them, measured no faster.)
// NWM: words per lane the caller supports (>= child_stride(S) / 64).
constexpr int kWriteUnroll = 4;

template <int NWM>
__device__ __forceinline__ void write_children(const LevelArgs& a, const Layout& Ly, int p, bool pact,
                                               GL const int* par, const JobView& v, const int (&slot)[2],
                                               const bool (&keep)[2], int lane, int4* rows) {
  const int S = Ly.S, B = a.B, m = Ly.m;
  const int CS = child_stride(S);
  const bool erase = (a.mode == 1);
  const int r = v.r;
  int jlast = -1;
  if (!erase) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const u64 bm = __ballot(v.rdy[h] && v.pos[h] == r - 1);
      if (bm) jlast = h * 64 + __ffsll((long long)bm) - 1;
    }
  }
  int pw[NWM], wid[NWM], wpos[NWM];
#pragma unroll
  for (int i = 0; i < NWM; ++i) {
    const int o = lane + 64 * i;
    pw[i] = (pact && o < S) ? par[o] : 0;
    int id = o < S ? 1024 : 2048;
    if (o < Ly.off_jr) id = o;
    else if (o < Ly.off_jw) id = 256 + (o - Ly.off_jr);
    else if (o < Ly.off_cmax) id = 512 + (o - Ly.off_jw);
    wid[i] = id;
    // erase order: jobs behind the removed position move up one
    const int ps = jw_pos(pw[i]);
    wpos[i] = (id >= 512 && id < 1024 && ps != kFinishedPos) ? ps : -1;
    // swap order: the last ready job takes the removed position
    if (!erase && id == 512 + jlast) wpos[i] = -2;
  }
  // Compact the kept children into the wave's rows.
  const u64 b0 = __ballot(keep[0]), b1 = __ballot(keep[1]);
  const int n0 = __popcll(b0), nk = n0 + __popcll(b1);
  const u64 below = lane ? (~0ull >> (64 - lane)) : 0ull;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (keep[h]) {
      const int rk = (h ? n0 : 0) + __popcll((h ? b1 : b0) & below);
      const int nx1 = v.nx[h] + 1;
      const int newjw = nx1 < m ? jw_make(nx1, r - 1, v.nmc[h]) : jw_make(nx1, kFinishedPos, 0);
      rows[rk] = make_int4(v.f[h], v.cm[h], newjw,
                           slot[h] | (v.mc[h] << 8) | (v.pos[h] << 16) | ((h * 64 + lane) << 24));
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  GL int* cbase = gp(a.cbuf) + (size_t)p * B * CS;
  for (int t = 0; t < nk; t += kWriteUnroll) {
    int4 c[kWriteUnroll];
#pragma unroll
    for (int u = 0; u < kWriteUnroll; ++u) c[u] = rows[min(t + u, nk - 1)];
#pragma unroll
    for (int u = 0; u < kWriteUnroll; ++u) {
      if (t + u < nk) {  // wave-uniform
        const int f = c[u].x, cm = c[u].y, newjw = c[u].z, misc = c[u].w;
        const int cs = misc & 0xFF, mc = (misc >> 8) & 0xFF, spos = (misc >> 16) & 0xFF;
        const int j = (int)((uint32_t)misc >> 24);
        const int idr = 256 + j, idw = 512 + j;
        GL int* out = cbase + (size_t)cs * CS;
#pragma unroll
        for (int i = 0; i < NWM; ++i) {
          const int o = lane + 64 * i;
          const int id = wid[i];
          int w = pw[i];
          w = (id == mc || id == idr) ? f : w;
          if (erase) w = (wpos[i] > spos) ? w - 256 : w;
          else w = (wpos[i] == -2) ? jw_setpos(w, spos) : w;
          w = (id == idw) ? newjw : w;
          w = (id == 1024) ? cm : w;
          if (o < CS) out[o] = w;
        }
      }
    }
  }
  // The parent words are used only inside the loop: a wave with no kept
  // child waits for them here, so no load is left in flight for the block's
  // tail to wait on (gfx9's one vector-memory counter also counts the
  // stores, so a wait there would hold the histogram flush behind every
  // child store).
#pragma unroll
  for (int i = 0; i < NWM; ++i) asm volatile("" ::"v"(pw[i]));
}

__device__ __forceinline__ void write_children_any(const LevelArgs& a, const Layout& Ly, int p, bool pact,
                                                   GL const int* par, const JobView& v, const int (&slot)[2],
                                                   const bool (&keep)[2], int lane, int4* rows) {
  const int CS = child_stride(Ly.S);
  if (CS <= 128) write_children<2>(a, Ly, p, pact, par, v, slot, keep, lane, rows);
  else if (CS <= 256) write_children<4>(a, Ly, p, pact, par, v, slot, keep, lane, rows);
  else write_children<kMaxWPL>(a, Ly, p, pact, par, v, slot, keep, lane, rows);
}

// Score (two-phase) or expand (single-phase) kernel body.  A group of G lanes
// (G = 16, 32 or 64, G >= n unless G = 64) handles one parent read from the
// beam in HBM; lane j of the group holds job j (and j + 64 when n > 64).
// Mode 1 = ICH_DET (every ready op, slot = ready position), mode 2 = PMS_GPU.
// A parent's cmax and, per lane, the job word and the predecessor's finish of
// its jobs gl and 64 + gl (lanes beyond n, or an absent parent: zeros).
struct ParentWords {
  int pc;
  int w[2], pred[2];
};

__device__ __forceinline__ ParentWords load_parent_words(const LevelArgs& a, const Layout& Ly, int p, bool ok,
                                                         int gl, int NH) {
  ParentWords r;
  GL const int* par = gp(a.beam_cur) + (size_t)(ok ? p : 0) * Ly.S;
  r.pc = ok ? par[Ly.off_cmax] : 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    r.w[h] = 0;
    r.pred[h] = 0;
    const int j = h * 64 + gl;
    if (ok && h < NH && j < Ly.n) {
      r.w[h] = par[Ly.off_jw + j];
      // gpu_snapshot.cuh:104 -- with M == 1 the predecessor is op-1 of the previous job.
      r.pred[h] = (Ly.m == 1) ? (j > 0 ? par[Ly.off_jr + j - 1] : 0) : par[Ly.off_jr + j];
    }
  }
  return r;
}

template <bool SINGLE>
__device__ void score_body(const LevelArgs& a, int local, int nblk) {
  __shared__ uint32_t s_h[kNB];
  // single-phase: per wave, the kept children of its parent (n <= 128)
  __shared__ int4 s_rows[SINGLE ? kScoreWaves * 128 : 1];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nwaves = (int)(blockDim.x >> 6);
  const Layout Ly = make_layout(a.n, a.m);
  const int n = a.n, m = a.m, S = Ly.S;
  const int G = a.G;
  const int lg = (G == 64) ? 6 : (G == 32 ? 5 : 4);
  const int PPW = 64 >> lg;
  const int sub = lane >> lg, gl = lane & (G - 1);
  const u64 gmask = (G == 64) ? ~0ull : (((1ull << G) - 1ull) << (sub * G));
  const int NH = (G == 64) ? ((n + 63) >> 6) : 1;
  const int pb0 = (local * nwaves + wave) * PPW;
  // The wave's first parent row is loaded before the level's parent count
  // arrives (every row below K is allocated; the words are used only when
  // the parent exists).
  ParentWords first = load_parent_words(a, Ly, pb0 + sub, pb0 + sub < a.K, gl, NH);
  const int P = gp(a.pcount)[a.level];
  const ScoreLevel sl = score_level(a, local);
  const ScoreDst dst = score_dst(a);
  GL const int2* ops = gp(a.ops);
  for (int i = threadIdx.x; i < kNB; i += blockDim.x) s_h[i] = 0;
  lds_barrier();  // the histogram (LDS); the parent loads stay in flight
  ScoreAcc acc;

  for (int pb = pb0; pb < P; pb += nblk * nwaves * PPW) {
    const int p = pb + sub;
    const bool pact = p < P;
    GL const int* par = gp(a.beam_cur) + (size_t)(pact ? p : 0) * S;
    const ParentWords pw = (pb == pb0) ? first : load_parent_words(a, Ly, p, pact, gl, NH);
    const int pc = pact ? pw.pc : 0;
    JobView v;
    v.r = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      v.rdy[h] = false; v.pos[h] = kFinishedPos; v.mc[h] = 0; v.f[h] = 0; v.cm[h] = 0; v.nx[h] = 0; v.nmc[h] = 0;
      const int j = h * 64 + gl;
      if (pact && h < NH && j < n) {
        const int w = pw.w[h];
        const int pred = pw.pred[h];
        v.pos[h] = jw_pos(w);
        if (v.pos[h] != kFinishedPos) {
          v.rdy[h] = true;
          const int nx = jw_next(w);
          v.nx[h] = nx;
          v.mc[h] = jw_mc(w);
          // Duration, the following op's machine and the machine's ready
          // time all depend on the job word only: one round of loads.
          const int dur = ops[j * m + nx].y;
          v.nmc[h] = next_machine(ops, j, m, nx);
          const int st = max(par[Ly.off_mr + v.mc[h]], pred);
          v.f[h] = st + dur;
          v.cm[h] = max(pc, v.f[h]);
        }
      }
      v.r += __popcll(__ballot(v.rdy[h]) & gmask);
    }
    int slot[2], nvalid;
    bool keep[2];
    emit_children(dst, sl, p, pact, G, gl, gmask, lane, v, s_h, acc, slot, keep, nvalid);
    if (SINGLE && !sl.last) write_children_any(a, Ly, p, pact, par, v, slot, keep, lane, s_rows + wave * 128);
  }
  // Loads whose values only some paths use are waited for here, so none is
  // left in flight for the block's tail (see write_children): the first
  // parent of a wave without a parent, the digit base of the last level.
  if (pb0 >= P) asm volatile("" ::"v"(first.pc), "v"(first.w[0]), "v"(first.w[1]), "v"(first.pred[0]), "v"(first.pred[1]));
  asm volatile("" ::"v"(sl.lo));
  finish_score_block(acc, s_h, a.bstat, gp(a.hist), local, sl.last, a.hist_copies);
}

// ---------------------------------------------------------------------------
// 8-bit MSD radix select over the full 64-bit key, by one block: keeps the
// `need` smallest of the t keys of R in R[0, need) (order unspecified).
__device__ __forceinline__ void refine_threshold(u64* R, uint32_t t, uint32_t need, u64* scratch) {
  __shared__ u64 s_mn[16], s_mx[16];
  __shared__ uint32_t s_h[256];
  __shared__ u64 s_want, s_mask;
  __shared__ uint32_t s_kth, s_cnt;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nw = (int)(blockDim.x >> 6);
  u64 mn = ~0ull, mx = 0;
  for (uint32_t i = tid; i < t; i += blockDim.x) {
    const u64 k = R[i];
    mn = k < mn ? k : mn;
    mx = k > mx ? k : mx;
  }
  mn = wave_min64(mn);
  mx = wave_max64(mx);
  if (lane == 0) { s_mn[wave] = mn; s_mx[wave] = mx; }
  __syncthreads();
  if (tid == 0) {
    u64 a = ~0ull, b = 0;
    for (int w = 0; w < nw; ++w) { a = s_mn[w] < a ? s_mn[w] : a; b = s_mx[w] > b ? s_mx[w] : b; }
    const int hb = 63 - __clzll((long long)(a ^ b));  // keys are distinct and t >= 2
    const int start = hb & ~7;
    const u64 mask = (start + 8 >= 64) ? 0ull : (~0ull << (start + 8));
    s_mask = mask;
    s_want = a & mask;
    s_kth = need - 1;
    s_cnt = (uint32_t)start;  // first digit position (reused below as a counter)
  }
  __syncthreads();
  const int start = (int)s_cnt;
  for (int shift = start; shift >= 0; shift -= 8) {
    for (int i = tid; i < 256; i += blockDim.x) s_h[i] = 0;
    __syncthreads();
    const u64 want = s_want, mask = s_mask;
    for (uint32_t i = tid; i < t; i += blockDim.x) {
      const u64 k = R[i];
      if ((k & mask) == want) atomicAdd(&s_h[(k >> shift) & 255], 1u);
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t kth = s_kth, cum = 0;
      int d = 0;
      while (d < 255 && cum + s_h[d] <= kth) { cum += s_h[d]; ++d; }
      s_kth = kth - cum;
      s_want = want | ((u64)d << shift);
      s_mask = mask | (255ull << shift);
    }
    __syncthreads();
  }
  const u64 thr = s_want;
  if (tid == 0) s_cnt = 0;
  __syncthreads();
  for (uint32_t i = tid; i < t; i += blockDim.x) {
    const u64 k = R[i];
    if (k <= thr) {
      const uint32_t at = atomicAdd(&s_cnt, 1u);
      if (at < need) scratch[at] = k;
    }
  }
  __threadfence_block();
  __syncthreads();
  for (uint32_t i = tid; i < need; i += blockDim.x) R[i] = scratch[i];
}

// Exact cut of the threshold bin, by one block.  Each round histograms the
// score bits of the remaining candidates into kSubBins bins spanning exactly
// their value range: the keys of the bins below the one holding the cut
// survive, the keys of that bin are the next round's candidates (the range
// shrinks 4096x per round, so after at most three rounds a bin holds a single
// score value).  Among the keys of the threshold score, the survivors are the
// smallest child indices: the need'-th set bit of an LDS bitmap over the child
// index.  Leaves the `need` smallest of the t keys of R in R[0, need).
constexpr int kSubBins = 4096;
constexpr int kLog2SubBins = 12;
__device__ __forceinline__ void refine_cut(u64* R, uint32_t t, uint32_t need, u64* bufA, u64* bufB, u64* outb,
                           uint32_t* bitmap, uint32_t bm_words, uint32_t N) {
  __shared__ uint32_t s_sub[kSubBins];
  __shared__ uint32_t s_mn[16], s_mx[16], s_w[16];
  __shared__ uint32_t s_b, s_needb, s_cntb, s_I, s_out, s_nxt;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nw = (int)(blockDim.x >> 6);
  const int per = kSubBins / (int)blockDim.x;
  if (tid == 0) s_out = 0;
  const u64* src = R;
  uint32_t cnt = t, nd = need;
  u64* nxt = bufA;
  for (int round = 0; round < 4; ++round) {
    uint32_t mn = 0xFFFFFFFFu, mx = 0;
    for_keys(src, cnt, [&](u64 k, bool in) {
      if (in) {
        const uint32_t h = (uint32_t)(k >> 32);
        mn = min(mn, h);
        mx = max(mx, h);
      }
    });
    mn = wave_min_u(mn);
    mx = wave_max(mx);
    if (lane == 0) { s_mn[wave] = mn; s_mx[wave] = mx; }
    for (int i = tid; i < kSubBins; i += blockDim.x) s_sub[i] = 0;
    __syncthreads();
    mn = 0xFFFFFFFFu; mx = 0;
    for (int w = 0; w < nw; ++w) { mn = min(mn, s_mn[w]); mx = max(mx, s_mx[w]); }
    const uint32_t range = mx - mn;
    const int bl = range ? 32 - __clz(range) : 0;
    const uint32_t sh = (uint32_t)max(0, bl - kLog2SubBins);
    for_keys(src, cnt, [&](u64 k, bool in) {
      const uint32_t b = in ? (((uint32_t)(k >> 32) - mn) >> sh) : 0u;
      lds_count(s_sub, b, in, lane);
    });
    __syncthreads();
    uint32_t loc = 0;
    for (int k = 0; k < per; ++k) loc += s_sub[tid * per + k];
    const uint32_t inc = wave_incl_scan(loc, lane);
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    uint32_t ex = inc - loc;
    for (int w = 0; w < wave; ++w) ex += s_w[w];
    for (int k = 0; k < per; ++k) {
      const uint32_t c = s_sub[tid * per + k];
      if (ex < nd && ex + c >= nd) { s_b = (uint32_t)(tid * per + k); s_needb = nd - ex; s_cntb = c; }
      ex += c;
    }
    __syncthreads();
    const uint32_t bstar = s_b, needb = s_needb, cntb = s_cntb;
    if (sh == 0 || needb == cntb) {
      // Final round: bins below bstar survive, and from bin bstar (one score
      // value when sh == 0) either all keys or the needb smallest indices.
      uint32_t I = 0xFFFFFFFFu;
      if (needb < cntb) {
        const uint32_t W = (N + 31) / 32;
        if (W > bm_words) {
          // No room for the bitmap: radix-select the nd smallest of src.
          u64* tmp = (src == bufA) ? bufB : bufA;
          refine_threshold(const_cast<u64*>(src), cnt, nd, tmp);
          const uint32_t o = s_out;
          for (uint32_t i = tid; i < nd; i += blockDim.x) outb[o + i] = src[i];
          __threadfence_block();
          __syncthreads();
          break;
        }
        const uint32_t V = mn + bstar;
        for (uint32_t i = tid; i < W; i += blockDim.x) bitmap[i] = 0;
        __syncthreads();
        for_keys(src, cnt, [&](u64 k, bool in) {
          if (in && (uint32_t)(k >> 32) == V) {
            const uint32_t x = (uint32_t)k;
            atomicOr(&bitmap[x >> 5], 1u << (x & 31));
          }
        });
        __syncthreads();
        const uint32_t wpt = (W + blockDim.x - 1) / blockDim.x;
        const uint32_t w0 = min(W, tid * wpt), w1 = min(W, w0 + wpt);
        uint32_t pc = 0;
        for (uint32_t w = w0; w < w1; ++w) pc += (uint32_t)__popc(bitmap[w]);
        const uint32_t pinc = wave_incl_scan(pc, lane);
        if (lane == 63) s_w[wave] = pinc;
        __syncthreads();
        uint32_t pex = pinc - pc;
        for (int w = 0; w < wave; ++w) pex += s_w[w];
        if (pex < needb && pex + pc >= needb) {
          uint32_t left = needb - pex;  // 1-based rank inside this thread's words
          for (uint32_t w = w0; w < w1; ++w) {
            uint32_t bits = bitmap[w];
            const uint32_t c = (uint32_t)__popc(bits);
            if (left > c) { left -= c; continue; }
            for (uint32_t q = 1; q < left; ++q) bits &= bits - 1;
            s_I = w * 32 + (uint32_t)(__ffs((int)bits) - 1);
            break;
          }
        }
        __syncthreads();
        I = s_I;
      }
      for_keys(src, cnt, [&](u64 k, bool in) {
        const uint32_t b = in ? ((uint32_t)(k >> 32) - mn) >> sh : 0u;
        const bool keep = in && (b < bstar || (b == bstar && (uint32_t)k <= I));
        const uint32_t at = wave_append(&s_out, keep, lane);
        if (keep) outb[at] = k;
      });
      __threadfence_block();
      __syncthreads();
      break;
    }
    // Narrow: keys of lower bins survive, keys of bin bstar go on.
    if (tid == 0) s_nxt = 0;
    __syncthreads();
    for_keys(src, cnt, [&](u64 k, bool in) {
      const uint32_t b = in ? ((uint32_t)(k >> 32) - mn) >> sh : 0u;
      const bool lower = in && b < bstar, same = in && b == bstar;
      const uint32_t at = wave_append(&s_out, lower, lane);
      if (lower) outb[at] = k;
      const uint32_t at2 = wave_append(&s_nxt, same, lane);
      if (same) nxt[at2] = k;
    });
    __threadfence_block();
    __syncthreads();
    src = nxt;
    cnt = cntb;
    nd = needb;
    nxt = (nxt == bufA) ? bufB : bufA;
  }
  for (uint32_t i = tid; i < need; i += blockDim.x) R[i] = outb[i];
}

// Threshold bin from the histogram, block-aggregated compaction of the keys
// below it (sure survivors): into cand[0, cum_lt) with one counter, or, when
// the threshold bin holds one score value (and K <= kRankLdsMax, B >= 8),
// into kSureSubs regions of cand, one counter each (block b takes counter
// b % kSureSubs), which the rank kernel gathers.  When the threshold bin
// holds one score value, its survivors are its `need` smallest child indices:
// each tile (one block pass of kFilterThreads children) writes its keys of
// that bin in index order at the tile's offset in ref_buf and their count in
// tcnt, and the rank kernel takes the first `need` of them tile by tile.
// Otherwise the bin's keys are compacted too and the last block to finish
// cuts them, so that cand[0, Kp) holds the Kp survivors.
// U: tiles a block pass covers (its threads hold one key of each): the
// pass's sure keys of all U tiles take one reservation, so a block with
// several tiles makes 1/U as many dependent round trips.  The single-solve
// form uses U = 1 (one tile per block at the benchmark sizes), the batch and
// server form U = kFilterTilesTp.
template <int U>
__device__ void filter_body(const LevelArgs& a, int local, int nblk) {
  __shared__ uint32_t s_ex[kNB];
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_e[16];
  __shared__ int s_km[16];
  __shared__ uint32_t s_T;
  __shared__ uint32_t s_off[2][U][16];
  __shared__ uint32_t s_base[2];
  __shared__ int s_last;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  GL SelState* sel = gp(a.sel);
  FILTER_STAMP(0);
  // The next level's histogram (read by the filter one level back) starts at
  // zero before its score pass, which follows this kernel.
  GL uint32_t* hist_next = gp(a.hist_next);
  const int copies = a.hist_copies;
  for (int i = local * (int)blockDim.x + tid; i < kNB * copies; i += nblk * (int)blockDim.x) hist_next[i] = 0;

  // Every load of the prologue goes out at once: the digit parameters, the
  // histogram copies, one score-block partial per thread (nblk_score <=
  // 512), the level's parents, and this block's first tile of key words
  // (read within the buffers' K*B capacity; slots past this level's count
  // are ignored below).  A slot x = p*B + s holds a key only below its
  // parent's valid count pvalid[p]; the first tile's keys are read
  // speculatively and masked.  A key's digit is computed from its key word.
  GL const uint32_t* keys = reinterpret_cast<GL const uint32_t*>(gp(a.keys));
  GL const uint8_t* pvalid = gp(a.pvalid);
  // (uniform: held in scalar registers, the throughput form is short of vector ones)
  const uint32_t dlo = __builtin_amdgcn_readfirstlane(sel->lo), ds1 = __builtin_amdgcn_readfirstlane(sel->s1);
  const uint32_t B = (uint32_t)a.B;
  const uint32_t cap = (uint32_t)a.K * B;
  constexpr uint32_t TS = kFilterThreads;  // slots per tile (blockDim.x)
  uint32_t c_key[U], c_pv[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const uint32_t xf = (uint32_t)local * (U * TS) + (uint32_t)u * TS + tid;
    c_key[u] = xf < cap ? keys[xf] : kNoScore;
    c_pv[u] = xf < cap ? (uint32_t)pvalid[xf / B] : 0u;
  }
  // (one branch on the copy count, so each path's loads go out together)
  uint32_t h0 = 0, h1 = 0;
  GL const u64* hsrc = reinterpret_cast<GL const u64*>(gp(a.hist)) + tid;
  auto sum_copies = [&](auto nc) {
    u64 hv[decltype(nc)::value];
#pragma unroll
    for (int c = 0; c < decltype(nc)::value; ++c) hv[c] = hsrc[(size_t)c * (kNB / 2)];
#pragma unroll
    for (int c = 0; c < decltype(nc)::value; ++c) {
      h0 += (uint32_t)hv[c];
      h1 += (uint32_t)(hv[c] >> 32);
    }
  };
  if (copies == kHistCopies) sum_copies(std::integral_constant<int, kHistCopies>());
  else if (copies == 2) sum_copies(std::integral_constant<int, 2>());
  else sum_copies(std::integral_constant<int, 1>());
  uint32_t ex = 0;
  int km = INT_MAX;
  if (tid < a.nblk_score) {
    ex = gp(a.bstat)[tid].expanded;
    km = gp(a.bstat)[tid].min_cmax;
  }
  const uint32_t N = (uint32_t)gp(a.pcount)[a.level] * (uint32_t)a.B;
  // Exclusive scan of the histogram (1024 threads x 2 bins); its total is the
  // number of kept children.
  const uint32_t sum = h0 + h1;
  const uint32_t inc = wave_incl_scan(sum, lane);
  ex = wave_sum(ex);
  km = wave_min_i(km);
  if (lane == 63) s_w[wave] = inc;
  if (lane == 0) { s_e[wave] = ex; s_km[wave] = km; }
  if (tid == 0) s_T = kNB - 1;
  lds_barrier();
  FILTER_STAMP(1);
  uint32_t wofs = 0, v = 0;
  for (int w = 0; w < 16; ++w) {
    wofs += (w < wave) ? s_w[w] : 0u;
    v += s_w[w];
  }
  const uint32_t Kp = min(v, (uint32_t)a.K);
  if (local == 0 && tid == 0) {
    ex = 0;
    km = INT_MAX;
    for (int w = 0; w < 16; ++w) { ex += s_e[w]; km = min(km, s_km[w]); }
    // Lower bound of the next level's parents' cmax (they are kept children).
    if (a.exact_digit && km != INT_MAX) gp(a.sel_next)->lo = (uint32_t)km;
    sel->V = v;
    sel->Kp = Kp;
    gp(a.pcount)[a.level + 1] = (int)Kp;
    *gp(a.expanded) += (u64)ex;
    gp(a.lvl_exp)[a.level] = ex;
  }
  if (Kp == 0) return;
  const bool exact = a.exact_digit != 0;
  const uint32_t s1 = exact ? 0u : ds1;
  const uint32_t ex0 = wofs + inc - sum;
  s_ex[2 * tid] = ex0;
  s_ex[2 * tid + 1] = ex0 + h0;
  if (ex0 < Kp && ex0 + h0 >= Kp) s_T = 2 * tid;
  if (ex0 + h0 < Kp && ex0 + h0 + h1 >= Kp) s_T = 2 * tid + 1;
  lds_barrier();
  FILTER_STAMP(2);
  const uint32_t T1 = s_T;
  const uint32_t cum_lt = s_ex[T1];
  const uint32_t t1 = (T1 + 1 < (uint32_t)kNB ? s_ex[T1 + 1] : v) - cum_lt;
  const uint32_t need = Kp - cum_lt;
  // A single score value in the threshold bin: exact digit whose offset is
  // not the clamped one, or generic digit at shift 0 below the clamp bin.
  const bool single = exact ? ((T1 & 7u) != 7u) : (s1 == 0 && T1 < (uint32_t)(kNB - 1));
  if (local == 0 && tid == 0) {
    sel->T1 = T1; sel->cum_lt = cum_lt; sel->t1 = t1; sel->need = need; sel->single = single ? 1u : 0u;
    sel->sorted_from = single ? cum_lt : Kp;
  }

  GL u64* cand = gp(a.cand);
  GL u64* tiles_keys = gp(a.ref_buf);
  const bool subs = sure_subs(single, a.K, a.B);
  const uint32_t sub = (uint32_t)local % kSureSubs;
  GL uint32_t* sctr = gp(a.sctr) + sub * kCtrStride;
  // A block with several passes (throughput shapes) loads each next pass's
  // key words and parent counts while it compacts the current one, so a
  // pass costs one round trip (its reservation atomic) instead of three; the
  // loads are speculative within the buffers' K*B capacity and masked when
  // used, as the first pass's (issued in the prologue).
  const uint32_t stride = (uint32_t)nblk * (U * TS);
  for (uint32_t base = (uint32_t)local * (U * TS); base < N; base += stride) {
    bool sure[U], tie[U];
    uint32_t kw[U], n_key[U], n_pv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t xn = base + stride + (uint32_t)u * TS + tid;
      n_key[u] = kNoScore;
      n_pv[u] = 0u;
      if (base + stride < N && xn < cap) {
        n_key[u] = keys[xn];
        n_pv[u] = (uint32_t)pvalid[xn / B];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t x = base + (uint32_t)u * TS + tid;
      // A slot holds a key only below its parent's valid count.
      const uint32_t px = x / B;
      const bool has = x < N && (x - px * B) < c_pv[u];
      kw[u] = has ? c_key[u] : kNoScore;
      c_key[u] = n_key[u];
      c_pv[u] = n_pv[u];
      const uint32_t d = kw[u] == kNoScore ? (uint32_t)kNB : (exact ? exact_digit(kw[u], dlo) : digit1(kw[u], dlo, ds1));
      sure[u] = d < T1;
      tie[u] = d == T1;
      const u64 bs = __ballot(sure[u]), bt = __ballot(tie[u]);
      if (lane == 0) {
        s_off[0][u][wave] = (uint32_t)__popcll(bs);
        s_off[1][u][wave] = (uint32_t)__popcll(bt);
      }
    }
    lds_barrier();
    // One reservation for the pass's sure keys (tile by tile, wave by wave);
    // the ties per tile (single-valued bin: at the tile's own offset, count in
    // tcnt) or after the sure keys with one reservation.
    if constexpr (U == 1) {
      if (tid == 0) {  // thread 0 alone: 16 counts
        uint32_t ts = 0, tt = 0;
        for (int w = 0; w < 16; ++w) {
          const uint32_t cs = s_off[0][0][w], ct = s_off[1][0][w];
          s_off[0][0][w] = ts;
          s_off[1][0][w] = tt;
          ts += cs;
          tt += ct;
        }
        if (subs) s_base[0] = ts ? sub * (uint32_t)a.K + dev_add(sctr, ts) : 0u;
        else s_base[0] = ts ? dev_add(&sel->sure_ctr, ts) : 0u;
        if (single) gp(a.tcnt)[base / TS] = tt;
        else s_base[1] = tt ? cum_lt + dev_add(&sel->tie_ctr, tt) : 0u;
      }
    } else if (wave == 0) {
      // A scan over the (tile, wave) counts, one lane each.
      static_assert(16 * U <= 64, "one lane per (tile, wave)");
      const int u = lane >> 4, w = lane & 15;
      const bool in = u < U;
      const uint32_t cs = in ? s_off[0][in ? u : 0][w] : 0u, ct = in ? s_off[1][in ? u : 0][w] : 0u;
      const uint32_t is = wave_incl_scan(cs, lane), it = wave_incl_scan(ct, lane);
      const uint32_t ts = (uint32_t)__builtin_amdgcn_readlane((int)is, 63);
      const uint32_t tt = (uint32_t)__builtin_amdgcn_readlane((int)it, 63);
      const uint32_t seg = (uint32_t)__shfl((int)it, max(0, 16 * u - 1));
      const uint32_t seg0 = u > 0 ? seg : 0u;  // ties of the pass's tiles before this one
      if (in) {
        s_off[0][u][w] = is - cs;
        s_off[1][u][w] = single ? it - ct - seg0 : it - ct;
        if (single && w == 15 && base + (uint32_t)u * TS < N) gp(a.tcnt)[(base + (uint32_t)u * TS) / TS] = it - seg0;
      }
      if (lane == 0) {
        if (subs) s_base[0] = ts ? sub * (uint32_t)a.K + dev_add(sctr, ts) : 0u;
        else s_base[0] = ts ? dev_add(&sel->sure_ctr, ts) : 0u;
        if (!single) s_base[1] = tt ? cum_lt + dev_add(&sel->tie_ctr, tt) : 0u;
      }
    }
    lds_barrier();
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t x = base + (uint32_t)u * TS + tid;
      const u64 key = ((u64)kw[u] << 32) | x;
      const u64 bs = __ballot(sure[u]), bt = __ballot(tie[u]);
      if (sure[u]) cand[s_base[0] + s_off[0][u][wave] + lanes_below(bs, lane)] = key;
      if (tie[u]) {
        // Single-valued bin: the tile's ties in child-index order at the
        // tile's own offset (no atomics); otherwise compacted after the sure keys.
        const uint32_t o = s_off[1][u][wave] + lanes_below(bt, lane);
        if (single) tiles_keys[base + (uint32_t)u * TS + o] = key;
        else cand[s_base[1] + o] = key;
      }
    }
    lds_barrier();
  }
  FILTER_STAMP(3);
  if (single) return;  // the rank kernel takes the ties tile by tile

  // The last block to finish sees every block's candidates and cuts the
  // threshold bin's keys cand[cum_lt, cum_lt + t1) to the `need` survivors.
  // Every thread's candidate stores complete before the block's ticket (the
  // tile loop's barriers order LDS only).
  __syncthreads();
  if (tid == 0) {
    __threadfence();
    const uint32_t ticket = dev_add(&sel->done_ctr, 1u);
    s_last = (ticket == (uint32_t)nblk - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  if (need < t1) {
    extern __shared__ uint32_t s_bitmap[];
    refine_cut(a.cand + cum_lt, t1, need, a.keys, a.ref_buf, a.ref_buf + N, s_bitmap, a.bm_words, N);
  }
}

// ---------------------------------------------------------------------------
// Word o of a snapshot whose words the wave holds as pw[i] = word lane + 64 i
// (o wave-uniform).
template <int NWM = kMaxWPL>
__device__ __forceinline__ int wave_word(const int (&pw)[kMaxWPL], int o) {
  o = __builtin_amdgcn_readfirstlane(o);
  const int i = o >> 6;
  int v = 0;
#pragma unroll
  for (int k = 0; k < NWM; ++k) v = (k == i) ? pw[k] : v;
  return __builtin_amdgcn_readlane(v, o & 63);
}

// Loads parent p's words into pw[i] = word lane + 64 i (zeros past S).
// NWM: words per lane the caller supports (S <= 64 * NWM).
template <int NWM = kMaxWPL>
__device__ __forceinline__ void load_parent(const LevelArgs& a, const Layout& Ly, int lane, int p,
                                            int (&pw)[kMaxWPL]) {
  const int S = Ly.S;
  const int NW = (S + 63) >> 6;
  GL const int* par = gp(a.beam_cur) + (size_t)p * S;
#pragma unroll
  for (int i = 0; i < kMaxWPL; ++i) {
    const int o = lane + 64 * i;
    pw[i] = (i < NWM && i < NW && o < S) ? par[o] : 0;
  }
}

// Writes survivor q's words nw (nw[i] = word lane + 64 i) to the next beam
// and its trail entry (parent p, job j).  One wave.
template <int NWM = kMaxWPL>
__device__ __forceinline__ void store_child(const LevelArgs& a, const Layout& Ly, int lane, int p, int j, int q,
                                            const int (&nw)[kMaxWPL]) {
  const int S = Ly.S;
  const int NW = (S + 63) >> 6;
  GL int* out = gp(a.beam_next) + (size_t)q * S;
#pragma unroll
  for (int i = 0; i < kMaxWPL; ++i) {
    const int o = lane + 64 * i;
    if (i < NWM && i < NW && o < S) out[o] = nw[i];
  }
  if (lane == 0) gp(a.trail)[(size_t)a.level * a.K + q] = (p << 8) | j;
}

// Builds survivor q of the next beam from its parent p's words pw and the
// child's record (two-phase), writes it and the trail entry.  One wave.
// Returns the child's cmax; nw[i] = the child's word lane + 64 i.  (Storing
// the words after the next level's op-table loads are issued, so that the
// wait for those loads does not also wait for the stores, measured no
// faster for single solves and slower for the server.)
template <int ORDER, int NWM = kMaxWPL>
__device__ __forceinline__ int build_child(const LevelArgs& a, const Layout& Ly, int lane, int p, u64 rec, int q,
                                           const int (&pw)[kMaxWPL], int (&nw)[kMaxWPL]) {
  const int S = Ly.S, m = a.m;
  const int NW = (S + 63) >> 6;
  int r = 0;
  int jlast = -1;
#pragma unroll
  for (int i = 0; i < NWM; ++i) {
    const int o = lane + 64 * i;
    const bool isjw = (i < NW) && o >= Ly.off_jw && o < Ly.off_cmax;
    r += __popcll(__ballot(isjw && jw_pos(pw[i]) != kFinishedPos));
  }
  if (ORDER == kOrderSwap) {
#pragma unroll
    for (int i = 0; i < NWM; ++i) {
      const int o = lane + 64 * i;
      const bool isjw = (i < NW) && o >= Ly.off_jw && o < Ly.off_cmax;
      const u64 bm = __ballot(isjw && jw_pos(pw[i]) == r - 1);
      if (bm) jlast = 64 * i + (__ffsll((long long)bm) - 1) - Ly.off_jw;
    }
  }
  const int j = rec_job(rec);
  const int jwj = wave_word<NWM>(pw, Ly.off_jw + j);
  const int nx = jw_next(jwj);
  ChildMod c;
  c.j = j;
  c.mc = rec_machine(rec);
  c.f = rec_finish(rec);
  c.cm = max(wave_word<NWM>(pw, Ly.off_cmax), c.f);
  c.spos = jw_pos(jwj);
  c.jlast = jlast;
  c.newjw = nx + 1 < m ? jw_make(nx + 1, r - 1, rec_next_machine(rec)) : jw_make(nx + 1, kFinishedPos, 0);
#pragma unroll
  for (int i = 0; i < kMaxWPL; ++i) {
    if (i < NWM && i < NW) {
      const int o = lane + 64 * i;
      nw[i] = child_word<ORDER>(o, pw[i], c, Ly);
    } else {
      nw[i] = 0;
    }
  }
  store_child<NWM>(a, Ly, lane, p, j, q, nw);
  return c.cm;
}

// Builds survivor q of the next beam from its parent p and the child's record
// (two-phase) and writes the trail entry.  One wave.  Returns the child's cmax;
// nw[i] = the child's word lane + 64 i.
// NWM: words per lane the caller supports (S <= 64 * NWM).
template <int ORDER, int NWM = kMaxWPL>
__device__ __forceinline__ int materialize_rec(const LevelArgs& a, const Layout& Ly, int lane, int p, u64 rec,
                                               int q, int (&nw)[kMaxWPL]) {
  int pw[kMaxWPL];
  load_parent<NWM>(a, Ly, lane, p, pw);
  return build_child<ORDER, NWM>(a, Ly, lane, p, rec, q, pw, nw);
}

// Key of the r-th (0-based) tie in child-index order: tile c holds tcnt[c]
// ties at ref_buf[c * kFilterThreads ...], tp[0..tiles] is the exclusive
// prefix of those counts.  Wave-uniform.
__device__ __forceinline__ u64 tie_key(GL const u64* tiles_keys, const uint32_t* tp, int tiles, uint32_t r) {
  int lo = 0, hi = tiles - 1;  // the last tile c with tp[c] <= r holds it
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tp[mid] <= r) lo = mid;
    else hi = mid - 1;
  }
  return tiles_keys[(size_t)lo * kFilterThreads + (r - tp[lo])];
}

// Number of the n keys of C below key, one wave; eight loads in flight per
// step (C is LDS or global memory, known after inlining).
template <class P>
__device__ __forceinline__ uint32_t count_below(P C, int n, u64 key, int lane) {
  uint32_t c = 0;
  int base = 0;
  for (; base + 512 <= n; base += 512) {
    u64 v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = C[base + 64 * k + lane];
#pragma unroll
    for (int k = 0; k < 8; ++k) c += (uint32_t)__popcll(__ballot(v[k] < key));
  }
  for (; base < n; base += 64) {
    const int i = base + lane;
    c += (uint32_t)__popcll(__ballot(i < n && C[i] < key));
  }
  return c;
}

// The same for NB wave-uniform keys at once (LDS keys): every LDS read is
// compared against all NB keys, so a wave ranking several survivors reads
// the staged keys once per NB of them.
template <int NB>
__device__ __forceinline__ void count_below_n(const u64* C, int n, const u64 (&key)[NB], uint32_t (&c)[NB],
                                              int lane) {
  u64 k[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)key[j]);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(key[j] >> 32));
    k[j] = ((u64)hi << 32) | lo;
    c[j] = 0;
  }
  int base = 0;
  for (; base + 512 <= n; base += 512) {
    u64 v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = C[base + 64 * t + lane];
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int j = 0; j < NB; ++j) c[j] += (uint32_t)__popcll(__ballot(v[t] < k[j]));
  }
  for (; base < n; base += 64) {
    const int i = base + lane;
    const u64 v = i < n ? C[i] : ~0ull;
#pragma unroll
    for (int j = 0; j < NB; ++j) c[j] += (uint32_t)__popcll(__ballot(i < n && v < k[j]));
  }
}

// Rank of every survivor among the Kp survivors, then materialise it
// (two-phase) or copy it from the child buffer (single-phase) into slot q of
// the next beam.  One wave per survivor.  The sf keys below the threshold
// bin (cand[0, sf), or the filter's kSureSubs regions gathered in order) are
// staged in LDS and each survivor's rank is the count of smaller keys (exact
// (score bits, child index) order -- the order CUB's stable sort gives the
// reference).
// Survivors u >= sf are the threshold bin's ties, which sort by child index
// after every key of cand: survivor u is the (u - sf)-th tie, at rank u.
// MODE: 0 = the instance's mode at run time (separate score kernel), 1 =
// ICH_DET, 2 = PMS_GPU (fused kernels are compiled per mode).  Fused: after
// materialising a survivor, the wave writes its words to its LDS view and
// scores its children for the next level.
// SZ (fused only; the snapshot size class, rank_size_class): 0 = S <= 128
// and n <= 64, so a snapshot is two words per lane and a job one slot per
// lane; 1 = S <= 256 (four words per lane); 2 = larger.  W: waves per block
// (kRankWaves for single solves, kRankWavesThroughput for batches and the
// server).
template <bool FUSED, int MODE, int SZ, int W, int NB, bool WIDE = false>
__device__ void rank_body(const LevelArgs& a, u64* crec_next, int local, int nblk) {
  constexpr int NWM = SZ == 0 ? 2 : (SZ == 1 ? 4 : kMaxWPL);
  constexpr int NHM = SZ == 0 ? 1 : 2;
  extern __shared__ u64 s_c[];
  __shared__ uint32_t s_h[kNB];
  __shared__ int s_mc[kRankWaves];
  __shared__ uint32_t s_mh[kRankWaves], s_ml[kRankWaves], s_w[kRankWaves];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tid = threadIdx.x;
  RANK_STAMP(0);
  if (FUSED)
    for (int i = tid; i < kNB; i += blockDim.x) s_h[i] = 0;
  if (local == 0 && tid == 0) {
    GL SelState* sn = gp(a.sel_next);
    sn->sure_ctr = 0;
    sn->tie_ctr = 0;
    sn->done_ctr = 0;
    for (int r = 0; r < kSureSubs; ++r) gp(a.sctr_next)[r * kCtrStride] = 0u;
    if (a.sel_log) store_global(a.sel_log + a.level, load_global(a.sel));
  }
  // The selection state, this thread's first tie count and the survivor keys
  // are loaded together; the keys go to LDS whatever the survivor count
  // (cand holds K*B entries, and entries past the survivors are never read).
  const bool in_lds = !WIDE;
  const int tiles_max = (int)(((uint32_t)a.K * (uint32_t)a.B + kFilterThreads - 1) / kFilterThreads);
  const int per = (tiles_max + (int)blockDim.x - 1) / (int)blockDim.x;  // tiles per thread (<= 8)
  const int t0 = min(tiles_max, tid * per);
  GL const u64* cand = gp(a.cand);
  GL const uint32_t* tcnt = gp(a.tcnt);
  GL const SelState* sel = gp(a.sel);
  const uint32_t kp_raw = sel->Kp, sf_raw = sel->sorted_from, single_raw = sel->single;
  // The op table, when staged, is loaded with this first round (and written
  // to LDS after the keys' loads are out).
  const bool ops_lds = FUSED && W == kRankWavesThroughput && rank_ops_in_lds(a.K, a.n, a.m, true, W);
  constexpr int kOpsPer = (kOpsLdsMax + 64 * W - 1) / (64 * W);
  int2 opv[kOpsPer];
#pragma unroll
  for (int k = 0; k < kOpsPer; ++k) {
    const int i = tid + k * 64 * W;
    opv[k] = (ops_lds && i < a.n * a.m) ? load_global(a.ops + i) : make_int2(0, 0);
  }
  const uint32_t N = (uint32_t)gp(a.pcount)[a.level] * (uint32_t)a.B;
  const uint32_t tc0 = t0 < tiles_max ? tcnt[t0] : 0u;
  // The keys below the threshold bin: one run per counter region when the
  // filter spread them (their counts add up to sf), else cand[0, sf).
  uint32_t sc[kSureSubs];
  GL const uint32_t* sctr = gp(a.sctr);
#pragma unroll
  for (int r = 0; r < kSureSubs; ++r) sc[r] = sctr[r * kCtrStride];
  const int Kp = (int)kp_raw;
  const int sf = min((int)sf_raw, Kp);
  if (in_lds) {
    if (sure_subs(single_raw != 0, a.K, a.B)) {
      // Entry i of the gathered run lies in region r (the last r whose runs
      // before it hold <= i keys), at cand[r*K + i - (keys of regions < r)]:
      // i plus (K - c[q-1]) for every region boundary q at or below i.
      uint32_t bound[kSureSubs], skip[kSureSubs];
      uint32_t run = 0;
#pragma unroll
      for (int q = 1; q < kSureSubs; ++q) {
        const uint32_t c = min(sc[q - 1], (uint32_t)a.K);
        run += c;
        bound[q] = run;
        skip[q] = (uint32_t)a.K - c;
      }
#pragma unroll 4
      for (int i = tid; i < sf; i += blockDim.x) {
        uint32_t src = (uint32_t)i;
#pragma unroll
        for (int q = 1; q < kSureSubs; ++q) src += (uint32_t)i >= bound[q] ? skip[q] : 0u;
        s_c[i] = cand[src];
      }
    } else {
#pragma unroll 4
      for (int i = tid; i < a.K; i += blockDim.x) s_c[i] = cand[i];
    }
  }
  const bool ties = sf < Kp;
  uint32_t* s_tp = reinterpret_cast<uint32_t*>(s_c + (in_lds ? a.K : 0));
  const int tiles = (int)((N + kFilterThreads - 1) / kFilterThreads);
  // Fused: per-wave snapshot views after the tile prefix (16-byte aligned),
  // then, where staged, the op table (8-byte aligned: W * S words, W = 8).
  int* s_view = reinterpret_cast<int*>(s_tp + ((tiles_max + 1 + 3) & ~3));
  int2* s_ops = reinterpret_cast<int2*>(s_view + W * (2 * a.n + a.m + 1));
  if (ops_lds) {
#pragma unroll
    for (int k = 0; k < kOpsPer; ++k) {
      const int i = tid + k * 64 * W;
      if (i < a.n * a.m) s_ops[i] = opv[k];
    }
  }
  if (ties) {
    // Exclusive prefix of the per-tile tie counts (tiles past this level's
    // count hold stale counts and are skipped).
    const int t1 = min(tiles, t0 + per);
    uint32_t loc = t0 < tiles ? tc0 : 0u;
    for (int t = t0 + 1; t < t1; ++t) loc += tcnt[t];
    const uint32_t inc = wave_incl_scan(loc, lane);
    if (lane == 63) s_w[wave] = inc;
    lds_barrier();
    uint32_t run = inc - loc;
    for (int w = 0; w < wave; ++w) run += s_w[w];
    for (int t = t0; t < t1; ++t) {
      s_tp[t] = run;
      run += (t == t0) ? tc0 : tcnt[t];
    }
    if (tid == (int)blockDim.x - 1) s_tp[tiles] = run;
  }
  lds_barrier();  // staged keys and tile prefix (LDS)
  RANK_STAMP(1);
  const Layout Ly = make_layout(a.n, a.m);
  const int S = Ly.S;
  const int NW = (S + 63) >> 6;
  // Fused: the next level's scoring parameters (exact digit: lo from this
  // level's filter).
  ScoreLevel sln;
  sln.lo = FUSED ? gp(a.sel_next)->lo : 0u;
  sln.s1 = kLevel0Shift;
  sln.ns = a.level + 2;
  sln.last = (a.level + 1 == a.L - 1);
  ScoreDst dn = score_dst(a);  // the next level's keys, digits, records
  dn.crec = gp(crec_next);
  dn.lseed = a.lseed + 1u;
  if (MODE != 0) dn.mode = MODE;
  GL const int2* ops = gp(a.ops);
  GL const u64* crec = gp(a.crec);
  ScoreAcc acc;
  int w_min = INT_MAX;
  uint32_t w_max = 0, w_minh = 0xFFFFFFFFu;
  const int nwaves = W;
  const bool single = !FUSED && a.cbuf != nullptr;
  const int order = MODE == 1 ? kOrderErase : (MODE == 2 ? kOrderSwap : ((a.mode == 1) ? kOrderErase : kOrderSwap));
  // A wave takes survivors u0, u0 + stride, u0 + 2*stride, ...; with NB > 1
  // (batches), when it holds two or more of the sorted ones, NB of them at a
  // time are ranked together from LDS, then materialised one by one in that
  // order.
  const int stride = nblk * nwaves;
  for (int ub = local * nwaves + wave; ub < Kp; ub += NB * stride) {
    u64 bkey[NB] = {};
    uint32_t bcnt[NB] = {};
    const bool bat = NB > 1 && in_lds && ub + stride < sf;
    if constexpr (NB > 1) {
      if (bat) {
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          const int u = ub + j * stride;
          bkey[j] = u < sf ? s_c[u] : ~0ull;
        }
        count_below_n<NB>(s_c, sf, bkey, bcnt, lane);
      }
    }
  for (int j = 0; j < NB; ++j) {
    const int u = ub + j * stride;
    if (u >= Kp) break;
    uint32_t cnt = (uint32_t)u;
    u64 key;
    bool count = false;  // rank by counting the smaller keys (below)
    if (u < sf) {
      if (bat) {
        key = bkey[0];
        cnt = bcnt[0];
      } else {
        key = in_lds ? s_c[u] : cand[u];
        count = true;
      }
    } else {
      key = tie_key(gp(a.ref_buf), s_tp, tiles, (uint32_t)(u - sf));
    }
    if constexpr (NB > 1) {
#pragma unroll
      for (int t = 0; t + 1 < NB; ++t) {
        bkey[t] = bkey[t + 1];
        bcnt[t] = bcnt[t + 1];
      }
    }
    const uint32_t x = (uint32_t)key;
    const uint32_t hi = (uint32_t)(key >> 32);
    const int p = (int)(x / (uint32_t)a.B);
    // Single solves request the child's record and its parent's words
    // before the count, which then runs (from LDS) while they travel; the
    // register-limited throughput blocks (batches, the server) cannot hold
    // the words across the count without spilling, so they load after it.
    constexpr bool kEarly = (W == kRankWaves && NB == 1);
    u64 rec = 0;
    int pw[kMaxWPL];
    if (kEarly) {
      rec = crec[x];
      if (!single) load_parent<NWM>(a, Ly, lane, p, pw);
    }
    if (count) cnt = in_lds ? count_below(s_c, sf, key, lane) : count_below(cand, sf, key, lane);
    if (!kEarly) {
      rec = crec[x];
      if (!single) load_parent<NWM>(a, Ly, lane, p, pw);
    }
    const int q = (int)cnt;
    RANK_STAMP(2);
    int cm;
    if (single) {
      GL const int* src = gp(a.cbuf) + (size_t)x * child_stride(S);
      GL int* dst = gp(a.beam_next) + (size_t)q * S;
      for (int o = lane; o < S; o += 64) dst[o] = src[o];
      cm = src[Ly.off_cmax];
      if (lane == 0) gp(a.trail)[(size_t)a.level * a.K + q] = (p << 8) | rec_job(rec);
    } else {
      int nw[kMaxWPL];
      cm = (order == kOrderErase) ? build_child<kOrderErase, NWM>(a, Ly, lane, p, rec, q, pw, nw)
                                  : build_child<kOrderSwap, NWM>(a, Ly, lane, p, rec, q, pw, nw);
      RANK_STAMP(3);
      if (FUSED) {
        // Score the survivor's children for the next level from its words.
        int* vw = s_view + wave * S;
#pragma unroll
        for (int i = 0; i < NWM; ++i) {
          const int o = lane + 64 * i;
          if (i < NW && o < S) vw[o] = nw[i];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        JobView v = ops_lds ? view_from_lds<NHM>(static_cast<const int2*>(s_ops), Ly, vw, cm, lane)
                            : view_from_lds<NHM>(ops, Ly, vw, cm, lane);
        int slot[2], nvalid;
        bool keep[2];
        emit_children<NHM>(dn, sln, q, true, 64, lane, ~0ull, lane, v, s_h, acc, slot, keep, nvalid);
      }
    }
    RANK_STAMP(4);
    w_min = min(w_min, cm);
    w_max = max(w_max, hi);
    w_minh = min(w_minh, hi);
  }
  }
  if (lane == 0) { s_mc[wave] = w_min; s_mh[wave] = w_max; s_ml[wave] = w_minh; }
  lds_barrier();  // LDS partials: the stores keep draining
  if (tid == 0) {
    RankStat rs;
    rs.min_cmax = INT_MAX;
    rs.max_hi = 0;
    rs.min_hi = 0xFFFFFFFFu;
    for (int w = 0; w < nwaves; ++w) {
      rs.min_cmax = min(rs.min_cmax, s_mc[w]);
      rs.max_hi = max(rs.max_hi, s_mh[w]);
      rs.min_hi = min(rs.min_hi, s_ml[w]);
    }
    store_global(a.rstat + local, rs);
  }
  RANK_STAMP(5);
  if (FUSED) finish_score_block(acc, s_h, a.bstat, gp(a.hist_next), local, sln.last, a.hist_copies);
  RANK_STAMP(6);
}

// ---------------------------------------------------------------------------
// Exact (host-replay) modes: per parent, the cmax of every child in
// erase-order ready position, the job and its machine, and the ready count.
__global__ __launch_bounds__(256) void k_exact_expand(LevelArgs a, int P) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const Layout Ly = make_layout(a.n, a.m);
  const int n = a.n, m = a.m, S = Ly.S, B = a.B;
  const int NH = (n + 63) >> 6;
  for (int p = blockIdx.x * 4 + wave; p < P; p += gridDim.x * 4) {
    const int* par = a.beam_cur + (size_t)p * S;
    const int pc = par[Ly.off_cmax];
    int r = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int j = h * 64 + lane;
      bool rdy = false;
      if (h < NH && j < n) {
        const int w = par[Ly.off_jw + j];
        const int pos = jw_pos(w);
        if (pos != kFinishedPos) {
          rdy = true;
          const int nx = jw_next(w);
          const int2 md = a.ops[j * m + nx];
          const int pred = (m == 1) ? (j > 0 ? par[Ly.off_jr + j - 1] : 0) : par[Ly.off_jr + j];
          const int f = max(par[Ly.off_mr + md.x], pred) + md.y;
          const size_t x = (size_t)p * B + pos;
          a.ex_cmax[x] = max(pc, f);
          a.ex_job[x] = (uint8_t)j;
          a.ex_mc[x] = (uint8_t)md.x;
          a.crec[x] = pack_rec(f, md.x, j, next_machine(a.ops, j, m, nx));
        }
      }
      r += __popcll(__ballot(rdy));
    }
    if (lane == 0) a.ex_rcount[p] = r;
  }
}

__global__ __launch_bounds__(256) void k_exact_materialize(LevelArgs a, int count) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const Layout Ly = make_layout(a.n, a.m);
  if (blockIdx.x == 0 && threadIdx.x == 0) a.pcount[a.level + 1] = count;
  for (int q = blockIdx.x * 4 + wave; q < count; q += gridDim.x * 4) {
    const int x = a.ex_order[q];
    const int p = x / a.B;
    int nw[kMaxWPL];
    (void)materialize_rec<kOrderErase>(a, Ly, lane, p, a.crec[x], q, nw);
  }
}

// ---------------------------------------------------------------------------
// Best complete child and its sequence, by walking the trail backwards.
__device__ void backtrack_body(const LevelArgs& a) {
  __shared__ uint8_t s_job[128 * 255];  // one job per level (n <= 128, m <= 255)
  __shared__ int s_seg[(128 * 255 + 31) / 32];  // path index at the top of each 32-level group
  __shared__ int s_pj[2];
  __shared__ uint32_t s_e[4];
  __shared__ u64 s_b[4];
  int p = a.bt_parent, j = a.bt_job;
  const int L = a.L, m = a.m;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (p < 0) {
    uint32_t ex = 0;
    u64 best = ~0ull;
    for (int i = threadIdx.x; i < a.nblk_score; i += blockDim.x) {
      ex += a.bstat[i].expanded;
      best = a.bstat[i].best < best ? a.bstat[i].best : best;
    }
    ex = wave_sum(ex);
    best = wave_min64(best);
    if (lane == 0) { s_e[wave] = ex; s_b[wave] = best; }
    __syncthreads();
    if (threadIdx.x == 0) {
      ex = 0; best = ~0ull;
      for (int w = 0; w < 4; ++w) { ex += s_e[w]; best = s_b[w] < best ? s_b[w] : best; }
      *a.expanded += (u64)ex;
      a.lvl_exp[L - 1] = ex;
      s_pj[0] = -1;
      if (best == ~0ull) {
        a.result[0] = a.best_known; a.result[1] = 0; a.result[2] = 0;
      } else {
        const uint32_t x = (uint32_t)best;
        const int cmax = (int)(best >> 32);
        s_pj[0] = (int)(x / (uint32_t)a.B);
        s_pj[1] = rec_job(a.crec[x]);
        a.result[0] = min(cmax, a.best_known);
        a.result[1] = 1;
        a.result[2] = cmax;
      }
    }
    __syncthreads();
    p = s_pj[0];
    j = s_pj[1];
    if (p < 0) return;  // no complete child (block-uniform)
  }
  // The trail is walked back into LDS (one job byte per level), then thread
  // j numbers job j's operations in sequence order.  With the jump table,
  // thread 0 first finds the path's index at the top of every 32-level group
  // (one load per group) and the groups are walked in parallel, one thread
  // each; without it (batches, exact modes) thread 0 walks every level.
  if (a.jump) {
    if (threadIdx.x == 0) {
      s_job[L - 1] = (uint8_t)j;
      int idx = p;
      for (int g = (L - 2) >> 5; g >= 0; --g) {
        s_seg[g] = idx;
        idx = a.jump[(size_t)g * a.K + idx];
      }
    }
    __syncthreads();
    const int groups = L >= 2 ? ((L - 2) >> 5) + 1 : 0;
    for (int g = threadIdx.x; g < groups; g += blockDim.x) {
      int idx = s_seg[g];
      for (int l = min(32 * g + 31, L - 2); l >= 32 * g; --l) {
        const int t = a.trail[(size_t)l * a.K + idx];
        s_job[l] = (uint8_t)(t & 0xFF);
        idx = t >> 8;
      }
    }
  } else if (threadIdx.x == 0) {
    s_job[L - 1] = (uint8_t)j;
    int r = p;
    for (int l = L - 2; l >= 0; --l) {
      const int t = a.trail[(size_t)l * a.K + r];
      s_job[l] = (uint8_t)(t & 0xFF);
      r = t >> 8;
    }
  }
  __syncthreads();
  int* seq = a.seq_out;
  for (int jj = threadIdx.x; jj < a.n; jj += blockDim.x) {
    int k = 0;
    for (int l = 0; l < L; ++l)
      if (s_job[l] == jj) seq[l] = jj * m + (k++) + 1;
  }
}


__global__ __launch_bounds__(256) void k_backtrack(LevelArgs a) { backtrack_body(a); }

// ---------------------------------------------------------------------------
// Step kernels over instance descriptors.
// (instance, local block) of this block; a null map means one instance whose
// local block is the grid block (saves the map load on single solves).
__device__ __forceinline__ BlockRef block_ref(const BlockRef* map) {
  return map ? map[blockIdx.x] : make_int2(0, (int)blockIdx.x);
}

// The descriptor of an instance, read with one coalesced load per wave (lane
// i holds its bytes [8i, 8i + 8)) and broadcast word by word: one memory
// round trip, where field-by-field scalar loads take several dependent ones
// (the level-indexed fields wait for the level counter).  Every lane of the
// wave must be active.  Fields the kernel stores are written through the
// descriptor pointer, never through the copy.
__device__ __forceinline__ InstDesc fetch_desc(const InstDesc* dp) {
  static_assert(sizeof(InstDesc) % 8 == 0 && sizeof(InstDesc) <= 64 * 8, "one load per lane");
  constexpr int NW = (int)(sizeof(InstDesc) / 8);
  const int lane = threadIdx.x & 63;
  GL const u64* src = reinterpret_cast<GL const u64*>(gp(dp));
  const u64 w = lane < NW ? src[lane] : 0ull;
  uint32_t words[2 * NW];
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    words[2 * i] = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)w, i);
    words[2 * i + 1] = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(w >> 32), i);
  }
  InstDesc d;
  __builtin_memcpy(&d, words, sizeof(d));
  return d;
}

__device__ __forceinline__ LevelArgs resolve(const InstDesc& d, int level) {
  LevelArgs a = d.base;
  a.level = level;
  a.lseed = d.base.lseed + (uint32_t)level;
  // selects, not indexing: the descriptor may be a register copy
  const bool odd = (level & 1) != 0;
  a.beam_cur = odd ? d.beam[1] : d.beam[0];
  a.beam_next = odd ? d.beam[0] : d.beam[1];
  a.hist = odd ? d.hist[1] : d.hist[0];
  a.hist_next = odd ? d.hist[0] : d.hist[1];
  a.sel = odd ? d.sel[1] : d.sel[0];
  a.sctr = odd ? d.sctr[1] : d.sctr[0];
  a.sctr_next = odd ? d.sctr[0] : d.sctr[1];
  a.sel_prev = odd ? d.sel[0] : d.sel[1];
  a.sel_next = odd ? d.sel[0] : d.sel[1];
  a.crec = odd ? d.crec[1] : d.crec[0];
  return a;
}

// Root of every instance mapped here; in the fused form, wave 0 also scores
// the root's children (level 0), which the rank kernel does for every later
// level.
__global__ __launch_bounds__(256) void k_init_b(InstDesc* desc, const BlockRef* map) {
  __shared__ uint32_t s_h[kNB];
  __shared__ int s_bound;
  InstDesc& d = desc[block_ref(map).x];
  const Layout Ly = make_layout(d.base.n, d.base.m);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < Ly.S; i += blockDim.x) {
    int v = 0;
    const int j = i - Ly.off_jw;
    if (i >= Ly.off_jw && i < Ly.off_cmax) v = jw_make(0, j, d.base.ops[j * d.base.m].x);  // ready = [0..n-1]
    d.beam[0][i] = v;
  }
  for (int i = threadIdx.x; i < kNB * kHistCopies; i += blockDim.x) {
    d.hist[0][i] = 0;
    d.hist[1][i] = 0;
  }
  for (int i = threadIdx.x; i < kNB; i += blockDim.x) s_h[i] = 0;
  for (int i = threadIdx.x; i < kSureSubs; i += blockDim.x) {
    d.sctr[0][i * kCtrStride] = 0u;
    d.sctr[1][i * kCtrStride] = 0u;
  }
  if (threadIdx.x == 0) {
    // Multi-seed passes prune with the shared bound as it stands when the
    // pass starts, so every beam of a pass sees the same bound.
    if (d.shared_bound) d.base.best_known = *d.shared_bound;
    s_bound = d.base.best_known;
    SelState z = {};
    z.s1 = kLevel0Shift;
    z.exact = d.base.exact_digit ? 1u : 0u;
    *d.sel[0] = z;
    *d.sel[1] = z;
    d.base.pcount[0] = 1;
    d.base.pcount[d.base.L] = 0;  // no beam after the last level (its filter returns at once)
    *d.base.expanded = 0ull;
    d.base.result[0] = d.base.best_known; d.base.result[1] = 0; d.base.result[2] = 0;
    d.lvl_score = 0;
    d.lvl_filter = 0;
    d.lvl_rank = -1;  // the first score (or, fused, filter) launch advances it to 0
    d.bt_done = 0;
  }
  __syncthreads();
  if (!d.fused) return;
  LevelArgs a = resolve(d, 0);
  a.best_known = s_bound;
  ScoreLevel sl;
  sl.lo = 0;  // the root's cmax
  sl.s1 = kLevel0Shift;
  sl.ns = 1;
  sl.last = (a.L == 1);
  ScoreAcc acc;
  if (wave == 0) {
    const int NW = (Ly.S + 63) >> 6;
    int w[kMaxWPL];
#pragma unroll
    for (int i = 0; i < kMaxWPL; ++i) {
      const int o = lane + 64 * i;
      const int j = o - Ly.off_jw;
      w[i] = (i < NW && o >= Ly.off_jw && o < Ly.off_cmax) ? jw_make(0, j, a.ops[j * a.m].x) : 0;
    }
    JobView v = view_from_words(gp(a.ops), Ly, w, NW, 0, lane);
    int slot[2], nvalid;
    bool keep[2];
    emit_children(score_dst(a), sl, 0, true, 64, lane, ~0ull, lane, v, s_h, acc, slot, keep, nvalid);
  }
  finish_score_block(acc, s_h, a.bstat, gp(a.hist), 0, sl.last, a.hist_copies);
  const BlockStat empty = {0u, 0u, INT_MAX, 0u, ~0ull};
  for (int i = 1 + (int)threadIdx.x; i < a.nblk_score; i += blockDim.x) a.bstat[i] = empty;
}

// Advances a level counter (block 0 of an instance, after its reader kernel
// has completed).  A global store: a flat one would also count against the
// wave's LDS waits until it completes.
__device__ __forceinline__ void advance(int& counter, int cur, int local, int L) {
  if (local == 0 && threadIdx.x == 0 && cur < L) *gp(&counter) = cur + 1;
}

// Two-phase score: keys and selection digits only.
__global__ __launch_bounds__(1024) void k_score_b(InstDesc* desc, const BlockRef* map) {
  const BlockRef r = block_ref(map);
  InstDesc& d = desc[r.x];
  const int level = d.lvl_score;
  // descriptor fields are read before the counter store (scalar loads)
  const LevelArgs a = resolve(d, level);
  const int nblk = d.nblk_score;
  advance(d.lvl_rank, d.lvl_rank, r.y, a.L);
  if (level >= a.L) return;
  score_body<false>(a, r.y, nblk);
}

// Single-phase expand: the same plus every kept child's snapshot.
__global__ __launch_bounds__(1024) void k_expand_b(InstDesc* desc, const BlockRef* map) {
  const BlockRef r = block_ref(map);
  InstDesc& d = desc[r.x];
  const int level = d.lvl_score;
  // descriptor fields are read before the counter store (scalar loads)
  const LevelArgs a = resolve(d, level);
  const int nblk = d.nblk_score;
  advance(d.lvl_rank, d.lvl_rank, r.y, a.L);
  if (level >= a.L) return;
  score_body<true>(a, r.y, nblk);
}

template <int U>
__device__ __forceinline__ void filter_step(InstDesc* desc, const BlockRef* map) {
  const BlockRef r = block_ref(map);
  const InstDesc d = fetch_desc(desc + r.x);
  const int level = d.lvl_filter;
  if (d.fused) advance(desc[r.x].lvl_rank, d.lvl_rank, r.y, d.base.L);
  else advance(desc[r.x].lvl_score, d.lvl_score, r.y, d.base.L);
  if (level >= d.base.L - 1) return;  // the last level only records the best child
  const LevelArgs a = resolve(d, level);
  filter_body<U>(a, r.y, d.nblk_filter);
}

__global__ __launch_bounds__(kFilterThreads) void k_filter_b(InstDesc* desc, const BlockRef* map) {
  filter_step<1>(desc, map);
}

// Batches and the server: registers held to 64 per lane (a few spills), so
// two blocks share a CU and one block's memory waits overlap another's work.
__global__ __launch_bounds__(kFilterThreads, 8) void k_filter_tp_b(InstDesc* desc, const BlockRef* map) {
  filter_step<kFilterTilesTp>(desc, map);
}

// Rank + materialise; instances of the other form return at once (a batch
// mixing both forms launches both kernels over the same map).
template <int W, int NB, bool WIDE>
__global__ __launch_bounds__(64 * W, W == kRankWavesThroughput ? 6 : 4) void k_rank_b(InstDesc* desc, const BlockRef* map) {
  const BlockRef r = block_ref(map);
  InstDesc& d = desc[r.x];
  if ((d.fused != 0) | ((d.base.K > kRankLdsMax) != WIDE)) return;
  const int level = d.lvl_rank;
  // descriptor fields are read before the counter store (scalar loads)
  const LevelArgs a = resolve(d, level);
  const int nblk = d.nblk_rank;
  advance(d.lvl_filter, d.lvl_filter, r.y, a.L);
  if (level >= a.L - 1) return;
  rank_body<false, 0, 2, W, NB, WIDE>(a, nullptr, r.y, nblk);
}

// Rank + materialise + score of the next level (fused form), per mode and
// snapshot size class.  Throughput blocks: four per CU for the small class
// (64 registers), three for the medium and large ones (80 registers: their
// spills at 64 cost more than the fourth block gains; the medium class, four
// snapshot words per lane instead of eight, then spills nothing).
template <int MODE, int SZ, int W, int NB>
__global__ __launch_bounds__(64 * W, W == kRankWavesThroughput ? (SZ == 0 ? 8 : 6) : 4) void k_rank_fused_b(
    InstDesc* desc, const BlockRef* map) {
  const BlockRef r = block_ref(map);
  InstDesc& d = desc[r.x];
  // one round of loads for the variant test (no short-circuit)
  const int sz = rank_size_class(d.base.S, d.base.n);
  if (!((d.fused != 0) & (d.base.mode == MODE) & (sz == SZ))) return;
  const int level = d.lvl_rank;
  // descriptor fields are read before the counter store (scalar loads)
  const LevelArgs a = resolve(d, level);
  u64* const crec_next = (level & 1) ? d.crec[0] : d.crec[1];
  const int nblk = d.nblk_rank;
  advance(d.lvl_filter, d.lvl_filter, r.y, a.L);
  if (level >= a.L - 1) return;
  rank_body<true, MODE, SZ, W, NB>(a, crec_next, r.y, nblk);
}

// Folds a served request's per-block partials into entry 0 and clears the
// rest (launch_fold_partials).  Every reader of the partials combines them
// by sums and minima / maxima, so entry 0 alone carries the level's values.
__global__ __launch_bounds__(256) void k_fold_partials(InstDesc* desc, const int4* jobs) {
  __shared__ uint32_t s_v[4], s_e[4], s_rmh[4], s_rlh[4];
  __shared__ int s_mc[4], s_rmc[4];
  __shared__ u64 s_b[4];
  const int4 job = jobs[blockIdx.x];
  const LevelArgs& a = desc[job.x].base;
  const int old = job.y, cap = job.z;
  const bool fused = job.w != 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t v = 0, ex = 0, mxh = 0, mnh = 0xFFFFFFFFu;
  int mc = INT_MAX, rmc = INT_MAX;
  u64 best = ~0ull;
  for (int i = threadIdx.x; i < old; i += blockDim.x) {
    if (fused) {
      const BlockStat b = a.bstat[i];
      v += b.valid;
      ex += b.expanded;
      mc = min(mc, b.min_cmax);
      best = b.best < best ? b.best : best;
    }
    const RankStat r = a.rstat[i];
    rmc = min(rmc, r.min_cmax);
    mxh = max(mxh, r.max_hi);
    mnh = min(mnh, r.min_hi);
  }
  v = wave_sum(v);
  ex = wave_sum(ex);
  mc = wave_min_i(mc);
  best = wave_min64(best);
  rmc = wave_min_i(rmc);
  mxh = wave_max(mxh);
  mnh = wave_min_u(mnh);
  if (lane == 0) {
    s_v[wave] = v; s_e[wave] = ex; s_mc[wave] = mc; s_b[wave] = best;
    s_rmc[wave] = rmc; s_rmh[wave] = mxh; s_rlh[wave] = mnh;
  }
  __syncthreads();  // every entry is read before any is written
  if (threadIdx.x == 0) {
    BlockStat b = {0u, 0u, INT_MAX, 0u, ~0ull};
    RankStat r = {INT_MAX, 0u, 0xFFFFFFFFu, 0u};
    for (int w = 0; w < 4; ++w) {
      b.valid += s_v[w];
      b.expanded += s_e[w];
      b.min_cmax = min(b.min_cmax, s_mc[w]);
      b.best = s_b[w] < b.best ? s_b[w] : b.best;
      r.min_cmax = min(r.min_cmax, s_rmc[w]);
      r.max_hi = max(r.max_hi, s_rmh[w]);
      r.min_hi = min(r.min_hi, s_rlh[w]);
    }
    if (fused) a.bstat[0] = b;
    a.rstat[0] = r;
  }
  const BlockStat eb = {0u, 0u, INT_MAX, 0u, ~0ull};
  const RankStat er = {INT_MAX, 0u, 0xFFFFFFFFu, 0u};
  for (int i = 1 + (int)threadIdx.x; i < cap; i += blockDim.x) {
    if (fused) a.bstat[i] = eb;
    a.rstat[i] = er;
  }
}

__global__ __launch_bounds__(256) void k_scatter(const uint32_t* block, const CopyJob* jobs) {
  const CopyJob j = jobs[blockIdx.x];
  uint32_t* dst = reinterpret_cast<uint32_t*>(j.dst);
  const uint32_t* src = block + j.src;
  for (uint32_t i = threadIdx.x; i < j.words; i += blockDim.x) dst[i] = src[i];
}

// Jump table of a single solve (base.jump): thread (g, q) follows survivor
// q after the last level of 32-level group g (level L-2 for the last group)
// down the trail to the group's first level and records its ancestor after
// level 32 g - 1.  All groups and survivors at once, so the backtrack's walk
// becomes one load per group plus 32 levels per group in parallel.
__global__ __launch_bounds__(256) void k_jump_b(InstDesc* desc) {
  const LevelArgs& b = desc[0].base;
  const int L = b.L, K = b.K;
  if (!b.jump || L < 2) return;
  const int groups = ((L - 2) >> 5) + 1;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)groups * K) return;
  const int g = (int)(t / K), q = (int)(t % K);
  const int top = min(32 * g + 31, L - 2);
  if (q >= b.pcount[top + 1]) return;  // survivors after level top
  int idx = q;
  for (int l = top; l >= 32 * g; --l) idx = b.trail[(size_t)l * K + idx] >> 8;
  b.jump[(size_t)g * K + q] = (uint16_t)idx;
}

// status (optional): status[instance] = whether the instance is finished
// after this launch (host memory the server reads after the launch's event).
__global__ __launch_bounds__(256) void k_backtrack_b(InstDesc* desc, const BlockRef* map, int* status) {
  const int inst = block_ref(map).x;
  InstDesc& d = desc[inst];
  if (d.lvl_filter < d.base.L || d.bt_done) {  // lvl_filter reaches L after the last rank
    if (status && threadIdx.x == 0) status[inst] = d.bt_done;
    return;
  }
  LevelArgs a = resolve(d, d.base.L - 1);
  a.bt_parent = -1;
  a.bt_job = -1;
  backtrack_body(a);
  if (threadIdx.x == 0) {
    d.bt_done = 1;
    if (d.shared_bound && a.result[1]) atomicMin(d.shared_bound, a.result[2]);
    if (status) status[inst] = 1;
  }
}
}  // namespace

// ---------------------------------------------------------------------------
hipError_t launch_init(const LevelArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_init, dim3(1), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_step_init(InstDesc* desc, const BlockRef* map, int nblk, hipStream_t s) {
  hipLaunchKernelGGL(k_init_b, dim3(nblk), dim3(256), 0, s, desc, map);
  return hipGetLastError();
}

// With a timing pair, the launch itself records the events around the
// kernel's execution (hipExtLaunchKernelGGL), so their interval is the
// kernel's duration rather than that of separate event packets.
template <class... A>
static void launch(void (*k)(A...), dim3 g, dim3 b, size_t lds, hipStream_t s, const EventPair* ev, A... args) {
  if (ev)
    hipExtLaunchKernelGGL(k, g, b, (uint32_t)lds, s, ev->start, ev->stop, 0, args...);
  else
    hipLaunchKernelGGL(k, g, b, lds, s, args...);
}

hipError_t launch_step_score(InstDesc* desc, const BlockRef* map, int nblk, bool single, hipStream_t s,
                             const EventPair* ev) {
  launch(single ? k_expand_b : k_score_b, dim3(nblk), dim3(64 * kScoreWaves), 0, s, ev, desc, map);
  return hipGetLastError();
}

hipError_t launch_step_filter(InstDesc* desc, const BlockRef* map, int nblk, size_t lds_bytes, bool throughput,
                              hipStream_t s, const EventPair* ev) {
  launch(throughput ? k_filter_tp_b : k_filter_b, dim3(nblk), dim3(kFilterThreads), lds_bytes, s, ev, desc, map);
  return hipGetLastError();
}

template <int W, int NB>
static void (*rank_kernel(int variant))(InstDesc*, const BlockRef*) {
  switch (variant) {
    case kRankFusedIch: return k_rank_fused_b<1, 2, W, NB>;
    case kRankFusedPms: return k_rank_fused_b<2, 2, W, NB>;
    case kRankFusedIchSmall: return k_rank_fused_b<1, 0, W, NB>;
    case kRankFusedPmsSmall: return k_rank_fused_b<2, 0, W, NB>;
    case kRankFusedIchMedium: return k_rank_fused_b<1, 1, W, NB>;
    case kRankFusedPmsMedium: return k_rank_fused_b<2, 1, W, NB>;
    case kRankUnfusedWide: return k_rank_b<W, NB, true>;
    default: return k_rank_b<W, NB, false>;
  }
}

hipError_t launch_step_rank(InstDesc* desc, const BlockRef* map, int nblk, size_t lds_bytes, int variant,
                            int waves, hipStream_t s, const EventPair* ev, int rank_batch) {
  constexpr int TW = kRankWavesThroughput;
  if (waves == TW && rank_batch > 1)
    launch(rank_kernel<TW, kRankBatch>(variant), dim3(nblk), dim3(64 * TW), lds_bytes, s, ev, desc, map);
  else if (waves == TW)
    launch(rank_kernel<TW, 1>(variant), dim3(nblk), dim3(64 * TW), lds_bytes, s, ev, desc, map);
  else
    launch(rank_kernel<kRankWaves, 1>(variant), dim3(nblk), dim3(64 * kRankWaves), lds_bytes, s, ev, desc, map);
  return hipGetLastError();
}

hipError_t launch_step_backtrack(InstDesc* desc, const BlockRef* map, int nblk, hipStream_t s, int* status) {
  hipLaunchKernelGGL(k_backtrack_b, dim3(nblk), dim3(256), 0, s, desc, map, status);
  return hipGetLastError();
}

hipError_t launch_fold_partials(InstDesc* desc, const int4* jobs, int count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fold_partials, dim3(count), dim3(256), 0, s, desc, jobs);
  return hipGetLastError();
}

hipError_t launch_scatter(const uint32_t* block, const CopyJob* jobs, int count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter, dim3(count), dim3(256), 0, s, block, jobs);
  return hipGetLastError();
}

hipError_t launch_step_jump(InstDesc* desc, int K, int L, hipStream_t s) {
  if (L < 2) return hipSuccess;
  const long long threads = (long long)(((L - 2) >> 5) + 1) * K;
  hipLaunchKernelGGL(k_jump_b, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, desc);
  return hipGetLastError();
}

hipError_t launch_exact_expand(const LevelArgs& a, int P, hipStream_t s) {
  const int nblk = max(1, min((P + 3) / 4, 2048));
  hipLaunchKernelGGL(k_exact_expand, dim3(nblk), dim3(256), 0, s, a, P);
  return hipGetLastError();
}

hipError_t launch_exact_materialize(const LevelArgs& a, int count, hipStream_t s) {
  const int nblk = max(1, min((count + 3) / 4, 2048));
  hipLaunchKernelGGL(k_exact_materialize, dim3(nblk), dim3(256), 0, s, a, count);
  return hipGetLastError();
}

hipError_t launch_backtrack(const LevelArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_backtrack, dim3(1), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace jssp

#ifdef JSSP_PHASE_STAMPS
// Phase stamps of the last solve (this build only): rank[kStampLevels][256][16][8],
// filter[kStampLevels][512][4], 100 MHz ticks, zero where not reached.
extern "C" int jssp_debug_phase_stamps(void* rank_out, void* filter_out) {
  if (hipMemcpyFromSymbol(rank_out, HIP_SYMBOL(jssp::g_rank_stamps), sizeof(jssp::g_rank_stamps)) != hipSuccess)
    return 1;
  if (hipMemcpyFromSymbol(filter_out, HIP_SYMBOL(jssp::g_filter_stamps), sizeof(jssp::g_filter_stamps)) !=
      hipSuccess)
    return 2;
  return 0;
}
#endif
