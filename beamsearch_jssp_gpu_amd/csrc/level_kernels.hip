// level_kernels.hip -- per-level kernels of the single-instance beam step on
// CDNA4 (gfx950, wave64).
//
// One level = expand/score -> select top-K -> materialise.  The reference does
// this with kernel_expand_beam_layer / kernel_score_candidates (one block per
// parent, one thread per branch, 17 KB of private arrays per thread,
// gpu/v1/solver.cu:9-157, gpu/v2_twophases/solver.cu:9-147), a full CUB radix
// sort of every slot (gpu/v1/solver.cu:255) and a one-thread-per-survivor copy
// (gpu/v1/solver.cu:160-176, gpu/v2_twophases/solver.cu:150-254).  Here:
//   * one wave per parent; lane j holds job j, so the ready list, the pivot
//     filter and the child slots come from ballots and popcounts, and every
//     child's start/finish/cmax is computed in registers;
//   * children are written as compact snapshots with lanes covering word
//     offsets (coalesced), only for valid children;
//   * top-K is a radix select: one histogram of a data-ranged 11-bit digit of
//     the score bits, one partition of the keys at or below the threshold bin
//     into per-bin buckets, an LDS sort of the threshold bin, and a rank by
//     counting inside each (small) bucket.  The output is the survivors in
//     exact (score bits, child index) order -- the order CUB's stable sort
//     gives the reference -- and the rank is the survivor's slot in the
//     next beam;
//   * survivors are rebuilt from their parent (two-phase) or copied from the
//     child buffer (single-phase) by one wave per survivor.
// All counts (parents per level, survivors) stay on the device, so the host
// enqueues a fixed sequence of launches with no readback.
#include <hip/hip_runtime.h>

#include "level_kernels.hpp"
#include "snapshot.hpp"

namespace jssp {
namespace {

typedef unsigned long long u64;

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ u64 wave_min64(u64 v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    u64 w = __shfl_xor(v, o);
    v = w < v ? w : v;
  }
  return v;
}
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  return v;
}

__device__ __forceinline__ uint32_t sel_digit(u64 key, uint32_t lo, uint32_t shift) {
  return ((uint32_t)(key >> 32) - lo) >> shift;
}

// What changes from a parent snapshot to one child.
struct ChildMod {
  int mc, f, j, cm, spos, jlast, newjw;
};

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
      v = (jj == c.jlast) ? jw_make(jw_next(v), c.spos) : v;
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
    if (i >= Ly.off_jw && i < Ly.off_cmax) v = jw_make(0, i - Ly.off_jw);  // ready = [0..n-1]
    root[i] = v;
  }
  for (int i = threadIdx.x; i < kNB; i += blockDim.x) {
    a.hist[i] = 0; a.fill[i] = 0; a.hist_next[i] = 0; a.fill_next[i] = 0;
  }
  if (threadIdx.x == 0) {
    a.pcount[0] = 1;
    *a.expanded = 0ull;
    a.result[0] = a.best_known; a.result[1] = 0; a.result[2] = 0;
  }
}

// ---------------------------------------------------------------------------
// Score (two-phase) or expand (single-phase): one wave per parent.
// MODE 1 = ICH_DET (every ready op, slot = ready position),
// MODE 2 = PMS_GPU (ops on the hash-chosen pivot machine, gpu/v1/solver.cu:47-70).
template <int MODE, bool SINGLE>
__global__ __launch_bounds__(1024) void k_score(LevelArgs a) {
  __shared__ uint32_t s_red[kScoreWaves][4];
  __shared__ u64 s_best[kScoreWaves];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const Layout Ly = make_layout(a.n, a.m);
  const int n = a.n, m = a.m, S = Ly.S, B = a.B;
  const int P = a.pcount[a.level];
  const bool last = (a.level == a.L - 1);
  const int ns = a.level + 1;
  const int NH = (n + 63) >> 6;
  const int NW = (S + 63) >> 6;
  constexpr int ORDER = (MODE == 1) ? kOrderErase : kOrderSwap;

  uint32_t w_cnt = 0, w_min = 0xFFFFFFFFu, w_max = 0, w_exp = 0;
  u64 w_best = ~0ull;

  for (int p = blockIdx.x * kScoreWaves + wave; p < P; p += gridDim.x * kScoreWaves) {
    const int* par = a.beam_cur + (size_t)p * S;
    const int pc = par[Ly.off_cmax];
    int pos[2], mc[2], f[2], cm[2], nxv[2];
    bool rdy[2];
    int r = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      rdy[h] = false; pos[h] = kFinishedPos; mc[h] = 0; f[h] = 0; cm[h] = 0; nxv[h] = 0;
      const int j = h * 64 + lane;
      if (h < NH && j < n) {
        const int w = par[Ly.off_jw + j];
        pos[h] = jw_pos(w);
        if (pos[h] != kFinishedPos) {
          rdy[h] = true;
          const int nx = jw_next(w);
          nxv[h] = nx;
          const int2 md = a.ops[j * m + nx];
          // gpu_snapshot.cuh:104 -- with M == 1 the predecessor is op-1 of the previous job.
          const int pred = (m == 1) ? (j > 0 ? par[Ly.off_jr + j - 1] : 0) : par[Ly.off_jr + j];
          const int st = max(par[Ly.off_mr + md.x], pred);
          mc[h] = md.x;
          f[h] = st + md.y;
          cm[h] = max(pc, f[h]);
        }
      }
      r += __popcll(__ballot(rdy[h]));
    }

    int slot[2];
    bool val[2];
    int nvalid;
    if (MODE == 1) {
      slot[0] = pos[0]; slot[1] = pos[1];
      val[0] = rdy[0]; val[1] = rdy[1];
      nvalid = r;
    } else {
      nvalid = 0;
      val[0] = val[1] = false;
      slot[0] = slot[1] = 0;
      if (r > 0) {
        const uint32_t hsh = pivot_hash(a.lseed + (uint32_t)p);
        const int tpos = (int)(hsh % (uint32_t)r);
        int tmc = -1;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const u64 bm = __ballot(rdy[h] && pos[h] == tpos);
          if (bm) tmc = __shfl(mc[h], __ffsll((long long)bm) - 1);
        }
        val[0] = rdy[0] && mc[0] == tmc;
        val[1] = rdy[1] && mc[1] == tmc;
        // Slot = rank of the ready position among the pivot machine's ops.
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          u64 bm = __ballot(val[h]);
          nvalid += __popcll(bm);
          while (bm) {
            const int l = __ffsll((long long)bm) - 1;
            bm &= bm - 1;
            const int pl = __shfl(pos[h], l);
            slot[0] += (pl < pos[0]) ? 1 : 0;
            slot[1] += (pl < pos[1]) ? 1 : 0;
          }
        }
      }
    }
    w_exp += (uint32_t)nvalid;

    bool keep[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      keep[h] = val[h] && cm[h] < a.best_known;
      if (!val[h]) continue;
      const uint32_t x = (uint32_t)(p * B + slot[h]);
      const int j = h * 64 + lane;
      if (keep[h]) {
        a.cjob[x] = (uint8_t)j;
        if (last) {
          const u64 c = ((u64)(uint32_t)cm[h] << 32) | x;
          w_best = c < w_best ? c : w_best;
        } else {
          const float sc = gpu_score(cm[h], ns, noise_hash(a.lseed, (uint32_t)p, (uint32_t)slot[h]));
          const uint32_t bits = __float_as_uint(sc);
          a.keys[x] = ((u64)bits << 32) | x;
          w_cnt++;
          w_min = min(w_min, bits);
          w_max = max(w_max, bits);
        }
      } else if (!last) {
        a.keys[x] = ~0ull;
      }
    }
    if (!last) {
      for (int s = lane; s < B; s += 64)
        if (s >= nvalid) a.keys[(size_t)p * B + s] = ~0ull;
    }

    if (SINGLE && !last) {
      int pw[kMaxWPL];
#pragma unroll
      for (int i = 0; i < kMaxWPL; ++i) {
        const int o = lane + 64 * i;
        pw[i] = (i < NW && o < S) ? par[o] : 0;
      }
      int jlast = -1;
      if (ORDER == kOrderSwap) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const u64 bm = __ballot(rdy[h] && pos[h] == r - 1);
          if (bm) jlast = h * 64 + __ffsll((long long)bm) - 1;
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        u64 bm = __ballot(keep[h]);
        while (bm) {
          const int l = __ffsll((long long)bm) - 1;
          bm &= bm - 1;
          ChildMod c;
          c.j = h * 64 + l;
          c.mc = __shfl(mc[h], l);
          c.f = __shfl(f[h], l);
          c.cm = __shfl(cm[h], l);
          c.spos = __shfl(pos[h], l);
          c.jlast = jlast;
          const int cs = __shfl(slot[h], l);
          const int nx1 = __shfl(nxv[h], l) + 1;
          c.newjw = jw_make(nx1, nx1 < m ? r - 1 : kFinishedPos);
          int* out = a.cbuf + ((size_t)p * B + cs) * S;
#pragma unroll
          for (int i = 0; i < kMaxWPL; ++i) {
            const int o = lane + 64 * i;
            if (i < NW && o < S) out[o] = child_word<ORDER>(o, pw[i], c, Ly);
          }
        }
      }
    }
  }

  // Block partials: kept count, score-bit range, expanded count, last-level best.
  w_cnt = wave_sum(w_cnt);
  w_min = wave_min(w_min);
  w_max = wave_max(w_max);
  w_best = wave_min64(w_best);
  if (lane == 0) {
    s_red[wave][0] = w_cnt; s_red[wave][1] = w_min; s_red[wave][2] = w_max; s_red[wave][3] = w_exp;
    s_best[wave] = w_best;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    BlockStat b;
    b.valid = 0; b.minhi = 0xFFFFFFFFu; b.maxhi = 0; b.expanded = 0; b.best = ~0ull;
    for (int w = 0; w < kScoreWaves; ++w) {
      b.valid += s_red[w][0];
      b.minhi = min(b.minhi, s_red[w][1]);
      b.maxhi = max(b.maxhi, s_red[w][2]);
      b.expanded += s_red[w][3];
      b.best = s_best[w] < b.best ? s_best[w] : b.best;
    }
    a.bstat[blockIdx.x] = b;
  }
}

// Reduces the score blocks' partials (every caller block does it redundantly,
// which avoids a separate launch and any same-address atomics).
struct Totals {
  uint32_t valid, minhi, maxhi, expanded;
  u64 best;
};

__device__ Totals reduce_bstat(const BlockStat* bs, int nb) {
  __shared__ uint32_t s_v[16], s_mn[16], s_mx[16], s_ex[16];
  __shared__ u64 s_b[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t v = 0, mn = 0xFFFFFFFFu, mx = 0, ex = 0;
  u64 b = ~0ull;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    const BlockStat s = bs[i];
    v += s.valid; mn = min(mn, s.minhi); mx = max(mx, s.maxhi); ex += s.expanded;
    b = s.best < b ? s.best : b;
  }
  v = wave_sum(v); mn = wave_min(mn); mx = wave_max(mx); ex = wave_sum(ex); b = wave_min64(b);
  if (lane == 0) { s_v[wave] = v; s_mn[wave] = mn; s_mx[wave] = mx; s_ex[wave] = ex; s_b[wave] = b; }
  __syncthreads();
  Totals t;
  t.valid = 0; t.minhi = 0xFFFFFFFFu; t.maxhi = 0; t.expanded = 0; t.best = ~0ull;
  const int nw = (blockDim.x + 63) >> 6;
  for (int w = 0; w < nw; ++w) {
    t.valid += s_v[w]; t.minhi = min(t.minhi, s_mn[w]); t.maxhi = max(t.maxhi, s_mx[w]);
    t.expanded += s_ex[w]; t.best = s_b[w] < t.best ? s_b[w] : t.best;
  }
  __syncthreads();
  return t;
}

// ---------------------------------------------------------------------------
// Histogram of the first digit over all kept keys of the level.
__global__ __launch_bounds__(1024) void k_hist(LevelArgs a) {
  __shared__ uint32_t s_h[kNB];
  const Totals t = reduce_bstat(a.bstat, a.nblk_score);
  const uint32_t range = t.maxhi - t.minhi;
  const int bitlen = (t.valid == 0 || range == 0) ? 0 : 32 - __clz(range);
  const uint32_t shift = (uint32_t)max(0, bitlen - kLog2NB);
  const uint32_t Kp = min(t.valid, (uint32_t)a.K);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    SelState s;
    s.V = t.valid; s.lo = t.minhi; s.hi = t.maxhi; s.shift = shift; s.Kp = Kp;
    s.T = 0; s.cumlt = 0; s.need = 0; s.t = 0;
    *a.sel = s;
    a.pcount[a.level + 1] = (int)Kp;
    *a.expanded += (u64)t.expanded;
    a.lvl_exp[a.level] = t.expanded;
  }
  if (t.valid == 0) return;
  for (int i = threadIdx.x; i < kNB; i += blockDim.x) s_h[i] = 0;
  __syncthreads();
  const size_t N = (size_t)a.pcount[a.level] * a.B;
  for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < N; x += (size_t)gridDim.x * blockDim.x) {
    const u64 key = a.keys[x];
    if (key != ~0ull) atomicAdd(&s_h[sel_digit(key, t.minhi, shift)], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kNB; i += blockDim.x)
    if (s_h[i]) atomicAdd(&a.hist[i], s_h[i]);
}

// ---------------------------------------------------------------------------
// Exclusive scan of the histogram (redundantly per block), threshold bin T,
// and the scatter of every key at or below T into its bin's bucket.
__global__ __launch_bounds__(1024) void k_partition(LevelArgs a) {
  __shared__ uint32_t s_ex[kNB];
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_T;
  const SelState sel = *a.sel;
  if (sel.Kp == 0) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) s_T = kNB - 1;
  // 1024 threads x 2 bins
  const uint32_t h0 = a.hist[2 * tid], h1 = a.hist[2 * tid + 1];
  const uint32_t sum = h0 + h1;
  const uint32_t inc = wave_incl_scan(sum, lane);
  if (lane == 63) s_w[wave] = inc;
  __syncthreads();
  uint32_t wofs = 0;
  for (int w = 0; w < wave; ++w) wofs += s_w[w];
  const uint32_t ex0 = wofs + inc - sum;
  s_ex[2 * tid] = ex0;
  s_ex[2 * tid + 1] = ex0 + h0;
  if (ex0 < sel.Kp && ex0 + h0 >= sel.Kp) s_T = 2 * tid;
  if (ex0 + h0 < sel.Kp && ex0 + h0 + h1 >= sel.Kp) s_T = 2 * tid + 1;
  __syncthreads();
  const uint32_t T = s_T;
  if (blockIdx.x == 0) {
    a.excl[2 * tid] = s_ex[2 * tid];
    a.excl[2 * tid + 1] = s_ex[2 * tid + 1];
    if (tid == 0) {
      SelState* g = a.sel;
      g->T = T;
      g->cumlt = s_ex[T];
      g->need = sel.Kp - s_ex[T];
      g->t = a.hist[T];
    }
  }
  const size_t N = (size_t)a.pcount[a.level] * a.B;
  for (size_t x = (size_t)blockIdx.x * blockDim.x + tid; x < N; x += (size_t)gridDim.x * blockDim.x) {
    const u64 key = a.keys[x];
    if (key == ~0ull) continue;
    const uint32_t d = sel_digit(key, sel.lo, sel.shift);
    if (d <= T) {
      const uint32_t pos = s_ex[d] + atomicAdd(&a.fill[d], 1u);
      a.bucketed[pos] = key;
    }
  }
}

// ---------------------------------------------------------------------------
// Bitonic sort of n (power of two) keys in LDS by one block.
__device__ void lds_bitonic(u64* s, int n) {
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const u64 x = s[i], y = s[ixj];
          const bool up = (i & k) == 0;
          if ((x > y) == up) { s[i] = y; s[ixj] = x; }
        }
      }
      __syncthreads();
    }
  }
}

// Same in global memory (one block; only for degenerate threshold bins).
__device__ void global_bitonic(u64* s, int n) {
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const u64 x = s[i], y = s[ixj];
          const bool up = (i & k) == 0;
          if ((x > y) == up) { s[i] = y; s[ixj] = x; }
        }
      }
      __threadfence_block();
      __syncthreads();
    }
  }
}

__device__ __forceinline__ int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Orders the threshold bin: its `need` smallest keys, sorted, are written to
// the front of its bucket.  One block.  Bins of up to kTieLds keys are sorted
// in LDS; larger ones (degenerate inputs) are first cut to `need` keys by an
// 8-bit MSD radix select over the full 64-bit key.
__global__ __launch_bounds__(1024) void k_tie(LevelArgs a) {
  __shared__ u64 s_k[kTieLds];
  __shared__ uint32_t s_h[256];
  __shared__ u64 s_want, s_mask;
  __shared__ uint32_t s_kth, s_cnt;
  const SelState sel = *a.sel;
  if (sel.Kp == 0) return;
  u64* region = a.bucketed + sel.cumlt;
  const int t = (int)sel.t, need = (int)sel.need;
  if (t <= kTieLds) {
    const int np = next_pow2(t);
    for (int i = threadIdx.x; i < np; i += blockDim.x) s_k[i] = (i < t) ? region[i] : ~0ull;
    __syncthreads();
    lds_bitonic(s_k, np);
    for (int i = threadIdx.x; i < need; i += blockDim.x) region[i] = s_k[i];
    return;
  }
  // Radix select of the need-th smallest key.
  if (threadIdx.x == 0) { s_want = 0; s_mask = 0; s_kth = (uint32_t)(need - 1); }
  __syncthreads();
  for (int shift = 56; shift >= 0; shift -= 8) {
    for (int i = threadIdx.x; i < 256; i += blockDim.x) s_h[i] = 0;
    __syncthreads();
    const u64 want = s_want, mask = s_mask;
    for (int i = threadIdx.x; i < t; i += blockDim.x) {
      const u64 k = region[i];
      if ((k & mask) == want) atomicAdd(&s_h[(k >> shift) & 255], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
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
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  u64* scratch = a.keys;
  for (int i = threadIdx.x; i < t; i += blockDim.x) {
    const u64 k = region[i];
    if (k <= thr) scratch[atomicAdd(&s_cnt, 1u)] = k;
  }
  __threadfence_block();
  __syncthreads();
  const int np = next_pow2(need);
  if (np <= kTieLds) {
    for (int i = threadIdx.x; i < np; i += blockDim.x) s_k[i] = (i < need) ? scratch[i] : ~0ull;
    __syncthreads();
    lds_bitonic(s_k, np);
    for (int i = threadIdx.x; i < need; i += blockDim.x) region[i] = s_k[i];
  } else {
    for (int i = need + threadIdx.x; i < np; i += blockDim.x) scratch[i] = ~0ull;
    __threadfence_block();
    __syncthreads();
    global_bitonic(scratch, np);
    for (int i = threadIdx.x; i < need; i += blockDim.x) region[i] = scratch[i];
  }
}

// ---------------------------------------------------------------------------
// Builds survivor q of the next beam from its parent (two-phase) and writes
// the trail entry.  One wave.
template <int ORDER>
__device__ __forceinline__ void materialize_one(const LevelArgs& a, const Layout& Ly, int lane,
                                                int p, int j, int q) {
  const int S = Ly.S, m = a.m;
  const int NW = (S + 63) >> 6;
  const int* par = a.beam_cur + (size_t)p * S;
  int pw[kMaxWPL];
  int r = 0;
  int jlast = -1;
#pragma unroll
  for (int i = 0; i < kMaxWPL; ++i) {
    const int o = lane + 64 * i;
    pw[i] = (i < NW && o < S) ? par[o] : 0;
    const bool isjw = (i < NW) && o >= Ly.off_jw && o < Ly.off_cmax;
    r += __popcll(__ballot(isjw && jw_pos(pw[i]) != kFinishedPos));
  }
  if (ORDER == kOrderSwap) {
#pragma unroll
    for (int i = 0; i < kMaxWPL; ++i) {
      const int o = lane + 64 * i;
      const bool isjw = (i < NW) && o >= Ly.off_jw && o < Ly.off_cmax;
      const u64 bm = __ballot(isjw && jw_pos(pw[i]) == r - 1);
      if (bm) jlast = 64 * i + (__ffsll((long long)bm) - 1) - Ly.off_jw;
    }
  }
  const int jwj = par[Ly.off_jw + j];
  const int nx = jw_next(jwj);
  const int2 md = a.ops[j * m + nx];
  const int pred = (m == 1) ? (j > 0 ? par[Ly.off_jr + j - 1] : 0) : par[Ly.off_jr + j];
  ChildMod c;
  c.j = j;
  c.mc = md.x;
  c.f = max(par[Ly.off_mr + md.x], pred) + md.y;
  c.cm = max(par[Ly.off_cmax], c.f);
  c.spos = jw_pos(jwj);
  c.jlast = jlast;
  c.newjw = jw_make(nx + 1, nx + 1 < m ? r - 1 : kFinishedPos);
  int* out = a.beam_next + (size_t)q * S;
#pragma unroll
  for (int i = 0; i < kMaxWPL; ++i) {
    const int o = lane + 64 * i;
    if (i < NW && o < S) out[o] = child_word<ORDER>(o, pw[i], c, Ly);
  }
  if (lane == 0) a.trail[(size_t)a.level * a.K + q] = (p << 8) | j;
}

// Rank of every survivor inside its bucket (count of smaller keys), then
// materialise (two-phase) or copy from the child buffer (single-phase) into
// slot q of the next beam.  One wave per survivor.
template <int ORDER, bool SINGLE>
__global__ __launch_bounds__(256) void k_rank(LevelArgs a) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kNB; i += gridDim.x * blockDim.x) {
    a.hist_next[i] = 0;
    a.fill_next[i] = 0;
  }
  const SelState sel = *a.sel;
  const int Kp = (int)sel.Kp;
  const Layout Ly = make_layout(a.n, a.m);
  const int S = Ly.S;
  for (int u = blockIdx.x * kRankWaves + wave; u < Kp; u += gridDim.x * kRankWaves) {
    const u64 key = a.bucketed[u];
    int q = u;
    if ((uint32_t)u < sel.cumlt) {
      const uint32_t d = sel_digit(key, sel.lo, sel.shift);
      const uint32_t bs = a.excl[d], b = a.hist[d];
      uint32_t cnt = 0;
      for (uint32_t base = 0; base < b; base += 64) {
        const uint32_t i = base + lane;
        const bool lt = (i < b) && a.bucketed[bs + i] < key;
        cnt += (uint32_t)__popcll(__ballot(lt));
      }
      q = (int)(bs + cnt);
    }
    const uint32_t x = (uint32_t)key;
    const int p = (int)(x / (uint32_t)a.B);
    const int j = a.cjob[x];
    if (SINGLE) {
      const int* src = a.cbuf + (size_t)x * S;
      int* dst = a.beam_next + (size_t)q * S;
      for (int o = lane; o < S; o += 64) dst[o] = src[o];
      if (lane == 0) a.trail[(size_t)a.level * a.K + q] = (p << 8) | j;
    } else {
      materialize_one<ORDER>(a, Ly, lane, p, j, q);
    }
  }
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
    const int j = a.ex_job[x];
    materialize_one<kOrderErase>(a, Ly, lane, p, j, q);
  }
}

// ---------------------------------------------------------------------------
// Best complete child and its sequence, by walking the trail backwards.
__global__ __launch_bounds__(256) void k_backtrack(LevelArgs a) {
  __shared__ int s_cnt[128];
  int p = a.bt_parent, j = a.bt_job;
  const int L = a.L, m = a.m;
  if (p < 0) {
    const Totals t = reduce_bstat(a.bstat, a.nblk_score);
    if (threadIdx.x != 0) return;
    *a.expanded += (u64)t.expanded;
    a.lvl_exp[L - 1] = t.expanded;
    if (t.best == ~0ull) {
      a.result[0] = a.best_known; a.result[1] = 0; a.result[2] = 0;
      return;
    }
    const uint32_t x = (uint32_t)t.best;
    const int cmax = (int)(t.best >> 32);
    p = (int)(x / (uint32_t)a.B);
    j = a.cjob[x];
    a.result[0] = min(cmax, a.best_known);
    a.result[1] = 1;
    a.result[2] = cmax;
  } else if (threadIdx.x != 0) {
    return;
  }
  int* seq = a.seq_out;
  seq[L - 1] = j;
  int r = p;
  for (int l = L - 2; l >= 0; --l) {
    const int t = a.trail[(size_t)l * a.K + r];
    seq[l] = t & 0xFF;
    r = t >> 8;
  }
  for (int i = 0; i < a.n; ++i) s_cnt[i] = 0;
  for (int l = 0; l < L; ++l) {
    const int jj = seq[l];
    seq[l] = jj * m + s_cnt[jj] + 1;
    s_cnt[jj]++;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
hipError_t launch_init(const LevelArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_init, dim3(1), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_score(const LevelArgs& a, int mode, bool single, hipStream_t s) {
  const dim3 g(a.nblk_score), b(64 * kScoreWaves);
  if (mode == 1) {
    if (single) hipLaunchKernelGGL((k_score<1, true>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_score<1, false>), g, b, 0, s, a);
  } else {
    if (single) hipLaunchKernelGGL((k_score<2, true>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_score<2, false>), g, b, 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_hist(const LevelArgs& a, int nblk, hipStream_t s) {
  hipLaunchKernelGGL(k_hist, dim3(nblk), dim3(1024), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_partition(const LevelArgs& a, int nblk, hipStream_t s) {
  hipLaunchKernelGGL(k_partition, dim3(nblk), dim3(1024), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_tie(const LevelArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_tie, dim3(1), dim3(1024), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_rank(const LevelArgs& a, int order, bool single, int nblk, hipStream_t s) {
  const dim3 g(nblk), b(64 * kRankWaves);
  if (order == kOrderErase) {
    if (single) hipLaunchKernelGGL((k_rank<kOrderErase, true>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_rank<kOrderErase, false>), g, b, 0, s, a);
  } else {
    if (single) hipLaunchKernelGGL((k_rank<kOrderSwap, true>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_rank<kOrderSwap, false>), g, b, 0, s, a);
  }
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
