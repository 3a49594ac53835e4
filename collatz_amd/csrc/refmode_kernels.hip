// Reference mode: the reference program's Steps 2 and 3 for the leaves of a case line.
//
//   k_bv_build       precompute() of the BV_i look-ahead bit vectors (collatz_cudav7.cu:423-480,
//                    collatz_openclv2.cpp:294-353, main.rs:125-181): one thread per residue
//                    n0 < 2^B, its B-step parity word and max dip; a bit at n0 * 3^{-f} mod 2^B
//                    in every BV_i whose cutoff it meets.  The first words of each vector are
//                    mirrored after the end so windows never wrap.
//   k_la_build       the same dips as a range-mode look-ahead table: per residue of number,
//                    the least odd-step count above the sieve's minimum that keeps it.
//   k_ref_fused      Steps 2 and 3 (collatz_cudav7.cu:219-235, :280-366): persistent waves
//                    take tiles of leaves, expand 64 words of their 2^A-bit candidate masks
//                    (BV_i window AND the mod 9 row of the base) at a time into a queue of
//                    (leaf, a) in LDS, and iterate the queued candidates with the schedule;
//                    hash, count and overflow count are folded with one atomic per wave and
//                    tile.
//   k_ref_iterate    Step 3 over a list of (leaf, a) candidates: the two- or three-limb rerun
//                    of the candidates that left 96 bits in k_ref_fused.
#include "schedule.hpp"

namespace cz {


// doubleeven(): main.rs:101-107.  The C++ programs evaluate `x & 0b11 == 0` as `x & 0`
// (collatz_cudav7.cu:418), which is never true; only main.rs tests the two bits.
__device__ __forceinline__ bool dbl_even(uint32_t upordown, uint32_t i, uint32_t B, bool rust) {
  if (!rust) return false;
  const uint32_t a = __builtin_ctz((upordown >> i) + 1u);
  if (i + a + 2 >= B) return false;
  return ((upordown >> (i + a)) & 3u) == 0;
}

__device__ __forceinline__ bool gt_sem(int32_t x, int32_t y, bool uns) {
  return uns ? ((uint32_t)x > (uint32_t)y) : (x > y);
}

// The B-step dip of residue n0 of number (collatz_cudav7.cu:436-467): the parity word of B
// steps of the shifted map, then the running maximum of cur = 306 (steps) - 485 (odd steps)
// with the path-merge terms (doubleeven and even-length runs of even steps).
__device__ __forceinline__ int32_t la_dip(uint32_t n0, uint32_t B, bool rust, bool uns) {
  uint32_t upordown = 0;
  uint64_t n = n0;
  for (uint32_t i = 0; i < B; ++i) {
    if ((n & 1) == 0) {
      n = n + (n >> 1);
      upordown |= 1u << i;
    } else {
      n = (n >> 1) + 1;
    }
  }
  int32_t cur = 0, mx = 0;
  uint32_t ne = 0;
  for (uint32_t i = 0; i < B; ++i) {
    if (upordown & (1u << i)) {
      if (gt_sem(cur + (int32_t)kNumerator, mx, uns) && dbl_even(upordown, i, B, rust))
        mx = cur + (int32_t)kNumerator;
      cur -= (int32_t)(kDenominator - kNumerator);
      ne = 0;
    } else {
      cur += (int32_t)kNumerator;
      ne += 1;
      if ((ne & 1) == 0 && gt_sem(cur + (int32_t)(kDenominator - kNumerator), mx, uns))
        mx = cur + (int32_t)(kDenominator - kNumerator);
    }
    if (cur > mx) mx = cur;
  }
  return mx;
}

__global__ __launch_bounds__(256) void k_bv_build(uint32_t B, uint32_t num_bv, int32_t first,
                                                  uint32_t sem, const uint32_t* magic,
                                                  uint32_t* bv, uint64_t words_per_bv,
                                                  uint32_t mirror_words) {
  const uint64_t n0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n0 >= (1ull << B)) return;
  const int32_t mx = la_dip((uint32_t)n0, B, sem == 1, sem == 2);
  const uint32_t mask = (1u << B) - 1u;
  for (uint32_t a = 0; a < num_bv; ++a) {
    if (mx <= first + (int32_t)(a * kDenominator)) {
      const uint32_t idx = ((uint32_t)n0 * magic[a]) & mask;
      uint32_t* v = bv + a * words_per_bv;
      atomicOr(&v[idx >> 5], 1u << (idx & 31));
      if ((idx >> 5) < mirror_words) atomicOr(&v[(1u << (B - 5)) + (idx >> 5)], 1u << (idx & 31));
    }
  }
}

// Range-mode look-ahead table of a depth-k sieve: for each residue n0 < 2^B of number, the
// least g >= 0 with 485 (e_min + g) >= dip(n0) + 306 k, where e_min = ceil(306 k / 485) is
// the least odd-step count a survivor can have; a candidate with f odd steps is kept iff
// f - e_min >= g.  Stored as nibbles, eight per word, clamped to 15 (for k <= 38,
// f - e_min <= 14, so the clamp changes no decision).  One thread per word.
__global__ __launch_bounds__(256) void k_la_build(uint32_t B, uint32_t k, uint32_t sem,
                                                  uint32_t* table) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= (1ull << (B - 3))) return;
  const int32_t base = (int32_t)(kNumerator * k) -
                       (int32_t)(kDenominator * ((kNumerator * k + kDenominator - 1) / kDenominator));
  uint32_t word = 0;
  for (uint32_t j = 0; j < 8; ++j) {
    const int32_t d = la_dip((uint32_t)(w * 8 + j), B, false, sem == 2) + base;
    uint32_t g = d <= 0 ? 0u : (uint32_t)((d + (int32_t)kDenominator - 1) / (int32_t)kDenominator);
    if (g > 15) g = 15;
    word |= g << (4 * j);
  }
  table[w] = word;
}

// Word w of a leaf's candidate mask: bit j set iff a = 32 w + j is a candidate.
__device__ __forceinline__ uint32_t leaf_mask_word(const RefTablesDev& T, uint64_t base,
                                                   uint64_t number, uint32_t meta, uint32_t w) {
  const uint32_t f = meta_f(meta);
  uint32_t m = T.mod9[meta_mod9(meta) * T.nw + w];
  const uint32_t b = f - T.tpo;
  if (b < T.num_bv) {
    const uint32_t idx = (T.magic[b] * (uint32_t)number) & ((1u << T.B) - 1u);
    const uint32_t* v = T.bv + (uint64_t)b * T.words_per_bv + (idx >> 5) + w;
    m &= __builtin_amdgcn_alignbit(v[1], v[0], idx & 31);
  }
  if (T.rust) {
    // main.rs:309-313, with a = 32 w + j, so a mod 4 = j mod 4.
    const uint32_t de = meta_de(meta);
    const uint32_t nl = (uint32_t)number;
    if (de == 2) m &= (nl & 1) ? 0xAAAAAAAAu : 0x55555555u;
    if (de == 1) {
      const uint32_t bad = (f % 2 == 0) ? ((1u - nl) & 3u) : ((3u * (1u - nl)) & 3u);
      m &= ~(0x11111111u << bad);
    }
  }
  return m;
}

__global__ __launch_bounds__(256) void k_leaf_bins(const uint32_t* meta, uint64_t n, uint32_t tpo,
                                                   unsigned long long* bins) {
  __shared__ unsigned long long lb[17];
  if (threadIdx.x < 17) lb[threadIdx.x] = 0;
  __syncthreads();
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t b = meta_f(meta[t]) - tpo;
    atomicAdd(&lb[b < 16 ? b : 16], 1ull);
  }
  __syncthreads();
  if (threadIdx.x < 17 && lb[threadIdx.x]) atomicAdd(&bins[threadIdx.x], lb[threadIdx.x]);
}

struct RefTotals {
  unsigned long long count;
  unsigned long long overflow;
  unsigned long long stall;
  unsigned int hash;
  unsigned int tile_counter;
  unsigned long long esc_count;  // candidates handed from 96 bits to the wide pass
  unsigned long long checks;     // checks run (counted per wave from its active-lane mask)
  unsigned long long aborts;     // tiles abandoned at the work-loop bound (Limits::max_loop)
};

// num >= 2^c, or (strict) num > 2^c.
template <int L>
__device__ __forceinline__ bool reaches(const Num<L>& v, uint32_t c, bool strict) {
  if (!ge_pow2<L>(v, c)) return false;
  if (!strict) return true;
  const uint32_t cw = c >> 5, cb = c & 31u;
  bool exact = true;
#pragma unroll
  for (int i = 0; i < L; ++i) exact = exact && v.w[i] == ((uint32_t)i == cw ? (1u << cb) : 0u);
  return !exact;
}

template <int L, bool LOWCUT>
__global__ __launch_bounds__(256) void k_ref_iterate(const uint64_t* cands, uint64_t ncand,
                                                     const uint64_t* base, const uint64_t* num,
                                                     const uint32_t* meta, uint32_t mbi,
                                                     Limits lim, const uint64_t* pow3,
                                                     uint32_t tile, RefTotals* tot,
                                                     uint64_t* ovf, uint64_t ovf_cap) {
  __shared__ Pow3Lds p3s;
  stage_pow3(p3s);
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t ntiles = (ncand + tile - 1) / tile;
  const uint32_t cutoff = lim.cutoff;
  const uint64_t thr = lim.fast_thr;
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(&tot->tile_counter, 1u);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= ntiles) break;
    uint64_t cur = (uint64_t)t * tile;
    const uint64_t end = cur + tile < ncand ? cur + tile : ncand;
    uint32_t h = 0, cnt = 0, ov = 0, stall = 0;
    bool act = false, strict = false;
    Num<L> v;
    uint32_t n[4];
    uint32_t checks = 0, wave_chk = 0;
    uint32_t iters = 0;
    bool aborted = false;
    for (;;) {
      if (__builtin_expect(++iters > lim.max_loop, 0)) {
        aborted = true;
        break;
      }
      const uint64_t need = __ballot(!act);
      if (need && cur < end) {
        const uint32_t want = __popcll(need);
        const uint64_t avail = end - cur;
        const uint32_t take = avail < want ? (uint32_t)avail : want;
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
            (uint32_t)(need >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
        if (!act && rank < take) {
          const uint64_t c = cands[cur + rank];
          const uint32_t leaf = (uint32_t)c;
          const uint64_t a = (c >> 32) & 0xFFFFu;
          strict = (c >> 48) & 1u;
          const uint64_t nn = base[leaf] + (a << mbi);
          const uint32_t f = meta_f(meta[leaf]);
          // number + a * 3^f (Lemma 2.3; collatz_cudav7.cu:262-278)
          const unsigned __int128 x = (unsigned __int128)a * pow3[f] + num[leaf];
          v.w[0] = (uint32_t)x;
          v.w[1] = (uint32_t)(x >> 32);
          v.w[2] = (uint32_t)(x >> 64);
          v.w[3] = (uint32_t)(x >> 96);
#pragma unroll
          for (int i = 4; i < L; ++i) v.w[i] = 0;
          n[0] = (uint32_t)nn;
          n[1] = (uint32_t)(nn >> 32);
          n[2] = n[3] = 0;
          checks = 0;
          act = true;
          cnt += 1;
        }
        cur += take;
      }
      const uint64_t running = __ballot(act);
      if (running == 0) break;
      wave_chk += (uint32_t)__popcll(running);
      if (act) {
        bool done = false;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
          odd_run<L>(v, p3s);
          if (over_fast<L, LOWCUT>(v, thr) && reaches<L>(v, cutoff, strict && half == 1)) {
            ov += 1;
            if (ovf) {
              const unsigned long long i = atomicAdd(&tot->overflow, 1ull);
              if (i < ovf_cap) {
                ovf[2 * i] = pair(n[0], n[1]);
                ovf[2 * i + 1] = pair(n[2], n[3]);
              }
            }
            v.w[0] = 2;
#pragma unroll
            for (int i = 1; i < L; ++i) v.w[i] = 0;
          }
          even_run<L>(v);
        }
        h ^= v.w[0];
        done = lt_base<L>(v, pair(n[0], n[1]), pair(n[2], n[3]));
        if (++checks >= lim.max_checks && !done) {
          stall += 1;
          done = true;
        }
        if (done) act = false;
      }
    }
    // One atomic per wave and tile.
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      h ^= __shfl_xor(h, m, 64);
      cnt += __shfl_xor(cnt, m, 64);
      stall += __shfl_xor(stall, m, 64);
      if (!ovf) ov += __shfl_xor(ov, m, 64);
    }
    if (lane == 0) {
      atomicXor(&tot->hash, h);
      atomicAdd(&tot->count, (unsigned long long)cnt);
      atomicAdd(&tot->checks, (unsigned long long)wave_chk);
      if (stall) atomicAdd(&tot->stall, (unsigned long long)stall);
      if (!ovf && ov) atomicAdd(&tot->overflow, (unsigned long long)ov);
      if (aborted) atomicAdd(&tot->aborts, 1ull);
    }
  }
}

// Steps 2 and 3 fused: persistent waves take tiles of leaves and keep a queue of
// candidates in LDS.  When the queue runs dry the wave expands the next 64 mask words of
// its tile at once (lane i takes word i: leaf = first + word / nw, its BV_i window AND mod 9
// row), counts their set bits, and lists every candidate (leaf, a) in order at its prefix
// offset; lanes that need a candidate take the next entries and load the leaf's base,
// number and f one refill ahead.  A candidate starts as number + a * 3^f with
// n = base + a * 2^MBI (Lemma 2.3) and runs up to U checks between refills; a lane whose
// candidate ends inside the U checks starts its prefetched one at once.  No candidate list
// is written to memory.  Hash, count and overflow count are folded with one atomic per
// wave and tile.
//
// L = 3 (96 bits, cutoff at or above 2^96) escalates instead of testing for overflow: a
// check multiplies by less than 2^9.4, so a candidate that starts every check below 2^86
// never reaches 2^96 (nor the cutoff) inside one; a candidate that ends a check above 2^86
// is listed as (leaf, a, CPU-leaf flag), its partial hash dropped, and rerun from its start
// by k_ref_iterate on two or three limbs.
constexpr uint32_t kRefQueue = 512;  // queue entries per wave: (leaf, a)

template <int L, bool LOWCUT>
__global__ __launch_bounds__(256) void k_ref_fused(RefTablesDev T, const uint64_t* base,
                                                   const uint64_t* num, const uint32_t* meta,
                                                   uint64_t nleaves, uint32_t mbi, Limits lim,
                                                   const uint64_t* pow3, uint32_t tile,
                                                   uint32_t U, RefTotals* tot, uint64_t* ovf,
                                                   uint64_t ovf_cap, uint64_t* esc,
                                                   uint64_t esc_cap) {
  __shared__ Pow3Lds p3s;
  __shared__ uint64_t p3f[kPow3Max];
  __shared__ uint2 queues[4][kRefQueue];
  stage_pow3(p3s);
  if (threadIdx.x < kPow3Max) p3f[threadIdx.x] = pow3[threadIdx.x];
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63;
  uint2* Q = queues[threadIdx.x >> 6];
  const uint64_t ntiles = (nleaves + tile - 1) / tile;
  const uint32_t cutoff = lim.cutoff;
  const uint64_t thr = lim.fast_thr;
  const uint32_t nw_log2 = __builtin_ctz(T.nw);
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(&tot->tile_counter, 1u);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= ntiles) break;
    const uint64_t first = (uint64_t)t * tile;
    const uint64_t end = first + tile < nleaves ? first + tile : nleaves;
    // Mask words of the tile, expanded 64 at a time; queue entries [qhead, qtail).
    const uint32_t words = (uint32_t)(end - first) << nw_log2;
    uint32_t wcur = 0, qhead = 0, qtail = 0;
    uint32_t h = 0, cnt = 0, ov = 0, stall = 0;
    uint32_t lane_chk = 0;  // checks of the lane's finished candidates
    // Prefetched candidate of the lane: its leaf (index, base, number, meta) and a.
    bool pf = false;
    uint32_t pf_leaf = 0, pf_a = 0, pf_meta = 0;
    uint64_t pf_lb = 0, pf_ln = 0;
    // Running candidate of the lane.
    bool act = false, strict = false;
    Num<L> v;
    uint64_t n_lo = 0;
    uint32_t checks = 0, ch = 0, cur_a = 0, cur_leaf = 0;
    auto start_prefetched = [&]() {
      pf = false;
      const uint32_t f = meta_f(pf_meta);
      // number + a * 3^f (Lemma 2.3; collatz_cudav7.cu:262-278)
      const unsigned __int128 x = (unsigned __int128)pf_a * p3f[f] + pf_ln;
      v.w[0] = (uint32_t)x;
      v.w[1] = (uint32_t)(x >> 32);
      v.w[2] = (uint32_t)(x >> 64);
      if constexpr (L > 3) v.w[3] = (uint32_t)(x >> 96);
#pragma unroll
      for (int i = 4; i < L; ++i) v.w[i] = 0;
      n_lo = pf_lb + ((uint64_t)pf_a << mbi);
      // Leaves without a bit vector run on the CPU in the reference (search2/search3,
      // collatz_cudav7.cu:517-519), whose second overflow test is '>' (:702).
      strict = (f - T.tpo) >= 16;
      cur_leaf = pf_leaf;
      cur_a = pf_a;
      checks = 0;
      ch = 0;
      act = true;
      cnt += 1;
      if constexpr (L == 3) {
        // A start at or above the 96-bit check bound goes to the wide pass at once
        // (a * 3^f < 2^76 and number < 2^64, so the start fits 96 bits).
        if (v.w[2] >= (1u << (kNarrowCheckLog2 - 64))) {
          const unsigned long long i = atomicAdd(&tot->esc_count, 1ull);
          if (i < esc_cap)
            esc[i] = cur_leaf | ((uint64_t)cur_a << 32) | (strict ? (1ull << 48) : 0ull);
          cnt -= 1;
          act = false;
        }
      }
    };
    // Iterations of the tile's loop (and of the take loop per refill), uniform: a wave
    // that reaches the bound abandons the tile; the abort is counted and the call fails.
    uint32_t iters = 0;
    bool aborted = false;
    for (;;) {
      if (__builtin_expect(++iters > lim.max_loop, 0)) {
        aborted = true;
        break;
      }
      // (1) Lanes without a candidate start their prefetched one.
      if (!act && pf) start_prefetched();
      // (2) Lanes without a prefetch take the next queue entries; an empty queue is
      //     refilled from the next 64 mask words of the tile.
      const uint64_t need = __ballot(!pf);
      if (need) {
        uint32_t want = __popcll(need);
        uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
        uint32_t takes = 0;
        for (;;) {
          if (__builtin_expect(++takes > lim.max_loop, 0)) {
            aborted = true;
            break;
          }
          const uint32_t avail = qtail - qhead;
          const uint32_t take = avail < want ? avail : want;
          if (!pf) {
            if (rank < take) {
              const uint2 e = Q[qhead + rank];
              pf_leaf = e.x;
              pf_a = e.y;
              pf_lb = base[e.x];
              pf_ln = num[e.x];
              pf_meta = meta[e.x];
              pf = true;
            } else {
              rank -= take;
            }
          }
          qhead += take;
          want -= take;
          if (want == 0 || wcur >= words) break;
          // The queue is empty here: expand words [wcur, wcur + 64).
          const uint32_t wi = wcur + lane;
          uint32_t m = 0, leaf = 0, w = 0;
          if (wi < words) {
            leaf = (uint32_t)first + (wi >> nw_log2);
            w = wi & (T.nw - 1u);
            m = leaf_mask_word(T, base[leaf], num[leaf], meta[leaf], w);
          }
          const uint32_t pc = __popc(m);
          // Inclusive prefix sum of the counts across the wave.
          uint32_t incl = pc;
#pragma unroll
          for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(incl, d, 64);
            if (lane >= (uint32_t)d) incl += y;
          }
          // Words whose entries fit the queue (a prefix of the lanes: at least 16 of them,
          // a word holds at most 32 entries) are listed now; the rest wait for the next
          // expansion.
          const bool fits = incl <= kRefQueue;
          const uint64_t fm = __ballot(fits);
          const uint32_t nfit = __popcll(fm);
          if (fits) {
            uint32_t pos = incl - pc;
            while (m) {
              const uint32_t j = __builtin_ctz(m);
              m &= m - 1u;
              Q[pos++] = make_uint2(leaf, 32u * w + j);
            }
          }
          // A wave's LDS operations execute in order; only the compiler must keep the
          // lanes' reads below after these writes.
          __atomic_signal_fence(__ATOMIC_SEQ_CST);
          qhead = 0;
          qtail = __shfl(incl, nfit - 1u, 64);
          wcur += nfit;
        }
      }
      if (__ballot(act || pf) == 0) break;
      // A candidate past the check bound is abandoned (tested once per refill: it may run
      // up to U checks beyond the bound first).
      if (act && __builtin_expect(checks >= lim.max_checks, 0)) {
        stall += 1;
        lane_chk += checks;
        act = false;
      }
      // (3) Up to U checks; a lane whose candidate ends before the last of them starts its
      //     prefetched candidate at once.
      for (uint32_t u = 0; u < U; ++u) {
        if (act) {
#pragma unroll
          for (int half = 0; half < 2; ++half) {
            odd_run<L>(v, p3s);
            if constexpr (L > 3) {
              if (over_fast<L, LOWCUT>(v, thr) && reaches<L>(v, cutoff, strict && half == 1)) {
                ov += 1;
                if (ovf) {
                  const unsigned long long i = atomicAdd(&tot->overflow, 1ull);
                  if (i < ovf_cap) {
                    ovf[2 * i] = n_lo;
                    ovf[2 * i + 1] = 0;
                  }
                }
                v.w[0] = 2;
#pragma unroll
                for (int i = 1; i < L; ++i) v.w[i] = 0;
              }
            }
            even_run<L>(v);
          }
          ch ^= v.w[0];
          bool done = lt_base<L>(v, n_lo, 0);
          bool wide = false;
          if constexpr (L == 3) wide = v.w[2] >= (1u << (kNarrowCheckLog2 - 64));
          ++checks;
          if (done || wide) {
            lane_chk += checks;
            if (wide) {
              const unsigned long long i = atomicAdd(&tot->esc_count, 1ull);
              if (i < esc_cap)
                esc[i] = cur_leaf | ((uint64_t)cur_a << 32) | (strict ? (1ull << 48) : 0ull);
              cnt -= 1;
            } else {
              h ^= ch;
            }
            act = false;
          }
        }
        if (u + 1 < U && !act && pf) start_prefetched();
      }
    }
    // One atomic per wave and tile.
#pragma unroll
    for (int k = 32; k >= 1; k >>= 1) {
      h ^= __shfl_xor(h, k, 64);
      cnt += __shfl_xor(cnt, k, 64);
      stall += __shfl_xor(stall, k, 64);
      lane_chk += __shfl_xor(lane_chk, k, 64);
      if (!ovf) ov += __shfl_xor(ov, k, 64);
    }
    if (lane == 0) {
      atomicXor(&tot->hash, h);
      atomicAdd(&tot->count, (unsigned long long)cnt);
      atomicAdd(&tot->checks, (unsigned long long)lane_chk);
      if (stall) atomicAdd(&tot->stall, (unsigned long long)stall);
      if (!ovf && ov) atomicAdd(&tot->overflow, (unsigned long long)ov);
      if (aborted) atomicAdd(&tot->aborts, 1ull);
    }
  }
}

static inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

hipError_t launch_bv_build(uint32_t B, uint32_t num_bv, int32_t first, uint32_t sem,
                           const uint32_t* magic, uint32_t* bv, uint64_t words_per_bv,
                           uint32_t mirror_words, hipStream_t s) {
  hipLaunchKernelGGL(k_bv_build, dim3(blocks_for(1ull << B, 256)), dim3(256), 0, s, B, num_bv,
                     first, sem, magic, bv, words_per_bv, mirror_words);
  return hipGetLastError();
}

hipError_t launch_la_build(uint32_t B, uint32_t k, uint32_t sem, uint32_t* table,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_la_build, dim3(blocks_for(1ull << (B - 3), 256)), dim3(256), 0, s, B, k,
                     sem, table);
  return hipGetLastError();
}

hipError_t launch_leaf_bins(const uint32_t* meta, uint64_t n, uint32_t tpo,
                            unsigned long long* bins, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_leaf_bins, dim3(512), dim3(256), 0, s, meta, n, tpo, bins);
  return hipGetLastError();
}

// limbs = 2 requires cutoff <= 123 (no escalation in this mode); three limbs otherwise.
hipError_t launch_ref_iterate(int limbs, const uint64_t* cands, uint64_t ncand,
                              const uint64_t* base, const uint64_t* num, const uint32_t* meta,
                              uint32_t mbi, const Limits& lim, const uint64_t* pow3, uint32_t tile,
                              uint32_t nblocks, void* tot, uint64_t* ovf, uint64_t ovf_cap,
                              hipStream_t s) {
  if (!ncand) return hipSuccess;
  if (limbs == 2)
    hipLaunchKernelGGL((k_ref_iterate<4, false>), dim3(nblocks), dim3(256), 0, s, cands, ncand,
                       base, num, meta, mbi, lim, pow3, tile, (RefTotals*)tot, ovf, ovf_cap);
  else if (lim.cutoff >= 128)
    hipLaunchKernelGGL((k_ref_iterate<6, false>), dim3(nblocks), dim3(256), 0, s, cands, ncand,
                       base, num, meta, mbi, lim, pow3, tile, (RefTotals*)tot, ovf, ovf_cap);
  else
    hipLaunchKernelGGL((k_ref_iterate<6, true>), dim3(nblocks), dim3(256), 0, s, cands, ncand,
                       base, num, meta, mbi, lim, pow3, tile, (RefTotals*)tot, ovf, ovf_cap);
  return hipGetLastError();
}

hipError_t launch_ref_fused(int limbs, const RefTablesDev& T, const uint64_t* base,
                            const uint64_t* num, const uint32_t* meta, uint64_t nleaves,
                            uint32_t mbi, const Limits& lim, const uint64_t* pow3, uint32_t tile,
                            uint32_t U, uint32_t nblocks, void* tot, uint64_t* ovf,
                            uint64_t ovf_cap, uint64_t* esc, uint64_t esc_cap, hipStream_t s) {
  if (!nleaves) return hipSuccess;
  if (limbs == 1)
    hipLaunchKernelGGL((k_ref_fused<3, false>), dim3(nblocks), dim3(256), 0, s, T, base, num,
                       meta, nleaves, mbi, lim, pow3, tile, U, (RefTotals*)tot, ovf, ovf_cap,
                       esc, esc_cap);
  else if (limbs == 2)
    hipLaunchKernelGGL((k_ref_fused<4, false>), dim3(nblocks), dim3(256), 0, s, T, base, num,
                       meta, nleaves, mbi, lim, pow3, tile, U, (RefTotals*)tot, ovf, ovf_cap,
                       esc, 0);
  else if (lim.cutoff >= 128)
    hipLaunchKernelGGL((k_ref_fused<6, false>), dim3(nblocks), dim3(256), 0, s, T, base, num,
                       meta, nleaves, mbi, lim, pow3, tile, U, (RefTotals*)tot, ovf, ovf_cap,
                       esc, 0);
  else
    hipLaunchKernelGGL((k_ref_fused<6, true>), dim3(nblocks), dim3(256), 0, s, T, base, num,
                       meta, nleaves, mbi, lim, pow3, tile, U, (RefTotals*)tot, ovf, ovf_cap,
                       esc, 0);
  return hipGetLastError();
}

}  // namespace cz
