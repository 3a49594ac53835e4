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
//                    of the candidates that left 96 bits in k_ref_fused, and of the
//                    candidates of spot-checked tiles.
//   k_ref_spot_list  the candidate list of the tiles a k_ref_fused launch sampled for a spot
//                    check (one thread per mask word, no queue).
#include <algorithm>

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
  unsigned long long numcases;   // main.rs semantics: 16-step blocks of finished candidates
  // Spot checks of a fused launch, set by the host: tiles t with t & spot_mask == spot_off
  // and t <= spot_last write their own results to spot_rec[t >> spot_shift] (spot_rec 0:
  // none).  Read from memory at the end of a tile, so they take no registers in the loop.
  unsigned long long spot_rec;
  unsigned int spot_mask, spot_off, spot_last, spot_shift;
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

// main.rs's search3 (main.rs:321-341) takes 16 steps of the shifted map per table lookup
// and tests the value after each such block.  Here a block is two 8-step steps from a
// 256-entry table in LDS: for r = number mod 2^8, entry r holds 3^t in bits 0-12 and
// S^8(r) above (t the odd steps among the 8, S the shifted map), and
// S^8(2^8 h + r) = 3^t h + S^8(r) gives number = (number >> 8) * 3^t + S^8(r).
struct Step8Lds {
  uint32_t e[256];
};

__device__ __forceinline__ void stage_step8(Step8Lds& t) {
  for (uint32_t r = threadIdx.x; r < 256; r += blockDim.x) {
    uint32_t n = r, m = 1;
    for (int i = 0; i < 8; ++i) {
      if ((n & 1u) == 0) {
        n = n + (n >> 1);
        m *= 3;
      } else {
        n = (n >> 1) + 1;
      }
    }
    t.e[r] = m | (n << 13);
  }
}

template <int L>
__device__ __forceinline__ void step8(Num<L>& v, const Step8Lds& T) {
  const uint32_t e = T.e[v.w[0] & 255u];
  const uint32_t m = e & 0x1FFFu, add = e >> 13;
  uint64_t p[L];
#pragma unroll
  for (int i = 0; i < L - 1; ++i) p[i] = (uint64_t)alignr(v.w[i + 1], v.w[i], 8) * m;
  p[L - 1] = (uint64_t)(v.w[L - 1] >> 8) * m;
  uint32_t c = 0;
  v.w[0] = addc((uint32_t)p[0], add, 0u, c);
#pragma unroll
  for (int i = 1; i < L; ++i) v.w[i] = addc((uint32_t)p[i], (uint32_t)(p[i - 1] >> 32), c, c);
}

// Start value number + a * 3^f of candidate a of a leaf (Lemma 2.3; collatz_cudav7.cu:262-278,
// main.rs:314), with 3^f from a table of (low, high) words: up to 2^111 at N - A = 62.
__device__ __forceinline__ unsigned __int128 ref_start(uint64_t num_lo, uint64_t num_hi,
                                                       uint64_t a, const uint64_t* p3w,
                                                       uint32_t f) {
  const unsigned __int128 p3 =
      (unsigned __int128)p3w[2 * f] | ((unsigned __int128)p3w[2 * f + 1] << 64);
  return (unsigned __int128)a * p3 + ((unsigned __int128)num_hi << 64 | num_lo);
}

template <int L>
__device__ __forceinline__ void set_words(Num<L>& v, unsigned __int128 x) {
  v.w[0] = (uint32_t)x;
  v.w[1] = (uint32_t)(x >> 32);
  v.w[2] = (uint32_t)(x >> 64);
  if constexpr (L > 3) v.w[3] = (uint32_t)(x >> 96);
#pragma unroll
  for (int i = 4; i < L; ++i) v.w[i] = 0;
}

// Step 3 over a list of (leaf, a) candidates on L words: the two- or three-limb rerun of the
// candidates that left 96 bits in k_ref_fused.  RUST selects main.rs's blocks (a check is
// one 16-step block; a value at or above 2^cutoff after a block stops the candidate and
// reports its start) instead of the reference schedule (a check is two odd and two even
// runs; a value at 2^cutoff is reported and replaced by 2).
template <int L, bool LOWCUT, bool RUST>
__global__ __launch_bounds__(256) void k_ref_iterate(const uint64_t* cands, uint64_t ncand,
                                                     RefLeaves lv, uint32_t mbi, Limits lim,
                                                     const uint64_t* pow3w, uint32_t tile,
                                                     RefTotals* tot, uint64_t* ovf,
                                                     uint64_t ovf_cap) {
  __shared__ Pow3Lds p3s;
  __shared__ Step8Lds st8;
  stage_pow3(p3s);
  if constexpr (RUST) stage_step8(st8);
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
    uint32_t h = 0, cnt = 0, ov = 0, stall = 0, nc = 0;
    bool act = false, strict = false;
    Num<L> v;
    uint64_t n_lo = 0;
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
          n_lo = lv.base[leaf] + (a << mbi);
          set_words<L>(v, ref_start(lv.num[leaf], lv.num_hi ? lv.num_hi[leaf] : 0, a, pow3w,
                                    meta_f(lv.meta[leaf])));
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
        if constexpr (RUST) {
          step8<L>(v, st8);
          step8<L>(v, st8);
          ++checks;
          done = lt_base<L>(v, n_lo, 0);
          if (!done && ge_pow2<L>(v, cutoff)) {
            // main.rs:336-338: the start is printed and the candidate stops.
            ov += 1;
            if (ovf) {
              const unsigned long long i = atomicAdd(&tot->overflow, 1ull);
              if (i < ovf_cap) {
                ovf[2 * i] = n_lo;
                ovf[2 * i + 1] = 0;
              }
            }
            done = true;
          }
          if (done) nc += checks;
        } else {
#pragma unroll
          for (int half = 0; half < 2; ++half) {
            odd_run<L>(v, p3s);
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
            even_run<L>(v);
          }
          h ^= v.w[0];
          done = lt_base<L>(v, n_lo, 0);
          ++checks;
        }
        if (checks >= lim.max_checks && !done) {
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
      nc += __shfl_xor(nc, m, 64);
      if (!ovf) ov += __shfl_xor(ov, m, 64);
    }
    if (lane == 0) {
      atomicXor(&tot->hash, h);
      atomicAdd(&tot->count, (unsigned long long)cnt);
      atomicAdd(&tot->checks, (unsigned long long)wave_chk);
      if (RUST) atomicAdd(&tot->numcases, (unsigned long long)nc);
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
// is written to memory.  Hash (reference schedule) or block tally (main.rs's, RUST), count
// and overflow count are folded with one atomic per wave and tile.
//
// L = 3 (96 bits) escalates instead of testing every value: a check multiplies by less than
// 2^9.4, so a candidate that starts every check below 2^86 never leaves 96 bits inside one;
// a start or a check end at or above 2^86 is listed as (leaf, a, CPU-leaf flag), its partial
// results dropped, and rerun from its start by k_ref_iterate on two or three limbs.  The
// reference schedule runs on 96 bits only with a cutoff at or above 2^96 (no overflow test
// below it); main.rs's blocks test their cutoff on 96 bits when it is below 2^86.
constexpr uint32_t kRefQueue = 512;  // queue entries per wave: (leaf, a)

// WIDE: leaf numbers or 3^f may exceed 64 bits (N - A > 40); otherwise a start is
// number + a 3^f from 64-bit words, with no high words loaded.
// BOUNDED: the tile loop and the take loop carry the host-set iteration bound
// (Limits::max_loop), only when a bound below the default is set (COLLATZ_LOOP_LIMIT): its
// counters take scalar registers that the kernel needs for eight waves per SIMD.  Without
// it the loops end: every iteration runs a check on each running candidate (whose check
// count ends it at the check bound), starts prefetched candidates, takes queue entries
// (the queue head advances), expands mask words (the word cursor advances by at least 16),
// or leaves; the exit tests are ballots (uniform).
// Eight waves per SIMD (the scalar registers this asks for are what the kernel had before
// its deep-leaf, main.rs and spot-check paths; without the limit it took seven).
template <int L, bool LOWCUT, bool RUST, bool WIDE, bool BOUNDED, int UC = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(L == 3 && !RUST ? 8 : 1))) void k_ref_fused(RefTablesDev T, RefLeaves lv,
                                                   uint64_t nleaves, uint32_t mbi, Limits lim,
                                                   const uint64_t* pow3w, uint32_t tile,
                                                   uint32_t U, RefTotals* tot, uint64_t* ovf,
                                                   uint64_t ovf_cap, uint64_t* esc,
                                                   uint64_t esc_cap) {
  __shared__ Pow3Lds p3s;
  __shared__ Step8Lds st8;
  __shared__ uint64_t p3w[2 * kPow3Max];
  __shared__ uint2 queues[4][kRefQueue];
  stage_pow3(p3s);
  if constexpr (RUST) stage_step8(st8);
  if (threadIdx.x < 2 * kPow3Max) p3w[threadIdx.x] = pow3w[threadIdx.x];
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63;
  uint2* Q = queues[threadIdx.x >> 6];
  const uint64_t ntiles = (nleaves + tile - 1) / tile;
  const uint32_t cutoff = lim.cutoff;
  const uint64_t thr = lim.fast_thr;
  const uint32_t nw_log2 = __builtin_ctz(T.nw);
  // main.rs's cutoff is tested on 96 bits only below the escalation bound.
  const bool rust_cut96 = RUST && L == 3 && cutoff < kNarrowCheckLog2;
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
    uint32_t h = 0, cnt = 0, ov = 0, stall = 0, nc = 0;
    uint32_t lane_chk = 0;  // checks of the lane's finished candidates
    // Prefetched candidate of the lane: its leaf (index, base, number, meta) and a, loaded
    // one refill ahead.
    bool pf = false;
    uint32_t pf_leaf = 0, pf_a = 0, pf_meta = 0;
    uint64_t pf_lb = 0, pf_ln = 0, pf_lnh = 0;
    // Converted candidate of the lane, ready to start: the prefetched one's start value, n,
    // leaf and a (bit 31 of ps_a: the strict overflow compare; a < 2^12).  A refill converts
    // the prefetched candidate of each lane that has none, one refill after its loads were
    // issued, so a start within the window is a few register moves for the lanes that start
    // (with per-lane flags, the code of a start runs whenever any lane starts one): the tree
    // below 2^48 1.72 -> 1.59 s.  main.rs's blocks (RUST) start the prefetched candidate
    // directly: there the converted one's registers cost more than the moves save (64-66
    // VGPRs, 2% slower).
    constexpr bool kConvert = !RUST;
    bool ps = false;
    Num<L> ps_v;
    uint64_t ps_n = 0;
    uint32_t ps_leaf = 0, ps_a = 0;
    // Running candidate of the lane.
    bool act = false, strict = false;
    Num<L> v;
    uint64_t n_lo = 0;
    uint32_t checks = 0, ch = 0, cur_a = 0, cur_leaf = 0;
    auto push_esc = [&](uint32_t leaf, uint32_t a, bool st) {
      const unsigned long long i = atomicAdd(&tot->esc_count, 1ull);
      if (i < esc_cap) esc[i] = leaf | ((uint64_t)a << 32) | (st ? (1ull << 48) : 0ull);
    };
    auto escalate = [&]() {
      push_esc(cur_leaf, cur_a, strict);
      cnt -= 1;
      act = false;
    };
    auto start_prefetched = [&]() {
      pf = false;
      const uint32_t f = meta_f(pf_meta);
      const unsigned __int128 x = WIDE ? ref_start(pf_ln, pf_lnh, pf_a, p3w, f)
                                       : (unsigned __int128)pf_a * p3w[2 * f] + pf_ln;
      set_words<L>(v, x);
      n_lo = pf_lb + ((uint64_t)pf_a << mbi);
      strict = (f - T.tpo) >= 16;
      cur_leaf = pf_leaf;
      cur_a = pf_a;
      checks = 0;
      ch = 0;
      act = true;
      cnt += 1;
      if constexpr (L == 3) {
        if ((x >> kNarrowCheckLog2) != 0) escalate();
      }
    };
    auto convert_prefetched = [&]() {
      pf = false;
      const uint32_t f = meta_f(pf_meta);
      const unsigned __int128 x = WIDE ? ref_start(pf_ln, pf_lnh, pf_a, p3w, f)
                                       : (unsigned __int128)pf_a * p3w[2 * f] + pf_ln;
      // Leaves without a bit vector run on the CPU in the reference (search2/search3,
      // collatz_cudav7.cu:517-519), whose second overflow test is '>' (:702).
      const bool st = (f - T.tpo) >= 16;
      // A start at or above the 96-bit check bound goes to the wide pass at once (it is not
      // counted here: the wide pass counts it).
      if constexpr (L == 3) {
        if ((x >> kNarrowCheckLog2) != 0) {
          push_esc(pf_leaf, pf_a, st);
          return;
        }
      }
      set_words<L>(ps_v, x);
      ps_n = pf_lb + ((uint64_t)pf_a << mbi);
      ps_leaf = pf_leaf;
      ps_a = pf_a | (st ? 0x80000000u : 0u);
      ps = true;
    };
    auto start_converted = [&]() {
      ps = false;
      v = ps_v;
      n_lo = ps_n;
      cur_leaf = ps_leaf;
      cur_a = ps_a & 0x7FFFFFFFu;
      strict = (ps_a >> 31) != 0;
      checks = 0;
      ch = 0;
      act = true;
      cnt += 1;
    };
    // Iterations of the tile's loop (and of the take loop per refill), uniform: a wave
    // that reaches the bound abandons the tile; the abort is counted and the call fails.
    uint32_t iters = 0;
    bool aborted = false;
    for (;;) {
      if (BOUNDED && __builtin_expect(++iters > lim.max_loop, 0)) {
        aborted = true;
        break;
      }
      // (1) Lanes without a candidate start their converted one; lanes without a converted
      //     one convert their prefetched one, and those still without a candidate (at the
      //     start of a tile, or after two candidates ended within the last window) start it.
      if constexpr (kConvert) {
        if (!act && ps) start_converted();
        if (!ps && pf) convert_prefetched();
        if (!act && ps) start_converted();
      } else {
        if (!act && pf) start_prefetched();
      }
      // (2) Lanes without a prefetch take the next queue entries; an empty queue is
      //     refilled from the next 64 mask words of the tile.
      const uint64_t need = __ballot(!pf);
      if (need) {
        uint32_t want = __popcll(need);
        uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
        uint32_t takes = 0;
        for (;;) {
          if (BOUNDED && __builtin_expect(++takes > lim.max_loop, 0)) {
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
              pf_lb = lv.base[e.x];
              pf_ln = lv.num[e.x];
              if constexpr (WIDE) pf_lnh = lv.num_hi ? lv.num_hi[e.x] : 0;
              pf_meta = lv.meta[e.x];
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
            m = leaf_mask_word(T, lv.base[leaf], lv.num[leaf], lv.meta[leaf], w);
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
      if (__ballot(act || ps || pf) == 0) break;
      // A candidate past the check bound is abandoned (tested once per refill: it may run
      // up to U checks beyond the bound first).
      if (act && __builtin_expect(checks >= lim.max_checks, 0)) {
        stall += 1;
        lane_chk += checks;
        act = false;
      }
      // (3) Up to U checks; a lane whose candidate ends before the last of them starts its
      //     prefetched candidate at once.
      const uint32_t Un = UC > 0 ? (uint32_t)UC : U;  // (a constant when UC is set)
#pragma unroll
      for (uint32_t u = 0; u < Un; ++u) {
        if (act) {
          bool done, stop = false;
          if constexpr (RUST) {
            // One block of main.rs's search3: 16 steps, then the two tests (main.rs:333-339).
            step8<L>(v, st8);
            step8<L>(v, st8);
            done = lt_base<L>(v, n_lo, 0);
            if (!done && (L > 3 || rust_cut96) && ge_pow2<L>(v, cutoff)) {
              ov += 1;
              if (ovf) {
                const unsigned long long i = atomicAdd(&tot->overflow, 1ull);
                if (i < ovf_cap) {
                  ovf[2 * i] = n_lo;
                  ovf[2 * i + 1] = 0;
                }
              }
              stop = true;
            }
          } else {
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
            done = lt_base<L>(v, n_lo, 0);
          }
          bool wide = false;
          if constexpr (L == 3) wide = !done && !stop && v.w[2] >= (1u << (kNarrowCheckLog2 - 64));
          ++checks;
          if (done || stop || wide) {
            lane_chk += checks;
            if (wide) {
              escalate();
            } else {
              if constexpr (RUST) nc += checks;
              else h ^= ch;
              act = false;
            }
          }
        }
        if constexpr (kConvert) {
          if (u + 1 < Un && !act && ps) start_converted();
        } else {
          if (u + 1 < Un && !act && pf) start_prefetched();
        }
      }
    }
    // A sampled tile (spot check) keeps its own results as well.  (t is uniform.)
    RefSpotRec* const spot_rec = (RefSpotRec*)tot->spot_rec;
    const bool spot_tile = spot_rec != nullptr && (t & tot->spot_mask) == tot->spot_off &&
                           t <= tot->spot_last;
    // One atomic per wave and tile.
#pragma unroll
    for (int k = 32; k >= 1; k >>= 1) {
      h ^= __shfl_xor(h, k, 64);
      cnt += __shfl_xor(cnt, k, 64);
      stall += __shfl_xor(stall, k, 64);
      lane_chk += __shfl_xor(lane_chk, k, 64);
      nc += __shfl_xor(nc, k, 64);
      if (!ovf || spot_tile) ov += __shfl_xor(ov, k, 64);
    }
    if (lane == 0) {
      atomicXor(&tot->hash, h);
      atomicAdd(&tot->count, (unsigned long long)cnt);
      atomicAdd(&tot->checks, (unsigned long long)lane_chk);
      if (RUST) atomicAdd(&tot->numcases, (unsigned long long)nc);
      if (stall) atomicAdd(&tot->stall, (unsigned long long)stall);
      if (!ovf && ov) atomicAdd(&tot->overflow, (unsigned long long)ov);
      if (aborted) atomicAdd(&tot->aborts, 1ull);
      if (spot_tile) {
        RefSpotRec r;
        r.hash = h;
        r.count = cnt;
        r.numcases = nc;
        r.overflow = ov;
        spot_rec[t >> tot->spot_shift] = r;
      }
    }
  }
}

// The candidates of the sampled tiles of a fused launch, in the entry format of the
// escalation list (leaf | a << 32 | no-bit-vector << 48), for a rerun by k_ref_iterate: one
// thread per (sampled tile, leaf, mask word), the words computed as k_ref_fused computes
// them.  Entries past cap are counted but not written.
__global__ __launch_bounds__(256) void k_ref_spot_list(RefTablesDev T, RefLeaves lv,
                                                       uint64_t nleaves, uint32_t tile,
                                                       RefSpot spot, uint64_t* list,
                                                       unsigned long long* count, uint64_t cap) {
  const uint64_t per_tile = (uint64_t)tile * T.nw;
  const uint64_t total = per_tile * spot.cap;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t j = g / per_tile;
    const uint64_t r = g - j * per_tile;
    const uint64_t t = spot.off + j * spot.freq;
    const uint64_t leaf = t * tile + r / T.nw;
    if (leaf >= nleaves) continue;
    const uint32_t w = (uint32_t)(r % T.nw);
    const uint32_t meta = lv.meta[leaf];
    uint32_t m = leaf_mask_word(T, lv.base[leaf], lv.num[leaf], meta, w);
    const uint64_t strict = (meta_f(meta) - T.tpo) >= 16 ? 1ull : 0ull;
    while (m) {
      const uint32_t b = __builtin_ctz(m);
      m &= m - 1u;
      const unsigned long long i = atomicAdd(count, 1ull);
      if (i < cap) list[i] = leaf | ((uint64_t)(32u * w + b) << 32) | (strict << 48);
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

template <bool RUST>
static void launch_iterate_sched(int limbs, const uint64_t* cands, uint64_t ncand,
                                 const RefLeaves& lv, uint32_t mbi, const Limits& lim,
                                 const uint64_t* pow3w, uint32_t tile, uint32_t nblocks,
                                 RefTotals* tot, uint64_t* ovf, uint64_t ovf_cap, hipStream_t s) {
  if (limbs == 2)
    hipLaunchKernelGGL((k_ref_iterate<4, false, RUST>), dim3(nblocks), dim3(256), 0, s, cands,
                       ncand, lv, mbi, lim, pow3w, tile, tot, ovf, ovf_cap);
  else if (lim.cutoff >= 128)
    hipLaunchKernelGGL((k_ref_iterate<6, false, RUST>), dim3(nblocks), dim3(256), 0, s, cands,
                       ncand, lv, mbi, lim, pow3w, tile, tot, ovf, ovf_cap);
  else
    hipLaunchKernelGGL((k_ref_iterate<6, true, RUST>), dim3(nblocks), dim3(256), 0, s, cands,
                       ncand, lv, mbi, lim, pow3w, tile, tot, ovf, ovf_cap);
}

// limbs = 2 requires cutoff <= 123 for the reference schedule and cutoff <= 118 for main.rs's
// blocks (a block multiplies by less than 2^9.4); three limbs otherwise.
hipError_t launch_ref_iterate(int limbs, bool rust, const uint64_t* cands, uint64_t ncand,
                              const RefLeaves& lv, uint32_t mbi, const Limits& lim,
                              const uint64_t* pow3w, uint32_t tile, uint32_t nblocks, void* tot,
                              uint64_t* ovf, uint64_t ovf_cap, hipStream_t s) {
  if (!ncand) return hipSuccess;
  if (rust)
    launch_iterate_sched<true>(limbs, cands, ncand, lv, mbi, lim, pow3w, tile, nblocks,
                               (RefTotals*)tot, ovf, ovf_cap, s);
  else
    launch_iterate_sched<false>(limbs, cands, ncand, lv, mbi, lim, pow3w, tile, nblocks,
                                (RefTotals*)tot, ovf, ovf_cap, s);
  return hipGetLastError();
}

template <bool RUST, bool WIDE, bool BOUNDED>
static void launch_fused_sched(int limbs, const RefTablesDev& T, const RefLeaves& lv,
                               uint64_t nleaves, uint32_t mbi, const Limits& lim,
                               const uint64_t* pow3w, uint32_t tile, uint32_t U,
                               uint32_t nblocks, RefTotals* tot, uint64_t* ovf,
                               uint64_t ovf_cap, uint64_t* esc, uint64_t esc_cap,
                               hipStream_t s) {
  const dim3 grid(nblocks), block(256);
  // On 96 bits the checks between refills are compiled in for the launches' usual five.
  if (limbs == 1 && U == kRefChecksPerRefill)
    hipLaunchKernelGGL((k_ref_fused<3, false, RUST, WIDE, BOUNDED, kRefChecksPerRefill>), grid,
                       block, 0, s, T, lv, nleaves, mbi, lim, pow3w, tile, U, tot, ovf, ovf_cap,
                       esc, esc_cap);
  else if (limbs == 1)
    hipLaunchKernelGGL((k_ref_fused<3, false, RUST, WIDE, BOUNDED>), grid, block, 0, s, T, lv,
                       nleaves, mbi, lim, pow3w, tile, U, tot, ovf, ovf_cap, esc, esc_cap);
  else if (limbs == 2)
    hipLaunchKernelGGL((k_ref_fused<4, false, RUST, WIDE, BOUNDED>), grid, block, 0, s, T, lv,
                       nleaves, mbi, lim, pow3w, tile, U, tot, ovf, ovf_cap, esc, 0);
  else if (lim.cutoff >= 128)
    hipLaunchKernelGGL((k_ref_fused<6, false, RUST, WIDE, BOUNDED>), grid, block, 0, s, T, lv,
                       nleaves, mbi, lim, pow3w, tile, U, tot, ovf, ovf_cap, esc, 0);
  else
    hipLaunchKernelGGL((k_ref_fused<6, true, RUST, WIDE, BOUNDED>), grid, block, 0, s, T, lv,
                       nleaves, mbi, lim, pow3w, tile, U, tot, ovf, ovf_cap, esc, 0);
}

hipError_t launch_ref_fused(int limbs, bool rust, const RefTablesDev& T, const RefLeaves& lv,
                            uint64_t nleaves, uint32_t mbi, const Limits& lim,
                            const uint64_t* pow3w, uint32_t tile, uint32_t U, uint32_t nblocks,
                            void* tot, uint64_t* ovf, uint64_t ovf_cap, uint64_t* esc,
                            uint64_t esc_cap, hipStream_t s) {
  if (!nleaves) return hipSuccess;
  RefTotals* t = (RefTotals*)tot;
  const int wide = lv.num_hi != nullptr || mbi > kWalk64MaxDepth ? 1 : 0;
  const int bounded = lim.max_loop < kMaxLoop ? 1 : 0;
  const int sel = (rust ? 4 : 0) | (wide ? 2 : 0) | bounded;
#define CZ_FUSED(R, W, B)                                                                 \
  launch_fused_sched<R, W, B>(limbs, T, lv, nleaves, mbi, lim, pow3w, tile, U, nblocks, t, \
                              ovf, ovf_cap, esc, esc_cap, s)
  switch (sel) {
    case 0: CZ_FUSED(false, false, false); break;
    case 1: CZ_FUSED(false, false, true); break;
    case 2: CZ_FUSED(false, true, false); break;
    case 3: CZ_FUSED(false, true, true); break;
    case 4: CZ_FUSED(true, false, false); break;
    case 5: CZ_FUSED(true, false, true); break;
    case 6: CZ_FUSED(true, true, false); break;
    default: CZ_FUSED(true, true, true); break;
  }
#undef CZ_FUSED
  return hipGetLastError();
}

hipError_t launch_ref_spot_list(const RefTablesDev& T, const RefLeaves& lv, uint64_t nleaves,
                                uint32_t tile, const RefSpot& spot, uint64_t* list,
                                unsigned long long* count, uint64_t cap, hipStream_t s) {
  if (!spot.cap || !nleaves) return hipSuccess;
  const uint64_t threads = (uint64_t)tile * T.nw * spot.cap;
  const uint32_t nb = (uint32_t)std::min<uint64_t>(4096, (threads + 255) / 256);
  hipLaunchKernelGGL(k_ref_spot_list, dim3(nb), dim3(256), 0, s, T, lv, nleaves, tile, spot,
                     list, count, cap);
  return hipGetLastError();
}

}  // namespace cz
