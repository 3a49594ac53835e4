// Range-verification kernels: the plain one-candidate-per-lane kernel, the three-limb pass
// for escalated candidates, and the per-chunk reduction.
//
// A candidate is n = (q << k) | r with r a sieve survivor; its start value is
// number = T^k(r) + 1 + q * 3^f (Lemma 2.3, paper p.4; the jump of
// collatz_cudav7.cu:262-278), and n mod 9 must be in {0,1,3,6,7}
// (collatz_cudav7.cu:482-494).  The schedule is in schedule.hpp.
#include <atomic>

#include "schedule.hpp"

namespace cz {

// q * 3^f for one q and f < 41, as six words; `big` tells whether it needs more than
// L words.
__device__ __forceinline__ void q_times_pow3(uint64_t q_lo, uint64_t q_hi, uint64_t p,
                                             uint32_t* out6) {
  const unsigned __int128 a = (unsigned __int128)q_lo * p;
  const unsigned __int128 b = (unsigned __int128)q_hi * p + (uint64_t)(a >> 64);
  out6[0] = (uint32_t)a;
  out6[1] = (uint32_t)((uint64_t)a >> 32);
  out6[2] = (uint32_t)b;
  out6[3] = (uint32_t)((uint64_t)b >> 32);
  out6[4] = (uint32_t)(b >> 64);
  out6[5] = (uint32_t)(b >> 96);
}

// Look-ahead filter of a sieve that carries one (collatz_cudav7.cu:219-235 at depth k):
// keep the candidate iff the dip of number mod 2^B is at most 485 f - 306 k, read from the
// nibble table as f - e_min >= g.  num_lo is the low word of the start value number.
__device__ __forceinline__ bool la_keep(const SieveArrays& sv, uint32_t f, uint32_t num_lo) {
  const uint32_t fe = f - sv.la_emin;
  if (sv.la1) {
    const uint32_t i1 = num_lo & sv.la1_mask;
    if (fe < ((sv.la1[i1 >> 3] >> ((i1 & 7u) << 2)) & 15u)) return false;
  }
  const uint32_t idx = num_lo & sv.la_mask;
  return fe >= ((sv.la[idx >> 3] >> ((idx & 7u) << 2)) & 15u);
}

// Waves of a 256-lane workgroup.
constexpr int kWavesPerBlock = 4;

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// State of one candidate in flight.
template <int L>
struct Cand {
  Num<L> v;
  uint64_t n_lo, n_hi;  // the start n
  uint32_t hash;
  uint32_t steps;
  uint32_t ovf;
  uint32_t slot;   // output record slot (for overflow and escalation reports)
};

// Where a candidate's values are offered for the max: one lane accumulator.  (The last
// argument, whether the lane holds a candidate, is always true where this sink is used.)
template <int L>
struct SinkOne {
  Acc<L>& acc;
  __device__ __forceinline__ void offer(const Num<L>& v, uint64_t n_lo, uint64_t n_hi, bool) {
    offer_max<L>(acc, v, n_lo, n_hi);
  }
};

__device__ __forceinline__ uint64_t readlane64(uint64_t x, uint32_t l) {
  return pair(__builtin_amdgcn_readlane((uint32_t)x, l),
              __builtin_amdgcn_readlane((uint32_t)(x >> 32), l));
}

// Adjacent 32-bit reads of a value's words would be merged into a straddling 64-bit read,
// which keeps the array out of registers (the compiler then keeps it in LDS and stores it
// after every run); an empty asm per word keeps each read on its own.
template <int L>
__device__ __forceinline__ void words_apart(uint32_t* x) {
#pragma unroll
  for (int i = 0; i < L; ++i) asm volatile("" : "+v"(x[i]));
}

// Max of a whole wave, for the persistent and plain kernels: the wave's best (value, start)
// is uniform (scalar registers).  The offer runs in uniform control flow on every lane: a
// lane's value is compared with the top 64 bits of the wave's best, and only when some lane
// of `live` (the lanes that hold a candidate, set by the caller before each check) reaches
// them, rare once the best is high, are the reaching lanes merged one at a time by uniform
// code: a larger value wins, an equal one with a smaller start.  A lane-private max would
// start from zero in every tile and take a divergent update path many times per tile; the
// wave's max needs the merge about log(offers) times.
// KEYMAX: merge many reaching lanes through a cross-lane max of their top-64-bit key first
// (the plain kernel; in the persistent kernel this code, though it rarely runs, slowed the
// check loop of long launches by 2-3%).
// LOWKEY (three or four words): for launches whose values stay small.  The key of the test
// above is the value's top two words, the value over 2^32 (over 2^64 on four words); while
// the best stays below about 2^36, most values share its key, nearly every offer reaches,
// and the wave merges its lanes one at a time twice per check ([1, 2^32) at k = 20 ran 1.3
// ms where [2^50, 2^50 + 2^32) ran 0.27).  LOWKEY tests the low two words and whether the
// upper words are nonzero: reach = upper != 0 or low pair >= thr, with thr the best's low
// pair while its upper words are zero, else all ones.  That is exact too, but costs time
// where ties are rare (one more VALU instruction per offer; the persistent kernel ran
// [1, 2^36) 15% slower with it, the plain one 4%), so launches take it only for ranges that
// end at or below 2^35 (kLowKeyLog2): [1, 2^32) then runs 0.28 ms (plain) and 0.44 ms
// (persistent) instead of 1.2 and 1.4, [1, 2^34) 1.03 and 1.34 instead of 1.31 and 1.44.
template <int L, bool KEYMAX = false, bool LOWKEY = false>
struct SinkWave {
  static_assert(!LOWKEY || L == 3 || L == 4, "the low key is for three or four words");
  uint32_t bw[L];         // the wave's best value (0: none yet; every value is >= 2)
  uint64_t bs_lo, bs_hi;  // its start
  uint64_t live;          // lanes with a candidate
  uint64_t thr;           // (LOWKEY) see above
  __device__ __forceinline__ void clear() {
#pragma unroll
    for (int i = 0; i < L; ++i) bw[i] = 0;
    bs_lo = bs_hi = 0;
    thr = 0;
  }
  __device__ __forceinline__ void offer(Num<L>& v, uint64_t n_lo, uint64_t n_hi, bool) {
    // Six words: the top pair is zero below 2^128, so the test there is the whole compare.
    // Elsewhere the key's two words pass an empty asm in place first, so that the key is
    // built from registers and not from a straddling 64-bit read of the value's words.
    bool reach;
    if constexpr (L == 6) {
      reach = ge_words<L>(v.w, bw);
    } else if constexpr (LOWKEY) {
      uint32_t up = v.w[2];
#pragma unroll
      for (int i = 3; i < L; ++i) up |= v.w[i];
      reach = (up != 0u) | (pair(v.w[0], v.w[1]) >= thr);
    } else {
      constexpr int k = L - 2;
      uint32_t lo = v.w[k], hi = v.w[k + 1];
      asm volatile("" : "+v"(lo), "+v"(hi));
      v.w[k] = lo;
      v.w[k + 1] = hi;
      reach = pair(lo, hi) >= pair(bw[k], bw[k + 1]);
    }
    const uint64_t pend0 = __builtin_amdgcn_ballot_w64(reach) & live;
    if (__builtin_expect(pend0 != 0, 0)) {
      uint32_t x[L];
#pragma unroll
      for (int i = 0; i < L; ++i) x[i] = v.w[i];
      words_apart<L>(x);
      uint64_t pend1 = pend0;
      if constexpr (KEYMAX && L != 6) {
        // Many reaching lanes (at the start of a tile, when the best is still low): only the
        // lanes holding the largest key can hold the largest value, so a cross-lane max of
        // the key (six shuffle rounds) cuts the one-at-a-time merge below to them, usually
        // one lane.  (The low key: a lane with nonzero upper words takes the largest key.)
        if (__popcll(pend0) > 2) {
          uint64_t own;
          if constexpr (LOWKEY) {
            uint32_t up = x[2];
#pragma unroll
            for (int i = 3; i < L; ++i) up |= x[i];
            own = up ? ~0ull : pair(x[0], x[1]);
          } else {
            own = pair(x[L - 2], x[L - 1]);
          }
          uint64_t key = __builtin_amdgcn_inverse_ballot_w64(pend0) ? own : 0;
#pragma unroll
          for (int m = 32; m >= 1; m >>= 1) {
            const uint64_t o = shfl_xor64(key, m);
            key = o > key ? o : key;
          }
          const uint64_t kmax = pair(__builtin_amdgcn_readfirstlane((uint32_t)key),
                                     __builtin_amdgcn_readfirstlane((uint32_t)(key >> 32)));
          pend1 = __builtin_amdgcn_ballot_w64(own == kmax) & pend0;
        }
      }
      for (uint64_t pend = pend1; pend; pend &= pend - 1) {
        const uint32_t l = (uint32_t)__builtin_ctzll(pend);
        uint32_t cw[L];
        if constexpr (LOWKEY) {
#pragma unroll
          for (int i = 0; i < L; ++i) cw[i] = __builtin_amdgcn_readlane(x[i], l);
        } else {
          // The lane's top two words first: below the best's, the lane cannot hold it.  When
          // every lane reaches (a tile's first check, the best still zero) only the running
          // maxima in lane order, about ln 64 of them, go on to the whole compare.
          cw[L - 1] = __builtin_amdgcn_readlane(x[L - 1], l);
          cw[L - 2] = __builtin_amdgcn_readlane(x[L - 2], l);
          if (pair(cw[L - 2], cw[L - 1]) < pair(bw[L - 2], bw[L - 1])) continue;
#pragma unroll
          for (int i = 0; i < L - 2; ++i) cw[i] = __builtin_amdgcn_readlane(x[i], l);
        }
        const uint64_t c_lo = readlane64(n_lo, l), c_hi = readlane64(n_hi, l);
        int cmp = 0;
#pragma unroll
        for (int i = L - 1; i >= 0; --i)
          if (cmp == 0 && cw[i] != bw[i]) cmp = cw[i] > bw[i] ? 1 : -1;
        if (cmp > 0 || (cmp == 0 && (c_hi < bs_hi || (c_hi == bs_hi && c_lo < bs_lo)))) {
#pragma unroll
          for (int i = 0; i < L; ++i) bw[i] = cw[i];
          bs_lo = c_lo;
          bs_hi = c_hi;
        }
      }
      if constexpr (LOWKEY) {
        uint32_t up = bw[2];
#pragma unroll
        for (int i = 3; i < L; ++i) up |= bw[i];
        thr = up ? ~0ull : pair(bw[0], bw[1]);
      }
    }
  }
};

// Set up a candidate: v = M + q3 (q3 has L words; q3_big says q * 3^f needs more), n =
// qb | r, and its step count at steps0.  Returns false if the start value does not fit L
// words (then the candidate must go to the three-limb pass).
template <int L>
__device__ __forceinline__ bool cand_start(Cand<L>& c, uint64_t M, const uint32_t* q3,
                                           bool q3_big, uint64_t qb_lo, uint64_t qb_hi,
                                           uint64_t r, uint32_t steps0 = 0) {
  uint32_t cy = 0;
  c.v.w[0] = addc(q3[0], (uint32_t)M, 0u, cy);
  c.v.w[1] = addc(q3[1], (uint32_t)(M >> 32), cy, cy);
#pragma unroll
  for (int i = 2; i < L; ++i) c.v.w[i] = addc(q3[i], 0u, cy, cy);
  const bool fits = cy == 0 && !q3_big;
  c.n_lo = qb_lo | r;
  c.n_hi = qb_hi;
  c.hash = 0;
  c.steps = steps0;
  c.ovf = 0;
  return fits;
}

// Overflow reports: (n_lo, n_hi, slot, 0) per event, in the order the events happen.
__device__ __forceinline__ void push_overflow(const DeviceLists& lists, uint64_t n_lo,
                                              uint64_t n_hi, uint32_t slot) {
  const unsigned long long i = atomicAdd(lists.ovf_count, 1ull);
  if (i < lists.ovf_cap) {
    lists.ovf[4 * i] = n_lo;
    lists.ovf[4 * i + 1] = n_hi;
    lists.ovf[4 * i + 2] = slot;
    lists.ovf[4 * i + 3] = 0;
  }
}

// A candidate abandoned at the check bound: counted, and its slot listed.
__device__ __forceinline__ void push_stall(const DeviceLists& lists, uint32_t slot) {
  const unsigned long long i = atomicAdd(lists.stall_count, 1ull);
  if (i < lists.stall_cap) lists.stall_slots[i] = slot;
}

__device__ __forceinline__ void push_escalate(const DeviceLists& lists, uint64_t n_lo,
                                              uint64_t n_hi, uint32_t slot) {
  const unsigned long long i = atomicAdd(lists.esc_count, 1ull);
  if (i < lists.esc_cap) {
    EscEntry e;
    e.n_lo = n_lo;
    e.n_hi = n_hi;
    e.slot = slot;
    e.pad = 0;
    lists.esc[i] = e;
  }
}

// Word widths that escalate (three 32-bit words; two u64 limbs with a cutoff at or above
// 2^123) test the bound once per check, at its end: a check multiplies by at most
// ((3/2)^8)^2 < 2^9.4 (two odd runs; even runs only decrease), so a check that starts
// below 2^86 (three words) or 2^118 (four words) never leaves the words, every value it
// offers is exact, and a candidate that ends a check at or above the bound goes to the
// three-limb pass, which reruns it from its start.
template <int L, bool LOWCUT>
struct Escalates {
  static constexpr bool value = L == 3 || (L == 4 && !LOWCUT);
};

template <int L>
__device__ __forceinline__ bool over_check_bound(const Num<L>& v) {
  if constexpr (L == 3) return v.w[2] >= (1u << (kNarrowCheckLog2 - 64));
  else return pair(v.w[2], v.w[3]) >= (1ull << (kWideCheckLog2 - 64));
}

// After an odd run: offer to the max; where the width overflows instead of escalating,
// the cutoff test (on four words LOWCUT means a cutoff below 2^123, where the top-pair
// compare is exact; on six words a cutoff below 2^128).
template <int L, bool LOWCUT, class Sink>
__device__ __forceinline__ void after_odd(Cand<L>& c, uint64_t n_hi, Sink& sink,
                                          const Limits& lim, const DeviceLists& lists,
                                          bool live) {
  sink.offer(c.v, c.n_lo, n_hi, live);
  if constexpr (!Escalates<L, LOWCUT>::value) {
    if (live && over_fast<L, LOWCUT>(c.v, lim.fast_thr) &&
        (L == 4 || ge_pow2<L>(c.v, lim.cutoff))) {
      // Overflow: report the start and continue from 2 (collatz_cudav7.cu:295-301).
      c.ovf += 1;
      push_overflow(lists, c.n_lo, n_hi, c.slot);
      c.v.w[0] = 2;
#pragma unroll
      for (int i = 1; i < L; ++i) c.v.w[i] = 0;
    }
  }
}

// One check.  `done`: number < n after it.  `esc`: the check ended at or above the
// width's bound, so the candidate goes to the three-limb pass (never together with done:
// the widths that escalate are used only for starts far below the bound).  n_hi is the
// candidate's c.n_hi, passed on its own so that a caller whose candidates share it (the
// persistent kernel: one q per tile) can keep it in scalar registers.  A lane with `live`
// false (no candidate) computes on whatever its state holds; its offers, overflow reports,
// esc and done are suppressed.
template <int L, bool LOWCUT, class Sink>
__device__ __forceinline__ void one_check(Cand<L>& c, uint64_t n_hi, Sink& sink,
                                          const Limits& lim, const DeviceLists& lists,
                                          const Pow3Lds& p3, bool& esc, bool& done,
                                          bool live = true) {
  const uint32_t a1 = odd_run<L>(c.v, p3);
  after_odd<L, LOWCUT>(c, n_hi, sink, lim, lists, live);
  const uint32_t b1 = even_run<L>(c.v);
  const uint32_t a2 = odd_run<L>(c.v, p3);
  after_odd<L, LOWCUT>(c, n_hi, sink, lim, lists, live);
  const uint32_t b2 = even_run<L>(c.v);
  c.steps += a1 + b1 + a2 + b2;
  c.hash ^= c.v.w[0];
  if constexpr (Escalates<L, LOWCUT>::value) esc = live && over_check_bound<L>(c.v);
  else esc = false;
  done = live && lt_base<L>(c.v, c.n_lo, n_hi);
}

template <int L>
__device__ __forceinline__ void finish(const Cand<L>& c, Acc<L>& acc, uint32_t k) {
  const uint32_t J = k + c.steps;
  acc.cnt += 1;
  acc.hash ^= c.hash;
  acc.ssum += J;
  acc.sxor ^= J;
  acc.ovf += c.ovf;
}

// Start-of-candidate test: a start at or above the width's per-check bound goes to the
// three-limb pass.  Four words with a cutoff below 2^123 do not escalate per check: every
// odd run starts at or below max(start, 2^cutoff) (an odd run that ends at or above the
// cutoff resets the value to 2; even runs only decrease) and multiplies by at most
// (3/2)^8 < 2^4.7, so values stay below 2^128 only for starts below 2^123; a larger start
// goes to the three-limb pass.  The start value itself is not offered to the max: the
// first odd run turns 2^a m into 3^a m, at least as large, for the same start, and every
// candidate that is not escalated runs at least one check.
template <int L, bool LOWCUT>
__device__ __forceinline__ bool cand_enter(const Cand<L>& c) {
  if constexpr (Escalates<L, LOWCUT>::value) return !over_check_bound<L>(c.v);
  if constexpr (L == 4) return c.v.w[3] < (1u << (kWideLog2 - 96));
  return true;
}

// Lane masks for the persistent kernel, each built from ballots of single compares (a
// ballot of one compare is that compare's mask; a ballot of a combined condition would be
// materialised in a vector register first).
__device__ __forceinline__ uint64_t ballot(bool b) { return __builtin_amdgcn_ballot_w64(b); }
__device__ __forceinline__ bool in_mask(uint64_t m) {
  return __builtin_amdgcn_inverse_ballot_w64(m);
}

// Lanes whose new start may run (cand_enter), where a start that did not fit its words has
// had its top word set to all ones (a value no start of these widths reaches).
template <int L, bool LOWCUT>
__device__ __forceinline__ uint64_t start_ok_mask(const Num<L>& v) {
  if constexpr (Escalates<L, LOWCUT>::value) {
    if constexpr (L == 3) return ballot(v.w[2] < (1u << (kNarrowCheckLog2 - 64)));
    else return ballot(pair(v.w[2], v.w[3]) < (1ull << (kWideCheckLog2 - 64)));
  } else if constexpr (L == 4) {
    return ballot(v.w[3] < (1u << (kWideLog2 - 96)));
  } else {
    return ballot(v.w[L - 1] != 0xFFFFFFFFu);
  }
}

// Lanes whose check ended at or above the width's bound (over_check_bound).
template <int L, bool LOWCUT>
__device__ __forceinline__ uint64_t esc_mask(const Num<L>& v) {
  if constexpr (!Escalates<L, LOWCUT>::value) return 0;
  else if constexpr (L == 3) return ballot(v.w[2] >= (1u << (kNarrowCheckLog2 - 64)));
  else return ballot(pair(v.w[2], v.w[3]) >= (1ull << (kWideCheckLog2 - 64)));
}

// Lanes with number < n (lt_base).
template <int L>
__device__ __forceinline__ uint64_t lt_base_mask(const Num<L>& v, uint64_t n_lo, uint64_t n_hi) {
  const uint64_t lo = ballot(pair(v.w[0], v.w[1]) < n_lo);
  if constexpr (L == 3) {
    // (A uniform branch on n's top word, to skip the < compare below 2^64, moved the tile
    // cursor and the check count of the persistent kernel into vector registers: slower.)
    const uint32_t nh = (uint32_t)n_hi;
    return ballot(v.w[2] < nh) | (ballot(v.w[2] == nh) & lo);
  } else {
    const uint64_t vh = pair(v.w[2], v.w[3]);
    const uint64_t m = ballot(vh < n_hi) | (ballot(vh == n_hi) & lo);
    if constexpr (L > 4) return m & ballot((v.w[4] | v.w[5]) == 0u);
    else return m;
  }
}

// Runs candidate c to completion.  Returns false if it was escalated.
template <int L, bool LOWCUT>
__device__ __forceinline__ bool run_candidate(Cand<L>& c, Acc<L>& acc, const Limits& lim,
                                              const DeviceLists& lists, uint32_t k,
                                              const Pow3Lds& p3) {
  SinkOne<L> sink{acc};
  if (!cand_enter<L, LOWCUT>(c)) return false;
  for (uint32_t guard = 0; guard < lim.max_checks; ++guard) {
    bool esc, done;
    one_check<L, LOWCUT>(c, c.n_hi, sink, lim, lists, p3, esc, done);
    acc.chk += 1;
    if (esc) return false;
    if (done) {
      finish<L>(c, acc, k);
      return true;
    }
  }
  push_stall(lists, c.slot);
  return true;
}

// ---- plain kernel: a fixed share of one row per wave, lanes refilled --------------------
// Wave g (= 4 blockIdx.x + its index in the block) takes share g: piece g - s of the row
// whose shares start at s (its block_start; a row is cut into equal pieces, most rows into
// one).  No state outlives the launch and no work moves between waves: this is the drop-in
// kernel of collatz_verify_range.  A lane runs one candidate at a time, two checks per loop
// iteration; a lane whose candidate has ended (descended, escalated, or abandoned at the
// step bound) starts the survivor record it loaded earlier, and the lanes without a loaded
// record take the next survivors of the wave's share by rank among them, so no lane waits
// for its wave's slowest trajectory until the share is used up.  As in the
// persistent kernel below, which lanes run a candidate and which hold a record are two
// uniform masks (scalar registers), every lane runs the check (a lane without a candidate
// on stale values, its effects dropped), the wave's max excursion is kept in scalar
// registers, the step count starts at k, and on three words a row at or above the start
// bound holds 2^95 in its q 3^f table.  The loop ends when no lane holds a candidate or a
// record and the share is used up, or at lim.max_loop iterations (counted in
// lists.loop_aborts: the host then fails the call).  Each wave writes its share's record.
template <int L, bool LOWCUT, bool LOWKEY>
__global__ __launch_bounds__(256) void k_verify_plain(SieveArrays sv, const RowDesc* rows,
                                                      uint32_t nrows, uint64_t nshares,
                                                      uint32_t mod9, Limits lim,
                                                      const uint64_t* pow3, DevRec* recs,
                                                      DeviceLists lists) {
  __shared__ uint32_t q3s[kWavesPerBlock][41][L];
  __shared__ uint32_t q3los[kWavesPerBlock][41];  // low word of q 3^f (the look-ahead index)
  __shared__ Pow3Lds p3s;
  stage_pow3(p3s);
  __syncthreads();
  // (The wave index through readfirstlane: the compiler does not know threadIdx.x >> 6 is
  // uniform in a wave, and would keep the share bounds, the masks and the cursor in vector
  // registers.)
  const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t gw = (uint64_t)blockIdx.x * kWavesPerBlock + wv;
  if (gw >= nshares) return;  // (the last block's spare waves; no barrier follows)
  uint32_t lo = 0, hi = nrows;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (rows[mid].block_start <= gw) lo = mid;
    else hi = mid;
  }
  const RowDesc row = rows[lo];
  const uint64_t pieces = (lo + 1 < nrows ? rows[lo + 1].block_start : nshares) - row.block_start;
  const uint64_t piece = gw - row.block_start;
  const uint64_t len = row.i_hi - row.i_lo;  // (< 2^32: one residue class of the sieve)
  uint64_t next = row.i_lo + len * piece / pieces;
  const uint64_t end = row.i_lo + len * (piece + 1) / pieces;
  uint32_t (*q3)[L] = q3s[wv];
  uint32_t* q3lo = q3los[wv];
  const uint32_t lane = threadIdx.x & 63u;
  if (lane < 41) {
    uint32_t w6[6];
    q_times_pow3(row.q_lo, row.q_hi, pow3[lane], w6);
    uint32_t big = 0;
#pragma unroll
    for (int i = L; i < 6; ++i) big |= w6[i];
    if constexpr (L == 3) {
      // (As in the persistent kernel: a start M + q 3^f then cannot carry out of three words
      // and fails the start test exactly when q 3^f reaches 2^86.)
      big |= (uint32_t)(w6[2] >= (1u << (kNarrowCheckLog2 - 64)));
      q3[lane][0] = big ? 0u : w6[0];
      q3[lane][1] = big ? 0u : w6[1];
      q3[lane][2] = big ? 0x80000000u : w6[2];
    } else {
#pragma unroll
      for (int i = 0; i < L; ++i) q3[lane][i] = big ? 0xFFFFFFFFu : w6[i];
    }
    q3lo[lane] = w6[0];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  Acc<L> acc;
  acc.init();
  SinkWave<L, true, LOWKEY> sink;
  sink.clear();
  uint32_t wave_chk = 0;
  const uint64_t qb_lo = row.qb_lo, qb_hi = row.qb_hi;
  const uint32_t slot = row.slot;
  Cand<L> c;
  c.n_hi = qb_hi;
  c.slot = slot;
  uint64_t actm = 0, pfm = 0;  // lanes with a running candidate, lanes with a loaded record
  uint4 pf = make_uint4(0, 0, 0, 0);
  const uint32_t step_lim = lim.max_steps + sv.k;  // (the step count starts at k)
  uint32_t iters = 0;
  bool aborted = false;
  // Starts the loaded records of the lanes in S (no running candidate, a loaded record); a
  // look-ahead sieve first tests each, and a record it drops is simply not started.
  auto start_loaded = [&](uint64_t S) {
    bool keep = true;
    if (in_mask(S)) {
      const uint32_t f = rec_f(pf);
      const uint64_t M = rec_M(pf);
      if (sv.la) keep = la_keep(sv, f, (uint32_t)M + q3lo[f]);
      const bool fits = cand_start<L>(c, M, q3[f], false, qb_lo, qb_hi, rec_r(pf), sv.k);
      if constexpr (L != 3) {
        if (!fits) c.v.w[L - 1] = 0xFFFFFFFFu;
      }
    }
    const uint64_t kept = ballot(keep) & S;
    const uint64_t ok = start_ok_mask<L, LOWCUT>(c.v) & kept;
    const uint64_t bad = kept & ~ok;
    if (__builtin_expect(bad != 0, 0)) {
      if (in_mask(bad)) push_escalate(lists, c.n_lo, qb_hi, slot);
    }
    actm |= ok;
    pfm &= ~S;
  };
  for (;;) {
    if (__builtin_expect(++iters > lim.max_loop, 0)) {
      aborted = true;
      break;
    }
    // (1) Lanes without a candidate start their loaded record.
    {
      const uint64_t S = ~actm & pfm;
      if (S) start_loaded(S);
    }
    // (2) Lanes without a loaded record take the next survivors of the share.
    const uint64_t need = ~pfm;
    if (need && next < end) {
      const uint32_t want = (uint32_t)__popcll(need);
      const uint32_t avail = (uint32_t)(end - next);  // (a share is part of one row: < 2^32)
      const uint32_t take = avail < want ? avail : want;
      const uint32_t rank = lane_rank(need);
      const uint64_t tk = ballot(rank < take) & need;
      if (in_mask(tk)) pf = sv.rec[next + rank];
      pfm |= tk;
      next += take;
    }
    // (Nothing running and nothing loaded means the share is used up: otherwise (2) loaded.)
    if ((actm | pfm) == 0) break;
    // A candidate past the step bound is abandoned (tested once per refill: it may run one
    // check beyond the bound first).
    const uint64_t stalled = ballot(c.steps >= step_lim) & actm;
    if (__builtin_expect(stalled != 0, 0)) {
      if (in_mask(stalled)) push_stall(lists, slot);
      actm &= ~stalled;
    }
    // (3) Two checks on every lane (a record loaded in (2) is then used two checks later, and
    //     the refill's cost is shared); a lane whose candidate ends at the first starts its
    //     loaded record before the second.
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      sink.live = actm;
      wave_chk += (uint32_t)__popcll(actm);
      bool esc_lane, done_lane;  // (unused here: the masks below are taken instead)
      one_check<L, LOWCUT>(c, qb_hi, sink, lim, lists, p3s, esc_lane, done_lane,
                           in_mask(actm));
      const uint64_t escm = esc_mask<L, LOWCUT>(c.v) & actm;
      const uint64_t donem = lt_base_mask<L>(c.v, c.n_lo, qb_hi) & actm & ~escm;
      if (__builtin_expect(escm != 0, 0)) {
        if (in_mask(escm)) push_escalate(lists, c.n_lo, qb_hi, slot);
      }
      if (donem) {
        if (in_mask(donem)) finish<L>(c, acc, 0);  // (the step count includes k)
      }
      actm &= ~(escm | donem);
      if (u == 0) {
        const uint64_t S = ~actm & pfm;
        if (S) start_loaded(S);
      }
    }
  }
  if (aborted && lane == 0) atomicAdd(lists.loop_aborts, 1ull);
  // The share's record: counts merged across the lanes, the max from the scalar best.
  DevRec r = acc_to_rec<L>(acc);
  wave_merge(r);
  r.checks = wave_chk;
  if (r.count) {
    r.maxnum[0] = pair(sink.bw[0], sink.bw[1]);
    if constexpr (L == 3) r.maxnum[1] = sink.bw[2];
    else r.maxnum[1] = pair(sink.bw[2], sink.bw[3]);
    if constexpr (L > 4) r.maxnum[2] = pair(sink.bw[4], sink.bw[5]);
    r.max_start[0] = sink.bs_lo;
    r.max_start[1] = sink.bs_hi;
  }
  if (lane == 0) recs[gw] = r;
}

// ---- escalated candidates on three limbs, one thread each -----------------------------
__global__ __launch_bounds__(256) void k_escalated(SieveArrays sv, const EscEntry* esc,
                                                   uint64_t n, Limits lim,
                                                   const uint64_t* pow3, DevRec* out,
                                                   DeviceLists lists) {
  __shared__ Pow3Lds p3s;
  stage_pow3(p3s);
  __syncthreads();
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const EscEntry e = esc[t];
  const uint32_t k = sv.k;
  // The survivor record of r = n mod 2^k is (T^k(r) + 1, the odd steps f among those k),
  // recomputed here: k < 40 steps of T from r < 2^k stay below 2^k (3/2)^k < 2^64.
  const uint64_t r = e.n_lo & ((1ull << k) - 1);
  uint64_t x = r;
  uint32_t f = 0;
  for (uint32_t i = 0; i < k; ++i) {
    if (x & 1u) {
      x += (x >> 1) + 1;  // (3x + 1) / 2
      ++f;
    } else {
      x >>= 1;
    }
  }
  const uint64_t M = x + 1;
  const uint64_t q_lo = k >= 64 ? (e.n_hi >> (k - 64)) : ((e.n_lo >> k) | (k ? e.n_hi << (64 - k) : 0));
  const uint64_t q_hi = k >= 64 ? 0 : (e.n_hi >> k);
  uint32_t q3[6];
  q_times_pow3(q_lo, q_hi, pow3[f], q3);
  const uint64_t mask = (k >= 64) ? ~0ull : ((1ull << k) - 1);
  const uint64_t qb_lo = e.n_lo & ~mask, qb_hi = e.n_hi;
  Acc<6> acc;
  acc.init();
  Cand<6> c;
  c.slot = e.slot;
  cand_start<6>(c, M, q3, false, qb_lo, qb_hi, r);
  if (lim.cutoff >= 128) run_candidate<6, false>(c, acc, lim, lists, k, p3s);
  else run_candidate<6, true>(c, acc, lim, lists, k, p3s);
  DevRec rec = acc_to_rec<6>(acc);
  rec.pad = e.slot;
  out[t] = rec;
}


// ---- persistent kernel: waves with lane refill and work stealing ------------------------
// Every wave is an independent worker.  It takes tiles (a piece of one row's survivors)
// from a global counter and keeps its 64 lanes busy: a lane whose candidate has descended
// starts the next residue of the tile at the next check, so inside a tile the wave waits
// for no lane; only the tile's last candidates run with some lanes idle.  Candidate data
// is prefetched one check ahead so the loads overlap the arithmetic.  Row data (q << k,
// the q * 3^f table) is the same for the whole tile and stays in SGPRs and LDS.

// q * 3^f for f < 41 on L words.  Where the product does not fit L words the row is all
// ones: adding any start value (>= 1) then carries out, which marks the candidate for the
// three-limb pass without a separate flag.  On three words a row at or above the start
// bound 2^86 holds 2^95 instead (see the tile set-up).
template <int L, bool LA>
struct WaveTables {
  uint32_t q3[41][L];
  uint32_t q3lo[LA ? 41 : 1];  // low word of q * 3^f (the look-ahead index), also where q3
                               // holds a marker
};

// Look-ahead form: a wave reads its tile 64 survivors at a time (one coalesced window,
// loaded one window ahead), tests them against the look-ahead table and keeps the passing
// records, in order, in LDS; lanes that need a candidate take the next
// one from there.
struct LaStage {
  uint4 rec[64];
};

// BOUNDED: the work loop carries the host-set iteration bound (Limits::max_loop).  The
// look-ahead form always does; the plain form only when a bound below the default is set
// (COLLATZ_LOOP_LIMIT), because the bound's counter costs 6% of its time (the loop is
// issue-bound, and in four forms tried the counter cost 5-12%), and the loop ends without
// it: every iteration either runs a check (which adds at least two T-steps to every lane's
// step count, so a running candidate is abandoned at the step bound), starts the
// prefetched records of idle lanes (clearing their prefetch bits), takes survivors (the
// tile cursor advances), or leaves the loop; the masks, the cursor and the exit test are
// uniform (scalar registers, scalar branches).
// UC: the checks per refill as a compile-time constant (1 to 4), or 0 to take them from
// U.  With a constant the check loop unrolls, and with one check per refill the record set-up
// for starts between checks (hoisted ahead of the loop) goes away.
template <int L, bool LOWCUT, bool LA, bool BOUNDED = LA, int UC = 0, bool LOWKEY = false>
__global__ __launch_bounds__(256) void k_verify_persistent(
    SieveArrays sv, const RowDesc* rows, const TileDesc* tiles, const uint32_t* order,
    Ctl* ctl, Limits lim, const uint64_t* pow3, DevRec* tile_recs, DeviceLists lists,
    uint32_t U) {
  __shared__ WaveTables<L, LA> tabs[kWavesPerBlock];
  __shared__ LaStage stages[LA ? kWavesPerBlock : 1];
  __shared__ Pow3Lds p3s;
  stage_pow3(p3s);
  __syncthreads();
  const uint32_t wv = threadIdx.x >> 6;
  const uint32_t lane = threadIdx.x & 63;
  WaveTables<L, LA>& T = tabs[wv];
  // The waves of this workgroup (on XCD blockIdx.x mod kXcds) take tiles from that XCD's
  // share of the tile order, then from the other shares in turn: the order puts tiles that
  // read the same survivor records next to each other, so the waves of one XCD read them
  // from its L2.  Every share is left at most once, so the loop takes at most kXcds - 1
  // empty turns besides its tiles.
  const uint32_t ntiles = ctl->ntiles;
  const uint32_t lg = ctl->share_log2 ? 3u : 0u;  // (kXcds = 2^3 shares or one)
  const uint32_t smask = (1u << lg) - 1u;
  uint32_t share = blockIdx.x & smask, turns = 0;
  for (;;) {
    const uint32_t s_lo = (uint32_t)(((uint64_t)ntiles * share) >> lg);
    const uint32_t s_n = (uint32_t)(((uint64_t)ntiles * (share + 1)) >> lg) - s_lo;
    uint32_t pos = 0;
    if (lane == 0) pos = atomicAdd(&ctl->share_next[share], 1u);
    pos = __builtin_amdgcn_readfirstlane(pos);
    if (pos >= s_n) {
      if (turns++ == smask) break;
      share = (share + 1) & smask;
      continue;
    }
    const uint32_t t = order[s_lo + pos];
    const TileDesc td = tiles[t];
    const RowDesc rw = rows[td.row];
    const uint64_t len = rw.i_hi - rw.i_lo;
    // The tile is survivors [first, first + n) of the row (n < 2^32: a residue class of the
    // deepest sieve holds fewer than 2^32 survivors); the cursor is an offset into it.
    const uint64_t first = rw.i_lo + len * td.piece / td.pieces;
    const uint32_t n = (uint32_t)(rw.i_lo + len * (td.piece + 1) / td.pieces - first);
    uint32_t cur = 0;
    if (lane < 41) {
      uint32_t w6[6];
      q_times_pow3(rw.q_lo, rw.q_hi, pow3[lane], w6);
      uint32_t big = 0;
#pragma unroll
      for (int i = L; i < 6; ++i) big |= w6[i];
      if constexpr (L == 3) {
        // On three words a row whose q 3^f reaches the start bound 2^86 holds 2^95 instead:
        // a start value M + q 3^f (M < 2^64) then stays below 2^96 without a carry out
        // (q 3^f < 2^86 leaves 2^96 - 2^86 of room; 2^95 + M < 2^96) and fails the start
        // test exactly when it should, so the start needs no carry-out test.
        big |= (uint32_t)(w6[2] >= (1u << (kNarrowCheckLog2 - 64)));
        T.q3[lane][0] = big ? 0u : w6[0];
        T.q3[lane][1] = big ? 0u : w6[1];
        T.q3[lane][2] = big ? 0x80000000u : w6[2];
      } else {
#pragma unroll
        for (int i = 0; i < L; ++i) T.q3[lane][i] = big ? 0xFFFFFFFFu : w6[i];
      }
      if constexpr (LA) T.q3lo[lane] = w6[0];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint64_t qb_lo = rw.qb_lo, qb_hi = rw.qb_hi;
    const uint32_t slot = rw.slot;

    // Per-lane counts and sums; the max is the wave's (SinkWave), cleared for each tile.
    Acc<L> acc;
    acc.init();
    SinkWave<L, false, LOWKEY> sink;
    sink.clear();
    // Checks of the tile, counted per wave from the mask of lanes that run one (scalar
    // work; the lane accumulators' chk stays 0).
    uint32_t wave_chk = 0;
    Cand<L> c;
    c.n_hi = qb_hi;  // the same for every candidate of the tile
    // Lane masks (uniform): lanes with a running candidate, lanes holding a prefetched
    // record.  Per-lane conditions are read from them with in_mask().
    uint64_t actm = 0, pfm = 0;
    uint4 pf_rec = make_uint4(0, 0, 0, 0);
    // A candidate's step count starts at k, so that it is J (k + the T-steps run) when the
    // candidate finishes; the step bound is moved by k with it.
    const uint32_t step_lim = lim.max_steps + sv.k;
    const uint4* tile_rec = sv.rec + first;
    // Look-ahead stage: entries [shead, scnt) are kept survivors not yet taken; nx holds the
    // records of the next window (survivors [cur, cur + 64)), loaded one window ahead.
    uint32_t scnt = 0, shead = 0;
    uint4 nx = make_uint4(0, 0, 0, 0);
    if constexpr (LA) {
      if (lane < n) nx = tile_rec[lane];
    }
    // Iterations of the tile's loop (BOUNDED): counted in a vector register (a scalar
    // counter in this loop costs more) and tested with the step bound, once per refill.  A
    // wave past the bound treats every running candidate as abandoned, drops its
    // prefetches and the rest of the tile, and counts the abort; the call then fails.  The
    // take loop of the look-ahead form has its own bound.
    uint32_t iters = 0;
    if constexpr (BOUNDED) asm volatile("" : "+v"(iters));
    bool aborted = false;
    for (;;) {
      if constexpr (BOUNDED) ++iters;
      // Starts the prefetched candidates of the lanes in S (no running candidate, a
      // prefetched record); a start that cannot run on these words is escalated.
      auto start_prefetched = [&](uint64_t S) {
        if (in_mask(S)) {
          const uint32_t f = rec_f(pf_rec);
          const bool fits = cand_start<L>(c, rec_M(pf_rec), T.q3[f], false, qb_lo, qb_hi,
                                          rec_r(pf_rec), sv.k);
          c.slot = slot;  // (overflow reports carry it)
          if constexpr (L != 3) {
            if (!fits) c.v.w[L - 1] = 0xFFFFFFFFu;
          }
        }
        const uint64_t ok = start_ok_mask<L, LOWCUT>(c.v) & S;
        const uint64_t bad = S & ~ok;
        if (__builtin_expect(bad != 0, 0)) {
          if (in_mask(bad)) push_escalate(lists, c.n_lo, qb_hi, slot);
        }
        actm |= ok;
        pfm &= ~S;
      };
      // (1) Lanes without a candidate start their prefetched one.
      {
        const uint64_t S = ~actm & pfm;
        if (S) start_prefetched(S);
      }
      // (2) Lanes without a prefetch take the next survivors of the tile.
      const uint64_t need = ~pfm;
      if constexpr (LA) {
        // ... the next kept ones, from the stage, refilled window by window.
        if (need) {
          LaStage& S = stages[wv];
          uint32_t want = __popcll(need);
          uint32_t rank = lane_rank(need);
          uint64_t open = need;
          uint32_t takes = 0;
          for (;;) {
            if (BOUNDED && __builtin_expect(++takes > lim.max_loop, 0)) {
              aborted = true;
              break;
            }
            const uint32_t avail = scnt - shead;
            const uint32_t take = avail < want ? avail : want;
            const uint64_t tk = ballot(rank < take) & open;
            if (in_mask(tk)) pf_rec = S.rec[shead + rank];
            pfm |= tk;
            open &= ~tk;
            rank -= take;  // (lanes that took, or need nothing, no longer look at it)
            shead += take;
            want -= take;
            if (want == 0 || cur >= n) break;
            const uint32_t o = cur + lane;
            bool keep = false;
            if (o < n) {
              const uint32_t f = rec_f(nx);
              keep = la_keep(sv, f, nx.z + T.q3lo[f]);
            }
            const uint64_t km = __ballot(keep);
            if (keep) S.rec[lane_rank(km)] = nx;
            // A wave's LDS operations execute in order, so other lanes' reads after this
            // point see the writes; only the compiler must keep them in order.  (A memory
            // fence would also wait for loads in flight.)
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            if (o + 64 < n) nx = tile_rec[o + 64];
            scnt = __popcll(km);
            shead = 0;
            cur += 64;
          }
        }
      } else if (need && cur < n) {
        const uint32_t want = __popcll(need);
        const uint32_t avail = n - cur;
        const uint32_t take = avail < want ? avail : want;
        const uint32_t rank = lane_rank(need);
        const uint64_t tk = ballot(rank < take) & need;
        if (in_mask(tk)) {
          pf_rec = tile_rec[cur + rank];
        }
        pfm |= tk;
        cur += take;
      }
      if ((actm | pfm) == 0) break;
      // A candidate past the step bound is abandoned (tested once per refill: it may run up
      // to U checks beyond the bound first).
      uint64_t stalled;
      if constexpr (BOUNDED) {
        const uint32_t pen = iters > lim.max_loop ? step_lim : 0u;
        stalled = ballot(c.steps + pen >= step_lim) & actm;
      } else {
        stalled = ballot(c.steps >= step_lim) & actm;
      }
      if (__builtin_expect(stalled != 0, 0)) {
        if (BOUNDED && __builtin_amdgcn_readfirstlane(iters) > lim.max_loop) {
          if (lane == 0) atomicAdd(lists.loop_aborts, 1ull);
          pfm = 0;
          cur = n;
        } else if (in_mask(stalled)) {
          push_stall(lists, slot);
        }
        actm &= ~stalled;
      }
      // (3) U checks; a lane whose candidate ends before the last of them starts its
      //     prefetched candidate at once instead of idling.
      //     Every lane runs the check (the SIMD issues it for all lanes either way); a lane
      //     without a candidate computes on stale values and its results are dropped, so
      //     the check's offers run in uniform control flow.
      const uint32_t Un = UC > 0 ? (uint32_t)UC : U;  // (a constant when UC is set)
      for (uint32_t u = 0; u < Un; ++u) {
        sink.live = actm;
        wave_chk += (uint32_t)__popcll(actm);
        bool esc_lane, done_lane;  // (unused here: the masks below are taken instead)
        one_check<L, LOWCUT>(c, qb_hi, sink, lim, lists, p3s, esc_lane, done_lane,
                             in_mask(actm));
        const uint64_t escm = esc_mask<L, LOWCUT>(c.v) & actm;
        const uint64_t donem = lt_base_mask<L>(c.v, c.n_lo, qb_hi) & actm & ~escm;
        if (__builtin_expect(escm != 0, 0)) {
          if (in_mask(escm)) push_escalate(lists, c.n_lo, qb_hi, slot);
        }
        if (donem) {
          if (in_mask(donem)) finish<L>(c, acc, 0);  // (the step count includes k)
        }
        actm &= ~(escm | donem);
        if (u + 1 < Un) {
          const uint64_t S = ~actm & pfm;
          if (S) start_prefetched(S);
        }
      }
    }
    if (BOUNDED && __builtin_expect(aborted, 0)) {
      if (lane == 0) atomicAdd(lists.loop_aborts, 1ull);
    }
    DevRec r = acc_to_rec<L>(acc);
    wave_merge(r);
    r.checks = wave_chk;
    if (r.count) {
      // A tile without candidates keeps max 0 (any it offered belonged to escalated
      // candidates, whose three-limb rerun offers them again).
      r.maxnum[0] = pair(sink.bw[0], sink.bw[1]);
      if constexpr (L == 3) r.maxnum[1] = sink.bw[2];
      else r.maxnum[1] = pair(sink.bw[2], sink.bw[3]);
      if constexpr (L > 4) r.maxnum[2] = pair(sink.bw[4], sink.bw[5]);
      r.max_start[0] = sink.bs_lo;
      r.max_start[1] = sink.bs_hi;
    }
    if (lane == 0) tile_recs[t] = r;
  }
}

// ---- per-chunk reduction of partial records -------------------------------------------
__global__ __launch_bounds__(256) void k_reduce_chunks(const DevRec* parts,
                                                       const uint64_t* part_start,
                                                       uint32_t nchunks, DevRec* out) {
  __shared__ DevRec lrec[4];
  const uint32_t c = blockIdx.x;
  if (c >= nchunks) return;
  DevRec r;
  rec_zero(r);
  for (uint64_t j = part_start[c] + threadIdx.x; j < part_start[c + 1]; j += 256)
    rec_merge(r, parts[j]);
  r = block_merge<4>(r, lrec);
  if (threadIdx.x == 0) out[c] = r;
}

hipError_t launch_verify_plain(int limbs, const SieveArrays& sv, const RowDesc* rows,
                               uint32_t nrows, uint64_t nshares, uint32_t mod9, bool lowkey,
                               const Limits& lim, const uint64_t* pow3, DevRec* share_recs,
                               DeviceLists lists, hipStream_t s) {
  if (nshares == 0) return hipSuccess;
  const uint64_t nblocks = (nshares + kWavesPerBlock - 1) / kWavesPerBlock;
  if (nblocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
  const dim3 grid((uint32_t)nblocks), block(256);
#define CZ_PLAIN(L, LC, LK)                                                                   \
  hipLaunchKernelGGL((k_verify_plain<L, LC, LK>), grid, block, 0, s, sv, rows, nrows, nshares, \
                     mod9, lim, pow3, share_recs, lists)
  if (limbs == 1) {
    if (lowkey) CZ_PLAIN(3, false, true);
    else CZ_PLAIN(3, false, false);
  } else if (limbs == 2 && lim.cutoff >= kWideLog2) {
    if (lowkey) CZ_PLAIN(4, false, true);
    else CZ_PLAIN(4, false, false);
  } else if (limbs == 2) {
    if (lowkey) CZ_PLAIN(4, true, true);
    else CZ_PLAIN(4, true, false);
  } else if (lim.cutoff >= 128) {
    CZ_PLAIN(6, false, false);
  } else {
    CZ_PLAIN(6, true, false);
  }
#undef CZ_PLAIN
  return hipGetLastError();
}


// Resident waves per CU of the persistent kernel that a launch with these arguments runs
// (its register use decides: a launch with more waves than fit leaves the extra waves to
// start late, as a tail).
template <int L, bool LOWCUT, bool LA>
static uint32_t waves_fit() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_verify_persistent<L, LOWCUT, LA>, 256,
                                                   0) != hipSuccess || nb <= 0)
    return 32;
  return (uint32_t)nb * kWavesPerBlock;
}

uint32_t persistent_waves_fit(int limbs, uint32_t cutoff, bool la) {
  static std::atomic<uint32_t> cache[2][5];
  const int w = limbs == 1 ? 0 : limbs == 2 ? (cutoff >= kWideLog2 ? 1 : 2) : (cutoff >= 128 ? 3 : 4);
  std::atomic<uint32_t>& c = cache[la][w];
  if (const uint32_t v = c.load(std::memory_order_relaxed)) return v;
  uint32_t v;
  if (la) {
    v = w == 0 ? waves_fit<3, false, true>() : w == 1 ? waves_fit<4, false, true>()
        : w == 2 ? waves_fit<4, true, true>() : w == 3 ? waves_fit<6, false, true>()
        : waves_fit<6, true, true>();
  } else {
    v = w == 0 ? waves_fit<3, false, false>() : w == 1 ? waves_fit<4, false, false>()
        : w == 2 ? waves_fit<4, true, false>() : w == 3 ? waves_fit<6, false, false>()
        : waves_fit<6, true, false>();
  }
  c.store(v, std::memory_order_relaxed);
  return v;
}

template <bool LA, bool BOUNDED>
static void launch_persistent_lb(int limbs, const SieveArrays& sv, const RowDesc* rows,
                                 const TileDesc* tiles, const uint32_t* order, Ctl* ctl,
                                 uint32_t nblocks, uint32_t U, bool lowkey, const Limits& lim,
                                 const uint64_t* pow3, DevRec* tile_recs, DeviceLists lists,
                                 hipStream_t s) {
  const dim3 grid(nblocks), block(256);
  // Launches whose ranges end at or below 2^35 on 96 bits: the low-key test of the wave max
  // (the forms without the look-ahead and the loop bound; elsewhere the top-word test, exact
  // too, only slower on such ranges).
  if constexpr (!LA && !BOUNDED) {
    if (lowkey && limbs == 1) {
#define CZ_LOWKEY(UC)                                                                        \
  hipLaunchKernelGGL((k_verify_persistent<3, false, false, false, UC, true>), grid, block, 0, s, \
                     sv, rows, tiles, order, ctl, lim, pow3, tile_recs, lists, U)
      if (U == 1) CZ_LOWKEY(1);
      else if (U == 2) CZ_LOWKEY(2);
      else if (U == 3) CZ_LOWKEY(3);
      else CZ_LOWKEY(0);
#undef CZ_LOWKEY
      return;
    }
  }
  // The hot forms (two-limb work below 2^68, with or without the look-ahead; without it only
  // where no loop bound is set): the tuner's checks per refill compiled in (its grid is 1..3,
  // and 1..4 with the look-ahead).
  constexpr bool kHot = LA || !BOUNDED;
  constexpr uint32_t kMaxUC = LA ? 4 : 3;
  if (kHot && limbs == 1 && U >= 1 && U <= kMaxUC) {
    if constexpr (!kHot) return;  // (no instantiations for the forms that are not hot)
    else if (U == 1)
      hipLaunchKernelGGL((k_verify_persistent<3, false, LA, BOUNDED, 1>), grid, block, 0, s,
                         sv, rows, tiles, order, ctl, lim, pow3, tile_recs, lists, U);
    else if (U == 2)
      hipLaunchKernelGGL((k_verify_persistent<3, false, LA, BOUNDED, 2>), grid, block, 0, s,
                         sv, rows, tiles, order, ctl, lim, pow3, tile_recs, lists, U);
    else if (U == 3)
      hipLaunchKernelGGL((k_verify_persistent<3, false, LA, BOUNDED, 3>), grid, block, 0, s,
                         sv, rows, tiles, order, ctl, lim, pow3, tile_recs, lists, U);
    else
      hipLaunchKernelGGL((k_verify_persistent<3, false, LA, BOUNDED, kMaxUC>), grid, block, 0,
                         s, sv, rows, tiles, order, ctl, lim, pow3, tile_recs, lists, U);
  } else if (limbs == 1)
    hipLaunchKernelGGL((k_verify_persistent<3, false, LA, BOUNDED>), grid, block, 0, s, sv,
                       rows, tiles, order, ctl, lim, pow3, tile_recs, lists, U);
  else if (limbs == 2 && lim.cutoff >= kWideLog2)
    hipLaunchKernelGGL((k_verify_persistent<4, false, LA, BOUNDED>), grid, block, 0, s, sv,
                       rows, tiles, order, ctl, lim, pow3, tile_recs, lists, U);
  else if (limbs == 2)
    hipLaunchKernelGGL((k_verify_persistent<4, true, LA, BOUNDED>), grid, block, 0, s, sv,
                       rows, tiles, order, ctl, lim, pow3, tile_recs, lists, U);
  else if (lim.cutoff >= 128)
    hipLaunchKernelGGL((k_verify_persistent<6, false, LA, BOUNDED>), grid, block, 0, s, sv,
                       rows, tiles, order, ctl, lim, pow3, tile_recs, lists, U);
  else
    hipLaunchKernelGGL((k_verify_persistent<6, true, LA, BOUNDED>), grid, block, 0, s, sv,
                       rows, tiles, order, ctl, lim, pow3, tile_recs, lists, U);
}

template <bool LA>
static void launch_persistent_la(int limbs, const SieveArrays& sv, const RowDesc* rows,
                                 const TileDesc* tiles, const uint32_t* order, Ctl* ctl,
                                 uint32_t nblocks, uint32_t U, bool lowkey, const Limits& lim,
                                 const uint64_t* pow3, DevRec* tile_recs, DeviceLists lists,
                                 hipStream_t s) {
  // (The look-ahead form always carries the bound.)
  if constexpr (LA) {
    launch_persistent_lb<true, true>(limbs, sv, rows, tiles, order, ctl, nblocks, U, lowkey,
                                     lim, pow3, tile_recs, lists, s);
  } else {
    if (lim.max_loop < kMaxLoop)
      launch_persistent_lb<false, true>(limbs, sv, rows, tiles, order, ctl, nblocks, U, lowkey,
                                        lim, pow3, tile_recs, lists, s);
    else
      launch_persistent_lb<false, false>(limbs, sv, rows, tiles, order, ctl, nblocks, U,
                                         lowkey, lim, pow3, tile_recs, lists, s);
  }
}

hipError_t launch_verify_persistent(int limbs, const SieveArrays& sv, const RowDesc* rows,
                                    const TileDesc* tiles, const uint32_t* order, Ctl* ctl,
                                    uint32_t nblocks, uint32_t U, bool lowkey,
                                    const Limits& lim, const uint64_t* pow3, DevRec* tile_recs,
                                    DeviceLists lists, hipStream_t s) {
  if (sv.la)
    launch_persistent_la<true>(limbs, sv, rows, tiles, order, ctl, nblocks, U, lowkey, lim,
                               pow3, tile_recs, lists, s);
  else
    launch_persistent_la<false>(limbs, sv, rows, tiles, order, ctl, nblocks, U, lowkey, lim,
                                pow3, tile_recs, lists, s);
  return hipGetLastError();
}

__global__ void k_engine_reset(Ctl* ctl, unsigned long long* cnt, uint32_t ncnt) {
  if (threadIdx.x < kXcds) ctl->share_next[threadIdx.x] = 0;
  if (threadIdx.x < ncnt) cnt[threadIdx.x] = 0;
}

hipError_t launch_engine_reset(Ctl* ctl, unsigned long long* cnt, uint32_t ncnt, hipStream_t s) {
  hipLaunchKernelGGL(k_engine_reset, dim3(1), dim3(64), 0, s, ctl, cnt, ncnt);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_make_tiles(const RowDesc* rows, const Ctl* ctl,
                                                    TileDesc* tiles) {
  const uint64_t n = ctl->nrows;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t b = rows[r].block_start;
    const uint64_t e = (r + 1 < n) ? rows[r + 1].block_start : ctl->ntiles;
    const uint32_t pieces = (uint32_t)(e - b);
    for (uint32_t j = 0; j < pieces; ++j) {
      TileDesc t;
      t.row = (uint32_t)r;
      t.piece = j;
      t.pieces = pieces;
      t.pad = 0;
      tiles[b + j] = t;
    }
  }
}

hipError_t launch_make_tiles(const RowDesc* rows, const Ctl* ctl, TileDesc* tiles, hipStream_t s) {
  hipLaunchKernelGGL(k_make_tiles, dim3(256), dim3(256), 0, s, rows, ctl, tiles);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_reduce_slots(const DevRec* parts,
                                                      const uint64_t* part_start, const Ctl* ctl,
                                                      DevRec* out) {
  __shared__ DevRec lrec[4];
  const uint64_t n = ctl->nslots;
  for (uint64_t c = blockIdx.x; c < n; c += gridDim.x) {
    DevRec r;
    rec_zero(r);
    for (uint64_t j = part_start[c] + threadIdx.x; j < part_start[c + 1]; j += 256)
      rec_merge(r, parts[j]);
    r = block_merge<4>(r, lrec);
    if (threadIdx.x == 0) out[c] = r;
    __syncthreads();
  }
}

hipError_t launch_reduce_slots(const DevRec* parts, const uint64_t* part_start, const Ctl* ctl,
                               DevRec* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_slots, dim3(1024), dim3(256), 0, s, parts, part_start, ctl, out);
  return hipGetLastError();
}

hipError_t launch_escalated(const SieveArrays& sv, const EscEntry* esc, uint64_t n,
                            uint32_t, const Limits& lim, const uint64_t* pow3, DevRec* out,
                            DeviceLists lists, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_escalated, dim3((uint32_t)blocks), dim3(256), 0, s, sv, esc, n, lim,
                     pow3, out, lists);
  return hipGetLastError();
}

hipError_t launch_reduce_chunks(const DevRec* parts, const uint64_t* part_start,
                                uint32_t nchunks, DevRec* out, hipStream_t s) {
  if (nchunks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_reduce_chunks, dim3(nchunks), dim3(256), 0, s, parts, part_start,
                     nchunks, out);
  return hipGetLastError();
}

}  // namespace cz
