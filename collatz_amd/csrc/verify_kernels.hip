// Range-verification kernels: the plain one-candidate-per-lane kernel, the three-limb pass
// for escalated candidates, and the per-chunk reduction.
//
// A candidate is n = (q << k) | r with r a sieve survivor; its start value is
// number = T^k(r) + 1 + q * 3^f (Lemma 2.3, paper p.4; the jump of
// collatz_cudav7.cu:262-278), and n mod 9 must be in {0,1,3,6,7}
// (collatz_cudav7.cu:482-494).  The schedule is in schedule.hpp.
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

// State of one candidate in flight.
template <int L>
struct Cand {
  Num<L> v;
  uint32_t n[4];
  uint32_t hash;
  uint32_t steps;
  uint32_t checks;
  uint32_t ovf;
  uint32_t slot;   // output record slot (for overflow and escalation reports)
};

// Where a candidate's values are offered for the max: one lane accumulator.
template <int L>
struct SinkOne {
  Acc<L>& acc;
  __device__ __forceinline__ void offer(const Num<L>& v, const uint32_t* n) {
    offer_max<L>(acc, v, n);
  }
};

// Two accumulator sets (two tiles in flight per wave); cm mirrors the max of the set the
// lane's candidate belongs to, so the common case costs one compare.
template <int L>
struct SinkTwo {
  Acc<L>& a0;
  Acc<L>& a1;
  uint32_t cset;
  uint32_t* cm;
  __device__ __forceinline__ void offer(const Num<L>& v, const uint32_t* n) {
    if (ge_words<L>(v.w, cm)) {
      uint32_t ms[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) ms[i] = cset ? a1.mstart[i] : a0.mstart[i];
      bool eq = true;
#pragma unroll
      for (int i = 0; i < L; ++i) eq = eq && v.w[i] == cm[i];
      if (!eq || !ge_words<4>(n, ms)) {
#pragma unroll
        for (int i = 0; i < L; ++i) cm[i] = v.w[i];
        if (cset) {
#pragma unroll
          for (int i = 0; i < L; ++i) a1.mx[i] = v.w[i];
#pragma unroll
          for (int i = 0; i < 4; ++i) a1.mstart[i] = n[i];
        } else {
#pragma unroll
          for (int i = 0; i < L; ++i) a0.mx[i] = v.w[i];
#pragma unroll
          for (int i = 0; i < 4; ++i) a0.mstart[i] = n[i];
        }
      }
    }
  }
};

// Outcome of running a candidate to its end.
enum : uint32_t { kDone = 0, kEscalate = 1 };

// Set up a candidate: v = M + q3 (q3 has L words; q3_big says q * 3^f needs more), n =
// qb | r.  Returns false if the start value does not fit L words (then the candidate must
// go to the three-limb pass).
template <int L>
__device__ __forceinline__ bool cand_start(Cand<L>& c, uint64_t M, const uint32_t* q3,
                                           bool q3_big, uint64_t qb_lo, uint64_t qb_hi,
                                           uint64_t r) {
  uint64_t t = (uint64_t)q3[0] + (uint32_t)M;
  c.v.w[0] = (uint32_t)t;
  t = (uint64_t)q3[1] + (uint32_t)(M >> 32) + (t >> 32);
  c.v.w[1] = (uint32_t)t;
#pragma unroll
  for (int i = 2; i < L; ++i) {
    t = (uint64_t)q3[i] + (t >> 32);
    c.v.w[i] = (uint32_t)t;
  }
  const bool fits = (t >> 32) == 0 && !q3_big;
  const uint64_t nlo = qb_lo | r;
  c.n[0] = (uint32_t)nlo;
  c.n[1] = (uint32_t)(nlo >> 32);
  c.n[2] = (uint32_t)qb_hi;
  c.n[3] = (uint32_t)(qb_hi >> 32);
  c.hash = 0;
  c.steps = 0;
  c.checks = 0;
  c.ovf = 0;
  return fits;
}

// Overflow reports: (n_lo, n_hi, slot, 0) per event, in the order the events happen.
__device__ __forceinline__ void push_overflow(const DeviceLists& lists, const uint32_t* n,
                                              uint32_t slot) {
  const unsigned long long i = atomicAdd(lists.ovf_count, 1ull);
  if (i < lists.ovf_cap) {
    lists.ovf[4 * i] = pair(n[0], n[1]);
    lists.ovf[4 * i + 1] = pair(n[2], n[3]);
    lists.ovf[4 * i + 2] = slot;
    lists.ovf[4 * i + 3] = 0;
  }
}

__device__ __forceinline__ void push_escalate(const DeviceLists& lists, const uint32_t* n,
                                              uint64_t res_index, uint32_t slot) {
  const unsigned long long i = atomicAdd(lists.esc_count, 1ull);
  if (i < lists.esc_cap) {
    EscEntry e;
    e.n_lo = pair(n[0], n[1]);
    e.n_hi = pair(n[2], n[3]);
    e.res_index = res_index;
    e.slot = slot;
    e.pad = 0;
    lists.esc[i] = e;
  }
}

// After an odd run: offer to the max, then the cutoff / limb limit.  Returns true if the
// candidate has to be escalated (two limbs only).
template <int L, bool LOWCUT, class Sink>
__device__ __forceinline__ bool after_odd(Cand<L>& c, Sink& sink, const Limits& lim,
                                          const DeviceLists& lists) {
  sink.offer(c.v, c.n);
  if (over_fast<L, LOWCUT>(c.v, lim.fast_thr)) {
    if (ge_pow2<L>(c.v, lim.cutoff)) {
      // Overflow: report the start and continue from 2 (collatz_cudav7.cu:295-301).
      c.ovf += 1;
      push_overflow(lists, c.n, c.slot);
      c.v.w[0] = 2;
#pragma unroll
      for (int i = 1; i < L; ++i) c.v.w[i] = 0;
    } else if (L == 4) {
      return true;
    }
  }
  return false;
}

// One check.  Returns 0 while running, 1 when the candidate finished (number < n), 2 when
// it has to be escalated.
template <int L, bool LOWCUT, class Sink>
__device__ __forceinline__ uint32_t one_check(Cand<L>& c, Sink& sink, const Limits& lim,
                                              const DeviceLists& lists) {
  uint32_t a1 = odd_run<L>(c.v);
  if (after_odd<L, LOWCUT>(c, sink, lim, lists)) return 2;
  uint32_t b1 = even_run<L>(c.v);
  uint32_t a2 = odd_run<L>(c.v);
  if (after_odd<L, LOWCUT>(c, sink, lim, lists)) return 2;
  uint32_t b2 = even_run<L>(c.v);
  c.steps += a1 + b1 + a2 + b2;
  c.checks += 1;
  c.hash ^= c.v.w[0];
  return lt_base<L>(c.v, c.n) ? 1u : 0u;
}

template <int L>
__device__ __forceinline__ void finish(const Cand<L>& c, Acc<L>& acc, uint32_t k) {
  const uint32_t J = k + c.steps;
  acc.cnt += 1;
  acc.hash ^= c.hash;
  acc.ssum += J;
  acc.sxor ^= J;
  acc.ovf += c.ovf;
  acc.chk += c.checks;
}

// Start-of-candidate test: the start value is offered to the max, and on two limbs a start
// at or above 2^123 goes to the three-limb pass (an odd run multiplies by at most
// (3/2)^8 < 2^5, so two limbs hold its product only below that).
template <int L, class Sink>
__device__ __forceinline__ bool cand_enter(const Cand<L>& c, Sink& sink) {
  sink.offer(c.v, c.n);
  return !(L == 4 && pair(c.v.w[2], c.v.w[3]) >= (1ull << (kWideLog2 - 64)));
}

// Runs candidate c to completion.  Returns false if it was escalated.
template <int L, bool LOWCUT>
__device__ __forceinline__ bool run_candidate(Cand<L>& c, Acc<L>& acc, const Limits& lim,
                                              const DeviceLists& lists, uint32_t k) {
  SinkOne<L> sink{acc};
  if (!cand_enter<L>(c, sink)) return false;
  for (uint32_t guard = 0; guard < kMaxChecks; ++guard) {
    const uint32_t s = one_check<L, LOWCUT>(c, sink, lim, lists);
    if (s == 1) {
      finish<L>(c, acc, k);
      return true;
    }
    if (s == 2) return false;
  }
  atomicAdd(lists.stall_count, 1ull);
  return true;
}

// ---- plain kernel: one lane per (residue, q) slot --------------------------------------
template <int L, bool LOWCUT>
__global__ __launch_bounds__(256) void k_verify_plain(SieveArrays sv, const RowDesc* rows,
                                                      uint32_t nrows, uint32_t R, uint32_t mod9,
                                                      Limits lim, const uint64_t* pow3,
                                                      DevRec* recs, DeviceLists lists) {
  __shared__ uint32_t q3[41][6];
  __shared__ DevRec lrec[4];
  const uint64_t blk = blockIdx.x;
  uint32_t lo = 0, hi = nrows;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (rows[mid].block_start <= blk) lo = mid;
    else hi = mid;
  }
  const RowDesc row = rows[lo];
  if (threadIdx.x < 41) q_times_pow3(row.q_lo, row.q_hi, pow3[threadIdx.x], q3[threadIdx.x]);
  __syncthreads();

  Acc<L> acc;
  acc.init();
  const uint64_t first = row.i_lo + (blk - row.block_start) * (256ull * R) + threadIdx.x;
  for (uint32_t it = 0; it < R; ++it) {
    const uint64_t i = first + (uint64_t)it * 256;
    if (i >= row.i_hi) break;
    const uint64_t r = sv.res[i];
    const uint64_t M = sv.num[i];
    const uint32_t meta = sv.meta[i];
    Cand<L> c;
    const uint32_t f = meta_f(meta);
    c.slot = row.slot;
    const bool big = L == 4 ? (q3[f][4] | q3[f][5]) != 0 : false;
    if (!cand_start<L>(c, M, q3[f], big, row.qb_lo, row.qb_hi, r) ||
        !run_candidate<L, LOWCUT>(c, acc, lim, lists, sv.k)) {
      push_escalate(lists, c.n, i, row.slot);
    }
  }
  DevRec rec = block_merge<4>(acc_to_rec<L>(acc), lrec);
  if (threadIdx.x == 0) recs[blk] = rec;
}

// ---- escalated candidates on three limbs, one thread each -----------------------------
__global__ __launch_bounds__(256) void k_escalated(SieveArrays sv, const EscEntry* esc,
                                                   uint64_t n, Limits lim,
                                                   const uint64_t* pow3, DevRec* out,
                                                   DeviceLists lists) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const EscEntry e = esc[t];
  const uint64_t r = sv.res[e.res_index];
  const uint64_t M = sv.num[e.res_index];
  const uint32_t f = meta_f(sv.meta[e.res_index]);
  const uint32_t k = sv.k;
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
  if (lim.cutoff >= 128) run_candidate<6, false>(c, acc, lim, lists, k);
  else run_candidate<6, true>(c, acc, lim, lists, k);
  DevRec rec = acc_to_rec<6>(acc);
  rec.pad = e.slot;
  out[t] = rec;
}


// ---- persistent kernel: waves with lane refill and work stealing ------------------------
// Every wave is an independent worker.  It takes tiles (sieve index ranges of one row) from
// a global counter, and keeps its 64 lanes busy: a lane whose candidate has descended takes
// the next residue of the tile at the next refill point, so the wave never waits for its
// slowest lane except at the very end.  Two tiles can be in flight per wave (two
// accumulator sets); a tile's record is written when its last candidate is done.  Candidate
// data is prefetched one refill ahead so the loads overlap the checks in between.
constexpr int kWavesPerBlock = 4;

template <int L>
struct WaveTables {
  uint32_t q3[2][41][L];   // q * 3^f per set
  uint32_t big[2][41];     // q * 3^f does not fit L words
  uint64_t qb[2][2];       // q << k per set
};

template <int L, bool LOWCUT>
__global__ __launch_bounds__(256) void k_verify_persistent(
    SieveArrays sv, const RowDesc* rows, const TileDesc* tiles, const uint32_t* ntiles_ptr,
    uint32_t* tile_counter, Limits lim, const uint64_t* pow3, DevRec* tile_recs,
    DeviceLists lists, uint32_t U) {
  __shared__ WaveTables<L> tabs[kWavesPerBlock];
  const uint32_t wv = threadIdx.x >> 6;
  const uint32_t lane = threadIdx.x & 63;
  WaveTables<L>& T = tabs[wv];
  const uint32_t ntiles = *ntiles_ptr;

  // Wave-uniform state.
  int32_t set_tile0 = -1, set_tile1 = -1;
  uint32_t set_slot0 = 0, set_slot1 = 0;
  int32_t fill = -1;
  uint64_t cur = 0, end = 0;
  bool exhausted = false;

  // Lane state.
  Acc<L> acc0, acc1;
  acc0.init();
  acc1.init();
  Cand<L> c;
  bool act = false;
  uint32_t cset = 0, cslot = 0;
  uint64_t cidx = 0;
  uint32_t cm[L];
  bool pf = false;
  uint64_t pf_r = 0, pf_M = 0, pf_idx = 0;
  uint32_t pf_meta = 0, pf_set = 0, pf_slot = 0;

  for (;;) {
    // (1) Lanes without a candidate start their prefetched one.
    if (!act && pf) {
      pf = false;
      const uint32_t f = meta_f(pf_meta);
      const bool fits = cand_start<L>(c, pf_M, T.q3[pf_set][f], T.big[pf_set][f] != 0,
                                      T.qb[pf_set][0], T.qb[pf_set][1], pf_r);
      cset = pf_set;
      cslot = pf_slot;
      cidx = pf_idx;
      c.slot = pf_slot;
#pragma unroll
      for (int i = 0; i < L; ++i) cm[i] = cset ? acc1.mx[i] : acc0.mx[i];
      SinkTwo<L> sink{acc0, acc1, cset, cm};
      if (fits && cand_enter<L>(c, sink)) {
        act = true;
      } else {
        push_escalate(lists, c.n, cidx, cslot);
      }
    }
    // (2) Assign new prefetches from the tile cursor; open tiles as needed.
    uint64_t need = __ballot(!pf);
    while (need && !exhausted) {
      if (cur >= end) {
        // The fill tile is used up: open the next tile in the other set, if it is free.
        const int32_t other = fill < 0 ? 0 : (fill ^ 1);
        const int32_t other_tile = other ? set_tile1 : set_tile0;
        if (other_tile >= 0) break;
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(tile_counter, 1u);
        t = __builtin_amdgcn_readfirstlane(t);
        if (t >= ntiles) {
          exhausted = true;
          fill = -1;
          break;
        }
        const TileDesc td = tiles[t];
        const RowDesc rw = rows[td.row];
        if (lane < 41) {
          uint32_t w6[6];
          q_times_pow3(rw.q_lo, rw.q_hi, pow3[lane], w6);
#pragma unroll
          for (int i = 0; i < L; ++i) T.q3[other][lane][i] = w6[i];
          uint32_t big = 0;
#pragma unroll
          for (int i = L; i < 6; ++i) big |= w6[i];
          T.big[other][lane] = big != 0;
        }
        if (lane == 0) {
          T.qb[other][0] = rw.qb_lo;
          T.qb[other][1] = rw.qb_hi;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (other) {
          set_tile1 = (int32_t)t;
          set_slot1 = rw.slot;
        } else {
          set_tile0 = (int32_t)t;
          set_slot0 = rw.slot;
        }
        fill = other;
        cur = td.i_lo;
        end = td.i_hi;
        continue;
      }
      const uint32_t want = __popcll(need);
      const uint64_t avail = end - cur;
      const uint32_t take = avail < want ? (uint32_t)avail : want;
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
      if (!pf && rank < take) {
        const uint64_t i = cur + rank;
        pf_r = sv.res[i];
        pf_M = sv.num[i];
        pf_meta = sv.meta[i];
        pf_idx = i;
        pf_set = (uint32_t)fill;
        pf_slot = fill ? set_slot1 : set_slot0;
        pf = true;
      }
      cur += take;
      need = __ballot(!pf);
    }
    // (3) Write the record of every set that takes no more candidates and has none left.
#pragma unroll
    for (uint32_t sset = 0; sset < 2; ++sset) {
      const int32_t tile = sset ? set_tile1 : set_tile0;
      if (tile < 0 || (int32_t)sset == fill) continue;
      const bool busy = (act && cset == sset) || (pf && pf_set == sset);
      if (__ballot(busy) != 0) continue;
      DevRec r = acc_to_rec<L>(sset ? acc1 : acc0);
      wave_merge(r);
      if (lane == 0) tile_recs[tile] = r;
      if (sset) {
        acc1.init();
        set_tile1 = -1;
      } else {
        acc0.init();
        set_tile0 = -1;
      }
    }
    if (exhausted && set_tile0 < 0 && set_tile1 < 0) break;
    // (4) Up to U checks.
    if (act) {
      SinkTwo<L> sink{acc0, acc1, cset, cm};
      for (uint32_t u = 0; u < U; ++u) {
        const uint32_t st = one_check<L, LOWCUT>(c, sink, lim, lists);
        if (st == 1) {
          if (cset) finish<L>(c, acc1, sv.k);
          else finish<L>(c, acc0, sv.k);
          act = false;
          break;
        }
        if (st == 2) {
          push_escalate(lists, c.n, cidx, cslot);
          act = false;
          break;
        }
        if (c.checks >= kMaxChecks) {
          atomicAdd(lists.stall_count, 1ull);
          act = false;
          break;
        }
      }
    }
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
                               uint32_t nrows, uint64_t nblocks, uint32_t R, uint32_t mod9,
                               const Limits& lim, const uint64_t* pow3, DevRec* block_recs,
                               DeviceLists lists, hipStream_t s) {
  if (nblocks == 0) return hipSuccess;
  if (nblocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
  const dim3 grid((uint32_t)nblocks), block(256);
  if (limbs == 2)
    hipLaunchKernelGGL((k_verify_plain<4, false>), grid, block, 0, s, sv, rows, nrows, R, mod9,
                       lim, pow3, block_recs, lists);
  else if (lim.cutoff >= 128)
    hipLaunchKernelGGL((k_verify_plain<6, false>), grid, block, 0, s, sv, rows, nrows, R, mod9,
                       lim, pow3, block_recs, lists);
  else
    hipLaunchKernelGGL((k_verify_plain<6, true>), grid, block, 0, s, sv, rows, nrows, R, mod9,
                       lim, pow3, block_recs, lists);
  return hipGetLastError();
}


hipError_t launch_verify_persistent(int limbs, const SieveArrays& sv, const RowDesc* rows,
                                    const TileDesc* tiles, const uint32_t* ntiles,
                                    uint32_t* tile_counter, uint32_t nblocks, uint32_t U,
                                    const Limits& lim, const uint64_t* pow3, DevRec* tile_recs,
                                    DeviceLists lists, hipStream_t s) {
  const dim3 grid(nblocks), block(256);
  if (limbs == 2)
    hipLaunchKernelGGL((k_verify_persistent<4, false>), grid, block, 0, s, sv, rows, tiles,
                       ntiles, tile_counter, lim, pow3, tile_recs, lists, U);
  else if (lim.cutoff >= 128)
    hipLaunchKernelGGL((k_verify_persistent<6, false>), grid, block, 0, s, sv, rows, tiles,
                       ntiles, tile_counter, lim, pow3, tile_recs, lists, U);
  else
    hipLaunchKernelGGL((k_verify_persistent<6, true>), grid, block, 0, s, sv, rows, tiles,
                       ntiles, tile_counter, lim, pow3, tile_recs, lists, U);
  return hipGetLastError();
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_zero_u64(unsigned long long* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

hipError_t launch_zero(uint32_t* p32, uint32_t n32, unsigned long long* p64, uint32_t n64,
                       hipStream_t s) {
  if (n32) hipLaunchKernelGGL(k_zero_u32, dim3((n32 + 255) / 256), dim3(256), 0, s, p32, n32);
  if (n64) hipLaunchKernelGGL(k_zero_u64, dim3((n64 + 255) / 256), dim3(256), 0, s, p64, n64);
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
