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
};

// Outcome of running a candidate to its end.
enum : uint32_t { kDone = 0, kEscalate = 1 };

// Set up a candidate: v = M + q3 (L words), n = qb | r.  Returns false if the start value
// does not fit L words (then the candidate must go to the three-limb pass).
template <int L>
__device__ __forceinline__ bool cand_start(Cand<L>& c, uint64_t M, const uint32_t* q3,
                                           uint64_t qb_lo, uint64_t qb_hi, uint64_t r) {
  uint64_t t = (uint64_t)q3[0] + (uint32_t)M;
  c.v.w[0] = (uint32_t)t;
  t = (uint64_t)q3[1] + (uint32_t)(M >> 32) + (t >> 32);
  c.v.w[1] = (uint32_t)t;
#pragma unroll
  for (int i = 2; i < L; ++i) {
    t = (uint64_t)q3[i] + (t >> 32);
    c.v.w[i] = (uint32_t)t;
  }
  bool fits = (t >> 32) == 0;
#pragma unroll
  for (int i = L; i < 6; ++i) fits = fits && q3[i] == 0;
  const uint64_t nlo = qb_lo | r;
  c.n[0] = (uint32_t)nlo;
  c.n[1] = (uint32_t)(nlo >> 32);
  c.n[2] = (uint32_t)qb_hi;
  c.n[3] = (uint32_t)(qb_hi >> 32);
  c.hash = 0;
  c.steps = 0;
  return fits;
}

__device__ __forceinline__ void push_overflow(const DeviceLists& lists, const uint32_t* n) {
  const unsigned long long i = atomicAdd(lists.ovf_count, 1ull);
  if (i < lists.ovf_cap) {
    lists.ovf[2 * i] = pair(n[0], n[1]);
    lists.ovf[2 * i + 1] = pair(n[2], n[3]);
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
template <int L, bool LOWCUT>
__device__ __forceinline__ bool after_odd(Cand<L>& c, Acc<L>& acc, const Limits& lim,
                                          const DeviceLists& lists) {
  offer_max<L>(acc, c.v, c.n);
  if (over_fast<L, LOWCUT>(c.v, lim.fast_thr)) {
    if (ge_pow2<L>(c.v, lim.cutoff)) {
      // Overflow: report the start and continue from 2 (collatz_cudav7.cu:295-301).
      acc.ovf += 1;
      push_overflow(lists, c.n);
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
template <int L, bool LOWCUT>
__device__ __forceinline__ uint32_t one_check(Cand<L>& c, Acc<L>& acc, const Limits& lim,
                                              const DeviceLists& lists) {
  uint32_t a1 = odd_run<L>(c.v);
  if (after_odd<L, LOWCUT>(c, acc, lim, lists)) return 2;
  uint32_t b1 = even_run<L>(c.v);
  uint32_t a2 = odd_run<L>(c.v);
  if (after_odd<L, LOWCUT>(c, acc, lim, lists)) return 2;
  uint32_t b2 = even_run<L>(c.v);
  c.steps += a1 + b1 + a2 + b2;
  acc.chk += 1;
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
}

// Runs candidate c to completion.  Returns false if it was escalated.
template <int L, bool LOWCUT>
__device__ __forceinline__ bool run_candidate(Cand<L>& c, Acc<L>& acc, const Limits& lim,
                                              const DeviceLists& lists, uint32_t k) {
  offer_max<L>(acc, c.v, c.n);
  // Two limbs hold an odd run's product only if the value entering it is below 2^123.
  if (L == 4 && pair(c.v.w[2], c.v.w[3]) >= (1ull << (kWideLog2 - 64))) return false;
  for (uint32_t guard = 0; guard < kMaxChecks; ++guard) {
    const uint32_t s = one_check<L, LOWCUT>(c, acc, lim, lists);
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
    uint32_t s9 = row.qb_mod9 + meta_mod9(meta);
    s9 -= (s9 >= 9u) ? 9u : 0u;
    if (!mod9 || ((kMod9Keep >> s9) & 1u)) {
      Cand<L> c;
      const uint32_t f = meta_f(meta);
      if (!cand_start<L>(c, M, q3[f], row.qb_lo, row.qb_hi, r) ||
          !run_candidate<L, LOWCUT>(c, acc, lim, lists, sv.k)) {
        push_escalate(lists, c.n, i, row.slot);
      }
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
  cand_start<6>(c, M, q3, qb_lo, qb_hi, r);
  if (lim.cutoff >= 128) run_candidate<6, false>(c, acc, lim, lists, k);
  else run_candidate<6, true>(c, acc, lim, lists, k);
  DevRec rec = acc_to_rec<6>(acc);
  rec.pad = e.slot;
  out[t] = rec;
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
