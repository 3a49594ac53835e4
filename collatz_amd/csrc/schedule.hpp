// The iteration schedule of the reference on L 32-bit words (L = 3: 96 bits, for two-limb
// work that stays below 2^91; L = 4: two u64 limbs; L = 6: three u64 limbs), plus the
// per-lane accumulators and wave reductions.
//
// The schedule follows compute_kernel (collatz_cudav7.cu:280-366):
//   check = odd run, even run, odd run, even run; then hash ^= low word; stop if number < n
//   odd run : a = ctz(number | 2^8), number = (number >> a) * 3^a
//   even run: b = ctz((number - 1) | 2^31), number = ((number - 1) >> b) + 1
// where number = T-value + 1.
#pragma once

#include "internal.hpp"

namespace cz {

template <int L>
struct Num {
  uint32_t w[L];
};

__device__ __forceinline__ uint32_t alignr(uint32_t hi, uint32_t lo, uint32_t s) {
  return __builtin_amdgcn_alignbit(hi, lo, s);  // low word of (hi:lo) >> (s & 31)
}

// 3^a for a in [0, 8], from a 16-entry table in LDS (every kernel that iterates stages it
// once per workgroup with stage_pow3()).
struct Pow3Lds {
  uint32_t v[16];
};

__device__ __forceinline__ void stage_pow3(Pow3Lds& t) {
  if (threadIdx.x < 16) {
    uint32_t x = 1;
    for (uint32_t i = 0; i < threadIdx.x && i < 8; ++i) x *= 3;
    t.v[threadIdx.x] = x;
  }
}

__device__ __forceinline__ uint32_t addc(uint32_t a, uint32_t b, uint32_t cin, uint32_t& cout) {
  return __builtin_addc(a, b, cin, &cout);
}

template <int L>
__device__ __forceinline__ uint32_t odd_run(Num<L>& v, const Pow3Lds& p3) {
  const uint32_t a = __builtin_ctz(v.w[0] | 0x100u);
  const uint32_t m = p3.v[a];
  // Products of the shifted words by m (< 2^13), then one carry chain.
  uint64_t p[L];
#pragma unroll
  for (int i = 0; i < L - 1; ++i) p[i] = (uint64_t)alignr(v.w[i + 1], v.w[i], a) * m;
  p[L - 1] = (uint64_t)(v.w[L - 1] >> a) * m;
  uint32_t c = 0;
  v.w[0] = (uint32_t)p[0];
#pragma unroll
  for (int i = 1; i < L; ++i) v.w[i] = addc((uint32_t)p[i], (uint32_t)(p[i - 1] >> 32), c, c);
  return a;
}

template <int L>
__device__ __forceinline__ uint32_t even_run(Num<L>& v) {
  // number - 1 is taken in the low word alone, as the reference does: after an odd run
  // that stopped below a = 8 the number is odd, so nothing borrows; after one that stopped
  // at the cap it may be even, and then b = 0 and the result is the number itself (a low
  // word of 0 wraps to 2^32 - 1, is shifted by 0 and wraps back, with no carry).
  const uint32_t w0 = v.w[0] - 1u;
  const uint32_t b = __builtin_ctz(w0 | 0x80000000u);
  const uint32_t x0 = alignr(v.w[1], w0, b) + 1u;
#pragma unroll
  for (int i = 1; i < L - 1; ++i) v.w[i] = alignr(v.w[i + 1], v.w[i], b);
  v.w[L - 1] = v.w[L - 1] >> b;
  v.w[0] = x0;
  // The + 1 carries out of the low word only if the shifted low word was 2^32 - 1: one
  // case in 2^31, taken as a branch rather than a carry chain on every run.  With b = 0
  // the low word wrapped without a carry (see above), so the carry is b != 0.
  if (__builtin_expect(x0 == 0u, 0)) {
    uint32_t c = b != 0u ? 1u : 0u;
#pragma unroll
    for (int i = 1; i < L; ++i) v.w[i] = addc(v.w[i], 0u, c, c);
  }
  return b;
}

__device__ __forceinline__ uint64_t pair(uint32_t lo, uint32_t hi) {
  return ((uint64_t)hi << 32) | lo;
}

// x >= y over L words, as 64-bit pair compares combined without branches (bitwise, not
// short-circuit, so every lane takes the same three or five compares).
template <int L>
__device__ __forceinline__ bool ge_words(const uint32_t* x, const uint32_t* y) {
  if constexpr (L == 3)
    return (x[2] > y[2]) | ((x[2] == y[2]) & (pair(x[0], x[1]) >= pair(y[0], y[1])));
  const uint64_t xh = pair(x[2], x[3]), yh = pair(y[2], y[3]);
  const bool ge = (xh > yh) | ((xh == yh) & (pair(x[0], x[1]) >= pair(y[0], y[1])));
  if constexpr (L == 4) {
    return ge;
  } else {
    const uint64_t xt = pair(x[4], x[5]), yt = pair(y[4], y[5]);
    return (xt > yt) | ((xt == yt) & ge);
  }
}

// number < n for n = n_lo + 2^64 n_hi.  Three words hold starts below 2^68 only, so there
// n_hi < 2^4 and the top compare is one 32-bit compare.
template <int L>
__device__ __forceinline__ bool lt_base(const Num<L>& v, uint64_t n_lo, uint64_t n_hi) {
  if constexpr (L == 3) {
    const uint32_t nh = (uint32_t)n_hi;
    return (v.w[2] < nh) | ((v.w[2] == nh) & (pair(v.w[0], v.w[1]) < n_lo));
  }
  const uint64_t vh = pair(v.w[2], v.w[3]);
  bool lt = (vh < n_hi) | ((vh == n_hi) & (pair(v.w[0], v.w[1]) < n_lo));
  if constexpr (L > 4) lt = lt & ((v.w[4] | v.w[5]) == 0);
  return lt;
}


// Fast test "number may be >= the limit": a 64-bit compare on the top word pair (plus, for
// three limbs with a cutoff below 2^128, the top two words being nonzero).
template <int L, bool LOWCUT>
__device__ __forceinline__ bool over_fast(const Num<L>& v, uint64_t thr) {
  if constexpr (L == 4) {
    return pair(v.w[2], v.w[3]) >= thr;
  } else if constexpr (!LOWCUT) {
    return pair(v.w[4], v.w[5]) >= thr;
  } else {
    return (v.w[4] | v.w[5]) != 0 || pair(v.w[2], v.w[3]) >= thr;
  }
}

template <int L>
__device__ __forceinline__ bool ge_pow2(const Num<L>& v, uint32_t c) {
  const uint32_t cw = c >> 5, cb = c & 31u;
  bool ge = false;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    if ((uint32_t)i > cw) ge |= v.w[i] != 0;
    else if ((uint32_t)i == cw) ge |= (v.w[i] >> cb) != 0;
  }
  return ge;
}

// Per-lane accumulators over the candidates one lane finishes.
template <int L>
struct Acc {
  uint32_t cnt, hash, ssum, sxor, ovf, chk;
  uint32_t mx[L];      // max number seen (T-value + 1), 0 if none
  uint64_t ms_lo, ms_hi;  // its start
  uint64_t mxh;        // L = 3: words 1 and 2 of mx as one register pair, for the filter
  __device__ void init() {
    cnt = hash = ssum = sxor = ovf = chk = 0;
    mxh = 0;
#pragma unroll
    for (int i = 0; i < L; ++i) mx[i] = 0;
    ms_lo = ms_hi = 0;
  }
};

// Offer (value, start) to the lane max: take it if larger, or equal with a smaller start.
template <int L>
__device__ __forceinline__ void offer_take(Acc<L>& acc, const Num<L>& v, uint64_t n_lo,
                                           uint64_t n_hi) {
  bool eq = true;
#pragma unroll
  for (int i = 0; i < L; ++i) eq = eq && v.w[i] == acc.mx[i];
  const bool smaller_start = (n_hi < acc.ms_hi) | ((n_hi == acc.ms_hi) & (n_lo < acc.ms_lo));
  const bool take = !eq || smaller_start;
  if (take) {
    uint32_t w[L];
#pragma unroll
    for (int i = 0; i < L; ++i) w[i] = v.w[i];
    // An element-wise copy of three words would be merged into a straddling 64-bit copy
    // that keeps the value's array out of registers; the empty asm keeps the copies apart.
    if constexpr (L == 3) asm("" : "+v"(w[0]), "+v"(w[1]), "+v"(w[2]));
#pragma unroll
    for (int i = 0; i < L; ++i) acc.mx[i] = w[i];
    acc.ms_lo = n_lo;
    acc.ms_hi = n_hi;
    if constexpr (L == 3) acc.mxh = ((uint64_t)w[2] << 32) | w[1];
  }
}

template <int L>
__device__ __forceinline__ void offer_max(Acc<L>& acc, const Num<L>& v, uint64_t n_lo,
                                          uint64_t n_hi) {
  if constexpr (L == 3) {
    // v >= max needs (v >> 32) >= (max >> 32): one 64-bit compare that a lane's running max
    // fails almost always; the exact test runs only behind it.
    const uint64_t vh = ((uint64_t)v.w[2] << 32) | v.w[1];
    const uint64_t mh = acc.mxh;
    if (__builtin_expect(vh >= mh, 0)) {
      if ((vh > mh) | (v.w[0] >= acc.mx[0])) offer_take<L>(acc, v, n_lo, n_hi);
    }
  } else if constexpr (L == 4) {
    // The same filter on the top word pair, which holds the max's magnitude on four words.
    if (__builtin_expect(pair(v.w[2], v.w[3]) >= pair(acc.mx[2], acc.mx[3]), 0)) {
      if (ge_words<L>(v.w, acc.mx)) offer_take<L>(acc, v, n_lo, n_hi);
    }
  } else {
    if (ge_words<L>(v.w, acc.mx)) offer_take<L>(acc, v, n_lo, n_hi);
  }
}

// ---- record merging ------------------------------------------------------------------
__device__ __forceinline__ int cmp3(const uint64_t* a, const uint64_t* b) {
  if (a[2] != b[2]) return a[2] < b[2] ? -1 : 1;
  if (a[1] != b[1]) return a[1] < b[1] ? -1 : 1;
  if (a[0] != b[0]) return a[0] < b[0] ? -1 : 1;
  return 0;
}

__device__ __forceinline__ bool less128(uint64_t alo, uint64_t ahi, uint64_t blo, uint64_t bhi) {
  return ahi < bhi || (ahi == bhi && alo < blo);
}

__device__ __forceinline__ void rec_zero(DevRec& r) {
  r.count = r.stop_sum = r.stop_xor = r.overflow = r.checks = 0;
  r.maxnum[0] = r.maxnum[1] = r.maxnum[2] = 0;
  r.max_start[0] = r.max_start[1] = 0;
  r.hash = 0;
  r.pad = 0;
}

// Merge b into a.  The max goes to the larger maxnum, ties to the smaller start; records
// without candidates carry maxnum 0, which every candidate exceeds (number >= 2).
__device__ __forceinline__ void rec_merge(DevRec& a, const DevRec& b) {
  const int c = cmp3(b.maxnum, a.maxnum);
  if (c > 0 || (c == 0 && b.count && less128(b.max_start[0], b.max_start[1], a.max_start[0],
                                             a.max_start[1]))) {
    a.maxnum[0] = b.maxnum[0];
    a.maxnum[1] = b.maxnum[1];
    a.maxnum[2] = b.maxnum[2];
    a.max_start[0] = b.max_start[0];
    a.max_start[1] = b.max_start[1];
  }
  a.count += b.count;
  a.stop_sum += b.stop_sum;
  a.stop_xor ^= b.stop_xor;
  a.overflow += b.overflow;
  a.checks += b.checks;
  a.hash ^= b.hash;
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t x, int m) {
  const uint32_t lo = __shfl_xor((uint32_t)x, m, 64);
  const uint32_t hi = __shfl_xor((uint32_t)(x >> 32), m, 64);
  return ((uint64_t)hi << 32) | lo;
}

// Butterfly merge across the 64 lanes of a wave; every lane ends with the wave's record.
__device__ __forceinline__ void wave_merge(DevRec& r) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    DevRec o;
    o.count = shfl_xor64(r.count, m);
    o.stop_sum = shfl_xor64(r.stop_sum, m);
    o.stop_xor = shfl_xor64(r.stop_xor, m);
    o.overflow = shfl_xor64(r.overflow, m);
    o.checks = shfl_xor64(r.checks, m);
    o.maxnum[0] = shfl_xor64(r.maxnum[0], m);
    o.maxnum[1] = shfl_xor64(r.maxnum[1], m);
    o.maxnum[2] = shfl_xor64(r.maxnum[2], m);
    o.max_start[0] = shfl_xor64(r.max_start[0], m);
    o.max_start[1] = shfl_xor64(r.max_start[1], m);
    o.hash = __shfl_xor(r.hash, m, 64);
    rec_merge(r, o);
  }
}

template <int L>
__device__ __forceinline__ DevRec acc_to_rec(const Acc<L>& a) {
  DevRec r;
  r.count = a.cnt;
  r.stop_sum = a.ssum;
  r.stop_xor = a.sxor;
  r.overflow = a.ovf;
  r.checks = a.chk;
  r.hash = a.hash;
  r.pad = 0;
  r.maxnum[0] = pair(a.mx[0], a.mx[1]);
  if constexpr (L == 3) r.maxnum[1] = a.mx[2];
  else r.maxnum[1] = pair(a.mx[2], a.mx[3]);
  if constexpr (L > 4) r.maxnum[2] = pair(a.mx[4], a.mx[5]);
  else r.maxnum[2] = 0;
  r.max_start[0] = a.ms_lo;
  r.max_start[1] = a.ms_hi;
  if (a.cnt == 0) {
    r.maxnum[0] = r.maxnum[1] = r.maxnum[2] = 0;
    r.max_start[0] = r.max_start[1] = 0;
  }
  return r;
}

// Block-wide merge for blocks of `nwaves` waves; thread 0 returns the block's record.
template <int NWAVES>
__device__ __forceinline__ DevRec block_merge(DevRec r, DevRec* lds) {
  wave_merge(r);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) lds[wv] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < NWAVES; ++i) rec_merge(r, lds[i]);
  }
  return r;
}

}  // namespace cz
