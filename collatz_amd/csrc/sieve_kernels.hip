// Sieve construction on the device: the reference's search1 (collatz_cudav7.cu:497-563)
// evaluated per residue instead of by recursion, then compaction of the resulting bit
// array into a sorted survivor list.
//
// Extensions of a root node at depth s0 are numbered h in [0, 2^e), e = D - s0; the
// residue is n = base + h * 2^s0 and bit j of h decides bit s0 + j of n.
//   1. prefix kernel : every h_low < 2^p walks p steps; its state (number, f, doubleeven,
//                      numevens) and an alive bit are stored.
//   2. expand kernel : one thread per 32-bit output word; the prefixes of the word that are
//                      alive continue from their stored state and walk the remaining
//                      e - p steps, exactly as search1 would below that node (number stays
//                      T^s(n mod 2^s) + 1, so the main.rs `number <= base` test sees the
//                      same values).
//   3. compaction    : popcount per block, exclusive scan, then each thread writes the
//                      indices of its set bits at its offset (128-bit loads, wave64
//                      prefix sums).
//   4. fill kernel   : every survivor recomputes its number, f and n mod 9.
#include "internal.hpp"

namespace cz {

// The walk's number type: 64 bits for depths up to 40 (T^40(n) + 1 < 3^40 + 1 < 2^64, all
// of range mode), 128 bits beyond (reference mode up to N - A = 62).
using u128d = unsigned __int128;

template <typename NT>
struct Walk {
  uint64_t base;  // n mod 2^s
  NT number;      // T^s(n) + 1
  NT p3;          // 3^f
  uint32_t f, de, ne;
};

// Take bit `bit` at depth s and apply the prunes of the child (collatz_cudav7.cu:506-511,
// :542-561; main.rs:216 when rust).  Returns false when the child is pruned or not taken.
template <typename NT>
__device__ __forceinline__ bool walk_step(Walk<NT>& x, uint32_t s, uint32_t bit, uint32_t rust) {
  const NT nb = x.number + (bit ? x.p3 : (NT)0);
  if (bit) x.base |= 1ull << s;
  const uint32_t inplay = (kNumerator * (s + 1) >= kDenominator * x.f) ? 1u : 0u;
  if ((nb & 1) == 0) {
    x.number = nb + (nb >> 1);
    x.f += 1;
    x.p3 *= 3;
    x.de = (x.de | inplay) & ~2u;
    x.ne = 0;
  } else {
    if (x.de >= 2) return false;
    x.number = (nb >> 1) + 1;
    x.de <<= 1;
    x.ne += 1;
  }
  const uint32_t s1 = s + 1;
  if (rust && x.number <= (NT)x.base) return false;
  if (kDenominator * x.f <= kNumerator * s1) return false;
  if (x.ne >= 2 && (x.ne & 1) == 0 && kNumerator * (s1 - 1) >= kDenominator * (x.f - 1))
    return false;
  return true;
}

__device__ __forceinline__ uint32_t pack_meta(uint32_t f, uint32_t mod9, uint32_t de,
                                              uint32_t ne, bool alive) {
  return f | (mod9 << 8) | (de << 12) | ((ne & 0xFFu) << 16) | (alive ? 0x80000000u : 0u);
}

template <typename NT>
__device__ __forceinline__ NT pow3_of(uint32_t f) {
  NT x = 1;
  for (uint32_t i = 0; i < f; ++i) x *= 3;
  return x;
}

// A walk's number as stored per prefix or leaf: the low word, and the high word in a second
// array for the 128-bit walk (null for the 64-bit one).
template <typename NT>
__device__ __forceinline__ NT load_num(const uint64_t* lo, const uint64_t* hi, uint64_t i) {
  if constexpr (sizeof(NT) > 8) return (NT)lo[i] | ((NT)hi[i] << 64);
  else return lo[i];
}

template <typename NT>
__device__ __forceinline__ void store_num(uint64_t* lo, uint64_t* hi, uint64_t i, NT v) {
  lo[i] = (uint64_t)v;
  if constexpr (sizeof(NT) > 8) hi[i] = (uint64_t)(v >> 64);
}

template <typename NT>
__device__ __forceinline__ NT root_num(const SieveRoot& root) {
  if constexpr (sizeof(NT) > 8) return (NT)root.number | ((NT)root.number_hi << 64);
  else return root.number;
}

constexpr int kWalkPow3 = 64;  // 3^f for f < 64 (f <= depth <= 62)

template <typename NT>
__global__ __launch_bounds__(256) void k_sieve_prefix(SieveRoot root, uint32_t p,
                                                      uint64_t* pnum, uint64_t* pnum_hi,
                                                      uint32_t* pmeta, uint64_t* alive_bits) {
  const uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t total = 1ull << p;
  bool alive = false;
  Walk<NT> x;
  x.base = root.base;
  x.number = root_num<NT>(root);
  x.p3 = pow3_of<NT>(root.threepower);
  x.f = root.threepower;
  x.de = root.doubleeven;
  x.ne = root.numevens;
  if (h < total) {
    alive = true;
    for (uint32_t j = 0; j < p && alive; ++j)
      alive = walk_step(x, root.sigfig + j, (uint32_t)(h >> j) & 1u, root.rust);
    store_num<NT>(pnum, pnum_hi, h, alive ? x.number : (NT)0);
    pmeta[h] = alive ? pack_meta(x.f, 0, x.de, x.ne, true) : 0u;
  }
  const uint64_t bal = __ballot(alive);
  if ((threadIdx.x & 63) == 0 && h < total) {
    if (total >= 64) alive_bits[h >> 6] = bal;
    else alive_bits[0] = bal & ((1ull << total) - 1);
  }
}

template <typename NT>
__global__ __launch_bounds__(256) void k_sieve_expand(SieveRoot root, uint32_t p,
                                                      const uint64_t* pnum,
                                                      const uint64_t* pnum_hi,
                                                      const uint32_t* pmeta,
                                                      const uint32_t* alive32, uint32_t* bits,
                                                      uint64_t nwords) {
  __shared__ NT p3[kWalkPow3];
  if (threadIdx.x < kWalkPow3) p3[threadIdx.x] = pow3_of<NT>(threadIdx.x);
  __syncthreads();
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwords) return;
  const uint64_t h0 = w << 5;
  const uint64_t hl0 = h0 & ((1ull << p) - 1);
  const uint64_t hh = h0 >> p;
  const uint32_t P = root.sigfig + p;
  const uint32_t rest = root.depth - P;
  uint32_t mask = alive32[hl0 >> 5];
  uint32_t out = 0;
  while (mask) {
    const uint32_t j = __builtin_ctz(mask);
    mask &= mask - 1;
    const uint32_t m = pmeta[hl0 + j];
    Walk<NT> x;
    x.f = meta_f(m);
    x.de = meta_de(m);
    x.ne = (m >> 16) & 0xFFu;
    x.p3 = p3[x.f];
    x.number = load_num<NT>(pnum, pnum_hi, hl0 + j);
    x.base = root.base + ((hl0 + j) << root.sigfig);
    bool alive = true;
    for (uint32_t t = 0; t < rest && alive; ++t)
      alive = walk_step(x, P + t, (uint32_t)(hh >> t) & 1u, root.rust);
    if (alive) out |= 1u << j;
  }
  bits[w] = out;
}

// ---- compaction ------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// Exclusive block scan for 256 threads; returns the thread's exclusive prefix and writes
// the block total.
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t* lds4,
                                                       uint32_t& total) {
  const uint32_t inc = wave_incl_scan(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 63) lds4[wv] = inc;
  __syncthreads();
  uint32_t before = 0;
  for (int i = 0; i < wv; ++i) before += lds4[i];
  total = lds4[0] + lds4[1] + lds4[2] + lds4[3];
  return before + inc - v;
}

__global__ __launch_bounds__(256) void k_compact_count(const uint4* bits, uint64_t nvec,
                                                       uint32_t* block_counts) {
  __shared__ uint32_t lds4[4];
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  uint32_t c = 0;
  if (i < nvec) {
    const uint4 v = bits[i];
    c = __popc(v.x) + __popc(v.y) + __popc(v.z) + __popc(v.w);
  }
  uint32_t total;
  block_excl_scan256(c, lds4, total);
  if (threadIdx.x == 0) block_counts[blockIdx.x] = total;
}

// Single-workgroup exclusive scan of n u32 values into u64 (out has n + 1 entries).
__global__ __launch_bounds__(1024) void k_scan(const uint32_t* in, uint64_t n, uint64_t* out) {
  __shared__ uint64_t wsum[16];
  __shared__ uint64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (uint64_t base = 0; base < n; base += 1024) {
    const uint64_t i = base + threadIdx.x;
    uint64_t v = i < n ? in[i] : 0;
    uint64_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t o = ((uint64_t)__shfl_up((uint32_t)(inc >> 32), d, 64) << 32) |
                         __shfl_up((uint32_t)inc, d, 64);
      if (lane >= d) inc += o;
    }
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    uint64_t before = carry;
    for (int j = 0; j < wv; ++j) before += wsum[j];
    if (i < n) out[i] = before + inc - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry = before + inc;
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = carry;
}

__global__ __launch_bounds__(256) void k_compact_write(const uint4* bits, uint64_t nvec,
                                                       const uint64_t* block_off,
                                                       uint64_t* out_h) {
  __shared__ uint32_t lds4[4];
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  uint4 v = make_uint4(0, 0, 0, 0);
  if (i < nvec) v = bits[i];
  const uint32_t c = __popc(v.x) + __popc(v.y) + __popc(v.z) + __popc(v.w);
  uint32_t total;
  const uint32_t ex = block_excl_scan256(c, lds4, total);
  uint64_t o = block_off[blockIdx.x] + ex;
  const uint64_t hbase = i << 7;
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    uint32_t m = w[q];
    while (m) {
      const uint32_t b = __builtin_ctz(m);
      m &= m - 1;
      out_h[o++] = hbase + (uint64_t)q * 32 + b;
    }
  }
}

template <typename NT>
__global__ __launch_bounds__(256) void k_sieve_fill(SieveRoot root, uint32_t p,
                                                    const uint64_t* pnum,
                                                    const uint64_t* pnum_hi,
                                                    const uint32_t* pmeta, uint64_t* res,
                                                    uint64_t* num, uint64_t* num_hi,
                                                    uint32_t* meta, uint64_t count) {
  __shared__ NT p3[kWalkPow3];
  if (threadIdx.x < kWalkPow3) p3[threadIdx.x] = pow3_of<NT>(threadIdx.x);
  __syncthreads();
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const uint64_t h = res[t];
  const uint64_t hl = h & ((1ull << p) - 1), hh = h >> p;
  const uint32_t m = pmeta[hl];
  const uint32_t P = root.sigfig + p;
  Walk<NT> x;
  x.f = meta_f(m);
  x.de = meta_de(m);
  x.ne = (m >> 16) & 0xFFu;
  x.p3 = p3[x.f];
  x.number = load_num<NT>(pnum, pnum_hi, hl);
  x.base = root.base + (hl << root.sigfig);
  for (uint32_t s = P; s < root.depth; ++s) walk_step(x, s, (uint32_t)(hh >> (s - P)) & 1u, root.rust);
  const uint64_t r = root.base + (h << root.sigfig);
  res[t] = r;
  store_num<NT>(num, num_hi, t, x.number);
  meta[t] = pack_meta(x.f, (uint32_t)(r % 9), x.de, x.ne, false);
}


// ---- stable partition of the survivors by n mod 9 ---------------------------------------
// The range kernels take the survivors class by class: for a row with (q << k) mod 9 = c,
// only the residue classes r mod 9 with (c + r) mod 9 in {0,1,3,6,7} are candidates, so
// every lane of a row holds a real candidate (no lane is lost to the mod 9 sieve).
constexpr uint32_t kPartItems = 2048;  // elements per block (256 threads x 8)

__global__ __launch_bounds__(256) void k_class_count(const uint32_t* meta, uint64_t n,
                                                     uint32_t cls, uint32_t* block_counts) {
  __shared__ uint32_t lds4[4];
  const uint64_t base = (uint64_t)blockIdx.x * kPartItems + threadIdx.x * 8;
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint64_t i = base + j;
    if (i < n && meta_mod9(meta[i]) == cls) ++c;
  }
  uint32_t total;
  block_excl_scan256(c, lds4, total);
  if (threadIdx.x == 0) block_counts[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_class_write(const uint64_t* res, const uint64_t* num,
                                                     const uint32_t* meta, uint64_t n,
                                                     uint32_t cls, const uint64_t* offs,
                                                     uint64_t out_base, uint4* orec) {
  __shared__ uint32_t lds4[4];
  const uint64_t base = (uint64_t)blockIdx.x * kPartItems + threadIdx.x * 8;
  uint32_t m[8];
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint64_t i = base + j;
    m[j] = (i < n) ? meta[i] : 0xFFFFFFFFu;
    if (i < n && meta_mod9(m[j]) == cls) ++c;
  }
  uint32_t total;
  uint64_t o = out_base + offs[blockIdx.x] + block_excl_scan256(c, lds4, total);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint64_t i = base + j;
    if (i < n && meta_mod9(m[j]) == cls) {
      const uint64_t r = res[i], M = num[i];
      orec[o] = make_uint4((uint32_t)r, meta_f(m[j]) | ((uint32_t)(r >> 32) << 8), (uint32_t)M,
                           (uint32_t)(M >> 32));
      ++o;
    }
  }
}

// keys[t] = residue bound, kcls[t] = class: searches class kcls[t] of a class-major array
// whose classes start at class_start[0..9].
__global__ __launch_bounds__(256) void k_lower_bound(const uint4* a,
                                                     const uint64_t* class_start,
                                                     const uint64_t* keys, const uint32_t* kcls,
                                                     uint64_t nkeys, uint64_t* out) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nkeys) return;
  const uint64_t key = keys[t];
  const uint32_t c = kcls[t];
  uint64_t lo = class_start[c], hi = class_start[c + 1];
  while (lo < hi) {
    const uint64_t mid = lo + ((hi - lo) >> 1);
    if (rec_r(a[mid]) < key) lo = mid + 1;
    else hi = mid;
  }
  out[t] = lo;
}

static inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

// The walk runs on 64-bit numbers up to depth 40 and on 128-bit ones beyond (the _hi
// arrays are used, and must be given, only then).
static bool wide_walk(const SieveRoot& root) { return root.depth > kWalk64MaxDepth; }

hipError_t launch_sieve_prefix(const SieveRoot& root, uint32_t p, uint64_t* pnum,
                               uint64_t* pnum_hi, uint32_t* pmeta, uint64_t* alive_bits,
                               hipStream_t s) {
  const uint64_t total = 1ull << p;
  if (wide_walk(root))
    hipLaunchKernelGGL(k_sieve_prefix<u128d>, dim3(blocks_for(total, 256)), dim3(256), 0, s,
                       root, p, pnum, pnum_hi, pmeta, alive_bits);
  else
    hipLaunchKernelGGL(k_sieve_prefix<uint64_t>, dim3(blocks_for(total, 256)), dim3(256), 0, s,
                       root, p, pnum, nullptr, pmeta, alive_bits);
  return hipGetLastError();
}

hipError_t launch_sieve_expand(const SieveRoot& root, uint32_t p, const uint64_t* pnum,
                               const uint64_t* pnum_hi, const uint32_t* pmeta,
                               const uint64_t* alive_bits, uint32_t* bits, hipStream_t s) {
  const uint32_t e = root.depth - root.sigfig;
  const uint64_t nwords = 1ull << (e - 5);
  if (wide_walk(root))
    hipLaunchKernelGGL(k_sieve_expand<u128d>, dim3(blocks_for(nwords, 256)), dim3(256), 0, s,
                       root, p, pnum, pnum_hi, pmeta, (const uint32_t*)alive_bits, bits, nwords);
  else
    hipLaunchKernelGGL(k_sieve_expand<uint64_t>, dim3(blocks_for(nwords, 256)), dim3(256), 0,
                       s, root, p, pnum, nullptr, pmeta, (const uint32_t*)alive_bits, bits,
                       nwords);
  return hipGetLastError();
}

hipError_t launch_compact_count(const uint32_t* bits, uint64_t nwords, uint32_t* block_counts,
                                uint64_t nblocks, hipStream_t s) {
  const uint64_t nvec = nwords / 4;
  hipLaunchKernelGGL(k_compact_count, dim3((uint32_t)nblocks), dim3(256), 0, s,
                     (const uint4*)bits, nvec, block_counts);
  return hipGetLastError();
}

hipError_t launch_scan(const uint32_t* in, uint64_t n, uint64_t* out_excl, hipStream_t s) {
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, in, n, out_excl);
  return hipGetLastError();
}

hipError_t launch_compact_write(const uint32_t* bits, uint64_t nwords,
                                const uint64_t* block_offsets, uint64_t* out_h, hipStream_t s) {
  const uint64_t nvec = nwords / 4;
  hipLaunchKernelGGL(k_compact_write, dim3(blocks_for(nvec, 256)), dim3(256), 0, s,
                     (const uint4*)bits, nvec, block_offsets, out_h);
  return hipGetLastError();
}

hipError_t launch_sieve_fill(const SieveRoot& root, uint32_t p, const uint64_t* pnum,
                             const uint64_t* pnum_hi, const uint32_t* pmeta,
                             uint64_t* res_inout, uint64_t* num, uint64_t* num_hi,
                             uint32_t* meta, uint64_t count, hipStream_t s) {
  if (count == 0) return hipSuccess;
  if (wide_walk(root))
    hipLaunchKernelGGL(k_sieve_fill<u128d>, dim3(blocks_for(count, 256)), dim3(256), 0, s, root,
                       p, pnum, pnum_hi, pmeta, res_inout, num, num_hi, meta, count);
  else
    hipLaunchKernelGGL(k_sieve_fill<uint64_t>, dim3(blocks_for(count, 256)), dim3(256), 0, s,
                       root, p, pnum, nullptr, pmeta, res_inout, num, nullptr, meta, count);
  return hipGetLastError();
}

hipError_t launch_lower_bound(const uint4* sorted, const uint64_t* class_start,
                              const uint64_t* keys, const uint32_t* kcls, uint64_t nkeys,
                              uint64_t* out, hipStream_t s) {
  if (nkeys == 0) return hipSuccess;
  hipLaunchKernelGGL(k_lower_bound, dim3(blocks_for(nkeys, 256)), dim3(256), 0, s, sorted,
                     class_start, keys, kcls, nkeys, out);
  return hipGetLastError();
}


__device__ __forceinline__ uint64_t lower_bound_dev(const uint4* a, uint64_t lo, uint64_t hi,
                                                    uint64_t key) {
  while (lo < hi) {
    const uint64_t mid = lo + ((hi - lo) >> 1);
    if (rec_r(a[mid]) < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_rows_resolve_ctl(RowDesc* rows, const Ctl* ctl,
                                                          const uint4* a,
                                                          const uint64_t* class_start,
                                                          const uint64_t* keys) {
  const uint64_t n = ctl->nrows;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (uint64_t)gridDim.x * blockDim.x) {
    RowDesc& r = rows[t];
    const uint64_t lo = class_start[r.cls], hi = class_start[r.cls + 1];
    uint64_t i_lo = r.i_lo, i_hi = r.i_hi;
    if (i_lo & kResolveFlag) i_lo = lower_bound_dev(a, lo, hi, keys[i_lo & ~kResolveFlag]);
    if (i_hi & kResolveFlag) i_hi = lower_bound_dev(a, lo, hi, keys[i_hi & ~kResolveFlag]);
    if (i_hi < i_lo) i_hi = i_lo;
    r.i_lo = i_lo;
    r.i_hi = i_hi;
  }
}

hipError_t launch_rows_resolve_ctl(RowDesc* rows, const Ctl* ctl, const uint4* sorted,
                                   const uint64_t* class_start, const uint64_t* keys,
                                   hipStream_t s) {
  hipLaunchKernelGGL(k_rows_resolve_ctl, dim3(512), dim3(256), 0, s, rows, ctl, sorted,
                     class_start, keys);
  return hipGetLastError();
}

hipError_t launch_class_partition(const uint64_t* res, const uint64_t* num, const uint32_t* meta,
                                  uint64_t n, uint4* orec, uint32_t* counts, uint64_t* offs, uint64_t* class_start_host,
                                  hipStream_t s) {
  const uint64_t nb = (n + kPartItems - 1) / kPartItems;
  uint64_t base = 0;
  for (uint32_t c = 0; c < 9; ++c) {
    class_start_host[c] = base;
    if (nb == 0) continue;
    hipLaunchKernelGGL(k_class_count, dim3((uint32_t)nb), dim3(256), 0, s, meta, n, c, counts);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, counts, nb, offs);
    hipLaunchKernelGGL(k_class_write, dim3((uint32_t)nb), dim3(256), 0, s, res, num, meta, n, c,
                       offs, base, orec);
    uint64_t cnt = 0;
    hipError_t e = hipMemcpyAsync(&cnt, offs + nb, 8, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    base += cnt;
  }
  class_start_host[9] = base;
  return hipGetLastError();
}

}  // namespace cz
