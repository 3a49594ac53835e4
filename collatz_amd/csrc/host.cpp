// C ABI: sieve handles and single-range verification (plain form).
//
// Replaces the reference's host driver around the kernel: main()'s table setup and
// search1 (collatz_cudav7.cu:95-183, :497-563) and launchkernel (:565-649).  Unlike the
// reference, which allocates, copies, launches and synchronises per batch of 2^16 leaves,
// the survivor list is built and kept on the device, and one call launches the whole range.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "engine.hpp"
#include "host_common.hpp"

namespace cz {

static thread_local std::string g_error;

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

bool trace_on() {
  static const int on = [] {
    const char* e = std::getenv("COLLATZ_TRACE");
    return (e && e[0] == '1') ? 1 : 0;
  }();
  return on != 0;
}

Span::Span(const char* n) : name(n), t0(trace_on() ? now_ms() : 0) {}
Span::~Span() {
  if (trace_on()) std::fprintf(stderr, "[collatz] %-22s %8.3f ms\n", name, now_ms() - t0);
}

void set_error(const std::string& msg) { g_error = msg; }

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

void chunk_from_rec(const DevRec& r, uint64_t chunk_index, uint64_t n_checked,
                    collatz_chunk* out) {
  std::memset(out, 0, sizeof(*out));
  out->chunk_index = chunk_index;
  out->n_checked = n_checked;
  out->count = r.count;
  out->stop_sum = r.stop_sum;
  out->stop_xor = r.stop_xor;
  out->overflow_count = r.overflow;
  out->hash = r.hash;
  if (r.count) {
    // T-value = number - 1.
    uint64_t m[3] = {r.maxnum[0], r.maxnum[1], r.maxnum[2]};
    for (int i = 0; i < 3; ++i) {
      if (m[i]-- != 0) break;
    }
    std::memcpy(out->max_exc, m, sizeof(m));
    out->max_start[0] = r.max_start[0];
    out->max_start[1] = r.max_start[1];
    const bool wide = r.maxnum[2] != 0 || (r.maxnum[1] >> (kWideLog2 - 64)) != 0;
    if (wide) out->flags |= COLLATZ_FLAG_WIDE;
  }
  if (r.overflow) out->flags |= COLLATZ_FLAG_OVERFLOW;
}

void chunk_merge(collatz_chunk* a, const collatz_chunk& b) {
  if (b.count) {
    const int c = cmp3h(b.max_exc, a->max_exc);
    const bool smaller =
        mk128(b.max_start[0], b.max_start[1]) < mk128(a->max_start[0], a->max_start[1]);
    if (a->count == 0 || c > 0 || (c == 0 && smaller)) {
      std::memcpy(a->max_exc, b.max_exc, sizeof(a->max_exc));
      std::memcpy(a->max_start, b.max_start, sizeof(a->max_start));
    }
  }
  a->n_checked += b.n_checked;
  a->count += b.count;
  a->stop_sum += b.stop_sum;
  a->stop_xor ^= b.stop_xor;
  a->overflow_count += b.overflow_count;
  a->hash ^= b.hash;
  a->flags |= b.flags;
}

int launch_width(int limbs, u128 max_hi, uint32_t cutoff) {
  if (limbs == 2 && cutoff >= kNarrowLog2 && max_hi <= ((u128)1 << kNarrowMaxStartLog2)) return 1;
  // Above 2^100 a start value (up to n (3/2)^k) can pass 2^118 (the four-word per-check
  // bound) or 2^123 (the four-word start bound under a lower cutoff): six words at once.
  if (limbs == 2 && max_hi > ((u128)1 << kWideMaxStartLog2)) return 3;
  return limbs;
}

Limits make_limits(int limbs, uint32_t cutoff) {
  Limits l;
  l.cutoff = cutoff;
  l.max_checks = kMaxChecks;
  l.max_steps = kMaxSteps;
  l.max_loop = kMaxLoop;
  if (const char* e = std::getenv("COLLATZ_LOOP_LIMIT")) {
    const unsigned long v = std::strtoul(e, nullptr, 10);
    if (v >= 1 && v < kMaxLoop) l.max_loop = (uint32_t)v;
  }
  if (const char* e = std::getenv("COLLATZ_CHECK_LIMIT")) {
    const unsigned long v = std::strtoul(e, nullptr, 10);
    if (v >= 1 && v < kMaxChecks) {
      l.max_checks = (uint32_t)v;
      l.max_steps = (uint32_t)(2 * v);
    }
  }
  if (limbs == 1) {
    l.fast_word = 2;
    l.fast_thr = 1ull << (kNarrowLog2 - 64);
  } else if (limbs == 2) {
    // (A cutoff below 2^64 happens only under main.rs semantics, whose test does not use
    // the fast threshold.)
    const uint32_t c = std::max<uint32_t>(std::min<uint32_t>(cutoff, kWideLog2), 64);
    l.fast_word = 2;
    l.fast_thr = 1ull << (c - 64);
  } else if (cutoff >= 128) {
    l.fast_word = 4;
    l.fast_thr = 1ull << (cutoff - 128);
  } else {
    l.fast_word = 2;
    l.fast_thr = 1ull << (std::max<uint32_t>(cutoff, 64) - 64);
  }
  return l;
}

int plan_rows(uint32_t k, u128 lo, u128 hi, uint32_t C, uint32_t slot_base, bool mod9,
              RowPlan& plan, uint64_t max_rows) {
  const u128 first = lo >> C, last = (hi - 1) >> C;
  const u128 nch = last - first + 1;
  if (nch > (u128)0x7FFFFFFF) return fail(COLLATZ_EINVAL, "range has too many chunks");
  plan.first_chunk = (uint64_t)first;
  plan.nchunks = (uint64_t)nch;
  const size_t base_rows = plan.rows.size();
  plan.chunk_first_row.reserve(plan.chunk_first_row.size() + plan.nchunks + 1);
  const u128 kspan = (u128)1 << k;
  for (uint64_t cj = 0; cj < plan.nchunks; ++cj) {
    const u128 j = first + cj;
    const u128 a = j << C;
    u128 c_lo = a < lo ? lo : a;
    const u128 width = (u128)1 << C;
    u128 c_hi = (hi - a > width) ? a + width : hi;
    plan.n_checked.push_back((uint64_t)(c_hi - c_lo));
    plan.chunk_first_row.push_back(plan.rows.size() - base_rows);
    for (u128 q = c_lo >> k; q <= (c_hi - 1) >> k; ++q) {
      const u128 qb = q << k;
      const u128 r_lo = c_lo > qb ? c_lo - qb : 0;
      const u128 r_hi = (c_hi - qb) < kspan ? (c_hi - qb) : kspan;
      const uint32_t qb9 = (uint32_t)(qb % 9);
      // n mod 9 = (qb + r) mod 9: only classes that the mod 9 sieve keeps get a row
      // (collatz_cudav7.cu:482-494).
      for (uint32_t c = 0; c < 9; ++c) {
        if (mod9 && !((kMod9Keep >> ((qb9 + c) % 9)) & 1u)) continue;
        if (plan.rows.size() - base_rows >= max_rows)
          return fail(COLLATZ_EINVAL, "range decomposes into too many rows");
        RowDesc row;
        std::memset(&row, 0, sizeof(row));
        row.q_lo = lo64(q);
        row.q_hi = hi64(q);
        row.qb_lo = lo64(qb);
        row.qb_hi = hi64(qb);
        row.slot = slot_base + (uint32_t)cj;
        row.cls = c;
        if (r_lo == 0) {
          plan.key_lo.push_back(-1);
        } else {
          plan.key_lo.push_back((int64_t)plan.keys.size());
          plan.keys.push_back((uint64_t)r_lo);
          plan.key_cls.push_back(c);
        }
        if (r_hi == kspan) {
          plan.key_hi.push_back(-1);
        } else {
          plan.key_hi.push_back((int64_t)plan.keys.size());
          plan.keys.push_back((uint64_t)r_hi);
          plan.key_cls.push_back(c);
        }
        plan.rows.push_back(row);
      }
    }
  }
  plan.chunk_first_row.push_back(plan.rows.size() - base_rows);
  return COLLATZ_OK;
}

static bool root_pruned(const SieveRoot& r) {
  if (r.rust && mk128(r.number, r.number_hi) <= r.base) return true;
  if (kDenominator * r.threepower <= kNumerator * r.sigfig) return true;
  if (r.numevens >= 2 && r.numevens % 2 == 0 &&
      kNumerator * (r.sigfig - 1) >= kDenominator * (r.threepower - 1))
    return true;
  return false;
}

static uint64_t round_up(uint64_t x, uint64_t m) { return (x + m - 1) / m * m; }

int build_leaves(const SieveRoot& root, hipStream_t s, LeafSet& out) {
  if (root.depth < root.sigfig || root.depth - root.sigfig > 36 || root.depth > kWalkMaxDepth)
    return fail(COLLATZ_EINVAL, "sieve depth out of range");
  if (root.depth <= kWalk64MaxDepth && root.number_hi)
    return fail(COLLATZ_EINVAL, "root number too large for its depth");
  out.count = 0;
  out.wide = root.depth > kWalk64MaxDepth;
  const uint32_t e = root.depth - root.sigfig;
  if (root_pruned(root) || e < 7) {
    // Tiny trees (and pruned roots) are not worth a device pass; the bit array is empty
    // and the leaves, if any, are produced by walking on the host.
    std::vector<uint64_t> r, n, nh;
    std::vector<uint32_t> m;
    if (!root_pruned(root)) {
      for (uint64_t h = 0; h < (1ull << e); ++h) {
        uint64_t base = root.base;
        u128 number = mk128(root.number, root.number_hi), p3 = 1;
        for (uint32_t i = 0; i < root.threepower; ++i) p3 *= 3;
        uint32_t f = root.threepower, de = root.doubleeven, ne = root.numevens;
        bool alive = true;
        for (uint32_t j = 0; j < e && alive; ++j) {
          const uint32_t sd = root.sigfig + j;
          const uint32_t bit = (h >> j) & 1;
          const u128 nb = number + (bit ? p3 : 0);
          if (bit) base |= 1ull << sd;
          const uint32_t inplay = kNumerator * (sd + 1) >= kDenominator * f;
          if ((nb & 1) == 0) {
            number = nb + nb / 2; f += 1; p3 *= 3; de = (de | inplay) & ~2u; ne = 0;
          } else if (de >= 2) {
            alive = false;
          } else {
            number = nb / 2 + 1; de <<= 1; ne += 1;
          }
          if (!alive) break;
          const uint32_t s1 = sd + 1;
          if (root.rust && number <= base) alive = false;
          else if (kDenominator * f <= kNumerator * s1) alive = false;
          else if (ne >= 2 && ne % 2 == 0 && kNumerator * (s1 - 1) >= kDenominator * (f - 1))
            alive = false;
        }
        if (alive) {
          r.push_back(base);
          n.push_back(lo64(number));
          nh.push_back(hi64(number));
          m.push_back(f | ((uint32_t)(base % 9) << 8) | (de << 12) | ((ne & 0xFF) << 16));
        }
      }
    }
    out.count = r.size();
    CZ_HIP(out.res.reserve(std::max<size_t>(1, r.size())));
    CZ_HIP(out.num.reserve(std::max<size_t>(1, r.size())));
    CZ_HIP(out.num_hi.reserve(out.wide ? std::max<size_t>(1, r.size()) : 1));
    CZ_HIP(out.meta.reserve(std::max<size_t>(1, r.size())));
    if (!r.empty()) {
      CZ_HIP(hipMemcpyAsync(out.res.p, r.data(), r.size() * 8, hipMemcpyHostToDevice, s));
      CZ_HIP(hipMemcpyAsync(out.num.p, n.data(), n.size() * 8, hipMemcpyHostToDevice, s));
      if (out.wide)
        CZ_HIP(hipMemcpyAsync(out.num_hi.p, nh.data(), nh.size() * 8, hipMemcpyHostToDevice, s));
      CZ_HIP(hipMemcpyAsync(out.meta.p, m.data(), m.size() * 4, hipMemcpyHostToDevice, s));
      CZ_HIP(hipStreamSynchronize(s));
    }
    out.bits_bits = 0;
    return COLLATZ_OK;
  }
  const uint32_t p = std::min<uint32_t>(e, 20);
  const uint64_t np = 1ull << p;
  DevBuf<uint64_t>& pnum = out.pnum;
  DevBuf<uint64_t>& pnum_hi = out.pnum_hi;
  DevBuf<uint32_t>& pmeta = out.pmeta;
  DevBuf<uint64_t>& alive = out.alive;
  CZ_HIP(pnum.reserve(np));
  CZ_HIP(pnum_hi.reserve(out.wide ? np : 1));
  CZ_HIP(pmeta.reserve(np));
  const uint64_t alive_words64 = round_up(std::max<uint64_t>(1, np / 64), 512);
  CZ_HIP(alive.reserve(alive_words64));
  CZ_HIP(hipMemsetAsync(alive.p, 0, alive_words64 * 8, s));
  CZ_HIP(launch_sieve_prefix(root, p, pnum.p, pnum_hi.p, pmeta.p, alive.p, s));
  const uint32_t* bits32;
  uint64_t nwords32;
  if (e > p) {
    const uint64_t words = 1ull << (e - 5);
    nwords32 = round_up(words, 1024);
    CZ_HIP(out.bits.reserve(nwords32));
    if (nwords32 > words)
      CZ_HIP(hipMemsetAsync(out.bits.p + words, 0, (nwords32 - words) * 4, s));
    CZ_HIP(launch_sieve_expand(root, p, pnum.p, pnum_hi.p, pmeta.p, alive.p, out.bits.p, s));
    bits32 = out.bits.p;
  } else {
    nwords32 = alive_words64 * 2;
    CZ_HIP(out.bits.reserve(nwords32));
    CZ_HIP(hipMemcpyAsync(out.bits.p, alive.p, nwords32 * 4, hipMemcpyDeviceToDevice, s));
    bits32 = out.bits.p;
  }
  out.bits_bits = 1ull << e;
  const uint64_t nblocks = nwords32 / 1024;
  DevBuf<uint32_t>& counts = out.counts;
  DevBuf<uint64_t>& offs = out.offs;
  CZ_HIP(counts.reserve(nblocks));
  CZ_HIP(offs.reserve(nblocks + 1));
  CZ_HIP(launch_compact_count(bits32, nwords32, counts.p, nblocks, s));
  CZ_HIP(launch_scan(counts.p, nblocks, offs.p, s));
  uint64_t total = 0;
  CZ_HIP(hipMemcpyAsync(&total, offs.p + nblocks, 8, hipMemcpyDeviceToHost, s));
  CZ_HIP(hipStreamSynchronize(s));
  out.count = total;
  CZ_HIP(out.res.reserve(std::max<uint64_t>(1, total)));
  CZ_HIP(out.num.reserve(std::max<uint64_t>(1, total)));
  CZ_HIP(out.num_hi.reserve(out.wide ? std::max<uint64_t>(1, total) : 1));
  CZ_HIP(out.meta.reserve(std::max<uint64_t>(1, total)));
  CZ_HIP(launch_compact_write(bits32, nwords32, offs.p, out.res.p, s));
  CZ_HIP(launch_sieve_fill(root, p, pnum.p, pnum_hi.p, pmeta.p, out.res.p, out.num.p,
                           out.num_hi.p, out.meta.p, total, s));
  CZ_HIP(hipStreamSynchronize(s));
  return COLLATZ_OK;
}

}  // namespace cz

using namespace cz;

extern "C" {

int collatz_version(void) { return COLLATZ_ABI_VERSION; }

const char* collatz_last_error(void) { return g_error.c_str(); }

int collatz_sieve_build(const collatz_sieve_params* params, collatz_sieve** out) {
  if (!params || !out) return fail(COLLATZ_EINVAL, "null argument");
  *out = nullptr;
  if (params->k < 8 || params->k > 38) return fail(COLLATZ_EINVAL, "k must be in [8, 38]");
  if (params->lookahead_bits &&
      (params->lookahead_bits < 8 || params->lookahead_bits > 24 ||
       (params->lookahead_semantics != COLLATZ_SEM_CUDA &&
        params->lookahead_semantics != COLLATZ_SEM_OPENCL)))
    return fail(COLLATZ_EINVAL,
                "lookahead_bits must be 0 or in [8, 24], with CUDA or OpenCL semantics");
  const auto t0 = std::chrono::steady_clock::now();
  HipDeviceGuard guard(params->device);
  collatz_sieve* sv = new (std::nothrow) collatz_sieve();
  if (!sv) return fail(COLLATZ_ENOMEM, "out of host memory");
  if (hipGetDevice(&sv->device) != hipSuccess) {
    delete sv;
    return fail(COLLATZ_EDEVICE, "no HIP device");
  }
  sv->k = params->k;
  sv->mod9 = params->mod9 ? 1 : 0;
  // The canonical root "3 2 2 9 1 0" (collatz_cudav7.cu:6, :173).
  SieveRoot root;
  root.base = 3;
  root.number = 9;
  root.number_hi = 0;
  root.sigfig = 2;
  root.threepower = 2;
  root.doubleeven = 1;
  root.numevens = 0;
  root.depth = params->k;
  root.rust = 0;
  hipStream_t s = hipStreamPerThread;
  LeafSet leaves;
  int rc = build_leaves(root, s, leaves);
  if (rc != COLLATZ_OK) {
    delete sv;
    return rc;
  }
  sv->count = leaves.count;
  sv->bits = std::move(leaves.bits);
  sv->bits_bits = leaves.bits_bits;
  {
    // Class-major layout by r mod 9, each class sorted by r.
    const uint64_t n = leaves.count;
    const uint64_t nb = (n + 2047) / 2048;
    DevBuf<uint32_t> counts;
    DevBuf<uint64_t> offs;
    const bool ok =
        sv->rec.alloc(std::max<uint64_t>(1, n)) == hipSuccess &&
        counts.alloc(std::max<uint64_t>(1, nb)) == hipSuccess &&
        offs.alloc(nb + 1) == hipSuccess && sv->class_start_dev.alloc(10) == hipSuccess;
    if (!ok) {
      delete sv;
      return fail(COLLATZ_ENOMEM, "device allocation failed (sieve)");
    }
    hipError_t e = launch_class_partition(leaves.res.p, leaves.num.p, leaves.meta.p, n,
                                          sv->rec.p, counts.p, offs.p,
                                          sv->class_start, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(sv->class_start_dev.p, sv->class_start, sizeof(sv->class_start),
                         hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      delete sv;
      return fail(COLLATZ_EDEVICE, std::string("class partition: ") + hipGetErrorString(e));
    }
  }
  uint64_t p3[kPow3Max] = {0};
  uint64_t x = 1;
  for (int i = 0; i <= 40; ++i) {
    p3[i] = x;
    x *= 3;
  }
  if (sv->pow3.alloc(kPow3Max) != hipSuccess ||
      hipMemcpyAsync(sv->pow3.p, p3, sizeof(p3), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    delete sv;
    return fail(COLLATZ_EDEVICE, "pow3 table upload failed");
  }
  if (params->lookahead_bits) {
    sv->la_bits = params->lookahead_bits;
    sv->la_sem = params->lookahead_semantics;
    hipError_t e = sv->la.alloc(1ull << (sv->la_bits - 3));
    if (e == hipSuccess) e = launch_la_build(sv->la_bits, sv->k, sv->la_sem, sv->la.p, s);
    // A deeper filter under the CUDA compares also gets a 16-bit first level (the OpenCL
    // compares can lower a dip later in the walk, so there the levels would not nest).
    if (e == hipSuccess && sv->la_bits > kLaLevel1Bits && sv->la_sem == COLLATZ_SEM_CUDA) {
      e = sv->la1.alloc(1ull << (kLaLevel1Bits - 3));
      if (e == hipSuccess) e = launch_la_build(kLaLevel1Bits, sv->k, sv->la_sem, sv->la1.p, s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      delete sv;
      return fail(COLLATZ_EDEVICE, std::string("look-ahead table: ") + hipGetErrorString(e));
    }
  }
  sv->build_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  *out = sv;
  return COLLATZ_OK;
}

int collatz_sieve_query(const collatz_sieve* sv, collatz_sieve_info* info) {
  if (!sv || !info) return fail(COLLATZ_EINVAL, "null argument");
  info->k = sv->k;
  info->mod9 = sv->mod9;
  info->survivors = sv->count;
  info->bitarray_bits = sv->bits_bits;
  info->device_bytes = sv->rec.bytes() + sv->bits.bytes() + sv->pow3.bytes() +
                       sv->class_start_dev.bytes() + sv->la.bytes() + sv->la1.bytes();
  info->build_ms = sv->build_ms;
  info->lookahead_bits = sv->la_bits;
  info->lookahead_semantics = sv->la_sem;
  return COLLATZ_OK;
}

int collatz_sieve_copy(const collatz_sieve* sv, uint64_t* r, uint64_t* number, uint8_t* f,
                       uint64_t cap) {
  if (!sv) return fail(COLLATZ_EINVAL, "null sieve");
  HipDeviceGuard guard(sv->device);
  // Survivors are stored class-major; return them sorted by residue.
  const uint64_t all = sv->count;
  std::vector<uint4> hrec(all);
  if (all) {
    CZ_HIP(hipMemcpyAsync(hrec.data(), sv->rec.p, all * sizeof(uint4), hipMemcpyDeviceToHost,
                          hipStreamPerThread));
    CZ_HIP(hipStreamSynchronize(hipStreamPerThread));
  }
  std::vector<uint64_t> hr(all);
  for (uint64_t i = 0; i < all; ++i) hr[i] = rec_r(hrec[i]);
  std::vector<uint64_t> order(all);
  for (uint64_t i = 0; i < all; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return hr[a] < hr[b]; });
  const uint64_t n = std::min(cap, all);
  for (uint64_t i = 0; i < n; ++i) {
    if (r) r[i] = hr[order[i]];
    if (number) number[i] = rec_M(hrec[order[i]]);
    if (f) f[i] = (uint8_t)rec_f(hrec[order[i]]);
  }
  return COLLATZ_OK;
}

int collatz_sieve_free(collatz_sieve* sv) {
  if (!sv) return COLLATZ_OK;
  HipDeviceGuard guard(sv->device);
  delete sv->engine;
  sv->engine = nullptr;
  delete sv;
  return COLLATZ_OK;
}

}  // extern "C"
