// C ABI: collatz_verify_batch, many intervals in one persistent launch.
//
// The intervals are decomposed into rows exactly as collatz_verify_range does (so each
// interval's chunk records are the same), the rows are cut into tiles, and one launch of
// k_verify_persistent works through all tiles with a shared atomic tile counter: waves
// that finish early steal the remaining tiles of every interval.  The launch sequence
// (counter reset, row resolution, persistent kernel, per-chunk reduction) can be captured
// in a hipGraph and is replayed for later batches.
#include <algorithm>
#include <cstring>
#include <optional>
#include <set>
#include <vector>

#include "engine.hpp"
#include "host_common.hpp"
#include "tuning.hpp"

namespace cz {

struct BatchArgs {
  uint32_t C = 24, limbs = 2, cutoff = 187, graph = 0;
  uint32_t tile = 0, waves = 0, U = 0;
  uint32_t width = 0;  // word width of the kernels; 0 = launch_width
};

static uint64_t estimate_candidates(const collatz_sieve* sv, u128 width) {
  const double density = (double)sv->count / (double)((u128)1 << sv->k) * (sv->mod9 ? 5.0 / 9 : 1.0);
  return (uint64_t)((double)width * density);
}

int verify_batch_impl(const collatz_sieve* sv, uint64_t n, const collatz_interval* iv,
                      const BatchArgs& a, hipStream_t s, collatz_range_result* outs,
                      bool allow_retry) {
  u128 max_hi = 0;
  uint64_t est = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const u128 lo = mk128(iv[i].lo[0], iv[i].lo[1]), hi = mk128(iv[i].hi[0], iv[i].hi[1]);
    if (lo >= hi) return fail(COLLATZ_EINVAL, "empty interval (lo >= hi)");
    if (hi > max_hi) max_hi = hi;
    est += estimate_candidates(sv, hi - lo);
  }
  const int width = a.width ? (int)a.width : launch_width((int)a.limbs, max_hi, a.cutoff);
  const Tuning tn = choose_tuning(sv, width, max_hi, est, a.tile, a.waves, a.U);
  Span sp_all("batch total");
  std::optional<Span> sp_plan(std::in_place, "batch plan");
  LaunchPlan lp;
  std::vector<uint64_t> nchk, first_chunk(n), slot_base(n + 1, 0);
  for (uint64_t i = 0; i < n; ++i) {
    const u128 lo = mk128(iv[i].lo[0], iv[i].lo[1]), hi = mk128(iv[i].hi[0], iv[i].hi[1]);
    uint64_t nch = 0;
    int rc = plan_interval(sv, lo, hi, a.C, slot_base[i], tn.tile, lp, nchk, first_chunk[i], nch);
    if (rc) return rc;
    slot_base[i + 1] = slot_base[i] + nch;
  }
  lp.part.push_back(lp.ntiles);
  if (lp.part.size() != lp.nslots + 1) return fail(COLLATZ_EDEVICE, "internal: slot without rows");
  // The launch's overflow list holds the starts of every interval that asks for them: room
  // for all their buffers together (up to 2^24 reports), and at least 1024.
  uint64_t ovf_cap = 0;
  for (uint64_t i = 0; i < n; ++i)
    if (outs[i].overflow_starts) ovf_cap = std::min<uint64_t>(ovf_cap + outs[i].overflow_cap, 1ull << 24);
  ovf_cap = std::max<uint64_t>(ovf_cap, 1024);

  sp_plan.reset();
  LaunchResult res;
  int rc;
  {
    Span sp("engine run");
    std::lock_guard<std::mutex> lk(sv->engine_mu);
    if (!sv->engine) sv->engine = new (std::nothrow) Engine(sv);
    if (!sv->engine) return fail(COLLATZ_ENOMEM, "out of host memory");
    rc = sv->engine->run(lp, tn, width, a.cutoff, a.graph != 0, ovf_cap, s, res);
  }
  if (rc == COLLATZ_ECAPACITY && width < 3 && allow_retry) {
    // More escalations than the list holds: the same batch one width up.
    BatchArgs b = a;
    b.width = (uint32_t)width + 1;
    return verify_batch_impl(sv, n, iv, b, s, outs, true);
  }
  if (rc) return rc;
  if (res.stalled) {
    // A trajectory that does not descend is a finding, not a result: the call fails and
    // names the first interval and chunk where it happened.
    const uint64_t sl = res.stall_slots.empty() ? 0 : res.stall_slots[0];
    const uint64_t i =
        std::upper_bound(slot_base.begin(), slot_base.end(), sl) - slot_base.begin() - 1;
    return fail(COLLATZ_EDEVICE,
                "a trajectory did not descend within the check bound (interval " +
                    std::to_string(i) + ", chunk " +
                    std::to_string(first_chunk[std::min<uint64_t>(i, n - 1)] + sl -
                                   slot_base[std::min<uint64_t>(i, n - 1)]) + ")");
  }

  bool short_buf = false;
  for (uint64_t i = 0; i < n; ++i) {
    collatz_range_result* out = &outs[i];
    const uint64_t nc = slot_base[i + 1] - slot_base[i];
    std::memset(&out->total, 0, sizeof(out->total));
    out->checks = 0;
    for (uint64_t c = 0; c < nc; ++c) {
      const uint64_t sl = slot_base[i] + c;
      collatz_chunk ch;
      chunk_from_rec(res.recs[sl], first_chunk[i] + c, nchk[sl], &ch);
      if (out->chunks && c < out->chunks_cap) out->chunks[c] = ch;
      chunk_merge(&out->total, ch);
      out->checks += res.recs[sl].checks;
    }
    out->total.chunk_index = first_chunk[i];
    out->chunks_written = nc;
    out->escalated = 0;
    out->spot_checks = 0;
    out->device_ms = res.device_ms;
    // The exact number of overflow reports is the interval's overflow count (every report
    // is one overflow event of a candidate that finished); the device list, which holds at
    // most ovf_cap reports of the whole launch, only supplies the starts.
    out->overflow_written = out->total.overflow_count;
    uint64_t listed = 0;
    for (size_t j = 0; j < res.ovf.size() / 4; ++j) {
      const uint64_t sl = res.ovf[4 * j + 2];
      if (sl < slot_base[i] || sl >= slot_base[i + 1]) continue;
      if (out->overflow_starts && listed < out->overflow_cap) {
        out->overflow_starts[2 * listed] = res.ovf[4 * j];
        out->overflow_starts[2 * listed + 1] = res.ovf[4 * j + 1];
      }
      listed++;
    }
    // The list of the launch was cut short, so this interval's starts may be incomplete.
    if (out->overflow_starts && listed < std::min(out->overflow_written, out->overflow_cap))
      short_buf = true;
    if ((out->chunks && nc > out->chunks_cap) ||
        (out->overflow_starts && out->overflow_written > out->overflow_cap))
      short_buf = true;
  }
  for (uint32_t sl : res.esc_slots) {
    const uint64_t i =
        std::upper_bound(slot_base.begin(), slot_base.end(), (uint64_t)sl) - slot_base.begin() - 1;
    if (i < n) outs[i].escalated++;
  }
  if (short_buf) return fail(COLLATZ_ECAPACITY, "output buffer smaller than the result");
  return COLLATZ_OK;
}

// One timed persistent launch over [lo, hi) with a fixed launch shape (for the tuner).
int verify_batch_tuned(const collatz_sieve* sv, u128 lo, u128 hi, int limbs, Tuning t,
                       hipStream_t s, double* ms) {
  collatz_interval iv;
  iv.lo[0] = lo64(lo);
  iv.lo[1] = hi64(lo);
  iv.hi[0] = lo64(hi);
  iv.hi[1] = hi64(hi);
  BatchArgs a;
  a.C = 32;
  a.limbs = limbs == 3 ? 3u : 2u;
  a.width = (uint32_t)limbs;
  a.tile = t.tile;
  a.waves = t.waves;
  a.U = t.U;
  collatz_range_result r;
  std::memset(&r, 0, sizeof(r));
  const int rc = verify_batch_impl(sv, 1, &iv, a, s, &r, false);
  *ms = r.device_ms;
  return rc == COLLATZ_ECAPACITY ? COLLATZ_OK : rc;
}

// Spot checks of a batch (the reference re-runs a random launch's candidates on the CPU and
// compares the hash, collatz_cudav7.cu:619-635): `count` distinct chunk records among those
// the caller stored (all of them if fewer), chosen by a SplitMix64 stream seeded from the
// intervals, are re-run by the plain kernel over that chunk's part of its interval and
// compared field by field.  outs[i].spot_checks counts interval i's records found equal; a
// difference fails the call and names the interval and chunk.
static int spot_check_batch(const collatz_sieve* sv, uint64_t n, const collatz_interval* iv,
                            const BatchArgs& a, uint32_t count, hipStream_t s,
                            collatz_range_result* outs) {
  std::vector<uint64_t> pre(n + 1, 0);
  for (uint64_t i = 0; i < n; ++i)
    pre[i + 1] = pre[i] + (outs[i].chunks ? std::min(outs[i].chunks_written, outs[i].chunks_cap)
                                          : 0);
  const uint64_t stored = pre[n];
  std::vector<uint64_t> picks;
  if (stored <= count) {
    for (uint64_t p = 0; p < stored; ++p) picks.push_back(p);
  } else {
    uint64_t z = 0xCBF29CE484222325ull ^ n;
    for (uint64_t i = 0; i < n; ++i)
      for (int w = 0; w < 2; ++w) z = (z ^ iv[i].lo[w] ^ (iv[i].hi[w] << 1)) * 0x100000001B3ull;
    std::set<uint64_t> seen;
    while (picks.size() < count) {
      z += 0x9E3779B97F4A7C15ull;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
      x ^= x >> 31;
      const uint64_t p = x % stored;
      if (seen.insert(p).second) picks.push_back(p);
    }
  }
  for (const uint64_t p : picks) {
    const uint64_t i = std::upper_bound(pre.begin(), pre.end(), p) - pre.begin() - 1;
    const collatz_chunk& want = outs[i].chunks[p - pre[i]];
    const u128 lo = mk128(iv[i].lo[0], iv[i].lo[1]), hi = mk128(iv[i].hi[0], iv[i].hi[1]);
    // (chunk_index holds the low 64 bits of the chunk number: recompute it in full.)
    const u128 c_lo = ((lo >> a.C) + (p - pre[i])) << a.C;
    const u128 sub_lo = std::max(lo, c_lo);
    const u128 sub_hi = hi - c_lo > ((u128)1 << a.C) ? c_lo + ((u128)1 << a.C) : hi;
    const uint64_t l2[2] = {lo64(sub_lo), hi64(sub_lo)}, h2[2] = {lo64(sub_hi), hi64(sub_hi)};
    collatz_verify_opts o;
    o.chunk_log2 = a.C;
    o.limbs = a.limbs;
    o.cutoff_log2 = a.cutoff;
    o.form = COLLATZ_FORM_PLAIN;
    collatz_chunk got;
    collatz_range_result r;
    std::memset(&r, 0, sizeof(r));
    r.chunks = &got;
    r.chunks_cap = 1;
    const int rc = collatz_verify_range(sv, l2, h2, &o, s, &r);
    if (rc != COLLATZ_OK && rc != COLLATZ_ECAPACITY) return rc;
    if (r.chunks_written != 1 || std::memcmp(&got, &want, sizeof(got)) != 0)
      return fail(COLLATZ_EDEVICE, "spot check failed: interval " + std::to_string(i) +
                                       ", chunk record " + std::to_string(p - pre[i]) +
                                       ": the persistent launch and the plain kernel differ");
    outs[i].spot_checks++;
  }
  return COLLATZ_OK;
}

}  // namespace cz

using namespace cz;

extern "C" int collatz_verify_batch(const collatz_sieve* sv, uint64_t n,
                                    const collatz_interval* intervals,
                                    const collatz_batch_opts* opts, void* stream,
                                    collatz_range_result* outs) {
  if (!sv || !intervals || !outs) return fail(COLLATZ_EINVAL, "null argument");
  if (n == 0) return COLLATZ_OK;
  BatchArgs a;
  if (opts) {
    if (opts->chunk_log2) a.C = opts->chunk_log2;
    if (opts->limbs) a.limbs = opts->limbs;
    if (opts->cutoff_log2) a.cutoff = opts->cutoff_log2;
    a.graph = opts->use_graph;
    a.tile = opts->tile_residues;
    a.waves = opts->waves_per_cu;
    a.U = opts->checks_per_refill;
  }
  if (a.C > 63) return fail(COLLATZ_EINVAL, "chunk_log2 must be <= 63");
  if (a.limbs != 2 && a.limbs != 3) return fail(COLLATZ_EINVAL, "limbs must be 2 or 3");
  if (a.cutoff < 64 || a.cutoff > 187) return fail(COLLATZ_EINVAL, "cutoff_log2 must be in [64, 187]");
  if (a.waves > 32) return fail(COLLATZ_EINVAL, "waves_per_cu must be <= 32");
  if (a.U > 64) return fail(COLLATZ_EINVAL, "checks_per_refill must be <= 64");
  if (a.tile > (1u << 24)) return fail(COLLATZ_EINVAL, "tile_residues must be <= 2^24");
  HipDeviceGuard guard(sv->device);
  const hipStream_t s = call_stream(stream);
  const int rc = verify_batch_impl(sv, n, intervals, a, s, outs, true);
  if ((rc != COLLATZ_OK && rc != COLLATZ_ECAPACITY) || !opts || !opts->spot_checks) return rc;
  const int src = spot_check_batch(sv, n, intervals, a, opts->spot_checks, s, outs);
  return src ? src : rc;
}
