// C ABI: collatz_verify_batch, many intervals in one persistent launch.
//
// The intervals are decomposed into rows exactly as collatz_verify_range does (so each
// interval's chunk records are the same), the rows are cut into tiles, and one launch of
// k_verify_persistent works through all tiles with a shared atomic tile counter: waves
// that finish early steal the remaining tiles of every interval.  The launch sequence
// (counter clear, persistent kernel, per-chunk reduction) can be captured in a hipGraph.
#include <algorithm>
#include <chrono>
#include <vector>

#include "host_common.hpp"
#include "tuning.hpp"

namespace cz {

int run_escalated(const collatz_sieve* sv, const DevBuf<EscEntry>& esc, uint64_t n,
                  uint32_t cutoff, DeviceLists lists, std::vector<DevRec>& recs,
                  uint32_t slot_base, hipStream_t s);

int device_cu_count(int device) {
  static int cached[64] = {0};
  if (device >= 0 && device < 64 && cached[device]) return cached[device];
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return 256;
  if (device >= 0 && device < 64) cached[device] = p.multiProcessorCount;
  return p.multiProcessorCount;
}

// Cuts rows into tiles of at most `tile` survivors; part_start[c] = first tile of chunk slot c.
void make_tiles(const RowPlan& plan, uint64_t nslots, uint64_t tile, std::vector<TileDesc>& tiles,
                std::vector<uint64_t>& part_start) {
  part_start.assign(nslots + 1, 0);
  std::vector<uint64_t> first(nslots, UINT64_MAX);
  for (uint32_t r = 0; r < plan.rows.size(); ++r) {
    const RowDesc& row = plan.rows[r];
    if (first[row.slot] == UINT64_MAX) first[row.slot] = tiles.size();
    uint64_t i = row.i_lo;
    do {
      TileDesc t;
      t.i_lo = i;
      t.i_hi = std::min(row.i_hi, i + tile);
      t.row = r;
      t.pad = 0;
      tiles.push_back(t);
      i = t.i_hi;
    } while (i < row.i_hi);
  }
  // Rows are generated slot by slot, so the tiles of one slot are contiguous.
  uint64_t next = tiles.size();
  for (uint64_t c = nslots; c-- > 0;) {
    if (first[c] == UINT64_MAX) first[c] = next;
    part_start[c] = first[c];
    next = first[c];
  }
  part_start[nslots] = tiles.size();
}

struct BatchArgs {
  uint32_t C = 24, limbs = 2, cutoff = 187, graph = 0;
  uint32_t tile = 0, waves = 0, U = 0;
};

int verify_batch_impl(const collatz_sieve* sv, uint64_t n, const collatz_interval* iv,
                      const BatchArgs& a, hipStream_t s, collatz_range_result* outs,
                      bool allow_retry) {
  RowPlan plan;
  std::vector<uint64_t> slot_base(n + 1, 0), first_chunk(n);
  std::vector<uint64_t> nchk;
  u128 max_hi = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const u128 lo = mk128(iv[i].lo[0], iv[i].lo[1]), hi = mk128(iv[i].hi[0], iv[i].hi[1]);
    if (lo >= hi) return fail(COLLATZ_EINVAL, "empty interval (lo >= hi)");
    if (hi > max_hi) max_hi = hi;
    RowPlan p;
    int rc = plan_rows(sv->k, lo, hi, a.C, (uint32_t)slot_base[i], sv->mod9 != 0, p, 1ull << 26);
    if (rc) return rc;
    if (slot_base[i] + p.nchunks > 0x7FFFFFFFull) return fail(COLLATZ_EINVAL, "too many chunks");
    slot_base[i + 1] = slot_base[i] + p.nchunks;
    first_chunk[i] = p.first_chunk;
    const size_t key_base = plan.keys.size();
    for (size_t r = 0; r < p.rows.size(); ++r) {
      plan.rows.push_back(p.rows[r]);
      plan.key_lo.push_back(p.key_lo[r] < 0 ? -1 : p.key_lo[r] + (int64_t)key_base);
      plan.key_hi.push_back(p.key_hi[r] < 0 ? -1 : p.key_hi[r] + (int64_t)key_base);
    }
    plan.keys.insert(plan.keys.end(), p.keys.begin(), p.keys.end());
    plan.key_cls.insert(plan.key_cls.end(), p.key_cls.begin(), p.key_cls.end());
    nchk.insert(nchk.end(), p.n_checked.begin(), p.n_checked.end());
    if (plan.rows.size() > (1ull << 27)) return fail(COLLATZ_EINVAL, "batch has too many rows");
  }
  const uint64_t nslots = slot_base[n];
  int rc = resolve_rows(sv, plan, s);
  if (rc) return rc;

  uint64_t cand = 0;
  for (const RowDesc& r : plan.rows) cand += r.i_hi - r.i_lo;
  const Tuning tn = choose_tuning(sv, (int)a.limbs, max_hi, cand, a.tile, a.waves, a.U);
  std::vector<TileDesc> tiles;
  std::vector<uint64_t> part_start;
  make_tiles(plan, nslots, tn.tile, tiles, part_start);
  if (tiles.size() > 0xFFFFFFFFull) return fail(COLLATZ_EINVAL, "batch has too many tiles");

  DevBuf<RowDesc> drows;
  DevBuf<TileDesc> dtiles;
  DevBuf<uint64_t> dpart;
  DevBuf<DevRec> dtile_recs, dchunk;
  DevBuf<uint32_t> dctl;  // [0] tile count, [1] tile counter
  CZ_HIP(drows.alloc(std::max<size_t>(1, plan.rows.size())));
  CZ_HIP(dtiles.alloc(std::max<size_t>(1, tiles.size())));
  CZ_HIP(dpart.alloc(part_start.size()));
  CZ_HIP(dtile_recs.alloc(std::max<size_t>(1, tiles.size())));
  CZ_HIP(dchunk.alloc(std::max<uint64_t>(1, nslots)));
  CZ_HIP(dctl.alloc(2));
  const uint32_t ctl[2] = {(uint32_t)tiles.size(), 0};
  if (!plan.rows.empty())
    CZ_HIP(hipMemcpyAsync(drows.p, plan.rows.data(), plan.rows.size() * sizeof(RowDesc),
                          hipMemcpyHostToDevice, s));
  if (!tiles.empty())
    CZ_HIP(hipMemcpyAsync(dtiles.p, tiles.data(), tiles.size() * sizeof(TileDesc),
                          hipMemcpyHostToDevice, s));
  CZ_HIP(hipMemcpyAsync(dpart.p, part_start.data(), part_start.size() * 8,
                        hipMemcpyHostToDevice, s));
  CZ_HIP(hipMemcpyAsync(dctl.p, ctl, sizeof(ctl), hipMemcpyHostToDevice, s));

  const uint64_t esc_cap = 1ull << 20;
  uint64_t ovf_cap = 1024;
  for (uint64_t i = 0; i < n; ++i) ovf_cap = std::max<uint64_t>(ovf_cap, outs[i].overflow_cap);
  DevBuf<EscEntry> desc;
  DevBuf<uint64_t> dovf;
  DevBuf<unsigned long long> dcnt;
  CZ_HIP(desc.alloc(a.limbs == 2 ? esc_cap : 1));
  CZ_HIP(dovf.alloc(4 * ovf_cap));
  CZ_HIP(dcnt.alloc(3));
  DeviceLists lists;
  lists.esc = desc.p;
  lists.esc_count = dcnt.p;
  lists.esc_cap = a.limbs == 2 ? esc_cap : 0;
  lists.ovf = dovf.p;
  lists.ovf_count = dcnt.p + 1;
  lists.ovf_cap = ovf_cap;
  lists.stall_count = dcnt.p + 2;
  const Limits lim = make_limits((int)a.limbs, a.cutoff);
  const uint32_t nblocks =
      (uint32_t)std::max<uint64_t>(1, (uint64_t)device_cu_count(sv->device) * tn.waves / 4);

  auto enqueue = [&](hipStream_t st) -> hipError_t {
    hipError_t e = launch_zero(dctl.p + 1, 1, dcnt.p, 3, st);
    if (e != hipSuccess) return e;
    e = launch_verify_persistent((int)a.limbs, sv->arrays(), drows.p, dtiles.p, dctl.p,
                                 dctl.p + 1, nblocks, tn.U, lim, sv->pow3.p, dtile_recs.p,
                                 lists, st);
    if (e != hipSuccess) return e;
    return launch_reduce_chunks(dtile_recs.p, dpart.p, (uint32_t)nslots, dchunk.p, st);
  };

  hipEvent_t e0, e1;
  CZ_HIP(hipEventCreate(&e0));
  CZ_HIP(hipEventCreate(&e1));
  if (a.graph) {
    hipStream_t cs;
    CZ_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    CZ_HIP(hipStreamSynchronize(s));
    hipGraph_t g;
    hipGraphExec_t ge;
    CZ_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    hipError_t ee = enqueue(cs);
    hipError_t ec = hipStreamEndCapture(cs, &g);
    if (ee != hipSuccess || ec != hipSuccess) {
      (void)hipStreamDestroy(cs);
      return fail(COLLATZ_EDEVICE, "graph capture failed");
    }
    CZ_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    CZ_HIP(hipEventRecord(e0, s));
    CZ_HIP(hipGraphLaunch(ge, s));
    CZ_HIP(hipEventRecord(e1, s));
    CZ_HIP(hipStreamSynchronize(s));
    (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(cs);
  } else {
    CZ_HIP(hipEventRecord(e0, s));
    CZ_HIP(enqueue(s));
    CZ_HIP(hipEventRecord(e1, s));
  }
  std::vector<DevRec> recs(nslots);
  unsigned long long cnt[3];
  if (nslots)
    CZ_HIP(hipMemcpyAsync(recs.data(), dchunk.p, nslots * sizeof(DevRec), hipMemcpyDeviceToHost, s));
  CZ_HIP(hipMemcpyAsync(cnt, dcnt.p, 24, hipMemcpyDeviceToHost, s));
  CZ_HIP(hipStreamSynchronize(s));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (a.limbs == 2 && cnt[0] > esc_cap && allow_retry) {
    BatchArgs b = a;
    b.limbs = 3;
    return verify_batch_impl(sv, n, iv, b, s, outs, false);
  }
  if (cnt[0]) {
    rc = run_escalated(sv, desc, cnt[0], a.cutoff, lists, recs, 0, s);
    if (rc) return rc;
    CZ_HIP(hipMemcpy(&cnt[1], dcnt.p + 1, 16, hipMemcpyDeviceToHost));
  }
  if (cnt[2])
    return fail(COLLATZ_EDEVICE, "a trajectory did not descend within the check bound");
  const uint64_t ovf_n = std::min<uint64_t>(cnt[1], ovf_cap);
  std::vector<uint64_t> ovf(4 * ovf_n);
  if (ovf_n) CZ_HIP(hipMemcpy(ovf.data(), dovf.p, ovf_n * 32, hipMemcpyDeviceToHost));
  bool short_buf = cnt[1] > ovf_cap;
  for (uint64_t i = 0; i < n; ++i) {
    collatz_range_result* out = &outs[i];
    const uint64_t nc = slot_base[i + 1] - slot_base[i];
    std::memset(&out->total, 0, sizeof(out->total));
    out->checks = 0;
    for (uint64_t c = 0; c < nc; ++c) {
      const uint64_t sl = slot_base[i] + c;
      collatz_chunk ch;
      chunk_from_rec(recs[sl], first_chunk[i] + c, nchk[sl], &ch);
      if (out->chunks && c < out->chunks_cap) out->chunks[c] = ch;
      chunk_merge(&out->total, ch);
      out->checks += recs[sl].checks;
    }
    out->total.chunk_index = first_chunk[i];
    out->chunks_written = nc;
    out->escalated = 0;
    out->device_ms = ms;
    out->overflow_written = 0;
    for (uint64_t j = 0; j < ovf_n; ++j) {
      const uint64_t sl = ovf[4 * j + 2];
      if (sl < slot_base[i] || sl >= slot_base[i + 1]) continue;
      if (out->overflow_starts && out->overflow_written < out->overflow_cap) {
        out->overflow_starts[2 * out->overflow_written] = ovf[4 * j];
        out->overflow_starts[2 * out->overflow_written + 1] = ovf[4 * j + 1];
      }
      out->overflow_written++;
    }
    if (out->total.overflow_count != out->overflow_written) short_buf = true;
    if ((out->chunks && nc > out->chunks_cap) || out->overflow_written > out->overflow_cap)
      short_buf = true;
  }
  if (cnt[0]) {
    // Attribute escalations to their intervals (for reporting).
    std::vector<EscEntry> h(cnt[0]);
    CZ_HIP(hipMemcpy(h.data(), desc.p, cnt[0] * sizeof(EscEntry), hipMemcpyDeviceToHost));
    for (const EscEntry& e : h) {
      const uint64_t i =
          std::upper_bound(slot_base.begin(), slot_base.end(), (uint64_t)e.slot) - slot_base.begin() - 1;
      if (i < n) outs[i].escalated++;
    }
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
  a.limbs = (uint32_t)limbs;
  a.tile = t.tile;
  a.waves = t.waves;
  a.U = t.U;
  collatz_range_result r;
  std::memset(&r, 0, sizeof(r));
  const int rc = verify_batch_impl(sv, 1, &iv, a, s, &r, false);
  *ms = r.device_ms;
  return rc == COLLATZ_ECAPACITY ? COLLATZ_OK : rc;
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
  HipDeviceGuard guard(sv->device);
  return verify_batch_impl(sv, n, intervals, a, (hipStream_t)stream, outs, true);
}
