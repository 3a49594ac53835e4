// Persistent-launch engine (see engine.hpp).
#include "engine.hpp"

#include <algorithm>
#include <cstring>
#include <unordered_map>

namespace cz {

int device_cu_count(int device) {
  static int cached[64] = {0};
  if (device >= 0 && device < 64 && cached[device]) return cached[device];
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return 256;
  if (device >= 0 && device < 64) cached[device] = p.multiProcessorCount;
  return p.multiProcessorCount;
}

int plan_interval(const collatz_sieve* sv, u128 lo, u128 hi, uint32_t C, uint64_t slot_base,
                  uint32_t tile, LaunchPlan& lp, std::vector<uint64_t>& n_checked,
                  uint64_t& first_chunk, uint64_t& nchunks) {
  if (slot_base > 0x7FFFFFFFull) return fail(COLLATZ_EINVAL, "too many chunks");
  RowPlan p;
  int rc = plan_rows(sv->k, lo, hi, C, (uint32_t)slot_base, sv->mod9 != 0, p, 1ull << 26);
  if (rc) return rc;
  if (slot_base + p.nchunks > 0x7FFFFFFFull) return fail(COLLATZ_EINVAL, "too many chunks");
  first_chunk = p.first_chunk;
  nchunks = p.nchunks;
  n_checked.insert(n_checked.end(), p.n_checked.begin(), p.n_checked.end());
  const double kspan = (double)((u128)1 << sv->k);
  uint32_t last_slot = UINT32_MAX;
  for (size_t r = 0; r < p.rows.size(); ++r) {
    RowDesc row = p.rows[r];
    const uint32_t c = row.cls;
    const double r_lo = p.key_lo[r] < 0 ? 0.0 : (double)p.keys[p.key_lo[r]];
    const double r_hi = p.key_hi[r] < 0 ? kspan : (double)p.keys[p.key_hi[r]];
    if (p.key_lo[r] < 0) {
      row.i_lo = sv->class_start[c];
    } else {
      row.i_lo = kResolveFlag | lp.keys.size();
      lp.keys.push_back(p.keys[p.key_lo[r]]);
    }
    if (p.key_hi[r] < 0) {
      row.i_hi = sv->class_start[c + 1];
    } else {
      row.i_hi = kResolveFlag | lp.keys.size();
      lp.keys.push_back(p.keys[p.key_hi[r]]);
    }
    // Pieces from the expected number of survivors in the window; the exact split of the
    // resolved index range happens on the device.
    const double est =
        (double)(sv->class_start[c + 1] - sv->class_start[c]) * (r_hi - r_lo) / kspan;
    lp.candidates_est += (uint64_t)est;
    uint64_t pieces = (uint64_t)(est / tile + 0.5);
    pieces = std::min<uint64_t>(std::max<uint64_t>(pieces, 1), 1u << 20);
    if (row.slot != last_slot) {
      lp.part.push_back(lp.ntiles);
      last_slot = row.slot;
    }
    row.block_start = lp.ntiles;
    lp.ntiles += pieces;
    lp.rows.push_back(row);
  }
  lp.nslots = slot_base + p.nchunks;
  if (lp.rows.size() > (1ull << 27) || lp.ntiles > 0x7FFFFFFFull)
    return fail(COLLATZ_EINVAL, "launch has too many rows or tiles");
  return COLLATZ_OK;
}

// Rows of a group interleaved at a time (a group of [2^60, 2^60 + 2^44) with k = 37 is its
// 128 q; a batch of 4096 intervals has groups of thousands of rows, and interleaving all of
// them at once put every wave of an XCD on the same records: 328 -> 339 ms per launch).
constexpr uint32_t kTileBand = 128;
// Tiles per resident wave from which a launch uses the per-XCD shares of the tile order
// (the 4096-interval batch has 65, [2^60, 2^60 + 2^44) at k = 37 has 38; a saturated served
// launch of 2^30 candidates has 8, and ran 5.3e13 n/s with shares against 5.8e13 without).
constexpr uint32_t kSharesMinTilesPerWave = 32;

uint64_t tile_order(const LaunchPlan& lp, uint32_t* order, uint32_t band) {
  const size_t n = lp.rows.size();
  // A row's survivor window: its class and its two bounds, each either a class boundary
  // (an index) or a residue key resolved on the device (its value); tagged apart.
  struct Key {
    uint64_t lo, hi;
    uint32_t cls;
    bool operator==(const Key& o) const { return lo == o.lo && hi == o.hi && cls == o.cls; }
  };
  struct KeyHash {
    size_t operator()(const Key& k) const {
      uint64_t h = k.lo * 0x9E3779B97F4A7C15ull;
      h ^= (k.hi + 0x632BE59BD9B4E019ull) * 0xBF58476D1CE4E5B9ull;
      h ^= (uint64_t)k.cls * 0x94D049BB133111EBull;
      return (size_t)(h ^ (h >> 31));
    }
  };
  auto bound = [&](uint64_t b) -> uint64_t {
    return (b & kResolveFlag) ? (lp.keys[b & ~kResolveFlag] << 1) | 1u : b << 1;
  };
  std::unordered_map<Key, uint32_t, KeyHash> gid;
  gid.reserve(n);
  std::vector<uint32_t> grp(n), gsize, gpieces;
  for (size_t r = 0; r < n; ++r) {
    const RowDesc& row = lp.rows[r];
    const Key k{bound(row.i_lo), bound(row.i_hi), row.cls};
    auto it = gid.emplace(k, (uint32_t)gsize.size()).first;
    if (it->second == gsize.size()) {
      gsize.push_back(0);
      gpieces.push_back(0);
    }
    grp[r] = it->second;
    gsize[it->second]++;
    const uint64_t e = r + 1 < n ? lp.rows[r + 1].block_start : lp.ntiles;
    gpieces[it->second] =
        std::max<uint32_t>(gpieces[it->second], (uint32_t)(e - lp.rows[r].block_start));
  }
  // The rows of each group, in row order (a counting sort by group).
  std::vector<uint32_t> gstart(gsize.size() + 1, 0), members(n);
  for (size_t g = 0; g < gsize.size(); ++g) gstart[g + 1] = gstart[g] + gsize[g];
  std::vector<uint32_t> fill(gstart.begin(), gstart.end() - 1);
  for (size_t r = 0; r < n; ++r) members[fill[grp[r]]++] = (uint32_t)r;
  uint64_t pos = 0;
  for (size_t g = 0; g < gsize.size(); ++g) {
    for (uint64_t m0 = gstart[g]; m0 < gstart[g + 1]; m0 += band) {
      const uint64_t m1 = std::min<uint64_t>(gstart[g + 1], m0 + band);
      for (uint32_t j = 0; j < gpieces[g]; ++j) {
        for (uint64_t m = m0; m < m1; ++m) {
          const uint32_t r = members[m];
          const uint64_t b = lp.rows[r].block_start;
          const uint64_t e = r + 1 < n ? lp.rows[r + 1].block_start : lp.ntiles;
          if (j < e - b) order[pos++] = (uint32_t)(b + j);
        }
      }
    }
  }
  return pos;
}

bool Engine::GraphKey::operator==(const GraphKey& o) const {
  return std::memcmp(bufs, o.bufs, sizeof(bufs)) == 0 && limbs == o.limbs &&
         cutoff == o.cutoff && nblocks == o.nblocks && U == o.U && max_steps == o.max_steps &&
         max_loop == o.max_loop && lowkey == o.lowkey;
}

Engine::~Engine() {
  if (gexec_) (void)hipGraphExecDestroy(gexec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  if (e0_) (void)hipEventDestroy(e0_);
  if (e1_) (void)hipEventDestroy(e1_);
  if (e2_) (void)hipEventDestroy(e2_);
  if (e3_) (void)hipEventDestroy(e3_);
}

int Engine::ensure(const LaunchPlan& lp, uint64_t ovf_cap) {
  const uint64_t esc_cap = 1ull << 20;
  CZ_HIP(rows_.reserve(std::max<size_t>(1, lp.rows.size())));
  CZ_HIP(keys_.reserve(std::max<size_t>(1, lp.keys.size())));
  CZ_HIP(tiles_.reserve(std::max<size_t>(1, lp.ntiles)));
  CZ_HIP(order_.reserve(std::max<size_t>(1, lp.ntiles)));
  CZ_HIP(h_order_.reserve(std::max<size_t>(1, lp.ntiles)));
  CZ_HIP(part_.reserve(lp.part.size()));
  CZ_HIP(tile_recs_.reserve(std::max<size_t>(1, lp.ntiles)));
  CZ_HIP(h_rows_.reserve(std::max<size_t>(1, lp.rows.size())));
  CZ_HIP(h_keys_.reserve(std::max<size_t>(1, lp.keys.size())));
  CZ_HIP(h_part_.reserve(lp.part.size()));
  CZ_HIP(slot_recs_.reserve(std::max<uint64_t>(1, lp.nslots)));
  CZ_HIP(esc_.reserve(esc_cap));
  CZ_HIP(ovf_.reserve(4 * ovf_cap));
  CZ_HIP(cnt_.reserve(4));
  CZ_HIP(stall_.reserve(kStallCap));
  CZ_HIP(ctl_.reserve(1));
  CZ_HIP(h_recs_.reserve(std::max<uint64_t>(1, lp.nslots)));
  CZ_HIP(h_cnt_.reserve(4));
  CZ_HIP(h_ctl_.reserve(1));
  if (!e0_) CZ_HIP(hipEventCreate(&e0_));
  if (!e1_) CZ_HIP(hipEventCreate(&e1_));
  if (!e2_) CZ_HIP(hipEventCreate(&e2_));
  if (!e3_) CZ_HIP(hipEventCreate(&e3_));
  return COLLATZ_OK;
}

hipError_t Engine::enqueue(const GraphKey& k, const Limits& lim, DeviceLists lists,
                           hipStream_t s) {
  hipError_t e = launch_engine_reset(ctl_.p, cnt_.p, 4, s);
  if (e != hipSuccess) return e;
  e = launch_rows_resolve_ctl(rows_.p, ctl_.p, sv_->rec.p, sv_->class_start_dev.p, keys_.p, s);
  if (e != hipSuccess) return e;
  e = launch_make_tiles(rows_.p, ctl_.p, tiles_.p, s);
  if (e != hipSuccess) return e;
  e = launch_verify_persistent(k.limbs, sv_->arrays(), rows_.p, tiles_.p, order_.p, ctl_.p,
                               k.nblocks, k.U, k.lowkey, lim, sv_->pow3.p, tile_recs_.p, lists,
                               s);
  if (e != hipSuccess) return e;
  return launch_reduce_slots(tile_recs_.p, part_.p, ctl_.p, slot_recs_.p, s);
}

int Engine::run(const LaunchPlan& lp, const Tuning& tn, int limbs, uint32_t cutoff, bool graph,
                uint64_t ovf_cap, hipStream_t s, LaunchResult& out) {
  ovf_cap = std::max<uint64_t>(ovf_cap, 1024);
  int rc = ensure(lp, ovf_cap);
  if (rc) return rc;
  Ctl& ctl = *h_ctl_.p;
  std::memset(&ctl, 0, sizeof(ctl));
  ctl.nrows = lp.rows.size();
  ctl.nslots = lp.nslots;
  ctl.ntiles = (uint32_t)lp.ntiles;
  // Staged through pinned buffers so the copies are asynchronous.  The previous call on this
  // engine has synchronised its stream, so the staging buffers are free.
  std::memcpy(h_rows_.p, lp.rows.data(), lp.rows.size() * sizeof(RowDesc));
  std::memcpy(h_keys_.p, lp.keys.data(), lp.keys.size() * 8);
  std::memcpy(h_part_.p, lp.part.data(), lp.part.size() * 8);
  if (tile_order(lp, h_order_.p, kTileBand) != lp.ntiles)
    return fail(COLLATZ_EDEVICE, "internal: the tile order does not list every tile once");
  if (!lp.rows.empty())
    CZ_HIP(hipMemcpyAsync(rows_.p, h_rows_.p, lp.rows.size() * sizeof(RowDesc),
                          hipMemcpyHostToDevice, s));
  if (!lp.keys.empty())
    CZ_HIP(hipMemcpyAsync(keys_.p, h_keys_.p, lp.keys.size() * 8, hipMemcpyHostToDevice, s));
  CZ_HIP(hipMemcpyAsync(part_.p, h_part_.p, lp.part.size() * 8, hipMemcpyHostToDevice, s));
  if (lp.ntiles)
    CZ_HIP(hipMemcpyAsync(order_.p, h_order_.p, lp.ntiles * 4, hipMemcpyHostToDevice, s));

  DeviceLists lists;
  lists.esc = esc_.p;
  lists.esc_count = cnt_.p;
  lists.esc_cap = limbs < 3 ? esc_.n : 0;
  lists.ovf = ovf_.p;
  lists.ovf_count = cnt_.p + 1;
  lists.ovf_cap = ovf_cap;
  lists.stall_count = cnt_.p + 2;
  lists.stall_slots = stall_.p;
  lists.stall_cap = kStallCap;
  lists.loop_aborts = cnt_.p + 3;
  const Limits lim = make_limits(limbs, cutoff);
  GraphKey key;
  std::memset(&key, 0, sizeof(key));
  key.bufs[0] = rows_.p;
  key.bufs[1] = keys_.p;
  key.bufs[2] = tiles_.p;
  key.bufs[3] = part_.p;
  key.bufs[4] = tile_recs_.p;
  key.bufs[5] = slot_recs_.p;
  key.bufs[6] = ovf_.p;
  key.bufs[7] = (const void*)(uintptr_t)ovf_cap;
  key.bufs[8] = order_.p;
  key.limbs = limbs;
  key.cutoff = cutoff;
  const uint32_t waves =
      std::min<uint32_t>(tn.waves, persistent_waves_fit(limbs, cutoff, sv_->la_bits != 0));
  key.nblocks =
      (uint32_t)std::max<uint64_t>(1, (uint64_t)device_cu_count(sv_->device) * waves / 4);
  // Per-XCD shares of the tile order only for launches that give every resident wave many
  // tiles: each share's waves then read their groups' records from their own L2 (on
  // [2^60, 2^60 + 2^44) with k = 37, 16 -> 1.5 B per candidate), while in a short launch the
  // extra turns of the exhausted shares' waves at the end cost more than the reads save
  // ([2^40, 2^40 + 2^36): 1.55 -> 2.17 ms; a served launch of a few requests: p50 0.26 ->
  // 0.95 ms).
  static_assert(kXcds == 8, "shares are 2^share_log2 with share_log2 = 3");
  ctl.share_log2 = lp.ntiles >= (uint64_t)kSharesMinTilesPerWave * key.nblocks * 4 ? 3 : 0;
  CZ_HIP(hipMemcpyAsync(ctl_.p, &ctl, sizeof(Ctl), hipMemcpyHostToDevice, s));
  key.U = tn.U;
  key.max_steps = lim.max_steps;
  key.max_loop = lim.max_loop;
  // Every start below 2^kLowKeyLog2 (a row's starts are below (q + 1) 2^k): the low-key
  // test of the wave max.
  key.lowkey = true;
  for (const RowDesc& r : lp.rows) {
    const uint64_t bound = 1ull << kLowKeyLog2;
    if (r.qb_hi || r.qb_lo >= bound || bound - r.qb_lo < (1ull << sv_->k)) {
      key.lowkey = false;
      break;
    }
  }

  CZ_HIP(hipEventRecord(e0_, s));
  if (graph) {
    if (!gexec_ || !(key == gkey_)) {
      Span sp("graph capture");
      if (gexec_) (void)hipGraphExecDestroy(gexec_);
      if (graph_) (void)hipGraphDestroy(graph_);
      gexec_ = nullptr;
      graph_ = nullptr;
      hipStream_t cs;
      CZ_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
      CZ_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      const hipError_t ee = enqueue(key, lim, lists, cs);
      const hipError_t ec = hipStreamEndCapture(cs, &graph_);
      (void)hipStreamDestroy(cs);
      if (ee != hipSuccess || ec != hipSuccess) return fail(COLLATZ_EDEVICE, "graph capture failed");
      CZ_HIP(hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0));
      gkey_ = key;
    }
    CZ_HIP(hipGraphLaunch(gexec_, s));
  } else {
    CZ_HIP(enqueue(key, lim, lists, s));
  }
  CZ_HIP(hipEventRecord(e1_, s));
  if (lp.nslots)
    CZ_HIP(hipMemcpyAsync(h_recs_.p, slot_recs_.p, lp.nslots * sizeof(DevRec),
                          hipMemcpyDeviceToHost, s));
  CZ_HIP(hipMemcpyAsync(h_cnt_.p, cnt_.p, 32, hipMemcpyDeviceToHost, s));
  {
    Span sp("engine wait");
    CZ_HIP(hipStreamSynchronize(s));
  }
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0_, e1_);
  out.device_ms = ms;
  out.recs.assign(h_recs_.p, h_recs_.p + lp.nslots);
  const uint64_t esc_n = h_cnt_.p[0];
  if (limbs < 3 && esc_n > esc_.n)
    return fail(COLLATZ_ECAPACITY, "escalation list overflowed; rerun one width up");
  out.escalated = esc_n;
  out.esc_slots.clear();
  if (esc_n) {
    CZ_HIP(hipEventRecord(e2_, s));
    rc = run_escalated(sv_, esc_, esc_n, cutoff, lists, out.recs, 0, esc_recs_, s);
    if (rc) return rc;
    CZ_HIP(hipEventRecord(e3_, s));
    CZ_HIP(hipEventSynchronize(e3_));
    float ems = 0;
    (void)hipEventElapsedTime(&ems, e2_, e3_);
    out.device_ms += ems;
    std::vector<EscEntry> h(esc_n);
    CZ_HIP(hipMemcpyAsync(h.data(), esc_.p, esc_n * sizeof(EscEntry), hipMemcpyDeviceToHost, s));
    CZ_HIP(hipMemcpyAsync(h_cnt_.p, cnt_.p, 32, hipMemcpyDeviceToHost, s));
    CZ_HIP(hipStreamSynchronize(s));
    for (const EscEntry& e : h) out.esc_slots.push_back(e.slot);
  }
  out.stalled = h_cnt_.p[2];
  out.loop_aborts = h_cnt_.p[3];
  if (out.loop_aborts)
    return fail(COLLATZ_EDEVICE, "a persistent work loop reached its iteration bound (" +
                                     std::to_string(out.loop_aborts) + " tiles abandoned)");
  out.stall_slots.clear();
  if (out.stalled) {
    out.stall_slots.resize(std::min<uint64_t>(out.stalled, kStallCap));
    CZ_HIP(hipMemcpyAsync(out.stall_slots.data(), stall_.p, out.stall_slots.size() * 4,
                          hipMemcpyDeviceToHost, s));
    CZ_HIP(hipStreamSynchronize(s));
  }
  out.ovf_total = h_cnt_.p[1];
  const uint64_t nov = std::min<uint64_t>(out.ovf_total, ovf_cap);
  out.ovf.resize(4 * nov);
  if (nov) {
    CZ_HIP(hipMemcpyAsync(out.ovf.data(), ovf_.p, nov * 32, hipMemcpyDeviceToHost, s));
    CZ_HIP(hipStreamSynchronize(s));
  }
  return COLLATZ_OK;
}

}  // namespace cz
