// Persistent-launch engine (see engine.hpp).
#include "engine.hpp"

#include <algorithm>
#include <cstring>

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

bool Engine::GraphKey::operator==(const GraphKey& o) const {
  return std::memcmp(bufs, o.bufs, sizeof(bufs)) == 0 && limbs == o.limbs &&
         cutoff == o.cutoff && nblocks == o.nblocks && U == o.U && max_steps == o.max_steps &&
         max_loop == o.max_loop;
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
  e = launch_verify_persistent(k.limbs, sv_->arrays(), rows_.p, tiles_.p, ctl_.p, k.nblocks,
                               k.U, lim, sv_->pow3.p, tile_recs_.p, lists, s);
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
  if (!lp.rows.empty())
    CZ_HIP(hipMemcpyAsync(rows_.p, h_rows_.p, lp.rows.size() * sizeof(RowDesc),
                          hipMemcpyHostToDevice, s));
  if (!lp.keys.empty())
    CZ_HIP(hipMemcpyAsync(keys_.p, h_keys_.p, lp.keys.size() * 8, hipMemcpyHostToDevice, s));
  CZ_HIP(hipMemcpyAsync(part_.p, h_part_.p, lp.part.size() * 8, hipMemcpyHostToDevice, s));
  CZ_HIP(hipMemcpyAsync(ctl_.p, &ctl, sizeof(Ctl), hipMemcpyHostToDevice, s));

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
  key.limbs = limbs;
  key.cutoff = cutoff;
  const uint32_t waves =
      std::min<uint32_t>(tn.waves, persistent_waves_fit(limbs, cutoff, sv_->la_bits != 0));
  key.nblocks =
      (uint32_t)std::max<uint64_t>(1, (uint64_t)device_cu_count(sv_->device) * waves / 4);
  key.U = tn.U;
  key.max_steps = lim.max_steps;
  key.max_loop = lim.max_loop;

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
