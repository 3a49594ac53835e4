// C ABI: collatz_verify_range, the plain single-range form.
#include <algorithm>
#include <vector>

#include "engine.hpp"
#include "host_common.hpp"

namespace cz {

struct RangeArgs {
  uint32_t C = 24, limbs = 2, cutoff = 187;
  uint32_t width = 2;  // word width of the kernels (launch_width)
};

static int check_opts(const collatz_verify_opts* o, RangeArgs& a) {
  if (o) {
    if (o->chunk_log2) a.C = o->chunk_log2;
    if (o->limbs) a.limbs = o->limbs;
    if (o->cutoff_log2) a.cutoff = o->cutoff_log2;
  }
  if (a.C > 63) return fail(COLLATZ_EINVAL, "chunk_log2 must be <= 63");
  if (a.limbs != 2 && a.limbs != 3) return fail(COLLATZ_EINVAL, "limbs must be 2 or 3");
  if (a.cutoff < 64 || a.cutoff > 187) return fail(COLLATZ_EINVAL, "cutoff_log2 must be in [64, 187]");
  return COLLATZ_OK;
}

// Fills i_lo / i_hi of every row from the sieve with one device lookup.
int resolve_rows(const collatz_sieve* sv, RowPlan& plan, PlainScratch& scr, hipStream_t s) {
  std::vector<uint64_t> idx(plan.keys.size());
  if (!plan.keys.empty()) {
    const size_t nk = plan.keys.size();
    CZ_HIP(scr.keys.reserve(nk));
    CZ_HIP(scr.key_cls.reserve(nk));
    CZ_HIP(scr.key_idx.reserve(nk));
    CZ_HIP(hipMemcpyAsync(scr.keys.p, plan.keys.data(), nk * 8, hipMemcpyHostToDevice, s));
    CZ_HIP(hipMemcpyAsync(scr.key_cls.p, plan.key_cls.data(), nk * 4, hipMemcpyHostToDevice,
                          s));
    CZ_HIP(launch_lower_bound(sv->rec.p, sv->class_start_dev.p, scr.keys.p, scr.key_cls.p, nk,
                              scr.key_idx.p, s));
    CZ_HIP(hipMemcpyAsync(idx.data(), scr.key_idx.p, nk * 8, hipMemcpyDeviceToHost, s));
    CZ_HIP(hipStreamSynchronize(s));
  }
  for (size_t i = 0; i < plan.rows.size(); ++i) {
    const uint32_t c = plan.rows[i].cls;
    plan.rows[i].i_lo = plan.key_lo[i] < 0 ? sv->class_start[c] : idx[plan.key_lo[i]];
    plan.rows[i].i_hi = plan.key_hi[i] < 0 ? sv->class_start[c + 1] : idx[plan.key_hi[i]];
    if (plan.rows[i].i_hi < plan.rows[i].i_lo) plan.rows[i].i_hi = plan.rows[i].i_lo;
  }
  return COLLATZ_OK;
}

// Runs the escalated candidates on three limbs and folds them into recs (by slot).
int run_escalated(const collatz_sieve* sv, const DevBuf<EscEntry>& esc, uint64_t n,
                  uint32_t cutoff, DeviceLists lists, std::vector<DevRec>& recs,
                  uint32_t slot_base, DevBuf<DevRec>& out, hipStream_t s) {
  if (n == 0) return COLLATZ_OK;
  // The list holds at most esc_cap entries; more escalations than that were dropped by the
  // kernels, so the range must be rerun on three limbs instead.
  if (n > esc.n || n > lists.esc_cap)
    return fail(COLLATZ_ECAPACITY, "escalation list overflowed; rerun with limbs = 3");
  CZ_HIP(out.reserve(n));
  const Limits lim3 = make_limits(3, cutoff);
  CZ_HIP(launch_escalated(sv->arrays(), esc.p, n, 0, lim3, sv->pow3.p, out.p, lists, s));
  std::vector<DevRec> h(n);
  CZ_HIP(hipMemcpyAsync(h.data(), out.p, n * sizeof(DevRec), hipMemcpyDeviceToHost, s));
  CZ_HIP(hipStreamSynchronize(s));
  for (const DevRec& r : h) {
    const uint32_t slot = r.pad - slot_base;
    if (slot >= recs.size()) return fail(COLLATZ_EDEVICE, "escalated record has a bad slot");
    DevRec c = r;
    c.pad = 0;
    rec_merge_host(recs[slot], c);
  }
  return COLLATZ_OK;
}

// The caller holds sv->plain_mu.
static int verify_plain(const collatz_sieve* sv, u128 lo, u128 hi, const RangeArgs& a,
                        hipStream_t s, collatz_range_result* out) {
  PlainScratch& scr = sv->plain;
  RowPlan plan;
  int rc = plan_rows(sv->k, lo, hi, a.C, 0, sv->mod9 != 0, plan, 1ull << 26);
  if (rc) return rc;
  rc = resolve_rows(sv, plan, scr, s);
  if (rc) return rc;
  // Blocks of 256 lanes; each lane takes R residues of its row in turn.
  uint64_t slots = 0;
  for (const RowDesc& r : plan.rows) slots += r.i_hi - r.i_lo;
  const uint64_t target_blocks = 1ull << 18;
  const uint32_t R = (uint32_t)std::min<uint64_t>(
      256, std::max<uint64_t>(1, (slots + 256 * target_blocks - 1) / (256 * target_blocks)));
  const uint64_t span = 256ull * R;
  uint64_t nblocks = 0;
  for (RowDesc& r : plan.rows) {
    r.block_start = nblocks;
    nblocks += (r.i_hi - r.i_lo + span - 1) / span;
  }
  std::vector<uint64_t> part_start(plan.nchunks + 1);
  for (uint64_t c = 0; c <= plan.nchunks; ++c) {
    const uint64_t row = plan.chunk_first_row[c];
    part_start[c] = row < plan.rows.size() ? plan.rows[row].block_start : nblocks;
  }
  DevBuf<RowDesc>& drows = scr.rows;
  DevBuf<uint64_t>& dpart = scr.part;
  DevBuf<DevRec>& dblock = scr.block;
  DevBuf<DevRec>& dchunk = scr.chunk;
  CZ_HIP(drows.reserve(plan.rows.size()));
  CZ_HIP(dpart.reserve(part_start.size()));
  CZ_HIP(dblock.reserve(std::max<uint64_t>(1, nblocks)));
  CZ_HIP(dchunk.reserve(plan.nchunks));
  CZ_HIP(hipMemcpyAsync(drows.p, plan.rows.data(), plan.rows.size() * sizeof(RowDesc),
                        hipMemcpyHostToDevice, s));
  CZ_HIP(hipMemcpyAsync(dpart.p, part_start.data(), part_start.size() * 8,
                        hipMemcpyHostToDevice, s));
  const uint64_t esc_cap = 1ull << 20;
  const uint64_t ovf_cap = std::max<uint64_t>(out->overflow_cap, 1024);
  DevBuf<EscEntry>& desc = scr.esc;
  DevBuf<uint64_t>& dovf = scr.ovf;
  DevBuf<unsigned long long>& dcnt = scr.cnt;
  CZ_HIP(desc.reserve(a.width < 3 ? esc_cap : 1));
  CZ_HIP(dovf.reserve(4 * ovf_cap));
  CZ_HIP(dcnt.reserve(3));
  CZ_HIP(hipMemsetAsync(dcnt.p, 0, 24, s));
  DeviceLists lists;
  lists.esc = desc.p;
  lists.esc_count = dcnt.p;
  lists.esc_cap = a.width < 3 ? esc_cap : 0;
  lists.ovf = dovf.p;
  lists.ovf_count = dcnt.p + 1;
  lists.ovf_cap = ovf_cap;
  lists.stall_count = dcnt.p + 2;
  const Limits lim = make_limits((int)a.width, a.cutoff);
  // Device time: the verification and reduction kernels, then the three-limb pass.
  for (hipEvent_t& e : scr.ev)
    if (!e) CZ_HIP(hipEventCreate(&e));
  struct {
    hipEvent_t a, b;
  } ev{scr.ev[0], scr.ev[1]};
  CZ_HIP(hipEventRecord(ev.a, s));
  CZ_HIP(launch_verify_plain((int)a.width, sv->arrays(), drows.p, (uint32_t)plan.rows.size(),
                             nblocks, R, sv->mod9, lim, sv->pow3.p, dblock.p, lists, s));
  CZ_HIP(launch_reduce_chunks(dblock.p, dpart.p, (uint32_t)plan.nchunks, dchunk.p, s));
  CZ_HIP(hipEventRecord(ev.b, s));
  std::vector<DevRec> recs(plan.nchunks);
  unsigned long long cnt[3];
  CZ_HIP(hipMemcpyAsync(recs.data(), dchunk.p, recs.size() * sizeof(DevRec),
                        hipMemcpyDeviceToHost, s));
  CZ_HIP(hipMemcpyAsync(cnt, dcnt.p, 24, hipMemcpyDeviceToHost, s));
  CZ_HIP(hipStreamSynchronize(s));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, ev.a, ev.b);
  if (a.width < 3 && cnt[0] > esc_cap) {
    // More escalations than the list holds: run the whole range one width up.
    RangeArgs b = a;
    b.width = a.width + 1;
    return verify_plain(sv, lo, hi, b, s, out);
  }
  uint64_t ovf_n = cnt[1];
  if (cnt[0]) {
    CZ_HIP(hipEventRecord(ev.a, s));
    rc = run_escalated(sv, desc, cnt[0], a.cutoff, lists, recs, 0, scr.esc_recs, s);
    if (rc) return rc;
    CZ_HIP(hipEventRecord(ev.b, s));
    CZ_HIP(hipEventSynchronize(ev.b));
    float ems = 0;
    (void)hipEventElapsedTime(&ems, ev.a, ev.b);
    ms += ems;
    CZ_HIP(hipMemcpy(&cnt[1], dcnt.p + 1, 16, hipMemcpyDeviceToHost));
    ovf_n = cnt[1];
  }
  if (cnt[2])
    return fail(COLLATZ_EDEVICE, "a trajectory did not descend within the check bound");
  out->escalated = cnt[0];
  out->device_ms = ms;
  out->checks = 0;
  for (const DevRec& r : recs) out->checks += r.checks;
  out->chunks_written = plan.nchunks;
  std::memset(&out->total, 0, sizeof(out->total));
  for (uint64_t c = 0; c < plan.nchunks; ++c) {
    collatz_chunk ch;
    chunk_from_rec(recs[c], plan.first_chunk + c, plan.n_checked[c], &ch);
    if (out->chunks && c < out->chunks_cap) out->chunks[c] = ch;
    chunk_merge(&out->total, ch);
  }
  out->total.chunk_index = plan.first_chunk;
  out->overflow_written = ovf_n;
  if (out->overflow_starts && out->overflow_cap) {
    const uint64_t n = std::min<uint64_t>(std::min(ovf_n, ovf_cap), out->overflow_cap);
    std::vector<uint64_t> h(4 * n);
    if (n) CZ_HIP(hipMemcpy(h.data(), dovf.p, n * 32, hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < n; ++i) {
      out->overflow_starts[2 * i] = h[4 * i];
      out->overflow_starts[2 * i + 1] = h[4 * i + 1];
    }
  }
  if ((out->chunks && plan.nchunks > out->chunks_cap) ||
      (out->overflow_starts && ovf_n > out->overflow_cap))
    return fail(COLLATZ_ECAPACITY, "output buffer smaller than the result");
  return COLLATZ_OK;
}

}  // namespace cz

using namespace cz;

extern "C" int collatz_verify_range(const collatz_sieve* sv, const uint64_t lo[2],
                                    const uint64_t hi[2], const collatz_verify_opts* opts,
                                    void* stream, collatz_range_result* out) {
  if (!sv || !lo || !hi || !out) return fail(COLLATZ_EINVAL, "null argument");
  RangeArgs a;
  int rc = check_opts(opts, a);
  if (rc) return rc;
  const u128 L = mk128(lo[0], lo[1]), H = mk128(hi[0], hi[1]);
  if (L >= H) return fail(COLLATZ_EINVAL, "empty range (lo >= hi)");
  HipDeviceGuard guard(sv->device);
  hipStream_t s = (hipStream_t)stream;
  if (opts && opts->form == COLLATZ_FORM_PERSISTENT) {
    collatz_interval iv;
    iv.lo[0] = lo[0];
    iv.lo[1] = lo[1];
    iv.hi[0] = hi[0];
    iv.hi[1] = hi[1];
    collatz_batch_opts bo;
    std::memset(&bo, 0, sizeof(bo));
    bo.chunk_log2 = a.C;
    bo.limbs = a.limbs;
    bo.cutoff_log2 = a.cutoff;
    return collatz_verify_batch(sv, 1, &iv, &bo, stream, out);
  }
  a.width = (uint32_t)launch_width((int)a.limbs, H, a.cutoff);
  std::lock_guard<std::mutex> lk(sv->plain_mu);
  return verify_plain(sv, L, H, a, s, out);
}
