// C ABI: collatz_verify_range, the plain single-range form.
#include <algorithm>
#include <optional>
#include <vector>

#include "engine.hpp"
#include "host_common.hpp"

namespace cz {

int range_args(const collatz_verify_opts* o, RangeArgs& a) {
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
static int resolve_rows(const collatz_sieve* sv, RowPlan& plan, PlainBuffers& B,
                        PlainView& v, hipStream_t s) {
  std::vector<uint64_t> idx(plan.keys.size());
  if (!plan.keys.empty()) {
    const size_t nk = plan.keys.size();
    int rc = B.keys(nk, v);
    if (rc) return rc;
    CZ_HIP(hipMemcpyAsync(v.keys, plan.keys.data(), nk * 8, hipMemcpyHostToDevice, s));
    CZ_HIP(hipMemcpyAsync(v.key_cls, plan.key_cls.data(), nk * 4, hipMemcpyHostToDevice, s));
    CZ_HIP(launch_lower_bound(sv->rec.p, sv->class_start_dev.p, v.keys, v.key_cls, nk,
                              v.key_idx, s));
    CZ_HIP(hipMemcpyAsync(idx.data(), v.key_idx, nk * 8, hipMemcpyDeviceToHost, s));
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
int run_escalated_into(const collatz_sieve* sv, const EscEntry* esc, uint64_t n,
                       uint32_t cutoff, DeviceLists lists, std::vector<DevRec>& recs,
                       uint32_t slot_base, DevRec* out, hipStream_t s) {
  if (n == 0) return COLLATZ_OK;
  // The list holds at most esc_cap entries; more escalations than that were dropped by the
  // kernels, so the range must be rerun on three limbs instead.
  if (n > lists.esc_cap)
    return fail(COLLATZ_ECAPACITY, "escalation list overflowed; rerun with limbs = 3");
  const Limits lim3 = make_limits(3, cutoff);
  CZ_HIP(launch_escalated(sv->arrays(), esc, n, 0, lim3, sv->pow3.p, out, lists, s));
  std::vector<DevRec> h(n);
  CZ_HIP(hipMemcpyAsync(h.data(), out, n * sizeof(DevRec), hipMemcpyDeviceToHost, s));
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

int run_escalated(const collatz_sieve* sv, const DevBuf<EscEntry>& esc, uint64_t n,
                  uint32_t cutoff, DeviceLists lists, std::vector<DevRec>& recs,
                  uint32_t slot_base, DevBuf<DevRec>& out, hipStream_t s) {
  if (n == 0) return COLLATZ_OK;
  if (n > esc.n || n > lists.esc_cap)
    return fail(COLLATZ_ECAPACITY, "escalation list overflowed; rerun with limbs = 3");
  CZ_HIP(out.reserve(n));
  return run_escalated_into(sv, esc.p, n, cutoff, lists, recs, slot_base, out.p, s);
}

int verify_plain_with(const collatz_sieve* sv, u128 lo, u128 hi, const RangeArgs& a,
                      hipStream_t s, collatz_range_result* out, PlainBuffers& B) {
  PlainView v;
  RowPlan& plan = B.plan();
  plan.clear();
  // Host steps under COLLATZ_TRACE=1 (each span ends where the next begins).
  std::optional<Span> step;
  step.emplace("plain: plan rows");
  int rc = plan_rows(sv->k, lo, hi, a.C, 0, sv->mod9 != 0, plan, B.max_rows());
  if (rc) return rc;
  if (plan.nchunks > B.max_chunks()) return fail(COLLATZ_EINVAL, "range has too many chunks");
  step.emplace("plain: resolve rows");
  rc = resolve_rows(sv, plan, B, v, s);
  if (rc) return rc;
  // One wave per share: each row is cut into equal pieces of at most S survivors.  A share
  // costs its wave a set-up (the row, its q 3^f table, the first records in flight), a first
  // check at which every lane offers its value for the max, and an end at which lanes idle
  // while the share's last trajectories finish, together about the work of a few hundred
  // survivors; so S is at least 16384 (most rows of a 2^24-wide chunk at k = 32 are one
  // share), but small enough that every resident wave has about four shares to take.  The
  // rows of the last quarter of the survivors, whose shares the launch's last blocks take
  // (blocks start in index order), are cut at S / 2, so that the launch ends on shorter
  // shares.  ([2^40, 2^40 + 2^36), k = 32: each row split over the four waves of a block
  // ran 1.55 ms; one share per row 1.38-1.43 ms; with the shorter last quarter 1.34 ms, and
  // the 4096 intervals' plain calls 5% less device time.)  The shares must also fit the
  // record buffer: at most 1.25 slots / S + rows of them.
  step.emplace("plain: shares, launch");
  uint64_t slots = 0;
  for (const RowDesc& r : plan.rows) slots += r.i_hi - r.i_lo;
  const uint64_t resident = (uint64_t)device_cu_count(sv->device) * 32;
  uint64_t S = std::max<uint64_t>(slots / (4 * resident),
                                  std::min<uint64_t>(16384, std::max<uint64_t>(1, slots / resident)));
  const uint64_t cap = B.max_shares() > plan.rows.size() ? B.max_shares() - plan.rows.size() : 1;
  S = std::max<uint64_t>({S, slots ? (slots * 5 / 4) / cap + 2 : 1, 2});
  const uint64_t tail_from = slots - slots / 4;
  uint64_t nshares = 0, seen = 0;
  for (RowDesc& r : plan.rows) {
    r.block_start = nshares;
    const uint64_t len = r.i_hi - r.i_lo;
    const uint64_t piece = seen >= tail_from ? S / 2 : S;
    nshares += (len + piece - 1) / piece;
    seen += len;
  }
  std::vector<uint64_t> part_start(plan.nchunks + 1);
  for (uint64_t c = 0; c <= plan.nchunks; ++c) {
    const uint64_t row = plan.chunk_first_row[c];
    part_start[c] = row < plan.rows.size() ? plan.rows[row].block_start : nshares;
  }
  const uint64_t ovf_cap = std::max<uint64_t>(out->overflow_cap, 1024);
  rc = B.launch(plan.rows.size(), nshares, plan.nchunks, ovf_cap, a.width < 3, v);
  if (rc) return rc;
  CZ_HIP(hipMemcpyAsync(v.rows, plan.rows.data(), plan.rows.size() * sizeof(RowDesc),
                        hipMemcpyHostToDevice, s));
  CZ_HIP(hipMemcpyAsync(v.part, part_start.data(), part_start.size() * 8,
                        hipMemcpyHostToDevice, s));
  CZ_HIP(hipMemsetAsync(v.cnt, 0, 32, s));
  DeviceLists lists;
  lists.esc = v.esc;
  lists.esc_count = v.cnt;
  lists.esc_cap = a.width < 3 ? v.esc_cap : 0;
  lists.ovf = v.ovf;
  lists.ovf_count = v.cnt + 1;
  lists.ovf_cap = v.ovf_cap;
  lists.stall_count = v.cnt + 2;
  lists.stall_slots = v.stall;
  lists.stall_cap = kStallCap;
  lists.loop_aborts = v.cnt + 3;
  const Limits lim = make_limits((int)a.width, a.cutoff);
  // Device time: the verification and reduction kernels, then the three-limb pass.
  CZ_HIP(hipEventRecord(v.ev[0], s));
  const bool lowkey = hi <= ((u128)1 << kLowKeyLog2);
  CZ_HIP(launch_verify_plain((int)a.width, sv->arrays(), v.rows, (uint32_t)plan.rows.size(),
                             nshares, sv->mod9, lowkey, lim, sv->pow3.p, v.block, lists, s));
  CZ_HIP(launch_reduce_chunks(v.block, v.part, (uint32_t)plan.nchunks, v.chunk, s));
  CZ_HIP(hipEventRecord(v.ev[1], s));
  step.emplace("plain: wait, records back");
  std::vector<DevRec> recs(plan.nchunks);
  unsigned long long cnt[4];
  CZ_HIP(hipMemcpyAsync(recs.data(), v.chunk, recs.size() * sizeof(DevRec),
                        hipMemcpyDeviceToHost, s));
  CZ_HIP(hipMemcpyAsync(cnt, v.cnt, 32, hipMemcpyDeviceToHost, s));
  CZ_HIP(hipStreamSynchronize(s));
  step.emplace("plain: chunk records");
  float ms = 0;
  (void)hipEventElapsedTime(&ms, v.ev[0], v.ev[1]);
  if (cnt[3])
    return fail(COLLATZ_EDEVICE, "a work loop of the plain kernel reached its iteration bound (" +
                                     std::to_string(cnt[3]) + " waves abandoned)");
  if (a.width < 3 && cnt[0] > lists.esc_cap) {
    // More escalations than the list holds: run the whole range one width up.
    RangeArgs b = a;
    b.width = a.width + 1;
    return verify_plain_with(sv, lo, hi, b, s, out, B);
  }
  uint64_t ovf_n = cnt[1];
  if (cnt[0]) {
    rc = B.esc_recs(cnt[0], v);
    if (rc) return rc;
    CZ_HIP(hipEventRecord(v.ev[0], s));
    rc = run_escalated_into(sv, v.esc, cnt[0], a.cutoff, lists, recs, 0, v.esc_recs, s);
    if (rc) return rc;
    CZ_HIP(hipEventRecord(v.ev[1], s));
    CZ_HIP(hipEventSynchronize(v.ev[1]));
    float ems = 0;
    (void)hipEventElapsedTime(&ems, v.ev[0], v.ev[1]);
    ms += ems;
    CZ_HIP(hipMemcpyAsync(&cnt[1], v.cnt + 1, 16, hipMemcpyDeviceToHost, s));
    CZ_HIP(hipStreamSynchronize(s));
    ovf_n = cnt[1];
  }
  if (cnt[2]) {
    uint32_t slot = 0;
    CZ_HIP(hipMemcpyAsync(&slot, v.stall, 4, hipMemcpyDeviceToHost, s));
    CZ_HIP(hipStreamSynchronize(s));
    return fail(COLLATZ_EDEVICE, "a trajectory did not descend within the check bound (chunk " +
                                     std::to_string(plan.first_chunk + slot) + ")");
  }
  out->escalated = cnt[0];
  out->spot_checks = 0;
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
  uint64_t listed = 0;
  if (out->overflow_starts && out->overflow_cap) {
    listed = std::min<uint64_t>(std::min<uint64_t>(ovf_n, v.ovf_cap), out->overflow_cap);
    std::vector<uint64_t> h(4 * listed);
    if (listed) {
      CZ_HIP(hipMemcpyAsync(h.data(), v.ovf, listed * 32, hipMemcpyDeviceToHost, s));
      CZ_HIP(hipStreamSynchronize(s));
    }
    for (uint64_t i = 0; i < listed; ++i) {
      out->overflow_starts[2 * i] = h[4 * i];
      out->overflow_starts[2 * i + 1] = h[4 * i + 1];
    }
  }
  if ((out->chunks && plan.nchunks > out->chunks_cap) ||
      (out->overflow_starts && (ovf_n > out->overflow_cap || listed < ovf_n)))
    return fail(COLLATZ_ECAPACITY, "output buffer smaller than the result");
  return COLLATZ_OK;
}

// The sieve handle's buffers: grown on demand and kept for later calls, so a call makes no
// device allocation unless it needs more room than any call before it.
class ScratchBuffers final : public PlainBuffers {
 public:
  explicit ScratchBuffers(PlainScratch& s) : scr_(s) {}
  RowPlan& plan() override { return scr_.plan; }
  uint64_t max_rows() const override { return 1ull << 26; }
  uint64_t max_chunks() const override { return 0x7FFFFFFFull; }
  uint64_t max_shares() const override { return ~0ull; }
  int keys(uint64_t n, PlainView& v) override {
    CZ_HIP(scr_.keys.reserve(n));
    CZ_HIP(scr_.key_cls.reserve(n));
    CZ_HIP(scr_.key_idx.reserve(n));
    v.keys = scr_.keys.p;
    v.key_cls = scr_.key_cls.p;
    v.key_idx = scr_.key_idx.p;
    return COLLATZ_OK;
  }
  int launch(uint64_t nrows, uint64_t nshares, uint64_t nchunks, uint64_t ovf_cap, bool esc,
             PlainView& v) override {
    const uint64_t esc_cap = 1ull << 20;
    CZ_HIP(scr_.rows.reserve(std::max<uint64_t>(1, nrows)));
    CZ_HIP(scr_.part.reserve(nchunks + 1));
    CZ_HIP(scr_.block.reserve(std::max<uint64_t>(1, nshares)));
    CZ_HIP(scr_.chunk.reserve(std::max<uint64_t>(1, nchunks)));
    CZ_HIP(scr_.esc.reserve(esc ? esc_cap : 1));
    CZ_HIP(scr_.ovf.reserve(4 * ovf_cap));
    CZ_HIP(scr_.cnt.reserve(4));
    CZ_HIP(scr_.stall.reserve(kStallCap));
    for (hipEvent_t& e : scr_.ev)
      if (!e) CZ_HIP(hipEventCreate(&e));
    v.rows = scr_.rows.p;
    v.part = scr_.part.p;
    v.block = scr_.block.p;
    v.chunk = scr_.chunk.p;
    v.esc = scr_.esc.p;
    v.esc_cap = esc ? esc_cap : 0;
    v.ovf = scr_.ovf.p;
    v.ovf_cap = ovf_cap;
    v.cnt = scr_.cnt.p;
    v.stall = scr_.stall.p;
    v.ev[0] = scr_.ev[0];
    v.ev[1] = scr_.ev[1];
    return COLLATZ_OK;
  }
  int esc_recs(uint64_t n, PlainView& v) override {
    CZ_HIP(scr_.esc_recs.reserve(n));
    v.esc_recs = scr_.esc_recs.p;
    return COLLATZ_OK;
  }

 private:
  PlainScratch& scr_;
};

}  // namespace cz

using namespace cz;

extern "C" int collatz_verify_range(const collatz_sieve* sv, const uint64_t lo[2],
                                    const uint64_t hi[2], const collatz_verify_opts* opts,
                                    void* stream, collatz_range_result* out) {
  if (!sv || !lo || !hi || !out) return fail(COLLATZ_EINVAL, "null argument");
  RangeArgs a;
  int rc = range_args(opts, a);
  if (rc) return rc;
  const u128 L = mk128(lo[0], lo[1]), H = mk128(hi[0], hi[1]);
  if (L >= H) return fail(COLLATZ_EINVAL, "empty range (lo >= hi)");
  HipDeviceGuard guard(sv->device);
  hipStream_t s = call_stream(stream);
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
  ScratchBuffers B(sv->plain);
  return verify_plain_with(sv, L, H, a, s, out, B);
}
