// C ABI: the served form -- a request loop with continuous batching.
//
// Requests are submitted at any time.  A worker thread owns an Engine (its own device
// buffers, stream and captured graph) and runs back-to-back persistent launches; each
// launch takes the remaining rows of every admitted request, oldest first, up to a work
// budget.  A request that arrives while a launch runs is admitted at the next launch as its
// footprint fits the budget, and its rows share that launch with the rows of earlier
// requests.  Per-launch records are folded into each request's chunk records on the host, so
// a request's result does not depend on what it was batched with.  The sieve stays resident.
// The reference has no counterpart: it is a single synchronous process per case line.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "host_common.hpp"
#include "tuning.hpp"

namespace cz {

struct Request {
  int64_t ticket = 0;
  u128 lo = 0, hi = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;
  RowPlan plan;
  std::vector<uint64_t> row_est;
  uint64_t est_left = 0;  // estimated candidates of the rows not yet launched
  size_t next_row = 0;
  std::vector<DevRec> recs;
  uint64_t escalated = 0;
  uint64_t footprint = 0;
  bool admitted = false, done = false;
  // Rows [launch_row0, next_row) belong to the launch numbered launch_id; a launch that fails
  // as a whole gives them back (next_row = launch_row0), at most kMaxRetries times.
  uint64_t launch_id = 0;
  size_t launch_row0 = 0;
  uint64_t launch_est0 = 0;
  uint32_t retries = 0;
  bool requeued = false;  // its rows of the last launch are to run again
  // Waiters: a result is handed out once; the last waiter to leave frees the request.
  uint32_t waiters = 0;
  bool collected = false;
  int status = COLLATZ_OK;
  std::string error;
};

// Launches that fail as a whole (a HIP error) are retried this many times per request.
constexpr uint32_t kMaxRetries = 2;

}  // namespace cz

struct collatz_server {
  const collatz_sieve* sv = nullptr;
  collatz_server_opts opts{};
  cz::Engine* engine = nullptr;
  hipStream_t stream = nullptr;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv, done_cv;
  std::deque<cz::Request*> pending;
  std::vector<cz::Request*> active;
  std::map<int64_t, cz::Request*> all;
  int64_t next_ticket = 0;
  bool stop = false;   // destroy called: no more submissions, the worker drains and exits
  bool dead = false;   // the worker has exited: waiters of unfinished requests give up
  uint32_t nwaiters = 0;
  uint64_t launches = 0;
  double t0 = 0;
  uint64_t bytes_in_use = 0;
  // Tuner sweeps the launches asked for (buckets without a tuned shape), run while the
  // server is idle, one timing launch at a time: bucket -> (word width, largest end, state).
  struct WantTune {
    int limbs;
    cz::u128 max_hi;
    cz::TuneSweep st;
  };
  std::map<uint32_t, WantTune> want_tune;
};

namespace cz {

static void server_loop(collatz_server* S) {
  (void)hipSetDevice(S->sv->device);
  for (;;) {
    LaunchPlan lp;
    u128 max_hi = 0;  // largest end of the requests in this launch (sets the word width)
    struct SlotRef {
      Request* r;
      uint64_t chunk;
    };
    std::vector<SlotRef> slots;
    std::vector<Request*> in_launch;   // requests with rows in this launch
    std::vector<uint64_t> row_est;     // expected survivors of each row of the launch
    std::vector<uint8_t> slot_first;   // 1 where a row starts a new slot
    uint64_t launch_id = 0;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv.wait(lk, [&] {
        return S->stop || !S->pending.empty() || !S->active.empty() || !S->want_tune.empty();
      });
      if (S->stop && S->pending.empty() && S->active.empty()) return;
      if (S->pending.empty() && S->active.empty()) {
        // Idle: one timing launch of a wanted tuner sweep (a few ms), then look for work
        // again, so a request that arrives meanwhile waits for one timing launch at most.
        auto it = S->want_tune.begin();
        const uint32_t bucket = it->first;
        collatz_server::WantTune w = it->second;
        lk.unlock();
        const int st = tune_step(S->sv, w.limbs, w.max_hi, w.st);
        lk.lock();
        if (st > 0) S->want_tune.erase(bucket);
        else S->want_tune[bucket] = w;
        // Another caller is tuning this sieve right now: wait for work or a moment.
        if (st < 0)
          S->cv.wait_for(lk, std::chrono::milliseconds(2),
                         [&] { return S->stop || !S->pending.empty(); });
        continue;
      }
      // Admission: oldest first, while the footprint budget and the in-flight cap allow.
      while (!S->pending.empty() && S->active.size() < S->opts.max_inflight) {
        Request* r = S->pending.front();
        if (!S->active.empty() && S->bytes_in_use + r->footprint > S->opts.max_device_bytes) break;
        S->pending.pop_front();
        r->admitted = true;
        r->t_admit = now_ms() - S->t0;
        S->bytes_in_use += r->footprint;
        S->active.push_back(r);
      }
      // Continuous batching: every admitted request gets an equal share of the launch's work
      // budget, and at least one row, so a short request finishes within one launch however
      // long the requests ahead of it are; what the shares leave goes to the oldest.  The
      // budget is slice_candidates, raised up to 64x while the admitted backlog is large
      // (a quarter of it per launch), so a saturated server runs fewer, longer launches
      // (the whole request stream of seed 1 submitted at once: 5.8e13 n/s with at most 16x,
      // 6.1e13 with 64x; at 500 requests/s, where the backlog stays small, p50 and p99 are
      // the same).
      uint64_t backlog = 0;
      for (const Request* r : S->active) backlog += r->est_left;
      const uint64_t budget = std::min<uint64_t>(
          std::max<uint64_t>(S->opts.slice_candidates, backlog / 4), 64 * S->opts.slice_candidates);
      uint64_t taken = 0;
      launch_id = ++S->launches;
      auto take = [&](Request* r, uint64_t want, bool one_row_always) {
        uint64_t mine = 0, last_chunk = UINT64_MAX;
        if (r->next_row < r->plan.rows.size() && r->launch_id != launch_id) {
          r->launch_id = launch_id;
          r->launch_row0 = r->next_row;
          r->launch_est0 = r->est_left;
          in_launch.push_back(r);
        }
        while (r->next_row < r->plan.rows.size() &&
               (mine == 0 && one_row_always ? true : (mine < want && taken < budget))) {
          const size_t i = r->next_row;
          RowDesc row = r->plan.rows[i];
          // The row's slot is its chunk within the request; rows of one chunk are contiguous.
          const uint64_t chunk = row.slot;
          const bool first_of_slot = chunk != last_chunk;
          if (first_of_slot) {
            slots.push_back({r, chunk});
            last_chunk = chunk;
          }
          row.slot = (uint32_t)(slots.size() - 1);
          const uint32_t c = row.cls;
          if (r->plan.key_lo[i] < 0) {
            row.i_lo = S->sv->class_start[c];
          } else {
            row.i_lo = kResolveFlag | lp.keys.size();
            lp.keys.push_back(r->plan.keys[r->plan.key_lo[i]]);
          }
          if (r->plan.key_hi[i] < 0) {
            row.i_hi = S->sv->class_start[c + 1];
          } else {
            row.i_hi = kResolveFlag | lp.keys.size();
            lp.keys.push_back(r->plan.keys[r->plan.key_hi[i]]);
          }
          lp.rows.push_back(row);
          row_est.push_back(r->row_est[i]);
          slot_first.push_back(first_of_slot ? 1 : 0);
          if (r->hi > max_hi) max_hi = r->hi;
          mine += r->row_est[i] + 1;
          taken += r->row_est[i] + 1;
          r->est_left -= std::min(r->est_left, r->row_est[i]);
          r->next_row++;
        }
      };
      const uint64_t share = std::max<uint64_t>(budget / std::max<size_t>(S->active.size(), 1), 1);
      for (Request* r : S->active) take(r, share, true);
      for (Request* r : S->active) {
        if (taken >= budget) break;
        take(r, budget - taken, false);
      }
    }
    // The launch shape: the autotuned waves per CU and checks per refill of the sieve for
    // this word width and magnitude, and tiles for this launch's size (a small launch
    // spreads over all resident waves); then each row's pieces and each slot's first tile.
    // A bucket without a tuned shape runs the default one, and its sweep is queued for the
    // server's idle moments: a sweep run here would hold every request of the launch back
    // by some 80 timing launches.
    uint64_t est_total = 0;
    for (const uint64_t e : row_est) est_total += e;
    int width = launch_width(2, max_hi, S->opts.cutoff_log2);
    bool want = false;
    const Tuning tn = lp.rows.empty() ? Tuning{4096, 32, 2}
                                      : tuned_or_default(S->sv, width, max_hi, est_total, want);
    if (want) {
      std::lock_guard<std::mutex> lk(S->mu);
      S->want_tune.emplace(tune_bucket(width, max_hi),
                           collatz_server::WantTune{width, max_hi, TuneSweep{}});
    }
    for (size_t i = 0; i < lp.rows.size(); ++i) {
      if (slot_first[i]) lp.part.push_back(lp.ntiles);
      uint64_t pieces = (uint64_t)((double)row_est[i] / tn.tile + 0.5);
      pieces = std::min<uint64_t>(std::max<uint64_t>(pieces, 1), 1u << 20);
      lp.rows[i].block_start = lp.ntiles;
      lp.ntiles += pieces;
    }
    lp.nslots = slots.size();
    lp.part.push_back(lp.ntiles);
    LaunchResult res;
    int rc = COLLATZ_OK;
    std::string launch_error;
    if (!lp.rows.empty()) {
      rc = S->engine->run(lp, tn, width, S->opts.cutoff_log2, true, 0, S->stream, res);
      // More escalations than the list holds: the same launch one width up.
      while (rc == COLLATZ_ECAPACITY && width < 3)
        rc = S->engine->run(lp, tn, ++width, S->opts.cutoff_log2, true, 0, S->stream, res);
      if (rc != COLLATZ_OK) launch_error = collatz_last_error();
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      if (rc != COLLATZ_OK) {
        // The launch failed as a whole, so none of its records can be used: each request
        // with rows in it takes them again in a later launch, and fails after kMaxRetries.
        for (Request* r : in_launch) {
          if (++r->retries > kMaxRetries) {
            r->status = rc;
            r->error = launch_error;
            r->next_row = r->plan.rows.size();
          } else {
            r->next_row = r->launch_row0;
            r->est_left = r->launch_est0;
          }
        }
      } else {
        // A candidate abandoned at the check bound fails only the request that owns it; the
        // other slots of the launch hold complete records and are folded.  If more were
        // abandoned than the device listed, the requests none of the listed ones name run
        // their rows of this launch again instead (their candidates may be among the
        // unlisted ones).
        std::vector<uint8_t> bad(slots.size(), 0);
        for (uint32_t sl : res.stall_slots)
          if (sl < slots.size()) bad[sl] = 1;
        for (size_t i = 0; i < slots.size(); ++i) {
          Request* r = slots[i].r;
          if (bad[i] && r->status == COLLATZ_OK) {
            r->status = COLLATZ_EDEVICE;
            r->error = "a trajectory did not descend within the check bound (chunk " +
                       std::to_string(r->plan.first_chunk + slots[i].chunk) + ")";
            r->next_row = r->plan.rows.size();
          }
        }
        const bool unlisted = res.stalled > res.stall_slots.size();
        for (Request* r : in_launch) {
          r->requeued = false;
          if (unlisted && r->status == COLLATZ_OK) {
            if (++r->retries > kMaxRetries) {
              r->status = COLLATZ_EDEVICE;
              r->error = "a trajectory did not descend within the check bound";
              r->next_row = r->plan.rows.size();
            } else {
              r->next_row = r->launch_row0;
              r->est_left = r->launch_est0;
              r->requeued = true;
            }
          }
        }
        for (size_t i = 0; i < slots.size(); ++i) {
          Request* r = slots[i].r;
          if (r->status == COLLATZ_OK && !r->requeued)
            rec_merge_host(r->recs[slots[i].chunk], res.recs[i]);
        }
        for (uint32_t sl : res.esc_slots)
          if (sl < slots.size()) slots[sl].r->escalated++;
      }
      for (auto it = S->active.begin(); it != S->active.end();) {
        Request* r = *it;
        if (r->next_row >= r->plan.rows.size()) {
          r->done = true;
          r->t_done = now_ms() - S->t0;
          S->bytes_in_use -= r->footprint;
          it = S->active.erase(it);
        } else {
          ++it;
        }
      }
    }
    S->done_cv.notify_all();
  }
}

}  // namespace cz

using namespace cz;

extern "C" {

int collatz_server_create(const collatz_sieve* sv, const collatz_server_opts* opts,
                          collatz_server** out) {
  if (!sv || !out) return fail(COLLATZ_EINVAL, "null argument");
  *out = nullptr;
  collatz_server* S = new (std::nothrow) collatz_server();
  if (!S) return fail(COLLATZ_ENOMEM, "out of host memory");
  S->sv = sv;
  if (opts) S->opts = *opts;
  if (!S->opts.chunk_log2) S->opts.chunk_log2 = 24;
  if (!S->opts.cutoff_log2) S->opts.cutoff_log2 = 187;
  if (!S->opts.max_device_bytes) S->opts.max_device_bytes = 1ull << 30;
  if (!S->opts.slice_candidates) S->opts.slice_candidates = 1ull << 26;
  if (!S->opts.max_inflight) S->opts.max_inflight = 1024;
  if (S->opts.chunk_log2 > 63 || S->opts.cutoff_log2 < 64 || S->opts.cutoff_log2 > 187) {
    delete S;
    return fail(COLLATZ_EINVAL, "bad server options");
  }
  HipDeviceGuard guard(sv->device);
  if (hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess) {
    delete S;
    return fail(COLLATZ_EDEVICE, "stream creation failed");
  }
  S->engine = new (std::nothrow) Engine(sv);
  if (!S->engine) {
    (void)hipStreamDestroy(S->stream);
    delete S;
    return fail(COLLATZ_ENOMEM, "out of host memory");
  }
  S->t0 = now_ms();
  S->worker = std::thread(server_loop, S);
  *out = S;
  return COLLATZ_OK;
}

int64_t collatz_server_submit(collatz_server* S, const collatz_interval* range) {
  if (!S || !range) return fail(COLLATZ_EINVAL, "null argument");
  Request* r = new (std::nothrow) Request();
  if (!r) return fail(COLLATZ_ENOMEM, "out of host memory");
  r->t_submit = now_ms() - S->t0;
  r->lo = mk128(range->lo[0], range->lo[1]);
  r->hi = mk128(range->hi[0], range->hi[1]);
  if (r->lo >= r->hi) {
    delete r;
    return fail(COLLATZ_EINVAL, "empty interval (lo >= hi)");
  }
  int rc = plan_rows(S->sv->k, r->lo, r->hi, S->opts.chunk_log2, 0, S->sv->mod9 != 0, r->plan,
                     1ull << 26);
  if (rc) {
    delete r;
    return rc;
  }
  // Expected survivors per row, from the row's residue window and its class size.
  const double kspan = (double)((u128)1 << S->sv->k);
  r->row_est.resize(r->plan.rows.size());
  for (size_t i = 0; i < r->plan.rows.size(); ++i) {
    const uint32_t c = r->plan.rows[i].cls;
    const double a = r->plan.key_lo[i] < 0 ? 0.0 : (double)r->plan.keys[r->plan.key_lo[i]];
    const double b = r->plan.key_hi[i] < 0 ? kspan : (double)r->plan.keys[r->plan.key_hi[i]];
    r->row_est[i] =
        (uint64_t)((double)(S->sv->class_start[c + 1] - S->sv->class_start[c]) * (b - a) / kspan);
    r->est_left += r->row_est[i];
  }
  r->recs.assign(r->plan.nchunks, DevRec{});
  for (DevRec& d : r->recs) std::memset(&d, 0, sizeof(d));
  r->footprint = r->plan.rows.size() * (sizeof(RowDesc) + 16) + r->plan.nchunks * sizeof(DevRec);
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->stop) {
    delete r;
    return fail(COLLATZ_EINVAL, "server is shutting down");
  }
  r->ticket = S->next_ticket++;
  S->all[r->ticket] = r;
  S->pending.push_back(r);
  S->cv.notify_one();
  return r->ticket;
}

int collatz_server_wait(collatz_server* S, int64_t ticket, double timeout_ms,
                        collatz_request_result* out, collatz_chunk* chunks, uint64_t cap) {
  if (!S || !out) return fail(COLLATZ_EINVAL, "null argument");
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->all.find(ticket);
  if (it == S->all.end()) return fail(COLLATZ_EINVAL, "unknown ticket");
  Request* r = it->second;
  if (r->collected) return fail(COLLATZ_EINVAL, "ticket already collected");
  // Several threads may wait on one ticket: the first to wake takes the result, the others
  // get EINVAL, and the last one to leave frees the request.  collatz_server_destroy waits
  // for every waiter to leave before it frees anything.
  r->waiters++;
  S->nwaiters++;
  auto ready = [&] { return r->done || r->collected || S->dead; };
  bool woke = true;
  if (timeout_ms < 0) S->done_cv.wait(lk, ready);
  else woke = S->done_cv.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready);
  r->waiters--;
  S->nwaiters--;
  if (S->nwaiters == 0) S->done_cv.notify_all();  // (destroy may be waiting for this)
  int rc = COLLATZ_OK;
  if (r->collected) {
    rc = fail(COLLATZ_EINVAL, "ticket already collected");
  } else if (!r->done) {
    rc = woke ? fail(COLLATZ_EINVAL, "server destroyed before the request finished")
              : fail(COLLATZ_ENOTREADY, "request not finished");
  } else {
    std::memset(out, 0, sizeof(*out));
    out->chunks = r->plan.nchunks;
    out->submit_ms = r->t_submit;
    out->admit_ms = r->t_admit;
    out->done_ms = r->t_done;
    for (uint64_t c = 0; c < r->plan.nchunks; ++c) {
      collatz_chunk ch;
      chunk_from_rec(r->recs[c], r->plan.first_chunk + c, r->plan.n_checked[c], &ch);
      if (chunks && c < cap) chunks[c] = ch;
      chunk_merge(&out->total, ch);
    }
    out->total.chunk_index = r->plan.first_chunk;
    r->collected = true;
    if (r->status != COLLATZ_OK) rc = fail(r->status, r->error);
    else if (chunks && out->chunks > cap) rc = fail(COLLATZ_ECAPACITY, "chunk buffer too small");
    S->done_cv.notify_all();  // other waiters of this ticket see it collected
  }
  if (r->collected && r->waiters == 0) {
    S->all.erase(r->ticket);
    delete r;
  }
  return rc;
}

int collatz_server_destroy(collatz_server* S) {
  if (!S) return COLLATZ_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv.notify_all();
  // The worker finishes the admitted and pending requests, then exits.
  if (S->worker.joinable()) S->worker.join();
  {
    std::unique_lock<std::mutex> lk(S->mu);
    S->dead = true;
    S->done_cv.notify_all();
    S->done_cv.wait(lk, [&] { return S->nwaiters == 0; });
  }
  HipDeviceGuard guard(S->sv->device);
  delete S->engine;
  if (S->stream) (void)hipStreamDestroy(S->stream);
  for (auto& kv : S->all) delete kv.second;
  delete S;
  return COLLATZ_OK;
}

}  // extern "C"
