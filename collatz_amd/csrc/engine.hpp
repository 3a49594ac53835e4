// The persistent-launch engine shared by the batched form (collatz_verify_batch) and the
// served form (collatz_server_*): reusable device buffers and pinned staging, and one launch
// sequence (counter reset, row resolution, persistent kernel, per-slot reduction) that can
// be captured once in a hipGraph and replayed for launches of any size.
#pragma once

#include <mutex>
#include <vector>

#include "host_common.hpp"
#include "tuning.hpp"

namespace cz {

// Pinned host buffer (hipHostMalloc), grow-only.
template <typename T>
struct HostBuf {
  T* p = nullptr;
  size_t n = 0;
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  ~HostBuf() { reset(); }
  hipError_t reserve(size_t count) {
    if (count <= n && p) return hipSuccess;
    reset();
    if (count == 0) count = 1;
    hipError_t e = hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault);
    if (e != hipSuccess) { p = nullptr; return e; }
    n = count;
    return hipSuccess;
  }
  void reset() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
};

// Work of one launch, as planned on the host.  Each row carries in block_start the index
// of its first tile; the tiles themselves are written on the device.
struct LaunchPlan {
  std::vector<RowDesc> rows;       // bounds flagged for device resolution where needed
  std::vector<uint64_t> keys;      // residue bounds referenced by flagged rows
  std::vector<uint64_t> part;      // per slot: first tile; nslots + 1 entries
  uint64_t ntiles = 0;
  uint64_t nslots = 0;
  uint64_t candidates_est = 0;
  void clear() {
    rows.clear();
    keys.clear();
    part.clear();
    ntiles = nslots = candidates_est = 0;
  }
};

struct LaunchResult {
  std::vector<DevRec> recs;        // per slot
  std::vector<uint64_t> ovf;       // overflow reports, 4 words each
  uint64_t ovf_total = 0;          // reports made (may exceed the stored ones)
  uint64_t escalated = 0;
  std::vector<uint32_t> esc_slots; // slot of each escalated candidate
  uint64_t stalled = 0;            // candidates abandoned at the check bound
  uint64_t loop_aborts = 0;        // tiles abandoned at the work-loop bound (records incomplete)
  std::vector<uint32_t> stall_slots;  // their slots (the first kStallCap of them)
  double device_ms = 0;
};

// The order in which a launch's tiles are taken (tile indices, lp.ntiles of them): rows that
// read the same survivors (same residue class and residue bounds, so they differ only in q)
// form a group, and a group's tiles are listed piece by piece across `band` of its rows at a
// time, so that the waves that take consecutive tiles read the same records.  Returns the
// number of tiles listed (lp.ntiles, each once).
uint64_t tile_order(const LaunchPlan& lp, uint32_t* order, uint32_t band);

// Appends the rows of [lo, hi) (chunk slots slot_base..) and their tiles to the plan.
int plan_interval(const collatz_sieve* sv, u128 lo, u128 hi, uint32_t C, uint64_t slot_base,
                  uint32_t tile, LaunchPlan& lp, std::vector<uint64_t>& n_checked,
                  uint64_t& first_chunk, uint64_t& nchunks);

class Engine {
 public:
  explicit Engine(const collatz_sieve* sv) : sv_(sv) {}
  ~Engine();
  // Runs one persistent launch over the plan (rows/keys/tiles/part) on stream s.  Candidates
  // abandoned at the check bound do not fail the launch: they are reported in out.stalled
  // and out.stall_slots, and the records of the other slots stay valid.
  int run(const LaunchPlan& lp, const Tuning& tn, int limbs, uint32_t cutoff, bool graph,
          uint64_t ovf_cap, hipStream_t s, LaunchResult& out);
  std::mutex mu;

 private:
  int ensure(const LaunchPlan& lp, uint64_t ovf_cap);
  const collatz_sieve* sv_;
  DevBuf<RowDesc> rows_;
  DevBuf<uint64_t> keys_, part_, ovf_;
  DevBuf<TileDesc> tiles_;
  DevBuf<uint32_t> order_;        // tile indices in the order waves take them
  DevBuf<DevRec> tile_recs_, slot_recs_, esc_recs_;
  DevBuf<EscEntry> esc_;
  DevBuf<unsigned long long> cnt_;
  DevBuf<uint32_t> stall_;
  DevBuf<Ctl> ctl_;
  HostBuf<DevRec> h_recs_;
  HostBuf<RowDesc> h_rows_;
  HostBuf<uint64_t> h_keys_, h_part_;
  HostBuf<uint32_t> h_order_;
  HostBuf<unsigned long long> h_cnt_;
  HostBuf<Ctl> h_ctl_;
  hipEvent_t e0_ = nullptr, e1_ = nullptr, e2_ = nullptr, e3_ = nullptr;
  // Captured launch sequence and what it was captured for.
  hipGraphExec_t gexec_ = nullptr;
  hipGraph_t graph_ = nullptr;
  struct GraphKey {
    const void* bufs[9];
    int limbs;
    uint32_t cutoff, nblocks, U, max_steps, max_loop;
    bool lowkey;
    bool operator==(const GraphKey& o) const;
  } gkey_{};
  hipError_t enqueue(const GraphKey& k, const Limits& lim, DeviceLists lists, hipStream_t s);
};

// Escalated candidates of a launch on three limbs, folded into recs by slot.
// `out` is the caller's scratch for the pass's records (grown as needed).
int run_escalated(const collatz_sieve* sv, const DevBuf<EscEntry>& esc, uint64_t n,
                  uint32_t cutoff, DeviceLists lists, std::vector<DevRec>& recs,
                  uint32_t slot_base, DevBuf<DevRec>& out, hipStream_t s);
// The same with the pass's records in out[0, n) (the caller provides room for n).
int run_escalated_into(const collatz_sieve* sv, const EscEntry* esc, uint64_t n,
                       uint32_t cutoff, DeviceLists lists, std::vector<DevRec>& recs,
                       uint32_t slot_base, DevRec* out, hipStream_t s);

int device_cu_count(int device);

}  // namespace cz
