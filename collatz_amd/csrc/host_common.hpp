// Host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/collatz.h"
#include "internal.hpp"

namespace cz {

using u128 = unsigned __int128;

void set_error(const std::string& msg);

// Host-side span timing, printed to stderr when COLLATZ_TRACE=1.
bool trace_on();
struct Span {
  const char* name;
  double t0;
  explicit Span(const char* n);
  ~Span();
};
double now_ms();
int fail(int code, const std::string& msg);

inline u128 mk128(uint64_t lo, uint64_t hi) { return ((u128)hi << 64) | lo; }

// The stream a call runs on: the caller's, or for NULL the calling thread's per-thread
// default stream, which orders itself after the caller's earlier default-stream work.  The
// library enqueues nothing on the legacy default stream and makes no synchronous copy (which
// would run there): the runtime refuses a legacy-stream operation while another thread
// captures a hipGraph (a server's engine does), with hipErrorStreamCaptureImplicit.
inline hipStream_t call_stream(void* s) { return s ? (hipStream_t)s : hipStreamPerThread; }
inline uint64_t lo64(u128 x) { return (uint64_t)x; }
inline uint64_t hi64(u128 x) { return (uint64_t)(x >> 64); }

#define CZ_HIP(expr)                                                                  \
  do {                                                                                \
    hipError_t e__ = (expr);                                                          \
    if (e__ != hipSuccess)                                                            \
      return ::cz::fail(COLLATZ_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e__)); \
  } while (0)

// Owning device buffer.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { reset(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { reset(); }
  hipError_t alloc(size_t count) {
    reset();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e != hipSuccess) { p = nullptr; return e; }
    n = count;
    return hipSuccess;
  }
  // Grows (never shrinks) a scratch buffer.
  hipError_t reserve(size_t count) {
    if (count <= n && p) return hipSuccess;
    return alloc(count);
  }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Host-side merge with the device semantics (larger maxnum wins, ties to the smaller start).
inline int cmp3h(const uint64_t* a, const uint64_t* b) {
  for (int i = 2; i >= 0; --i)
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  return 0;
}

inline void rec_merge_host(DevRec& a, const DevRec& b) {
  if (b.count) {
    const int c = cmp3h(b.maxnum, a.maxnum);
    const bool smaller =
        mk128(b.max_start[0], b.max_start[1]) < mk128(a.max_start[0], a.max_start[1]);
    if (a.count == 0 || c > 0 || (c == 0 && smaller)) {
      std::memcpy(a.maxnum, b.maxnum, sizeof(a.maxnum));
      std::memcpy(a.max_start, b.max_start, sizeof(a.max_start));
    }
  }
  a.count += b.count;
  a.stop_sum += b.stop_sum;
  a.stop_xor ^= b.stop_xor;
  a.overflow += b.overflow;
  a.checks += b.checks;
  a.hash ^= b.hash;
}

void chunk_from_rec(const DevRec& r, uint64_t chunk_index, uint64_t n_checked,
                    collatz_chunk* out);
void chunk_merge(collatz_chunk* into, const collatz_chunk& from);

// Decomposition of [lo, hi) into rows (one q, one chunk, a residue window).
struct RowPlan {
  std::vector<RowDesc> rows;
  std::vector<uint64_t> keys;        // residue bounds to look up
  std::vector<uint32_t> key_cls;     // class of each key
  std::vector<int64_t> key_lo, key_hi;  // per row: index into keys, or -1 for 0 / S
  std::vector<uint64_t> chunk_first_row;  // per chunk: first row; size nchunks + 1
  std::vector<uint64_t> n_checked;  // per chunk
  uint64_t first_chunk = 0;
  uint64_t nchunks = 0;
  // Empties the plan and keeps its vectors' capacity (a plan kept between calls reallocates
  // nothing once it has held the largest).
  void clear() {
    rows.clear();
    keys.clear();
    key_cls.clear();
    key_lo.clear();
    key_hi.clear();
    chunk_first_row.clear();
    n_checked.clear();
    first_chunk = nchunks = 0;
  }
};

int plan_rows(uint32_t k, u128 lo, u128 hi, uint32_t C, uint32_t slot_base, bool mod9,
              RowPlan& plan, uint64_t max_rows);

// Device buffers and events of the plain form, grown on demand and kept for later calls
// (a call then makes no device allocation unless it needs more room than any before it).
struct PlainScratch {
  DevBuf<uint64_t> keys, key_idx, part, ovf;
  DevBuf<uint32_t> key_cls;
  DevBuf<RowDesc> rows;
  DevBuf<DevRec> block, chunk, esc_recs;
  DevBuf<EscEntry> esc;
  DevBuf<unsigned long long> cnt;
  DevBuf<uint32_t> stall;
  hipEvent_t ev[2] = {nullptr, nullptr};
  RowPlan plan;  // host-side row plan, kept for its capacity
  PlainScratch() = default;
  PlainScratch(const PlainScratch&) = delete;
  PlainScratch& operator=(const PlainScratch&) = delete;
  ~PlainScratch() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

// Options of one plain-form call (collatz_verify_opts with defaults applied).
struct RangeArgs {
  uint32_t C = 24, limbs = 2, cutoff = 187;
  uint32_t width = 2;  // word width of the kernels (launch_width)
};
int range_args(const collatz_verify_opts* o, RangeArgs& a);

// Device memory of one plain-form call, as handed out by a PlainBuffers provider.
struct PlainView {
  uint64_t* keys = nullptr;      // row bounds to look up
  uint32_t* key_cls = nullptr;
  uint64_t* key_idx = nullptr;   // their indices in the sieve
  RowDesc* rows = nullptr;
  uint64_t* part = nullptr;      // first share of each chunk (nchunks + 1)
  DevRec* block = nullptr;       // one record per share
  DevRec* chunk = nullptr;       // one record per chunk
  EscEntry* esc = nullptr;       // escalation list
  uint64_t esc_cap = 0;
  DevRec* esc_recs = nullptr;    // records of the three-limb pass
  uint64_t* ovf = nullptr;       // overflow reports, 4 words each
  uint64_t ovf_cap = 0;
  unsigned long long* cnt = nullptr;  // escalations, overflow reports, stalls, loop aborts
  uint32_t* stall = nullptr;     // slots of abandoned candidates (kStallCap)
  hipEvent_t ev[2] = {nullptr, nullptr};
};

// Where a plain-form call gets its device memory: the sieve's grow-on-demand buffers
// (collatz_verify_range) or a fixed static array (collatz_verify_range_noalloc).
class PlainBuffers {
 public:
  virtual ~PlainBuffers() = default;
  virtual uint64_t max_rows() const = 0;
  virtual uint64_t max_chunks() const = 0;
  // Share records a launch may use, at most (the plain kernel's shares follow from it).
  virtual uint64_t max_shares() const = 0;
  virtual int keys(uint64_t n, PlainView& v) = 0;
  // Everything a launch of these sizes needs; ovf_cap is the overflow reports wanted.
  virtual int launch(uint64_t nrows, uint64_t nshares, uint64_t nchunks, uint64_t ovf_cap,
                     bool esc, PlainView& v) = 0;
  virtual int esc_recs(uint64_t n, PlainView& v) = 0;
  // A row plan to fill (cleared by the caller); the default is the provider's own.
  virtual RowPlan& plan() { return plan_; }

 private:
  RowPlan plan_;
};

// The plain form over [lo, hi) with the given buffers (the caller serialises calls that
// share them).
int verify_plain_with(const struct ::collatz_sieve* sv, u128 lo, u128 hi, const RangeArgs& a,
                      hipStream_t s, collatz_range_result* out, PlainBuffers& B);

Limits make_limits(int limbs, uint32_t cutoff);
// Word width of a launch for a two- or three-limb request: two-limb work whose starts are
// below 2^68 and whose cutoff is at least 2^91 runs on 96 bits (width 1); two-limb work
// with starts above 2^100 on six words (width 3); otherwise the width is the limb count.
// A launch whose escalations overflow the list is rerun one width up.
int launch_width(int limbs, u128 max_hi, uint32_t cutoff);

// Leaves of the search1 tree below a root at root.depth, sorted by residue: the device
// sieve (sieve_kernels.hip) behind collatz_sieve_build and reference mode.
struct LeafSet {
  DevBuf<uint64_t> res, num;
  DevBuf<uint64_t> num_hi;  // high words of num (a one-entry placeholder unless wide)
  bool wide = false;        // depth above 40: numbers take 128 bits
  DevBuf<uint32_t> meta;
  DevBuf<uint32_t> bits;   // dense bit array over the extensions h (padded)
  uint64_t bits_bits = 0;
  uint64_t count = 0;
  // Build scratch (prefix states, alive bits, per-block counts and offsets).  Every buffer
  // of a LeafSet only grows, so a LeafSet reused for many builds allocates nothing once it
  // has held the largest; sizes are the build's, not the buffers'.
  DevBuf<uint64_t> pnum, pnum_hi, alive, offs;
  DevBuf<uint32_t> pmeta, counts;
};
int build_leaves(const SieveRoot& root, hipStream_t s, LeafSet& out);

struct HipDeviceGuard {
  int prev = -1;
  bool changed = false;
  explicit HipDeviceGuard(int dev) {
    if (dev >= 0 && hipGetDevice(&prev) == hipSuccess && prev != dev) {
      changed = hipSetDevice(dev) == hipSuccess;
    }
  }
  ~HipDeviceGuard() {
    if (changed) (void)hipSetDevice(prev);
  }
};

class Engine;
}  // namespace cz

struct collatz_sieve {
  int device = 0;
  uint32_t k = 0, mod9 = 1;
  uint64_t count = 0;
  cz::DevBuf<uint4> rec;               // packed survivor records (SieveArrays)
  cz::DevBuf<uint32_t> bits;
  uint64_t bits_bits = 0;
  cz::DevBuf<uint64_t> pow3;
  uint64_t class_start[10] = {0};       // survivors are stored class-major by r mod 9
  cz::DevBuf<uint64_t> class_start_dev;
  uint32_t la_bits = 0, la_sem = 0;     // look-ahead filter (0 bits: none)
  cz::DevBuf<uint32_t> la;              // its table: 2^B nibbles
  cz::DevBuf<uint32_t> la1;             // first-level table at 16 bits (B > 16, CUDA compares)
  double build_ms = 0;
  // Buffers and captured graph of the persistent form, created on first use; calls that
  // share a sieve handle are serialised on engine_mu.
  mutable cz::Engine* engine = nullptr;
  mutable std::mutex engine_mu;
  // Buffers of the plain form; plain calls that share a sieve handle are serialised on
  // plain_mu.
  mutable cz::PlainScratch plain;
  mutable std::mutex plain_mu;
  // Launch shapes the autotuner chose for this sieve, per (word width, magnitude bucket):
  // {waves per CU, checks per refill}.  The sieve's depth, mod 9 and look-ahead are fixed
  // for the handle, so they need no place in the key.  Tuning of one sieve holds tune_mu.
  mutable std::map<uint32_t, std::array<uint32_t, 2>> tuned;
  mutable std::mutex tune_mu;
  cz::SieveArrays arrays() const {
    cz::SieveArrays a;
    a.rec = rec.p;
    a.count = count;
    a.k = k;
    a.la = la_bits ? la.p : nullptr;
    a.la_mask = la_bits ? (1u << la_bits) - 1u : 0u;
    a.la_emin = (306u * k + 484u) / 485u;
    a.la1 = la_bits && la1.p ? la1.p : nullptr;
    a.la1_mask = a.la1 ? (1u << cz::kLaLevel1Bits) - 1u : 0u;
    return a;
  }
};
