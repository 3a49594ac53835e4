// C ABI: collatz_verify_range_noalloc, the plain form without device allocation.
//
// For callers that cannot allocate device memory per call (tight serving loops, harnesses
// that forbid hipMalloc): the same computation as collatz_verify_range's plain form, with
// every buffer the kernels need (row descriptors, bound lookups, share and chunk records,
// escalation and overflow lists, counters) carved out of one file-scope device array of
// fixed size.  Each device holds its own copy of the array (about 20 MiB) from the moment
// the library is loaded until the process ends.  Host code never uses the array's name as
// an address: the device address comes from hipGetSymbolAddress, once per device, with that
// device current, and every kernel argument, copy and memset below is derived from it.
//
// Calls on one device share the array; this file serialises them with a lock per device,
// and each call clears its counters before its kernels run.  Work that another caller
// enqueues on the array's memory by other means would race with a call.
#include <algorithm>
#include <mutex>

#include "engine.hpp"
#include "host_common.hpp"

namespace cz {
namespace {

// Limits of one call.
constexpr uint64_t kNaRows = 32768;               // row descriptors (q, chunk, residue class)
constexpr uint64_t kNaKeys = 2 * kNaRows;         // residue bounds to look up
constexpr uint64_t kNaChunks = 4096;              // chunk records
// Share records (one per wave of the plain kernel): a launch sizes its shares to fit them.
constexpr uint64_t kNaShares = 98304;
static_assert(kNaShares > kNaRows, "share records too few");
constexpr uint64_t kNaEsc = 1ull << 16;           // escalated candidates (more: one width up)
constexpr uint64_t kNaOvf = 4096;                 // overflow starts listed

constexpr uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

// Byte offsets of the buffers in the array.
constexpr uint64_t kOffKeys = 0;
constexpr uint64_t kOffKeyCls = align256(kOffKeys + kNaKeys * 8);
constexpr uint64_t kOffKeyIdx = align256(kOffKeyCls + kNaKeys * 4);
constexpr uint64_t kOffRows = align256(kOffKeyIdx + kNaKeys * 8);
constexpr uint64_t kOffPart = align256(kOffRows + kNaRows * sizeof(RowDesc));
constexpr uint64_t kOffBlock = align256(kOffPart + (kNaChunks + 1) * 8);
constexpr uint64_t kOffChunk = align256(kOffBlock + kNaShares * sizeof(DevRec));
constexpr uint64_t kOffEsc = align256(kOffChunk + kNaChunks * sizeof(DevRec));
constexpr uint64_t kOffEscRecs = align256(kOffEsc + kNaEsc * sizeof(EscEntry));
constexpr uint64_t kOffOvf = align256(kOffEscRecs + kNaEsc * sizeof(DevRec));
constexpr uint64_t kOffCnt = align256(kOffOvf + kNaOvf * 32);
constexpr uint64_t kOffStall = align256(kOffCnt + 4 * 8);
constexpr uint64_t kScratchBytes = align256(kOffStall + kStallCap * 4);
static_assert(kScratchBytes < (24ull << 20), "scratch array larger than intended");

}  // namespace

// The array itself.  Its name is a host-side stand-in in host code, not a device address.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;

struct DeviceScratch {
  std::mutex mu;          // held for a whole call
  unsigned char* ptr = nullptr;  // device address of `scratch` on this device
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool ready = false;
};

DeviceScratch g_dev[kMaxDevices];

// Set-up on the first call on a device (the device is current): the array's address there
// and the two timing events.  The caller holds d.mu.
int setup(DeviceScratch& d) {
  if (d.ready) return COLLATZ_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(COLLATZ_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                     hipGetErrorString(e));
  d.ptr = static_cast<unsigned char*>(p);
  for (hipEvent_t& ev : d.ev)
    if (!ev) CZ_HIP(hipEventCreate(&ev));
  d.ready = true;
  return COLLATZ_OK;
}

class StaticBuffers final : public PlainBuffers {
 public:
  explicit StaticBuffers(const DeviceScratch& d) : d_(d) {}
  uint64_t max_rows() const override { return kNaRows; }
  uint64_t max_chunks() const override { return kNaChunks; }
  uint64_t max_shares() const override { return kNaShares; }
  int keys(uint64_t n, PlainView& v) override {
    if (n > kNaKeys) return fail(COLLATZ_EINVAL, "noalloc: too many row bounds");
    v.keys = at<uint64_t>(kOffKeys);
    v.key_cls = at<uint32_t>(kOffKeyCls);
    v.key_idx = at<uint64_t>(kOffKeyIdx);
    return COLLATZ_OK;
  }
  int launch(uint64_t nrows, uint64_t nshares, uint64_t nchunks, uint64_t, bool esc,
             PlainView& v) override {
    if (nrows > kNaRows || nchunks > kNaChunks)
      return fail(COLLATZ_EINVAL, "noalloc: range has more rows or chunks than the limit");
    if (nshares > kNaShares)
      return fail(COLLATZ_EINVAL, "noalloc: range needs more shares than the limit (" +
                                      std::to_string(kNaShares) + ")");
    v.rows = at<RowDesc>(kOffRows);
    v.part = at<uint64_t>(kOffPart);
    v.block = at<DevRec>(kOffBlock);
    v.chunk = at<DevRec>(kOffChunk);
    v.esc = at<EscEntry>(kOffEsc);
    v.esc_cap = esc ? kNaEsc : 0;
    v.ovf = at<uint64_t>(kOffOvf);
    v.ovf_cap = kNaOvf;
    v.cnt = at<unsigned long long>(kOffCnt);
    v.stall = at<uint32_t>(kOffStall);
    v.ev[0] = d_.ev[0];
    v.ev[1] = d_.ev[1];
    return COLLATZ_OK;
  }
  int esc_recs(uint64_t n, PlainView& v) override {
    if (n > kNaEsc) return fail(COLLATZ_EINVAL, "noalloc: too many escalated candidates");
    v.esc_recs = at<DevRec>(kOffEscRecs);
    return COLLATZ_OK;
  }

 private:
  template <typename T>
  T* at(uint64_t off) const {
    return reinterpret_cast<T*>(d_.ptr + off);
  }
  const DeviceScratch& d_;
};

}  // namespace
}  // namespace cz

using namespace cz;

extern "C" int collatz_verify_range_noalloc(const collatz_sieve* sv, const uint64_t lo[2],
                                            const uint64_t hi[2],
                                            const collatz_verify_opts* opts, void* stream,
                                            collatz_range_result* out) {
  if (!sv || !lo || !hi || !out) return fail(COLLATZ_EINVAL, "null argument");
  RangeArgs a;
  int rc = range_args(opts, a);
  if (rc) return rc;
  if (opts && opts->form != COLLATZ_FORM_PLAIN)
    return fail(COLLATZ_EINVAL, "noalloc: only the plain form");
  const u128 L = mk128(lo[0], lo[1]), H = mk128(hi[0], hi[1]);
  if (L >= H) return fail(COLLATZ_EINVAL, "empty range (lo >= hi)");
  if (sv->device < 0 || sv->device >= kMaxDevices)
    return fail(COLLATZ_EINVAL, "noalloc: device ordinal out of range");
  HipDeviceGuard guard(sv->device);
  DeviceScratch& d = g_dev[sv->device];
  std::lock_guard<std::mutex> lk(d.mu);
  rc = setup(d);
  if (rc) return rc;
  a.width = (uint32_t)launch_width((int)a.limbs, H, a.cutoff);
  StaticBuffers B(d);
  return verify_plain_with(sv, L, H, a, call_stream(stream), out, B);
}
