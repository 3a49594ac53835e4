// C ABI: reference mode -- the reference program itself, per case line, on the device.
//
// Replaces main() / search1 / launchkernel / compute_kernel of collatz_cudav7.cu:95-411 and
// its OpenCL twin (collatz_openclv2.cpp:91-593): the leaves of the case line's subtree at
// depth N - A come from the device sieve (search1 evaluated per residue), the candidates of
// each leaf from its BV_i window and mod 9 row (Step 2), and the complete hash from the
// schedule (Step 3).  The launch count it reports is the one the reference would make:
// one per full batch of 2^16 leaves of a bin plus one flush per non-empty bin
// (collatz_cudav7.cu:532-533, :177-179).
#include <algorithm>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "engine.hpp"
#include "host_common.hpp"

namespace cz {




struct RefTotalsHost {
  unsigned long long count;
  unsigned long long overflow;
  unsigned long long stall;
  unsigned int hash;
  unsigned int tile_counter;
  unsigned long long esc_count;
  unsigned long long checks;
  unsigned long long aborts;
  unsigned long long numcases;
  unsigned long long spot_rec;
  unsigned int spot_mask, spot_off, spot_last, spot_shift;
};

struct RefTables {
  DevBuf<uint32_t> bv, mod9, magic;
  DevBuf<uint64_t> pow3w;  // 3^f for f < 64 as (low, high) word pairs
  RefTablesDev dev;
};

static std::mutex g_ref_mu;
static std::map<std::tuple<int, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>,
                std::unique_ptr<RefTables>>
    g_ref_cache;

// Tables alone can be built for deeper runs (N - A up to 120, e.g. main.rs's N = 72); a run
// needs N - A <= 62 and N <= 64 (every start n = base + a 2^(N - A) in one u64; leaf numbers
// take 128 bits beyond depth 40).
static int check_params(const collatz_ref_params* p, bool run) {
  if (!p) return fail(COLLATZ_EINVAL, "null params");
  if (p->A < 1 || p->A > 12) return fail(COLLATZ_EINVAL, "A (LASTBITS) must be in [1, 12]");
  const uint32_t max_mbi = run ? kWalkMaxDepth : 120;
  if (p->N < p->A + 4 || p->N - p->A > max_mbi || (run && p->N > 64))
    return fail(COLLATZ_EINVAL,
                "N - A out of range (4..62 with N <= 64 for a run, 4..120 for tables)");
  // main.rs's blocks multiply by up to 2^9.4 between tests, which six words hold below 2^182;
  // its test is a plain compare after each block, so it also takes cutoffs below 2^64.
  const bool rust = p->semantics == COLLATZ_SEM_RUST;
  if (rust ? (p->cutoff_log2 < 32 || p->cutoff_log2 > 180)
           : (p->cutoff_log2 < 64 || p->cutoff_log2 > 187))
    return fail(COLLATZ_EINVAL, rust ? "cutoff_log2 must be in [32, 180] under main.rs semantics"
                                     : "cutoff_log2 must be in [64, 187]");
  if (p->B < 8 || p->B > 28) return fail(COLLATZ_EINVAL, "B must be in [8, 28]");
  // A leaf's 2^A-bit candidate window is read from its bit vector with the vector's first
  // words mirrored after its end (collatz_cudav7.cu:475-477), which wraps a window correctly
  // only if the window is no longer than the vector.
  if (p->A > p->B) return fail(COLLATZ_EINVAL, "A (LASTBITS) must not exceed B");
  if (p->num_bv < 1 || p->num_bv > 32) return fail(COLLATZ_EINVAL, "num_bv must be in [1, 32]");
  if (p->semantics > 2) return fail(COLLATZ_EINVAL, "unknown semantics");
  return COLLATZ_OK;
}

static uint32_t tpo_of(uint32_t mbi) { return mbi * kNumerator / kDenominator + 1; }

static int get_tables(const collatz_ref_params* p, int device, RefTables*& out) {
  std::lock_guard<std::mutex> lk(g_ref_mu);
  const auto key = std::make_tuple(device, p->N, p->A, p->B, p->num_bv, p->semantics);
  auto it = g_ref_cache.find(key);
  if (it != g_ref_cache.end()) {
    out = it->second.get();
    return COLLATZ_OK;
  }
  auto t = std::make_unique<RefTables>();
  const uint32_t mbi = p->N - p->A, tpo = tpo_of(mbi);
  const uint32_t nw = std::max<uint32_t>(1, (1u << p->A) / 32);
  const uint64_t wpb = (1ull << (p->B - 5)) + nw;
  // 3^{-(tpo + i)} mod 2^32 (collatz_cudav7.cu:122-128).
  std::vector<uint32_t> magic(p->num_bv);
  for (uint32_t i = 0; i < p->num_bv; ++i) {
    uint32_t m = 1;
    for (uint32_t e = 0; e < tpo + i; ++e) m *= 2863311531u;
    magic[i] = m;
  }
  // cutoff_0 = 485 tpo - 306 mbi (collatz_cudav7.cu:424, held in an i16 there).
  const int32_t first = (int32_t)(int16_t)(kDenominator * tpo - kNumerator * mbi);
  // Mod 9 rows: bit i of row r set iff (r + i 2^mbi) mod 9 in {0,1,3,6,7} (:482-494).
  std::vector<uint32_t> mod9(9 * nw, 0);
  const uint32_t step = (uint32_t)(((u128)1 << mbi) % 9);
  for (uint32_t r = 0; r < 9; ++r) {
    uint32_t red = r;
    for (uint32_t i = 0; i < (1u << p->A); ++i) {
      if ((kMod9Keep >> red) & 1u) mod9[r * nw + i / 32] |= 1u << (i % 32);
      red += step;
      if (red >= 9) red -= 9;
    }
  }
  uint64_t p3w[2 * kPow3Max];
  u128 x = 1;
  for (int i = 0; i < kPow3Max; ++i) {
    p3w[2 * i] = lo64(x);
    p3w[2 * i + 1] = hi64(x);
    x *= 3;
  }
  CZ_HIP(t->bv.alloc(p->num_bv * wpb));
  CZ_HIP(t->mod9.alloc(mod9.size()));
  CZ_HIP(t->magic.alloc(p->num_bv));
  CZ_HIP(t->pow3w.alloc(2 * kPow3Max));
  const hipStream_t s = hipStreamPerThread;
  CZ_HIP(hipMemsetAsync(t->bv.p, 0, t->bv.bytes(), s));
  CZ_HIP(hipMemcpyAsync(t->mod9.p, mod9.data(), mod9.size() * 4, hipMemcpyHostToDevice, s));
  CZ_HIP(hipMemcpyAsync(t->magic.p, magic.data(), magic.size() * 4, hipMemcpyHostToDevice, s));
  CZ_HIP(hipMemcpyAsync(t->pow3w.p, p3w, sizeof(p3w), hipMemcpyHostToDevice, s));
  CZ_HIP(launch_bv_build(p->B, p->num_bv, first, p->semantics, t->magic.p, t->bv.p, wpb, nw, s));
  CZ_HIP(hipStreamSynchronize(s));
  t->dev.bv = t->bv.p;
  t->dev.mod9 = t->mod9.p;
  t->dev.magic = t->magic.p;
  t->dev.words_per_bv = wpb;
  t->dev.nw = nw;
  t->dev.num_bv = p->num_bv;
  t->dev.tpo = tpo;
  t->dev.B = p->B;
  t->dev.A = p->A;
  t->dev.rust = p->semantics == COLLATZ_SEM_RUST ? 1u : 0u;
  out = t.get();
  g_ref_cache[key] = std::move(t);
  return COLLATZ_OK;
}

// Frontier of the subtree below `root` at depth root.sigfig + d, found by walking on the
// host (search1's prunes, collatz_cudav7.cu:506-561).
static void frontier(const SieveRoot& root, uint32_t d, std::vector<SieveRoot>& out) {
  struct St {
    uint64_t base;
    u128 number, p3;
    uint32_t s, f, de, ne;
  };
  std::vector<St> stack;
  u128 p3 = 1;
  for (uint32_t i = 0; i < root.threepower; ++i) p3 *= 3;
  stack.push_back({root.base, mk128(root.number, root.number_hi), p3, root.sigfig,
                   root.threepower, root.doubleeven, root.numevens});
  const uint32_t target = root.sigfig + d;
  while (!stack.empty()) {
    St x = stack.back();
    stack.pop_back();
    if (x.s == target) {
      SieveRoot r = root;
      r.base = x.base;
      r.number = lo64(x.number);
      r.number_hi = hi64(x.number);
      r.sigfig = x.s;
      r.threepower = x.f;
      r.doubleeven = x.de;
      r.numevens = x.ne;
      out.push_back(r);
      continue;
    }
    for (uint32_t bit = 0; bit < 2; ++bit) {
      St y = x;
      const u128 nb = x.number + (bit ? x.p3 : 0);
      if (bit) y.base |= 1ull << x.s;
      const uint32_t inplay = kNumerator * (x.s + 1) >= kDenominator * x.f;
      if ((nb & 1) == 0) {
        y.number = nb + nb / 2;
        y.f += 1;
        y.p3 *= 3;
        y.de = (x.de | inplay) & ~2u;
        y.ne = 0;
      } else {
        if (x.de >= 2) continue;
        y.number = nb / 2 + 1;
        y.de = x.de << 1;
        y.ne = x.ne + 1;
      }
      y.s = x.s + 1;
      if (root.rust && y.number <= y.base) continue;
      if (kDenominator * y.f <= kNumerator * y.s) continue;
      if (y.ne >= 2 && y.ne % 2 == 0 && kNumerator * (y.s - 1) >= kDenominator * (y.f - 1)) continue;
      stack.push_back(y);
    }
  }
}

// Device buffers of one run (the leaf arrays and their build scratch, totals, lists), kept
// between calls on a device and grown as needed: a case line at N = 60 builds about eight
// million leaves (a few hundred MB with the bit array), and allocating and freeing them on
// every call took about 2 ms of a 16.5 ms line.  A call that finds its device's workspace
// in use by another thread takes a workspace of its own for that call.
struct RefWork {
  LeafSet ls;
  DevBuf<unsigned long long> dbins, dcount;
  DevBuf<RefTotalsHost> dtot, dstot;
  DevBuf<uint64_t> dovf, desc, dlist, dsel;
  DevBuf<RefSpotRec> dspot;
  bool busy = false;
};

static std::map<int, std::unique_ptr<RefWork>> g_ref_work;  // guarded by g_ref_mu

class RefWorkLease {
 public:
  explicit RefWorkLease(int device) {
    std::lock_guard<std::mutex> lk(g_ref_mu);
    std::unique_ptr<RefWork>& w = g_ref_work[device];
    if (!w) w = std::make_unique<RefWork>();
    if (!w->busy) {
      w->busy = true;
      shared_ = w.get();
    } else {
      own_ = std::make_unique<RefWork>();
    }
  }
  ~RefWorkLease() {
    if (shared_) {
      std::lock_guard<std::mutex> lk(g_ref_mu);
      shared_->busy = false;
    }
  }
  RefWorkLease(const RefWorkLease&) = delete;
  RefWorkLease& operator=(const RefWorkLease&) = delete;
  RefWork& get() { return shared_ ? *shared_ : *own_; }

 private:
  RefWork* shared_ = nullptr;
  std::unique_ptr<RefWork> own_;
};

}  // namespace cz

using namespace cz;

extern "C" int collatz_reference_tables(const collatz_ref_params* params, uint32_t* bv,
                                        uint32_t* mod9) {
  int rc = check_params(params, false);
  if (rc) return rc;
  HipDeviceGuard guard(params->device);
  int dev = 0;
  CZ_HIP(hipGetDevice(&dev));
  RefTables* t = nullptr;
  rc = get_tables(params, dev, t);
  if (rc) return rc;
  const uint64_t words = 1ull << (params->B - 5);
  if (bv)
    for (uint32_t i = 0; i < params->num_bv; ++i)
      CZ_HIP(hipMemcpyAsync(bv + i * words, t->bv.p + i * t->dev.words_per_bv, words * 4,
                            hipMemcpyDeviceToHost, hipStreamPerThread));
  if (mod9)
    CZ_HIP(hipMemcpyAsync(mod9, t->mod9.p, t->mod9.bytes(), hipMemcpyDeviceToHost,
                          hipStreamPerThread));
  CZ_HIP(hipStreamSynchronize(hipStreamPerThread));
  return COLLATZ_OK;
}

extern "C" int collatz_reference_run(const collatz_ref_params* params, const collatz_node* root,
                                     void* stream, collatz_ref_result* out) {
  int rc = check_params(params, true);
  if (rc) return rc;
  if (!root || !out) return fail(COLLATZ_EINVAL, "null argument");
  const uint32_t mbi = params->N - params->A;
  // The root's base is below 2^sigfig and its number below 3^threepower + 1.
  if (root->base[1] || root->sigfig > mbi || root->threepower >= kPow3Max ||
      (root->sigfig <= kWalk64MaxDepth && root->number[1]))
    return fail(COLLATZ_EINVAL, "case line out of range");
  HipDeviceGuard guard(params->device);
  hipStream_t s = call_stream(stream);
  int dev = 0;
  CZ_HIP(hipGetDevice(&dev));
  RefTables* T = nullptr;
  rc = get_tables(params, dev, T);
  if (rc) return rc;
  const double t0 = now_ms();

  SieveRoot r;
  r.base = root->base[0];
  r.number = root->number[0];
  r.number_hi = root->number[1];
  r.sigfig = root->sigfig;
  r.threepower = root->threepower;
  r.doubleeven = root->doubleeven;
  r.numevens = root->numevens;
  r.depth = mbi;
  r.rust = params->semantics == COLLATZ_SEM_RUST ? 1u : 0u;
  std::vector<SieveRoot> roots;
  const uint32_t e = mbi - r.sigfig;
  if (e > 30) frontier(r, e - 30, roots);
  else roots.push_back(r);

  // Word widths: the fused pass runs on 96 bits where it can and the candidates that leave
  // them are rerun on two limbs or three.  The reference schedule tests its cutoff after
  // each odd run (which multiplies by less than 2^4.7), main.rs's after each 16-step block
  // (less than 2^9.4), so two limbs hold cutoffs up to 2^123 and 2^118 respectively.
  const bool rust = params->semantics == COLLATZ_SEM_RUST;
  const int limbs = params->cutoff_log2 <= (rust ? 118u : kWideLog2) ? 2 : 3;
  const Limits lim = make_limits(limbs, params->cutoff_log2);
  RefWorkLease lease(dev);
  RefWork& W = lease.get();
  DevBuf<unsigned long long>& dbins = W.dbins;
  DevBuf<RefTotalsHost>& dtot = W.dtot;
  DevBuf<uint64_t>& dovf = W.dovf;
  DevBuf<uint64_t>& desc = W.desc;
  const uint64_t ovf_cap = out->overflow_starts ? std::max<uint64_t>(out->overflow_cap, 1) : 0;
  const uint64_t esc_cap = 1ull << 20;
  CZ_HIP(dbins.reserve(17));
  CZ_HIP(dtot.reserve(1));
  CZ_HIP(dovf.reserve(std::max<uint64_t>(1, 2 * ovf_cap)));
  CZ_HIP(desc.reserve(esc_cap));
  const uint32_t nblocks = (uint32_t)device_cu_count(dev) * 32 / 4;
  uint64_t leaves = 0;
  // Spot checks: sampled tiles of each fused launch keep their own results (dspot), and the
  // same tiles' candidates are listed (dlist) and rerun by k_ref_iterate (dstot), together
  // with a rerun of the sampled tiles' escalated candidates (dsesc, whose results the fused
  // pass leaves to the wide pass).
  constexpr uint32_t kTile = 256;
  constexpr uint64_t kSpotEntries = 1ull << 23;
  // The stride: the largest power of two not above spot_check_freq.
  const uint32_t freq = params->spot_check_freq
                            ? 1u << (31 - __builtin_clz(params->spot_check_freq)) : 0u;
  const uint32_t max_spot = (uint32_t)(kSpotEntries / ((uint64_t)kTile << params->A));
  DevBuf<RefSpotRec>& dspot = W.dspot;
  DevBuf<uint64_t>& dlist = W.dlist;
  DevBuf<uint64_t>& dsel = W.dsel;
  DevBuf<RefTotalsHost>& dstot = W.dstot;
  DevBuf<unsigned long long>& dcount = W.dcount;
  if (freq) CZ_HIP(dspot.reserve(max_spot));
  uint64_t spot_checks = 0, spot_failures = 0;
  std::string spot_msg;
  // Runs the spot check of one root's fused launch: sp describes its sampled tiles, nesc
  // the length of its escalation list.
  auto spot_check = [&](size_t root_index, const RefLeaves& lv, uint64_t nleaves,
                        const RefSpot& sp, uint64_t nesc) -> int {
    CZ_HIP(dlist.reserve(kSpotEntries));
    CZ_HIP(dstot.reserve(2));
    CZ_HIP(dcount.reserve(1));
    CZ_HIP(hipMemsetAsync(dcount.p, 0, 8, s));
    CZ_HIP(hipMemsetAsync(dstot.p, 0, 2 * sizeof(RefTotalsHost), s));
    CZ_HIP(launch_ref_spot_list(T->dev, lv, nleaves, kTile, sp, dlist.p, dcount.p,
                                kSpotEntries, s));
    unsigned long long nlist = 0;
    CZ_HIP(hipMemcpyAsync(&nlist, dcount.p, 8, hipMemcpyDeviceToHost, s));
    CZ_HIP(hipStreamSynchronize(s));
    if (nlist > kSpotEntries) return fail(COLLATZ_ECAPACITY, "spot-check list overflowed");
    CZ_HIP(launch_ref_iterate(limbs, rust, dlist.p, nlist, lv, mbi, lim, T->pow3w.p, 4096,
                              nblocks, dstot.p, nullptr, 0, s));
    // The sampled tiles' escalated candidates, from the fused pass's list.
    std::vector<uint64_t> sel;
    if (nesc) {
      std::vector<uint64_t> e(nesc);
      CZ_HIP(hipMemcpyAsync(e.data(), desc.p, nesc * 8, hipMemcpyDeviceToHost, s));
      CZ_HIP(hipStreamSynchronize(s));
      for (uint64_t x : e) {
        const uint64_t t = (uint32_t)x / kTile;
        if (t >= sp.off && (t - sp.off) % sp.freq == 0 && (t - sp.off) / sp.freq < sp.cap)
          sel.push_back(x);
      }
    }
    if (!sel.empty()) {
      CZ_HIP(dsel.reserve(sel.size()));
      CZ_HIP(hipMemcpyAsync(dsel.p, sel.data(), sel.size() * 8, hipMemcpyHostToDevice, s));
      CZ_HIP(launch_ref_iterate(limbs, rust, dsel.p, sel.size(), lv, mbi, lim, T->pow3w.p,
                                4096, nblocks, dstot.p + 1, nullptr, 0, s));
    }
    std::vector<RefSpotRec> recs(sp.cap);
    RefTotalsHost st[2];
    CZ_HIP(hipMemcpyAsync(recs.data(), dspot.p, sp.cap * sizeof(RefSpotRec),
                          hipMemcpyDeviceToHost, s));
    CZ_HIP(hipMemcpyAsync(st, dstot.p, sizeof(st), hipMemcpyDeviceToHost, s));
    CZ_HIP(hipStreamSynchronize(s));
    // The fused pass's tiles plus their escalated candidates against the rerun of all.
    uint32_t h = st[1].hash;
    uint64_t cnt = st[1].count, nc = st[1].numcases, ov = st[1].overflow;
    for (const RefSpotRec& r : recs) {
      h ^= r.hash;
      cnt += r.count;
      nc += r.numcases;
      ov += r.overflow;
    }
    spot_checks += sp.cap;
    if (h != st[0].hash || cnt != st[0].count || nc != st[0].numcases ||
        ov != st[0].overflow || st[0].aborts || st[1].aborts) {
      spot_failures += sp.cap;
      if (spot_msg.empty())
        spot_msg = "spot check failed: root " + std::to_string(root_index) + ", tiles " +
                   std::to_string(sp.off) + " + j * " + std::to_string(sp.freq) + " (j < " +
                   std::to_string(sp.cap) + "): fused pass and rerun differ";
    }
    return COLLATZ_OK;
  };
  // The whole run at one word width of the fused pass: 96 bits (escalating to `limbs`
  // above 2^86) when the cutoff is at least 2^96, `limbs` otherwise.
  auto run_all = [&](int width) -> int {
    CZ_HIP(hipMemsetAsync(dbins.p, 0, 17 * 8, s));
    CZ_HIP(hipMemsetAsync(dtot.p, 0, sizeof(RefTotalsHost), s));
    leaves = 0;
    spot_checks = spot_failures = 0;
    spot_msg.clear();
    for (size_t ri = 0; ri < roots.size(); ++ri) {
      const SieveRoot& sr = roots[ri];
      LeafSet& ls = W.ls;
      int rc2 = build_leaves(sr, s, ls);
      if (rc2) return rc2;
      leaves += ls.count;
      CZ_HIP(launch_leaf_bins(ls.meta.p, ls.count, T->dev.tpo, dbins.p, s));
      // Steps 2 and 3 in one persistent launch over the leaves.
      CZ_HIP(hipMemsetAsync(&dtot.p->tile_counter, 0, 4, s));
      CZ_HIP(hipMemsetAsync(&dtot.p->esc_count, 0, 8, s));
      // Tiles of 256 leaves (the wave's candidate queue evens out the leaves' uneven
      // candidate counts, so small tiles only shorten the launch's tail) and five checks
      // between refills (candidates that pass the look-ahead filters run long); both
      // measured best at N = 48.
      RefLeaves lv;
      lv.base = ls.res.p;
      lv.num = ls.num.p;
      lv.num_hi = ls.wide ? ls.num_hi.p : nullptr;
      lv.meta = ls.meta.p;
      // Sampled tiles: t = off + j freq, off from a SplitMix64 step of the root's index.
      RefSpot sp{};
      const uint64_t ntiles = (ls.count + kTile - 1) / kTile;
      if (freq) {
        uint64_t z = 0x9E3779B97F4A7C15ull * (ri + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        sp.freq = freq;
        sp.off = (uint32_t)(z % freq);
        sp.cap = sp.off < ntiles
                     ? (uint32_t)std::min<uint64_t>(max_spot, (ntiles - 1 - sp.off) / freq + 1)
                     : 0;
        sp.rec = dspot.p;
        if (!sp.cap) sp.freq = 0;
      }
      {
        // The fused pass reads its spot-check fields from the totals block.
        struct {
          unsigned long long rec;
          unsigned int mask, off, last, shift;
        } f = {0, 0, 0, 0, 0};
        if (sp.cap) {
          f.rec = (unsigned long long)(uintptr_t)sp.rec;
          f.mask = sp.freq - 1;
          f.off = sp.off;
          f.last = sp.off + (sp.cap - 1) * sp.freq;
          f.shift = (uint32_t)__builtin_ctz(sp.freq);
        }
        static_assert(sizeof(f) == 24, "spot fields");
        CZ_HIP(hipMemcpyAsync(&dtot.p->spot_rec, &f, sizeof(f), hipMemcpyHostToDevice, s));
      }
      CZ_HIP(launch_ref_fused(width, rust, T->dev, lv, ls.count, mbi, lim, T->pow3w.p, kTile,
                              kRefChecksPerRefill, nblocks, dtot.p, ovf_cap ? dovf.p : nullptr,
                              ovf_cap, desc.p, esc_cap, s));
      unsigned long long nesc = 0;
      CZ_HIP(hipMemcpyAsync(&nesc, &dtot.p->esc_count, 8, hipMemcpyDeviceToHost, s));
      CZ_HIP(hipStreamSynchronize(s));
      if (nesc > esc_cap) return fail(COLLATZ_ECAPACITY, "escalation list overflowed");
      if (sp.cap) {
        rc2 = spot_check(ri, lv, ls.count, sp, nesc);
        if (rc2) return rc2;
      }
      if (nesc) {
        // Candidates that left 96 bits: rerun from their start on two or three limbs.
        CZ_HIP(hipMemsetAsync(&dtot.p->tile_counter, 0, 4, s));
        CZ_HIP(launch_ref_iterate(limbs, rust, desc.p, nesc, lv, mbi, lim, T->pow3w.p, 4096,
                                  nblocks, dtot.p, ovf_cap ? dovf.p : nullptr, ovf_cap, s));
        CZ_HIP(hipStreamSynchronize(s));
      }
    }
    return COLLATZ_OK;
  };
  // 96 bits for the reference schedule needs a cutoff of at least 2^96 (it has no overflow
  // test there); main.rs's blocks test a cutoff below 2^86 on 96 bits and escalate above.
  const bool narrow = rust || params->cutoff_log2 >= 96;
  rc = run_all(narrow ? 1 : limbs);
  if (rc == COLLATZ_ECAPACITY && narrow) rc = run_all(limbs);
  if (rc) return rc;
  RefTotalsHost tot;
  unsigned long long bins[17];
  CZ_HIP(hipMemcpyAsync(&tot, dtot.p, sizeof(tot), hipMemcpyDeviceToHost, s));
  CZ_HIP(hipMemcpyAsync(bins, dbins.p, sizeof(bins), hipMemcpyDeviceToHost, s));
  CZ_HIP(hipStreamSynchronize(s));
  if (tot.aborts)
    return fail(COLLATZ_EDEVICE, "a persistent work loop reached its iteration bound (" +
                                     std::to_string(tot.aborts) + " tiles abandoned)");
  if (tot.stall) return fail(COLLATZ_EDEVICE, "a trajectory did not descend within the check bound");
  out->complete_hash = tot.hash;
  out->numcases = tot.numcases;
  out->leaves = leaves;
  out->candidates = tot.count;
  out->checks = tot.checks;
  out->overflow_count = tot.overflow;
  out->launches = 0;
  for (int b = 0; b < 17; ++b) out->bin_leaves[b] = bins[b];
  for (int b = 0; b < 16; ++b) out->launches += (bins[b] + 65535) / 65536;
  out->overflow_written = tot.overflow;
  if (ovf_cap) {
    const uint64_t nov = std::min<uint64_t>(tot.overflow, ovf_cap);
    if (nov) {
      CZ_HIP(hipMemcpyAsync(out->overflow_starts, dovf.p, nov * 16, hipMemcpyDeviceToHost, s));
      CZ_HIP(hipStreamSynchronize(s));
    }
  }
  out->spot_checks = spot_checks;
  out->spot_check_failures = spot_failures;
  out->device_ms = now_ms() - t0;
  if (spot_failures) return fail(COLLATZ_EDEVICE, spot_msg);
  if (ovf_cap && tot.overflow > ovf_cap)
    return fail(COLLATZ_ECAPACITY, "overflow buffer smaller than the result");
  return COLLATZ_OK;
}
