// Autotuner for the persistent kernel.  For each (limbs, magnitude bucket) of a sieve it
// times a short synthetic range of that magnitude under a small grid of launch shapes
// (waves per CU, checks per refill) and keeps the fastest; the tile size (residues per
// wave visit) follows the size of each launch.  Results are kept on the sieve handle (and
// go with it when it is freed), per bucket; COLLATZ_AUTOTUNE=0 disables tuning.  A sweep is
// a list of timing launches that can be run all at once (choose_tuning: range and batch
// calls) or one at a time in idle moments (tune_step: the server, which never holds a
// request back for a sweep).
#include <cstdlib>
#include <mutex>

#include "engine.hpp"
#include "tuning.hpp"

namespace cz {

struct BatchArgs;
int verify_batch_tuned(const collatz_sieve* sv, u128 lo, u128 hi, int limbs, Tuning t,
                       hipStream_t s, double* ms);

static uint32_t bitlen(u128 x) {
  uint32_t b = 0;
  while (x) {
    ++b;
    x >>= 1;
  }
  return b;
}

// Waves per CU: 32 is the occupancy limit of the kernel (under 64 VGPRs, eight waves per
// SIMD).  Checks between refills: 2.
static const uint32_t kDefaultWaves = 32, kDefaultU = 2;

// Survivors per tile: the largest power of two in [8192, 131072] that still gives every
// resident wave about 8 tiles.  Each tile costs its wave a set-up (its row, the q 3^f table,
// the first records in flight) and ends with a tail in which lanes fall idle while its last
// trajectories finish, so larger tiles waste less; but the tiles of a launch that mixes
// ranges of very different sizes differ in size too (a row shorter than the tile is a tile
// of its own), and with few tiles per wave the largest ones set the launch's end.
// (Measured against 16 tiles per wave and at least 4096: [2^40, 2^40 + 2^36) 1.50 -> 1.37
// ms, the saturated served stream 5.3e13 -> 5.8e13 n/s with the per-XCD shares kept for
// launches of 32 tiles per wave; 4 tiles per wave ran the stream at 5.7e13.)  A look-ahead
// keeps about a third of the survivors, so tiles of such a sieve may be twice as large.  A
// launch too small to give every resident wave a tile of 8192 takes smaller tiles (down to
// 256) until it does: its duration is about one tile per wave, so this is its latency.
uint32_t adaptive_tile(const collatz_sieve* sv, uint64_t candidates, uint32_t waves) {
  const uint64_t nwaves = (uint64_t)device_cu_count(sv->device) * waves;
  uint32_t t = sv->la_bits ? 262144 : 131072;
  while (t > 8192 && candidates / t < 8 * nwaves) t >>= 1;
  while (t > 256 && candidates / t < nwaves) t >>= 1;
  return t;
}

// The tuner's timing launches: 2^29 candidates in tiles of 4096, about 16 tiles per
// resident wave, as the large launches the tuning is for have.  (With 2^27 candidates,
// four tiles per wave, the end of the launch weighed so much that 28 waves per CU won
// over 32, which runs the 4096-interval batch 4% faster.)  Launches of at most 2^22
// candidates are not worth a sweep: they take the defaults.
static constexpr uint32_t kTuneTile = 4096;
static constexpr uint32_t kTuneLog2 = 29;
static constexpr uint64_t kTuneMinCandidates = 1u << 22;
// The grid: waves per CU x checks per refill, each shape timed three times (the fastest
// counts).  A look-ahead sieve's refill is dearer (a window of the tile is tested first),
// so its grid goes one check further between refills.
static const uint32_t kWavesGrid[] = {24, 28, 32};
static const uint32_t kUGrid[] = {1, 2, 3, 4};
static constexpr uint32_t kReps = 3;

static bool tuning_enabled() {
  const char* env = std::getenv("COLLATZ_AUTOTUNE");
  return !(env && env[0] == '0');
}

uint32_t tune_bucket(int limbs, u128 max_hi) {
  return (uint32_t)limbs * 256 + bitlen(max_hi) / 8;
}

// Runs timing launches of the sweep until it ends or `budget` launches have run; the caller
// holds sv->tune_mu.  On the end, stores the fastest shape and returns true.
static bool sweep_run(const collatz_sieve* sv, int limbs, u128 max_hi, TuneSweep& st,
                      uint32_t budget) {
  // A range of the bucket's magnitude with about 2^29 candidates (three times as many
  // survivors for a look-ahead sieve, which keeps about a third).
  const uint32_t bits = std::max<uint32_t>(bitlen(max_hi), sv->k + 1);
  const double density = (double)sv->count / (double)((u128)1 << sv->k) * 5.0 / 9.0;
  double w = (double)(1u << kTuneLog2) / (density > 0 ? density : 1e-3);
  if (sv->la_bits) w *= 3;
  u128 width = (u128)w;
  if (width < ((u128)1 << 20)) width = (u128)1 << 20;
  const u128 lo = (u128)1 << (bits - 1);
  const u128 hi = lo + width;
  const uint32_t nU = sv->la_bits ? 4 : 3;
  const uint32_t total = 3 * nU * kReps;
  for (; st.next < total && budget > 0; --budget) {
    const uint32_t shape = st.next / kReps, rep = st.next % kReps;
    const Tuning cand{kTuneTile, kWavesGrid[shape / nU], kUGrid[shape % nU]};
    double m = 0;
    const bool ok = verify_batch_tuned(sv, lo, hi, limbs, cand, nullptr, &m) == 0;
    st.shape_ms = ok ? std::min(st.shape_ms, m) : 1e30;
    // A failed launch ends its shape's repetitions.
    const bool shape_done = !ok || rep + 1 == kReps;
    st.next = shape_done ? (shape + 1) * kReps : st.next + 1;
    if (shape_done) {
      if (st.shape_ms < st.best_ms) {
        st.best_ms = st.shape_ms;
        st.best_waves = cand.waves;
        st.best_U = cand.U;
      }
      st.shape_ms = 1e30;
    }
  }
  if (st.next < total) return false;
  if (st.best_waves) sv->tuned[tune_bucket(limbs, max_hi)] = {st.best_waves, st.best_U};
  else sv->tuned[tune_bucket(limbs, max_hi)] = {kDefaultWaves, kDefaultU};
  return true;
}

Tuning choose_tuning(const collatz_sieve* sv, int limbs, u128 max_hi, uint64_t candidates,
                     uint32_t tile, uint32_t waves, uint32_t U) {
  Tuning t{0, kDefaultWaves, kDefaultU};
  const bool tune = tuning_enabled() && candidates > kTuneMinCandidates && !(waves && U);
  if (tune) {
    // Callers on other sieves are not held up; callers on this one wait for its sweep.
    std::lock_guard<std::mutex> lk(sv->tune_mu);
    const uint32_t bucket = tune_bucket(limbs, max_hi);
    if (sv->tuned.find(bucket) == sv->tuned.end()) {
      TuneSweep st;
      sweep_run(sv, limbs, max_hi, st, ~0u);
    }
    const auto& e = sv->tuned.at(bucket);
    t.waves = e[0];
    t.U = e[1];
  }
  if (waves) t.waves = waves;
  if (U) t.U = U;
  t.tile = tile ? tile : adaptive_tile(sv, candidates, t.waves);
  return t;
}

Tuning tuned_or_default(const collatz_sieve* sv, int limbs, u128 max_hi, uint64_t candidates,
                        bool& want) {
  Tuning t{0, kDefaultWaves, kDefaultU};
  want = false;
  if (tuning_enabled() && candidates > kTuneMinCandidates) {
    std::lock_guard<std::mutex> lk(sv->tune_mu);
    auto it = sv->tuned.find(tune_bucket(limbs, max_hi));
    if (it != sv->tuned.end()) {
      t.waves = it->second[0];
      t.U = it->second[1];
    } else {
      want = true;
    }
  }
  t.tile = adaptive_tile(sv, candidates, t.waves);
  return t;
}

int tune_step(const collatz_sieve* sv, int limbs, u128 max_hi, TuneSweep& st) {
  std::unique_lock<std::mutex> lk(sv->tune_mu, std::try_to_lock);
  if (!lk.owns_lock()) return -1;
  if (sv->tuned.count(tune_bucket(limbs, max_hi))) return 1;
  return sweep_run(sv, limbs, max_hi, st, 1) ? 1 : 0;
}

}  // namespace cz
