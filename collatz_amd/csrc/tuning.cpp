// Autotuner for the persistent kernel.  For each (limbs, magnitude bucket) of a sieve it
// times a short synthetic range of that magnitude under a small grid of launch shapes
// (tile size = residues per wave visit, waves per CU, checks per refill) and keeps the
// fastest.  Results are cached in the sieve handle; COLLATZ_AUTOTUNE=0 disables tuning.
#include <cstdlib>
#include <map>
#include <mutex>

#include "tuning.hpp"

namespace cz {

struct BatchArgs;
int verify_batch_tuned(const collatz_sieve* sv, u128 lo, u128 hi, int limbs, Tuning t,
                       hipStream_t s, double* ms);

static std::mutex g_mu;
static std::map<std::pair<const void*, uint32_t>, Tuning> g_cache;

static uint32_t bitlen(u128 x) {
  uint32_t b = 0;
  while (x) {
    ++b;
    x >>= 1;
  }
  return b;
}

static const Tuning kDefault = {16384, 24, 2};

Tuning choose_tuning(const collatz_sieve* sv, int limbs, u128 max_hi, uint64_t candidates,
                     uint32_t tile, uint32_t waves, uint32_t U) {
  if (tile && waves && U) return Tuning{tile, waves, U};
  Tuning t = kDefault;
  const char* env = std::getenv("COLLATZ_AUTOTUNE");
  const bool tune = !(env && env[0] == '0') && candidates > (1u << 22);
  const uint32_t bucket = (uint32_t)limbs * 256 + bitlen(max_hi) / 8;
  const auto key = std::make_pair((const void*)sv, bucket);
  if (tune) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
      t = it->second;
    } else {
      // A range of the bucket's magnitude with about 2^26 candidates (enough tiles for the
      // largest tile size to show its tail).
      const uint32_t bits = std::max<uint32_t>(bitlen(max_hi), sv->k + 1);
      const double density = (double)sv->count / (double)((u128)1 << sv->k) * 5.0 / 9.0;
      double w = (double)(1u << 26) / (density > 0 ? density : 1e-3);
      u128 width = (u128)w;
      if (width < ((u128)1 << 20)) width = (u128)1 << 20;
      const u128 lo = (u128)1 << (bits - 1);
      const u128 hi = lo + width;
      double best = 1e30;
      static const uint32_t tiles[] = {4096, 16384, 65536};
      static const uint32_t wavess[] = {16, 24, 28};
      static const uint32_t Us[] = {1, 2, 3};
      for (uint32_t a : tiles)
        for (uint32_t b : wavess)
          for (uint32_t c : Us) {
            double ms = 0;
            if (verify_batch_tuned(sv, lo, hi, limbs, Tuning{a, b, c}, nullptr, &ms) != 0) continue;
            if (ms < best) {
              best = ms;
              t = Tuning{a, b, c};
            }
          }
      g_cache[key] = t;
    }
  }
  if (tile) t.tile = tile;
  if (waves) t.waves = waves;
  if (U) t.U = U;
  return t;
}

}  // namespace cz
