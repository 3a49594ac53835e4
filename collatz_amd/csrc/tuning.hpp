// Launch-shape selection for the persistent kernel, per (sieve depth, limbs, magnitude).
#pragma once

#include "host_common.hpp"

namespace cz {

struct Tuning {
  uint32_t tile;   // survivors per work tile (residues handed to a wave at a time)
  uint32_t waves;  // persistent waves per CU
  uint32_t U;      // checks between lane refills
};

// Tile size for a launch of about `candidates` candidates at `waves` waves per CU.
uint32_t adaptive_tile(const collatz_sieve* sv, uint64_t candidates, uint32_t waves);

// Explicit (non-zero) arguments win; otherwise the tuned entry for the bucket of max_hi is
// used, running the tuner on first use (the caller waits for the whole sweep).
Tuning choose_tuning(const collatz_sieve* sv, int limbs, u128 max_hi, uint64_t candidates,
                     uint32_t tile, uint32_t waves, uint32_t U);

// The tuner's bucket of a launch: word width and bit length of its largest end / 8.
uint32_t tune_bucket(int limbs, u128 max_hi);

// Never sweeps: the tuned shape of the bucket if it has one, else the defaults (32 waves
// per CU, two checks per refill), with the tile for this launch.  `want` is set when the
// launch is one the tuner would tune for (large enough, tuning enabled) and the bucket has
// no entry yet: the caller may then advance a sweep with tune_step when it has time.
Tuning tuned_or_default(const collatz_sieve* sv, int limbs, u128 max_hi, uint64_t candidates,
                        bool& want);

// A sweep over the tuner's grid of launch shapes, advanced one timing launch at a time.
struct TuneSweep {
  uint32_t next = 0;       // next timing: shape index x 3 + repetition
  double shape_ms = 1e30;  // fastest repetition of the current shape so far
  double best_ms = 1e30;   // fastest shape so far
  uint32_t best_waves = 0, best_U = 0;
};

// Runs the next timing launch of the sweep for (limbs, max_hi) on the sieve.  Returns 1 when
// the sweep has ended (the fastest shape is stored for the bucket; also at once if the
// bucket has been tuned meanwhile), 0 when more timing launches remain, and -1 without a
// launch if another caller is tuning this sieve right now.
int tune_step(const collatz_sieve* sv, int limbs, u128 max_hi, TuneSweep& st);

}  // namespace cz
