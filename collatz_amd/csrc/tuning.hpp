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
// used, running the tuner on first use.
Tuning choose_tuning(const collatz_sieve* sv, int limbs, u128 max_hi, uint64_t candidates,
                     uint32_t tile, uint32_t waves, uint32_t U);

}  // namespace cz
