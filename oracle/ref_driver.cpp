// Builds the reference's own CPU code path into oracle/_ref/libref.so so that the oracle
// can be checked against it.  TEST INFRASTRUCTURE ONLY.
//
// The reference source is compiled where it lies (its path comes in through
// REFERENCE_SOURCE, set by oracle/Makefile); nothing of it is copied here.  Only its
// CPU functions are called: precompute (collatz_openclv2.cpp:294-366), search1 (:368-438),
// search2/search3 (:598-659).  Its main() (which needs an OpenCL GPU and reads
// compute_kernel2.cl) is renamed out of the way and never called.  launchkernel is only
// reached when a batch bin fills (2^16 leaves); with no OpenCL context it throws, and
// ref_case reports that case as too large instead of running it.
//
// The setup below follows main() (collatz_openclv2.cpp:99-127): THREEPOWERS, the magic
// numbers, the table sizes, then precompute().
#define main collatz_reference_main_unused
#include REFERENCE_SOURCE
#undef main

#include <exception>

static PrecomputedData* g_data = nullptr;

extern "C" {

int ref_constants(uint32_t* out) {
  out[0] = MAXBITS;
  out[1] = LASTBITS;
  out[2] = STEPSIZEINITIAL;
  out[3] = CUTOFFPOWER;
  out[4] = THREEPOWEROFFSET;
  out[5] = INITIALBITARRAYSIZE;
  out[6] = LASTBITSTEPS;
  out[7] = BATCHSIZE;
  return 0;
}

int ref_init() {
  if (g_data) return 0;
  u128 a = 1;
  THREEPOWERS[0] = a;
  for (isize i = 1; i < 80; i++) {
    a *= 3;
    THREEPOWERS[i] = a;
  }
  g_data = new PrecomputedData{std::array<std::vector<u32>, 16>{}, std::array<u32, 16>{},
                               std::array<std::vector<u32>, 9>{},
                               std::array<std::vector<u32>, 16>{}, 0, 0};
  for (auto& v : g_data->initialbitarray) v.resize(INITIALBITARRAYSIZE);
  for (auto& v : g_data->mod9sieve) v.resize(LASTBITSTEPS);
  u32 magicnumber = 1;
  for (isize i = 1; i < 128; i++) {
    magicnumber *= ONETHIRD;
    if (i >= THREEPOWEROFFSET && i - THREEPOWEROFFSET < 16)
      g_data->magicnumbers[i - THREEPOWEROFFSET] = magicnumber;
  }
  precompute(g_data);
  return 0;
}

int ref_bv(uint32_t i, uint32_t* out) {
  if (!g_data || i >= 16) return -1;
  std::memcpy(out, g_data->initialbitarray[i].data(), INITIALBITARRAYSIZE * sizeof(u32));
  return 0;
}

int ref_mod9(uint32_t* out) {
  if (!g_data) return -1;
  for (int r = 0; r < 9; ++r)
    std::memcpy(out + r * LASTBITSTEPS, g_data->mod9sieve[r].data(), LASTBITSTEPS * sizeof(u32));
  return 0;
}

int ref_magic(uint32_t* out) {
  if (!g_data) return -1;
  for (int i = 0; i < 16; ++i) out[i] = g_data->magicnumbers[i];
  return 0;
}

uint32_t ref_search3(uint64_t base_lo, uint64_t base_hi, uint64_t num_lo, uint64_t num_hi) {
  u128 base = ((u128)base_hi << 64) | base_lo;
  u128 number = ((u128)num_hi << 64) | num_lo;
  return search3(base, number, g_data);
}

uint32_t ref_search2(uint64_t base_lo, uint64_t base_hi, uint32_t threepower, uint64_t num_lo,
                     uint64_t num_hi, uint32_t bigindex) {
  u128 base = ((u128)base_hi << 64) | base_lo;
  u128 number = ((u128)num_hi << 64) | num_lo;
  return search2(base, threepower, number, g_data, bigindex);
}

// Runs one case line (base sigfig threepower number doubleeven numevens).  Writes the
// complete hash, the launch count the reference would report, the leaves per bin, and
// up to `cap` leaves as (base, number, bin) triples.  Returns the number of binned
// leaves, -1 if not initialised, -2 if a bin reached the batch size.
int64_t ref_case(const uint64_t* t, uint32_t* hash, uint64_t* launches, uint64_t* bins,
                 uint64_t* leaves, int64_t cap) {
  if (!g_data) return -1;
  for (auto& v : g_data->bigpairvector2) v.clear();
  g_data->complete_hash = 0;
  try {
    search1((u128)t[0], (u32)t[1], (u32)t[2], (u128)t[3], (u32)t[4], (u32)t[5], g_data,
            nullptr, nullptr, nullptr, nullptr, nullptr);
  } catch (const std::exception&) {
    for (auto& v : g_data->bigpairvector2) v.clear();
    return -2;
  }
  u32 h = g_data->complete_hash;
  uint64_t nl = 0, nlaunch = 0;
  for (u32 b = 0; b < 16; ++b) {
    const std::vector<u32>& v = g_data->bigpairvector2[b];
    const u32 count = v.size() / 8;
    bins[b] = count;
    if (count) nlaunch++;
    for (u32 i = 0; i < count; ++i) {
      u128 base = (u128)v[8 * i] | ((u128)v[8 * i + 1] << 32) | ((u128)v[8 * i + 2] << 64);
      u128 number = (u128)v[8 * i + 4] | ((u128)v[8 * i + 5] << 32) |
                    ((u128)v[8 * i + 6] << 64) | ((u128)v[8 * i + 7] << 96);
      h ^= search2(base, THREEPOWEROFFSET + b, number, g_data, b);
      if ((int64_t)nl < cap) {
        leaves[3 * nl] = (uint64_t)base;
        leaves[3 * nl + 1] = (uint64_t)number;
        leaves[3 * nl + 2] = b;
      }
      nl++;
    }
  }
  for (auto& v : g_data->bigpairvector2) v.clear();
  *hash = h;
  *launches = nlaunch;
  return (int64_t)nl;
}

}  // extern "C"
