// Definitions shared by the host orchestration (host.cpp) and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace cz {

constexpr uint32_t kNumerator = 306;    // 306/485 ~ log 2 / log 3 (collatz_cudav7.cu:50-52)
constexpr uint32_t kDenominator = 485;
constexpr uint32_t kWideLog2 = 123;     // values at or above 2^123 need a third u64 limb
constexpr uint32_t kMod9Keep = 0xCBu;   // bits {0,1,3,6,7}: residues mod 9 that stay
constexpr uint32_t kLaLevel1Bits = 16;  // first-level look-ahead table of deeper filters
constexpr int kPow3Max = 64;            // 3^f tables cover f < 64

// Root node of a sieve walk (a node of the reference's search1 tree,
// collatz_cudav7.cu:497-505): n = base + h * 2^sigfig for the extension h.
struct SieveRoot {
  uint64_t base;        // n mod 2^sigfig
  uint64_t number;      // T^sigfig(base) + 1, low 64 bits
  uint64_t number_hi;   // its high 64 bits (0 unless the walk is 128-bit wide)
  uint32_t sigfig;      // depth of the root
  uint32_t threepower;  // f at the root
  uint32_t doubleeven;  // odd-even-even state
  uint32_t numevens;    // consecutive even steps
  uint32_t depth;       // target depth D
  uint32_t rust;        // main.rs:216 prune (number <= base) on
};

// Survivor / leaf meta word: f in bits 0-7, n mod 9 in bits 8-11, doubleeven in 12-13,
// numevens in 16-23.
__host__ __device__ inline uint32_t meta_f(uint32_t m) { return m & 0xFFu; }
__host__ __device__ inline uint32_t meta_mod9(uint32_t m) { return (m >> 8) & 0xFu; }
__host__ __device__ inline uint32_t meta_de(uint32_t m) { return (m >> 12) & 0x3u; }

// One row of range work: residues with sieve indices [i_lo, i_hi) and one q, all inside
// one output chunk.  n = (q << k) | r.
struct RowDesc {
  uint64_t q_lo, q_hi;        // q
  uint64_t qb_lo, qb_hi;      // q << k (low k bits zero)
  uint64_t i_lo, i_hi;        // sieve index range
  uint64_t block_start;       // first plain-kernel share (or persistent tile) of the row
  uint32_t slot;              // output record slot (chunk of an interval)
  uint32_t cls;               // residue class r mod 9 of this row's survivors
};

// A unit of persistent work: piece `piece` of `pieces` equal parts of one row's survivors
// (the row's index range is resolved on the device, so the host only fixes the split).
struct TileDesc {
  uint32_t row;
  uint32_t piece;
  uint32_t pieces;
  uint32_t pad;
};

// Row bounds still to be looked up: i_lo / i_hi hold kResolveFlag | key index.
constexpr uint64_t kResolveFlag = 1ull << 63;

// Sizes of one persistent launch, read by the kernels from device memory so that one
// captured graph serves launches of any size.
// XCDs of an MI355X, each with its own L2.  The workgroups of a launch are dispatched to
// them round-robin: workgroup w runs on XCD w mod kXcds.
constexpr uint32_t kXcds = 8;

struct Ctl {
  uint64_t nrows;
  uint64_t nslots;
  uint32_t ntiles;
  // The tile order is cut into 2^share_log2 equal shares (1 or kXcds): share x is positions
  // [ntiles x >> share_log2, ntiles (x + 1) >> share_log2).  The waves of XCD
  // x mod 2^share_log2 take from share x first, then from the others.  (A power of two, so
  // that the kernel needs no division: a 64-bit one is a long VALU routine whose registers
  // count for the whole kernel.)
  uint32_t share_log2;
  // Next position of each share, on a cache line of their own: every wave's atomics hit
  // them, and the fields above are read by the kernels.
  alignas(64) uint32_t share_next[kXcds];
};

// Partial or final record of one chunk, as the kernels produce it.  maxnum is the max of
// the shifted value number = T-value + 1, so a record with count == 0 has maxnum == 0.
struct DevRec {
  uint64_t count;
  uint64_t stop_sum;
  uint64_t stop_xor;
  uint64_t maxnum[3];
  uint64_t max_start[2];
  uint64_t overflow;
  uint64_t checks;   // checks executed (for the roofline accounting)
  uint32_t hash;
  uint32_t pad;
};

// A candidate handed from the two-limb kernels to the three-limb pass: its start n and its
// output slot (the pass recomputes the survivor record of n mod 2^k).
struct EscEntry {
  uint64_t n_lo, n_hi;
  uint32_t slot;
  uint32_t pad;
};

struct DeviceLists {
  EscEntry* esc;              // escalation list
  unsigned long long* esc_count;
  uint64_t esc_cap;
  uint64_t* ovf;              // overflow reports, 4 words each (n_lo, n_hi, slot, 0)
  unsigned long long* ovf_count;
  uint64_t ovf_cap;
  unsigned long long* stall_count;  // candidates abandoned after kMaxChecks checks
  uint32_t* stall_slots;            // output slot of each abandoned candidate (the first
  uint64_t stall_cap;               // stall_cap of them), so a caller can name the range
  unsigned long long* loop_aborts;  // tiles abandoned at the work-loop bound (Limits)
};

// Abandoned candidates reported per launch.
constexpr uint64_t kStallCap = 256;

// Safety bound on checks per candidate.  Stopping times of all n < 2^128 that have been
// computed are a few thousand T-steps; a candidate still running after this many checks
// means a defect, and is counted instead of looping forever.
// Two-limb work whose starts stay below 2^68 runs on three 32-bit words (96 bits) and
// escalates to three limbs above 2^86 (see the per-check bounds below).  Survivors climb by
// 2^x above their start with probability about 2^-(x-5), so below 2^68 about one candidate
// in 10^6 escalates.
constexpr uint32_t kNarrowLog2 = 91;
constexpr uint32_t kNarrowMaxStartLog2 = 68;
// Per-check escalation bounds (a check multiplies by less than 2^9.4): 96 - 10 and 128 - 10.
constexpr uint32_t kNarrowCheckLog2 = 86;
constexpr uint32_t kWideCheckLog2 = 118;
// Two-limb work whose starts are above 2^100 runs on three limbs from the start (at the
// four-word check bound of 2^118, or the four-word start bound of 2^123 under a lower
// cutoff, too many of its candidates would escalate).
constexpr uint32_t kWideMaxStartLog2 = 100;
constexpr uint32_t kMaxChecks = 1u << 16;
// The same bound in T-steps for the persistent kernel: every half-check (odd run, even run)
// takes at least one step, so 2^18 steps mean at most 2^17 checks.
constexpr uint32_t kMaxSteps = 1u << 18;
constexpr uint32_t kMaxLoop = 1u << 31;

// Limits of the schedule for one launch (cutoff = CUTOFFPOWER, collatz_cudav7.cu:44-46).
struct Limits {
  uint32_t cutoff;       // overflow at number >= 2^cutoff
  uint32_t fast_word;    // index of the low word of the 64-bit pair checked on the fast path
  uint64_t fast_thr;     // fast path: pair >= fast_thr, or any word above the pair nonzero
  // A candidate that has not descended after max_checks checks (the persistent kernel
  // counts max_steps T-steps instead) is abandoned and the call fails: kMaxChecks and
  // kMaxSteps, or less under COLLATZ_CHECK_LIMIT (a test hook that forces the failure
  // path from the host).
  uint32_t max_checks;
  uint32_t max_steps;
  // Iterations of a kernel's work loop per tile or wave share (and of a take loop per
  // refill): kMaxLoop, or less under COLLATZ_LOOP_LIMIT (a test hook).  Every such loop makes
  // progress on each iteration, so real work never comes near the bound; a wave that reaches
  // it abandons its work, counts the loop abort, and exits normally, and the call fails.  The
  // plain kernel and the look-ahead form of the persistent kernel always carry the bound;
  // the persistent kernel's plain form and the fused reference-mode kernel carry it only
  // when it is below kMaxLoop (their counters cost 6% and an occupancy step respectively).
  uint32_t max_loop;
};

// One survivor as the range kernels read it, 16 bytes in one load:
//   x = r bits 0..31, y = f | (r bits 32..55) << 8, z:w = T^k(r) + 1.
// Survivors are stored class-major by r mod 9, each class sorted by r.
__host__ __device__ inline uint64_t rec_r(const uint4& v) {
  return v.x | ((uint64_t)(v.y >> 8) << 32);
}
__host__ __device__ inline uint32_t rec_f(const uint4& v) { return v.y & 0xFFu; }
__host__ __device__ inline uint64_t rec_M(const uint4& v) { return v.z | ((uint64_t)v.w << 32); }

struct SieveArrays {
  const uint4* rec;
  uint64_t count;
  uint32_t k;
  // Look-ahead filter (null: none): nibble g per residue of number mod 2^B, keep iff
  // f - la_emin >= g.
  const uint32_t* la;
  uint32_t la_mask;   // 2^B - 1
  uint32_t la_emin;   // ceil(306 k / 485)
  // Optional first level (null: none): the same table at 16 bits.  Under the CUDA
  // compares a dip over 16 steps is at most the dip over B, so a candidate the 16-bit
  // table drops the B-bit one drops too, and only the survivors read the large table.
  const uint32_t* la1;
  uint32_t la1_mask;
};

// Leaves of a reference-mode run on the device: base, number = T^(N-A)(base) + 1 (its high
// words in num_hi when the walk was 128-bit wide, else null) and the meta word.
struct RefLeaves {
  const uint64_t* base;
  const uint64_t* num;
  const uint64_t* num_hi;
  const uint32_t* meta;
};

// Spot checks of a reference-mode launch (the reference re-runs a random launch's
// candidates on the CPU and compares the hash, collatz_cudav7.cu:619-635): tiles
// t = off + j freq with j < cap (freq a power of two) record their own results in rec[j],
// and the host compares them with a rerun of the same tiles' candidates by the list kernel.
// The fused kernel reads these from its totals block in device memory.
struct RefSpotRec {
  uint32_t hash, count, numcases, overflow;
};
struct RefSpot {
  uint32_t freq;  // 0: no tile is sampled
  uint32_t off;   // < freq
  uint32_t cap;
  RefSpotRec* rec;
};

// Look-ahead tables of reference mode on the device.
struct RefTablesDev {
  const uint32_t* bv;      // num_bv vectors of words_per_bv words (2^B bits + mirrored tail)
  const uint32_t* mod9;    // 9 rows of nw words
  const uint32_t* magic;   // 3^{-(tpo+i)} mod 2^32
  uint64_t words_per_bv;
  uint32_t nw;             // words per window (ceil(2^A / 32))
  uint32_t num_bv;
  uint32_t tpo;
  uint32_t B;
  uint32_t A;
  uint32_t rust;
};

// ---- launchers (defined in the .hip files) ------------------------------------------
// Walks below depth kWalk64MaxDepth keep numbers in 64 bits (T^40(n) + 1 < 2^64); deeper
// ones (reference mode up to N - A = 62) in 128, with the high words in the _hi arrays.
constexpr uint32_t kWalk64MaxDepth = 40;
constexpr uint32_t kWalkMaxDepth = 62;
hipError_t launch_sieve_prefix(const SieveRoot& root, uint32_t p, uint64_t* pnum,
                               uint64_t* pnum_hi, uint32_t* pmeta, uint64_t* alive_bits,
                               hipStream_t s);
hipError_t launch_sieve_expand(const SieveRoot& root, uint32_t p, const uint64_t* pnum,
                               const uint64_t* pnum_hi, const uint32_t* pmeta,
                               const uint64_t* alive_bits, uint32_t* bits, hipStream_t s);
hipError_t launch_compact_count(const uint32_t* bits, uint64_t nwords, uint32_t* block_counts,
                                uint64_t nblocks, hipStream_t s);
hipError_t launch_scan(const uint32_t* in, uint64_t n, uint64_t* out_excl, hipStream_t s);
hipError_t launch_compact_write(const uint32_t* bits, uint64_t nwords,
                                const uint64_t* block_offsets, uint64_t* out_h, hipStream_t s);
hipError_t launch_sieve_fill(const SieveRoot& root, uint32_t p, const uint64_t* pnum,
                             const uint64_t* pnum_hi, const uint32_t* pmeta,
                             uint64_t* res_inout, uint64_t* num, uint64_t* num_hi,
                             uint32_t* meta, uint64_t count, hipStream_t s);
hipError_t launch_lower_bound(const uint4* sorted, const uint64_t* class_start,
                              const uint64_t* keys, const uint32_t* kcls, uint64_t nkeys,
                              uint64_t* out, hipStream_t s);
// Replaces flagged row bounds by their lower_bound in the row's class.
// Same, for ctl->nrows rows, with a fixed grid (graph-capturable).
hipError_t launch_rows_resolve_ctl(RowDesc* rows, const Ctl* ctl, const uint4* sorted,
                                   const uint64_t* class_start, const uint64_t* keys,
                                   hipStream_t s);
hipError_t launch_class_partition(const uint64_t* res, const uint64_t* num, const uint32_t* meta,
                                  uint64_t n, uint4* orec, uint32_t* counts, uint64_t* offs, uint64_t* class_start_host,
                                  hipStream_t s);

// Range kernels: `limbs` is the word width, 1 = three 32-bit words (96 bits, escalating at
// 2^91), 2 = two u64 limbs (escalating at 2^123), 3 = three u64 limbs.
// The plain kernel runs one wave per share: row r owns shares [rows[r].block_start,
// rows[r + 1].block_start) (equal pieces of its survivors), and share g's record goes to
// share_recs[g].  `lowkey`: the range ends at or below 2^kLowKeyLog2 (the wave max's filter
// then tests the values' low words; three- and four-word launches).
constexpr unsigned kLowKeyLog2 = 35;
hipError_t launch_verify_plain(int limbs, const SieveArrays& sv, const RowDesc* rows,
                               uint32_t nrows, uint64_t nshares, uint32_t mod9, bool lowkey,
                               const Limits& lim, const uint64_t* pow3, DevRec* share_recs,
                               DeviceLists lists, hipStream_t s);
// `order` lists the ctl->ntiles tile indices in the order they are to be taken (tile t's
// record goes to tile_recs[t] whatever its position).  `lowkey` as for the plain kernel.
hipError_t launch_verify_persistent(int limbs, const SieveArrays& sv, const RowDesc* rows,
                                    const TileDesc* tiles, const uint32_t* order, Ctl* ctl,
                                    uint32_t nblocks, uint32_t U, bool lowkey,
                                    const Limits& lim, const uint64_t* pow3, DevRec* tile_recs,
                                    DeviceLists lists, hipStream_t s);
// Clears the tile counters and the list counters with a kernel (no memset nodes, so the
// sequence can be graph-captured).
// Waves per CU that stay resident for the persistent kernel of this width and cutoff.
uint32_t persistent_waves_fit(int limbs, uint32_t cutoff, bool la);
hipError_t launch_engine_reset(Ctl* ctl, unsigned long long* cnt, uint32_t ncnt, hipStream_t s);
// Expands rows into tiles: row r owns tiles [rows[r].block_start, rows[r+1].block_start)
// (the last row up to ctl->ntiles).
hipError_t launch_make_tiles(const RowDesc* rows, const Ctl* ctl, TileDesc* tiles, hipStream_t s);
// Per-slot reduction for ctl->nslots slots, fixed grid.
hipError_t launch_reduce_slots(const DevRec* parts, const uint64_t* part_start, const Ctl* ctl,
                               DevRec* out, hipStream_t s);
hipError_t launch_bv_build(uint32_t B, uint32_t num_bv, int32_t first, uint32_t sem,
                           const uint32_t* magic, uint32_t* bv, uint64_t words_per_bv,
                           uint32_t mirror_words, hipStream_t s);
hipError_t launch_la_build(uint32_t B, uint32_t k, uint32_t sem, uint32_t* table,
                           hipStream_t s);
hipError_t launch_leaf_bins(const uint32_t* meta, uint64_t n, uint32_t tpo,
                            unsigned long long* bins, hipStream_t s);
// Reference mode: Steps 2 and 3 (rust: Step 3 is main.rs's search3), and the wide rerun of
// escalated candidates.  pow3w holds 3^f for f < 64 as (low, high) word pairs.  U checks run
// between refills (kRefChecksPerRefill, compiled in on 96 bits, unless told otherwise).
constexpr uint32_t kRefChecksPerRefill = 5;
hipError_t launch_ref_fused(int limbs, bool rust, const RefTablesDev& T, const RefLeaves& lv,
                            uint64_t nleaves, uint32_t mbi, const Limits& lim,
                            const uint64_t* pow3w, uint32_t tile, uint32_t U, uint32_t nblocks,
                            void* tot, uint64_t* ovf, uint64_t ovf_cap, uint64_t* esc,
                            uint64_t esc_cap, hipStream_t s);
// Lists the candidates of the sampled tiles (tile leaves each) of a fused launch as
// (leaf | a << 32 | no-bit-vector << 48) entries, appending at *count (at most cap kept).
hipError_t launch_ref_spot_list(const RefTablesDev& T, const RefLeaves& lv, uint64_t nleaves,
                                uint32_t tile, const RefSpot& spot, uint64_t* list,
                                unsigned long long* count, uint64_t cap, hipStream_t s);
hipError_t launch_ref_iterate(int limbs, bool rust, const uint64_t* cands, uint64_t ncand,
                              const RefLeaves& lv, uint32_t mbi, const Limits& lim,
                              const uint64_t* pow3w, uint32_t tile, uint32_t nblocks, void* tot,
                              uint64_t* ovf, uint64_t ovf_cap, hipStream_t s);
hipError_t launch_reduce_chunks(const DevRec* parts, const uint64_t* part_start,
                                uint32_t nchunks, DevRec* out, hipStream_t s);
hipError_t launch_escalated(const SieveArrays& sv, const EscEntry* esc, uint64_t n,
                            uint32_t mod9_unused, const Limits& lim, const uint64_t* pow3,
                            DevRec* out, DeviceLists lists, hipStream_t s);

}  // namespace cz
