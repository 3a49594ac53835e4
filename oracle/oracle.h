/*
 * CPU oracle for the Collatz range verifier.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (collatz_amd/, include/) may link,
 * load or call this code.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker.
 *
 * It restates, in plain C with 64/128-bit integers, the algorithm of the reference
 * (vigleik0/collatz):
 *   - search1 (Step 1, the 2^k residue sieve)        collatz_cudav7.cu:497-563
 *   - precompute (BV_i look-ahead bit vectors, mod 9) collatz_cudav7.cu:423-495
 *   - doubleeven (odd-even-even test)                collatz_cudav7.cu:413-421, main.rs:101-107
 *   - search2/search3 (per-leaf candidates, schedule) collatz_cudav7.cu:654-715
 *   - compute_kernel schedule and hash               collatz_cudav7.cu:280-409
 * and the range-mode outputs that the reference does not have (count, stopping-time
 * checksums, max excursion and its start, overflow flags per aligned chunk).
 *
 * Pins (each is checked by tests/test_oracle_*.py):
 *   or_traj            against the compiled reference search3 (oracle/_ref), and against
 *                      naive step-by-step T iteration with Python integers.
 *   or_precompute_bv   against the compiled reference precompute (C++ semantics, N=45),
 *                      and against the paper's Fig 4.1 popcounts (Rust semantics, N=72).
 *   or_la_dip          through or_precompute_bv, which is built on it.
 *   or_verify_range_la over [1, 2^N) at k = N - A: the same candidate set, count and hash
 *                      as the reference-mode run of the whole tree (or_case_run on the
 *                      root), which is pinned to the compiled reference.
 *   or_mod9_rows       against the compiled reference, and against n mod 9 computed
 *                      directly.
 *   or_search1         against the compiled reference search1 leaves (case lines, N=45),
 *                      the case-file partition identity, and the paper's Fig 5.1 counts.
 *   or_leaf_hash       against the compiled reference search2.
 *   or_rust_search3    against the shifted map iterated step by step with Python integers
 *                      (the tally is the first multiple of 16 steps at which the value is
 *                      below the start or at least 2^cutoff), including starts near the
 *                      cutoff; the 16-step table itself against the same step-by-step map.
 *   or_verify_range    against brute-force T iteration (final value, max excursion,
 *                      descent) and a brute-force candidate set on small ranges.
 *   or_sieve_windows   against the windows of or_sieve_sorted (single windows and sets of
 *                      them at k = 16 and k = 20), which is pinned by the sieve pins above.
 */
#ifndef COLLATZ_ORACLE_H
#define COLLATZ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_NUMERATOR 306u   /* 306/485 ~ log 2 / log 3, collatz_cudav7.cu:50-52 */
#define OR_DENOMINATOR 485u
#define OR_WIDE_LOG2 123u   /* trajectories at or above 2^123 need a third u64 limb */

/* OR_SEM_CXX: collatz_cudav7.cu.  OR_SEM_OPENCL: collatz_openclv2.cpp, whose precompute
 * compares `curval + const > maxval` in unsigned arithmetic (collatz_openclv2.cpp:319,331),
 * so a negative path-merge term can lower maxval.  OR_SEM_RUST: main.rs. */
enum { OR_SEM_CXX = 0, OR_SEM_RUST = 1, OR_SEM_OPENCL = 2 };

/* A node of the search1 tree: n0 = base (mod 2^sigfig), f = threepower,
 * number = T^sigfig(base) + 1, and the two sieve states. */
typedef struct {
  uint64_t base_lo, base_hi;
  uint64_t number_lo, number_hi;
  uint32_t sigfig, threepower, doubleeven, numevens;
} or_node;

/* Knobs of one reference-mode run (collatz_cudav7.cu:43-55, main.rs:18-31). */
typedef struct {
  uint32_t N;          /* MAXBITS */
  uint32_t A;          /* LASTBITS */
  uint32_t B;          /* STEPSIZEINITIAL */
  uint32_t num_bv;     /* 16 for the C++ programs, 8 for main.rs */
  uint32_t semantics;  /* OR_SEM_CXX, OR_SEM_RUST or OR_SEM_OPENCL */
  uint32_t cutoff;     /* CUTOFFPOWER */
} or_knobs;

/* Per-trajectory result of the reference schedule. */
typedef struct {
  uint32_t hash;        /* XOR of low 32 bits of number at every check point */
  uint32_t overflow;    /* number of times the value reached 2^cutoff */
  uint64_t steps;       /* T-steps taken after the start value, up to the passing check */
  uint64_t maxnum[3];   /* max of number (= T-value + 1) over start and odd-run ends */
  uint64_t final[3];    /* number at the passing check */
} or_traj_out;

/* Range-mode record of one aligned chunk [j*2^C, (j+1)*2^C) intersected with the range. */
typedef struct {
  uint64_t chunk_index;     /* j */
  uint64_t n_checked;       /* integers of the range in this chunk */
  uint64_t count;           /* candidates iterated (survived both sieves) */
  uint64_t stop_sum;        /* sum of stopping steps J */
  uint64_t stop_xor;        /* xor of stopping steps J */
  uint64_t max_exc[3];      /* max T-value seen, 192-bit */
  uint64_t max_start[2];    /* smallest start reaching max_exc */
  uint64_t overflow_count;  /* candidates whose value reached 2^cutoff */
  uint32_t hash;            /* xor of the per-candidate reference hashes */
  uint32_t flags;           /* bit0: overflow_count > 0; bit1: max_exc >= 2^123 - 1 */
} or_chunk;

/* --- arithmetic helpers ------------------------------------------------------ */
void or_pow3(uint32_t f, uint64_t out[2]);
uint32_t or_magic(uint32_t e);                /* 3^{-e} mod 2^32 */
uint32_t or_tpo(uint32_t mbi);                /* THREEPOWEROFFSET */

/* --- Step 1: search1 --------------------------------------------------------- */
/* Enumerate the leaves at depth `depth` below `root`; returns the number of leaves, or -1
 * on bad input.  If `leaves` is non-null up to `cap` leaves are stored in DFS order. */
int64_t or_search1(const or_node* root, uint32_t depth, uint32_t semantics,
                   or_node* leaves, int64_t cap);
/* Leaves from the canonical root (3, 2, 2, 9, 1, 0) at depth k, sorted by residue. */
int64_t or_sieve_sorted(uint32_t k, uint64_t* r, uint64_t* number, uint8_t* f, int64_t cap);
/* The survivors with r >> w in tops[0..ntops) (2 <= w <= k), sorted by r, walked on
 * `threads` threads; written only if their number (the return value) is at most cap. */
int64_t or_sieve_windows(uint32_t k, uint32_t w, uint32_t ntops, const uint64_t* tops,
                         int threads, uint64_t* r, uint64_t* number, uint8_t* f, int64_t cap);

/* --- precompute -------------------------------------------------------------- */
/* bv: num_bv arrays of 2^B bits each, as 2^(B-6) u64 words per array (logical bit order:
 * bit j of array i is word j/64, bit j%64). */
/* B-step dip ("maxval") of residue n0 of number, as precompute() computes it per residue
 * (collatz_cudav7.cu:436-467; unsigned compares for OR_SEM_OPENCL). */
int32_t or_la_dip(uint64_t n0, uint32_t B, uint32_t semantics);
int or_precompute_bv(const or_knobs* kn, uint64_t* bv);
/* rows: 9 rows of 2^A bits as 2^(A-6) (at least 1) u64 words per row. */
int or_mod9_rows(uint32_t mbi, uint32_t A, uint64_t* rows);
int or_doubleeven(uint32_t upordown, uint32_t i, uint32_t B, uint32_t semantics);

/* --- Step 3: schedule -------------------------------------------------------- */
/* start value num0 (3 limbs), start integer base (2 limbs).  second_gt selects the
 * search3 variant that compares the second odd run with '>' (collatz_cudav7.cu:702). */
void or_traj(const uint64_t base[2], const uint64_t num0[3], uint32_t cutoff,
             int second_gt, or_traj_out* out);

/* --- reference mode ---------------------------------------------------------- */
typedef struct {
  uint32_t complete_hash;
  uint32_t pad;
  uint64_t launches;        /* kernel launches the reference would make */
  uint64_t leaves;          /* leaves at depth N - A */
  uint64_t candidates;      /* candidates iterated */
  uint64_t bin_leaves[17];  /* leaves per bigindex 0..15, and >= 16 */
  uint64_t overflow_count;
  uint64_t numcases;        /* main.rs semantics: the printed "numcases" (16-step blocks) */
} or_case_out;

/* One leaf (search2): hash of all candidates of the leaf. */
uint32_t or_leaf_hash(const or_knobs* kn, const uint64_t* bv, const uint64_t* mod9,
                      const or_node* leaf, uint64_t* ncand);
/* A whole case line.  bv/mod9 as produced above for the same knobs.  Under OR_SEM_RUST
 * Step 3 is main.rs's search3 (or_rust_search3): complete_hash stays 0, numcases is the
 * sum of its tallies and cutoff is its overflow threshold (main.rs: 120). */
int or_case_run(const or_knobs* kn, const uint64_t* bv, const uint64_t* mod9,
                const or_node* root, or_case_out* out);
/* The same, also storing the start n of up to ovf_cap overflow reports (two words each). */
int or_case_run_starts(const or_knobs* kn, const uint64_t* bv, const uint64_t* mod9,
                       const or_node* root, or_case_out* out, uint64_t* ovf_starts,
                       uint64_t ovf_cap);
/* or_case_run_starts on `threads` threads (the subtree split at a frontier 10 levels below
 * the root); the same results, the listed starts in another order. */
int or_case_run_mt(const or_knobs* kn, const uint64_t* bv, const uint64_t* mod9,
                   const or_node* root, or_case_out* out, uint64_t* ovf_starts,
                   uint64_t ovf_cap, int threads);
/* main.rs's search3 (main.rs:321-341) from start value num0 for start base: returns 1 if a
 * block ended at or above 2^cutoff (main.rs prints the start then), 0 if the value fell
 * below base; *tally = blocks of 16 steps run.  cutoff <= 180. */
int or_rust_search3(const uint64_t base[2], const uint64_t num0[3], uint32_t cutoff,
                    uint64_t* tally);
/* main.rs semantics: number of candidates that reach search3 (paper Fig 5.1). */
int64_t or_rust_count(const or_knobs* kn, const uint64_t* bv, const uint64_t* mod9,
                      const or_node* root);

/* --- range mode -------------------------------------------------------------- */
/* Verify [lo, hi) with a sorted sieve (r, number, f) of depth k.  chunks must hold
 * ((hi-1)>>C) - (lo>>C) + 1 records.  total receives the merge of all chunks.
 * threads <= 1 runs on the calling thread. */
int or_verify_range(uint32_t k, int64_t nres, const uint64_t* r, const uint64_t* number,
                    const uint8_t* f, int mod9, const uint64_t lo[2], const uint64_t hi[2],
                    uint32_t C, uint32_t cutoff, int threads, or_chunk* chunks,
                    or_chunk* total);
/* The same with the Step 2 look-ahead at depth k (la_bits = B > 0): a candidate with
 * survivor (r, f) and start value number is iterated only if
 * or_la_dip(number mod 2^B, B, la_sem) <= 485 f - 306 k.  la_sem: OR_SEM_CXX or
 * OR_SEM_OPENCL. */
int or_verify_range_la(uint32_t k, int64_t nres, const uint64_t* r, const uint64_t* number,
                       const uint8_t* f, int mod9, uint32_t la_bits, uint32_t la_sem,
                       const uint64_t lo[2], const uint64_t hi[2], uint32_t C,
                       uint32_t cutoff, int threads, or_chunk* chunks, or_chunk* total);
void or_chunk_merge(or_chunk* into, const or_chunk* from);

#ifdef __cplusplus
}
#endif
#endif
