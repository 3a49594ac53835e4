/*
 * collatz.h -- C ABI of the MI355X Collatz range verifier (libcollatz_amd.so).
 *
 * Plain C types only: integers, pointers, sizes.  Streams are HIP streams passed as
 * `void*`; NULL selects the calling thread's per-thread default stream (hipStreamPerThread),
 * which runs after the caller's earlier default-stream work: the library enqueues nothing
 * on the legacy default stream, which the runtime blocks while another thread captures a
 * hipGraph (a server does).  Every entry point returns COLLATZ_OK (0) or a
 * negative status; collatz_last_error() returns a message for the calling thread.
 * No exceptions cross this boundary and no kernel prints.
 *
 * The reference (vigleik0/collatz) has no library interface; each entry point below
 * replaces a piece of its host program, cited at the declaration.
 */
#ifndef COLLATZ_H
#define COLLATZ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define COLLATZ_ABI_VERSION 2

enum collatz_status {
  COLLATZ_OK = 0,
  COLLATZ_EINVAL = -1,    /* bad argument */
  COLLATZ_ENOMEM = -2,    /* device or host allocation failed */
  COLLATZ_EDEVICE = -3,   /* a HIP call failed */
  COLLATZ_ECAPACITY = -4, /* an output buffer was too small (the count is still reported) */
  COLLATZ_ENOTREADY = -5  /* server: the request is not finished yet */
};

enum collatz_semantics {
  COLLATZ_SEM_CUDA = 0,   /* collatz_cudav7.cu */
  COLLATZ_SEM_RUST = 1,   /* main.rs: odd-even-even term in the bit vectors, extra filters */
  COLLATZ_SEM_OPENCL = 2  /* collatz_openclv2.cpp: unsigned compares in precompute */
};

enum collatz_flags {
  COLLATZ_FLAG_OVERFLOW = 1u, /* a trajectory reached 2^cutoff_log2 (reported, then reset) */
  COLLATZ_FLAG_WIDE = 2u      /* a trajectory reached 2^123 and ran on three u64 limbs */
};

/* ---- sieve ---------------------------------------------------------------------------
 * Replaces the reference's Step 1: search1() from the root "3 2 2 9 1 0" down to depth k
 * (collatz_cudav7.cu:497-563, :173), evaluated on the device for every residue class,
 * stored as a bit array and compacted into a sorted survivor list (r, T^k(r)+1, f) that
 * stays resident in device memory.
 *
 * With lookahead_bits = B the sieve also carries the reference's Step 2 filter
 * (precompute() BV_i, collatz_cudav7.cu:423-480, applied at :219-235) at depth k: a
 * candidate n with survivor (r, f) and start value number = T^k(n)+1 is iterated only if
 * the B-step dip of number mod 2^B (the BV_i "maxval") is at most 485 f - 306 k.  At
 * k = N - A over [1, 2^N) this is exactly the reference's candidate set.  Every
 * verification with such a sieve applies it; counts and checksums cover the candidates it
 * keeps. */
typedef struct collatz_sieve collatz_sieve;

typedef struct {
  uint32_t k;       /* sieve depth: residues mod 2^k, 8 <= k <= 38 */
  uint32_t mod9;    /* 1: apply the mod 9 preimage sieve (collatz_cudav7.cu:482-494) */
  int32_t device;   /* HIP device ordinal, -1 = current device */
  uint32_t lookahead_bits;      /* 0 = none, or B in [8, 24]: a B-step look-ahead filter */
  uint32_t lookahead_semantics; /* COLLATZ_SEM_CUDA or COLLATZ_SEM_OPENCL */
  uint32_t reserved;
} collatz_sieve_params;

typedef struct {
  uint32_t k, mod9;
  uint64_t survivors;      /* |S_k| */
  uint64_t bitarray_bits;  /* bits of the dense array (classes r = 3 mod 4) */
  uint64_t device_bytes;   /* device memory held by the handle */
  double build_ms;         /* wall time of collatz_sieve_build */
  uint32_t lookahead_bits, lookahead_semantics;
} collatz_sieve_info;

int collatz_sieve_build(const collatz_sieve_params* params, collatz_sieve** out);
int collatz_sieve_query(const collatz_sieve* sieve, collatz_sieve_info* info);
/* Copies up to cap survivors (residue, T^k(r)+1, f) to host memory; for tests and tools. */
int collatz_sieve_copy(const collatz_sieve* sieve, uint64_t* r, uint64_t* number, uint8_t* f,
                       uint64_t cap);
int collatz_sieve_free(collatz_sieve* sieve);

/* ---- range verification --------------------------------------------------------------
 * One record per aligned chunk [j*2^C, (j+1)*2^C) intersected with the range.  Every
 * n in the chunk whose residue mod 2^k survives the sieve and (with mod9) whose n mod 9
 * is in {0,1,3,6,7} is iterated with the reference schedule (collatz_cudav7.cu:280-366)
 * from T^k(n)+1 (Lemma 2.3) until the check number < n passes. */
typedef struct {
  uint64_t chunk_index;    /* j (its low 64 bits, for chunks beyond 2^(64 + C)) */
  uint64_t n_checked;      /* integers of the range in this chunk */
  uint64_t count;          /* candidates iterated */
  uint64_t stop_sum;       /* sum over candidates of J = k + T-steps to the passing check */
  uint64_t stop_xor;       /* xor of the same J */
  uint64_t max_exc[3];     /* largest T-value seen (at the start or after an odd run) */
  uint64_t max_start[2];   /* smallest start n that reached max_exc */
  uint64_t overflow_count; /* times a value reached 2^cutoff_log2 */
  uint32_t hash;           /* xor of the per-candidate reference hashes */
  uint32_t flags;          /* collatz_flags */
} collatz_chunk;

typedef struct {
  collatz_chunk total;        /* merge of all chunks; chunk_index = first chunk */
  collatz_chunk* chunks;      /* caller-owned array, or NULL */
  uint64_t chunks_cap;
  uint64_t chunks_written;    /* number of chunks of the range (even if > chunks_cap) */
  uint64_t* overflow_starts;  /* caller-owned, two words (lo, hi) per start, or NULL */
  uint64_t overflow_cap;
  uint64_t overflow_written;  /* starts reported (even if > overflow_cap) */
  uint64_t escalated;         /* candidates moved to three limbs */
  uint64_t checks;            /* checks (two odd and two even runs) executed */
  double device_ms;           /* device time of the launch: the kernels that set up, verify
                                 and reduce (inputs already in device memory) plus the
                                 three-limb pass of escalated candidates */
  uint64_t spot_checks;       /* chunk records of this result re-run by the other kernel form
                                 and found equal (collatz_batch_opts.spot_checks) */
} collatz_range_result;

enum collatz_form {
  COLLATZ_FORM_PLAIN = 0,     /* fixed blocks over the range's survivors, lanes refilled within
                                 each wave's share: the drop-in kernel */
  COLLATZ_FORM_PERSISTENT = 1 /* persistent waves with lane refill and work stealing */
};

typedef struct {
  uint32_t chunk_log2;   /* C, 0 = 24 */
  uint32_t limbs;        /* 2 (three on demand) or 3; 0 = 2 */
  uint32_t cutoff_log2;  /* CUTOFFPOWER, 64..187; 0 = 187 */
  uint32_t form;         /* collatz_form */
} collatz_verify_opts;

/* Replaces launchkernel() + compute_kernel (collatz_cudav7.cu:565-649, :185-411):
 * [lo, hi) with 128-bit endpoints lo = lo[0] + 2^64 lo[1].  The call returns when its
 * results are in `out`.  Calls on one sieve handle from several threads are safe: the
 * plain form keeps its device buffers on the handle (grown on demand, reused by later
 * calls) and runs one call at a time; the persistent form and collatz_verify_batch share
 * the handle's launch engine the same way. */
int collatz_verify_range(const collatz_sieve* sieve, const uint64_t lo[2], const uint64_t hi[2],
                         const collatz_verify_opts* opts, void* stream,
                         collatz_range_result* out);

/* No counterpart in the reference.  For callers that cannot allocate device memory per call
 * (tight serving loops, harnesses that forbid hipMalloc): the plain form of
 * collatz_verify_range (opts->form must be COLLATZ_FORM_PLAIN) with identical results, whose
 * buffers come from a fixed device array of about 20 MiB per device, held from library load
 * until the process ends; the call itself allocates no device memory.  Limits per call: at
 * most 4096 chunk records and 32768 (q, chunk, residue class) rows (a range past them fails
 * with COLLATZ_EINVAL; the kernel's 98304 share records are sized to fit), and at most 4096
 * overflow starts listed (overflow_written is exact; more starts than that return
 * COLLATZ_ECAPACITY).  Calls on one device share the array and are serialised by a lock;
 * work enqueued on it by other means would race. */
int collatz_verify_range_noalloc(const collatz_sieve* sieve, const uint64_t lo[2],
                                 const uint64_t hi[2], const collatz_verify_opts* opts,
                                 void* stream, collatz_range_result* out);

/* ---- batched form --------------------------------------------------------------------
 * Many intervals in one persistent launch (no counterpart in the reference, which runs one
 * batch of 2^16 leaves per synchronous launch, collatz_cudav7.cu:532-533).  outs[i] gets
 * interval i's records exactly as collatz_verify_range would produce them. */
typedef struct {
  uint64_t lo[2];
  uint64_t hi[2];
} collatz_interval;

typedef struct {
  uint32_t chunk_log2;       /* 0 = 24 */
  uint32_t limbs;            /* 0 = 2 */
  uint32_t cutoff_log2;      /* 0 = 187 */
  uint32_t use_graph;        /* 1: capture the launch sequence in a hipGraph */
  uint32_t tile_residues;    /* residues per work tile; 0 = autotuned */
  uint32_t waves_per_cu;     /* persistent waves per CU; 0 = autotuned */
  uint32_t checks_per_refill;/* checks between lane refills; 0 = autotuned */
  uint32_t spot_checks;      /* chunk records (seeded choice among the stored ones) re-run by
                                the plain kernel after the launch and compared field by field
                                (the reference's spot check, collatz_cudav7.cu:619-635); a
                                difference returns COLLATZ_EDEVICE with the outputs filled in */
} collatz_batch_opts;

int collatz_verify_batch(const collatz_sieve* sieve, uint64_t n,
                         const collatz_interval* intervals, const collatz_batch_opts* opts,
                         void* stream, collatz_range_result* outs);

/* ---- reference mode ------------------------------------------------------------------
 * The reference's own full program for one case line "base sigfig threepower number
 * doubleeven numevens" (collatz_cudav7.cu:95-183): search1 to depth N - A, the BV_i
 * look-ahead and mod 9 filter of the top A bits (Step 2, :219-235), the schedule, the
 * complete hash and the launch count it would print.  The library keeps a run's device
 * buffers (leaf arrays, totals, lists) per device between calls and grows them as needed;
 * calls from several threads are safe, and one that finds its device's buffers in use
 * allocates its own for that call. */
typedef struct {
  uint64_t base[2];
  uint64_t number[2];
  uint32_t sigfig, threepower, doubleeven, numevens;
} collatz_node;

typedef struct {
  uint32_t N;            /* MAXBITS: N <= 64 and N - A <= 62 */
  uint32_t A;            /* LASTBITS, 1..12 */
  uint32_t B;            /* STEPSIZEINITIAL, 24; 8..28 and at least A */
  uint32_t num_bv;       /* bit vectors, 16 (CUDA/OpenCL) or 8 (Rust) */
  uint32_t semantics;    /* collatz_semantics */
  uint32_t cutoff_log2;  /* CUTOFFPOWER, 64..187 (main.rs semantics: 64..180, main.rs uses
                            120: the threshold of its (number >> 64) >= 2^56 test) */
  int32_t device;
  uint32_t spot_check_freq; /* 0: none.  Else spot checks (collatz_cudav7.cu:619-635): one
                               tile of 256 leaves in every P, P the largest power of two not
                               above spot_check_freq, at a seeded offset, is re-run by a second
                               kernel and its hash, count, block tally and overflow count
                               compared (at most 2^23 candidates per root launch).  The
                               reference checks one launch of 2^16 leaves in 10000
                               (SPOTCHECKFREQUENCY, :55), i.e. one tile in 2,560,000. */
} collatz_ref_params;

/* Under COLLATZ_SEM_RUST Step 3 is main.rs's own search3 (main.rs:321-341): each candidate
 * runs in blocks of 16 steps until it falls below its start, or until a block ends at or
 * above 2^cutoff_log2, which reports its start (main.rs:336-338 prints it in decimal) and
 * stops it.  numcases is the sum of the blocks (main.rs:97-98, "numcases = %d") and
 * complete_hash is 0 (main.rs has no hash). */
typedef struct {
  uint32_t complete_hash;     /* "Complete hash = %X" */
  uint32_t pad;
  uint64_t launches;          /* "Number of kernel launches = %ld" */
  uint64_t leaves;            /* leaves at depth N - A */
  uint64_t candidates;        /* candidates iterated */
  uint64_t checks;            /* checks run (a candidate rerun wider counts both runs) */
  uint64_t bin_leaves[17];    /* leaves per bigindex 0..15, and >= 16 */
  uint64_t overflow_count;
  uint64_t* overflow_starts;  /* caller-owned, two words per start, or NULL */
  uint64_t overflow_cap;
  uint64_t overflow_written;
  double device_ms;
  uint64_t numcases;          /* main.rs semantics: its "numcases"; 0 otherwise */
  uint64_t spot_checks;       /* tiles re-run and compared ("Spot checks = %u") */
  uint64_t spot_check_failures; /* of them, tiles whose rerun differed ("Spot check failed");
                                   the call then returns COLLATZ_EDEVICE with every output
                                   filled in */
} collatz_ref_result;

int collatz_reference_run(const collatz_ref_params* params, const collatz_node* root,
                          void* stream, collatz_ref_result* out);

/* The look-ahead tables reference mode uses, built on the device: precompute()
 * (collatz_cudav7.cu:423-495).  bv receives num_bv vectors of 2^B bits as 2^(B-5) u32 words
 * each (the mirrored tail is not copied), mod9 nine rows of ceil(2^A / 32) words. */
int collatz_reference_tables(const collatz_ref_params* params, uint32_t* bv, uint32_t* mod9);

/* ---- served form ---------------------------------------------------------------------
 * A request loop: submit ranges at any time; a host thread packs the pending work of all
 * admitted requests into back-to-back persistent launches (continuous batching) while the
 * sieve stays resident.  Requests are admitted while their device footprint fits. */
typedef struct collatz_server collatz_server;

typedef struct {
  uint32_t chunk_log2;       /* 0 = 24 */
  uint32_t cutoff_log2;      /* 0 = 187 */
  uint64_t max_device_bytes; /* admission budget for per-request buffers; 0 = 1 GiB */
  uint64_t slice_candidates; /* candidates per launch (up to 64x more while the backlog is
                                large); 0 = 2^26 */
  uint32_t max_inflight;     /* admitted requests at once; 0 = 1024 */
  uint32_t reserved;
} collatz_server_opts;

typedef struct {
  collatz_chunk total;
  uint64_t chunks;           /* chunk records of the request */
  double submit_ms;          /* host clock at submission (ms since server start) */
  double admit_ms;           /* when it was admitted */
  double done_ms;            /* when its last record was folded */
} collatz_request_result;

int collatz_server_create(const collatz_sieve* sieve, const collatz_server_opts* opts,
                          collatz_server** out);
/* Returns a ticket >= 0, or a negative status. */
int64_t collatz_server_submit(collatz_server* srv, const collatz_interval* range);
/* Waits up to timeout_ms (< 0: forever) for the ticket; COLLATZ_ENOTREADY on timeout. */
int collatz_server_wait(collatz_server* srv, int64_t ticket, double timeout_ms,
                        collatz_request_result* out, collatz_chunk* chunks, uint64_t cap);
int collatz_server_destroy(collatz_server* srv);

/* ---- misc ---------------------------------------------------------------------------- */
int collatz_version(void);
const char* collatz_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* COLLATZ_H */
