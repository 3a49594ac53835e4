/*
 * CPU oracle for the Collatz range verifier -- TEST INFRASTRUCTURE ONLY.
 * See oracle.h for what each function restates and how it is pinned.
 *
 * Plain C11 with unsigned __int128.  Deliberately written as straightforward scalar code
 * (u64 limbs, one T-run at a time) so that it shares no arithmetic with the HIP kernels
 * (u32 words, funnel shifts, wave-wide reductions).
 */
#include "oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

static inline u128 mk128(uint64_t lo, uint64_t hi) { return ((u128)hi << 64) | lo; }

/* 3^f as u128, f <= 80 (THREEPOWERS, collatz_cudav7.cu:63,104-105). */
static u128 pow3_128(uint32_t f) {
  u128 x = 1;
  while (f--) x *= 3;
  return x;
}

void or_pow3(uint32_t f, uint64_t out[2]) {
  u128 x = pow3_128(f);
  out[0] = (uint64_t)x;
  out[1] = (uint64_t)(x >> 64);
}

/* 3^{-e} mod 2^32: ONETHIRD = 2863311531 raised to e (collatz_cudav7.cu:66,122-128). */
uint32_t or_magic(uint32_t e) {
  uint32_t m = 1;
  while (e--) m *= 2863311531u;
  return m;
}

/* THREEPOWEROFFSET = floor(MBI * 306 / 485) + 1 (collatz_cudav7.cu:67). */
uint32_t or_tpo(uint32_t mbi) { return mbi * OR_NUMERATOR / OR_DENOMINATOR + 1; }

/* ------------------------------------------------------------------------------------
 * search1: collatz_cudav7.cu:497-563 (C++), main.rs:206-278 (Rust).
 * number = T^sigfig(base) + 1.  Shifted map: number even -> odd step number*3/2,
 * number odd -> even step number/2 + 1 (collatz_cudav7.cu:545-561).
 * ---------------------------------------------------------------------------------- */
typedef void (*leaf_fn)(void* ctx, const or_node* leaf);

typedef struct {
  uint32_t depth, sem;
  leaf_fn fn;
  void* ctx;
  int64_t count;
  u128 p3[128];
  /* Residue windows (or_sieve_windows): at depths s >= win_bits, bits win_bits..s of base
   * must be the low bits of one of the ntops values tops[] (base >> win_bits == top at the
   * leaves).  ntops = 0: no window. */
  uint32_t win_bits;
  uint32_t ntops;
  const uint64_t* tops;
} s1_ctx;

/* Whether a path whose bits win_bits..s-1 are `pre` may take bit s = b. */
static int s1_win_ok(const s1_ctx* c, u128 base, uint32_t s, uint32_t b) {
  if (!c->ntops || s < c->win_bits) return 1;
  const uint32_t d = s - c->win_bits;
  const uint64_t m = d ? (d >= 64 ? ~0ull : (1ull << d) - 1) : 0;
  const uint64_t pre = (uint64_t)(base >> c->win_bits) & m;
  for (uint32_t i = 0; i < c->ntops; ++i)
    if ((c->tops[i] & m) == pre && ((c->tops[i] >> d) & 1) == b) return 1;
  return 0;
}

static void s1_rec(s1_ctx* c, u128 base, uint32_t s, uint32_t f, u128 number, uint32_t de,
                   uint32_t ne) {
  /* main.rs:216, Rust only (never removes a leaf at the depths used here). */
  if (c->sem == OR_SEM_RUST && number <= base) return;
  /* Induction sieve: 485 f <= 306 k (collatz_cudav7.cu:507). */
  if (OR_DENOMINATOR * f <= OR_NUMERATOR * s) return;
  /* Mod 3 preimage / path merging sieve (collatz_cudav7.cu:509-511). */
  if (ne >= 2 && ne % 2 == 0) {
    if (OR_NUMERATOR * (s - 1) >= OR_DENOMINATOR * (f - 1)) return;
  }
  if (s == c->depth) {
    c->count++;
    if (c->fn) {
      or_node n;
      n.base_lo = (uint64_t)base;
      n.base_hi = (uint64_t)(base >> 64);
      n.number_lo = (uint64_t)number;
      n.number_hi = (uint64_t)(number >> 64);
      n.sigfig = s;
      n.threepower = f;
      n.doubleeven = de;
      n.numevens = ne;
      c->fn(c->ctx, &n);
    }
    return;
  }
  /* Odd-even-even state (collatz_cudav7.cu:538-543). */
  uint32_t inplay = (OR_NUMERATOR * (s + 1) >= OR_DENOMINATOR * f) ? 1u : 0u;
  u128 nA = number;
  u128 nB = number + c->p3[f];
  u128 bit = (u128)1 << s;
  /* In residue windows only the children on the path of some window are walked. */
  const int w0 = s1_win_ok(c, base, s, 0);
  const int w1 = s1_win_ok(c, base, s, 1);
  if ((nA & 1) == 0) {
    if (w0) s1_rec(c, base, s + 1, f + 1, nA + nA / 2, (de | inplay) & ~2u, 0);
    if (w1 && de < 2) s1_rec(c, base | bit, s + 1, f, nB / 2 + 1, de << 1, ne + 1);
  } else {
    if (w0 && de < 2) s1_rec(c, base, s + 1, f, nA / 2 + 1, de << 1, ne + 1);
    if (w1) s1_rec(c, base | bit, s + 1, f + 1, nB + nB / 2, (de | inplay) & ~2u, 0);
  }
}

static int64_t s1_walk_win(const or_node* root, uint32_t depth, uint32_t sem, leaf_fn fn,
                           void* ctx, uint32_t win_bits, uint32_t ntops,
                           const uint64_t* tops) {
  if (!root || depth > 100 || root->sigfig > depth) return -1;
  s1_ctx* c = (s1_ctx*)malloc(sizeof(s1_ctx));
  if (!c) return -1;
  c->win_bits = win_bits;
  c->ntops = ntops;
  c->tops = tops;
  c->depth = depth;
  c->sem = sem;
  c->fn = fn;
  c->ctx = ctx;
  c->count = 0;
  u128 x = 1;
  for (int i = 0; i < 128; ++i) {
    c->p3[i] = x;
    x *= 3;
  }
  s1_rec(c, mk128(root->base_lo, root->base_hi), root->sigfig, root->threepower,
         mk128(root->number_lo, root->number_hi), root->doubleeven, root->numevens);
  int64_t n = c->count;
  free(c);
  return n;
}

static int64_t s1_walk(const or_node* root, uint32_t depth, uint32_t sem, leaf_fn fn,
                       void* ctx) {
  return s1_walk_win(root, depth, sem, fn, ctx, 0, 0, NULL);
}

typedef struct {
  or_node* out;
  int64_t cap, n;
} store_ctx;

static void store_leaf(void* vctx, const or_node* leaf) {
  store_ctx* s = (store_ctx*)vctx;
  if (s->n < s->cap) s->out[s->n] = *leaf;
  s->n++;
}

int64_t or_search1(const or_node* root, uint32_t depth, uint32_t semantics, or_node* leaves,
                   int64_t cap) {
  store_ctx s = {leaves, leaves ? cap : 0, 0};
  return s1_walk(root, depth, semantics, leaves ? store_leaf : NULL, &s);
}

/* Leaves from the canonical root, which the reference starts from: "3 2 2 9 1 0"
 * (collatz_cudav7.cu:6, :173), sorted by residue with an LSD radix sort. */
typedef struct {
  uint64_t* r;
  uint64_t* num;
  uint8_t* f;
  int64_t cap, n;
} sorted_ctx;

static void sorted_leaf(void* vctx, const or_node* leaf) {
  sorted_ctx* s = (sorted_ctx*)vctx;
  if (s->n < s->cap) {
    s->r[s->n] = leaf->base_lo;
    s->num[s->n] = leaf->number_lo;
    s->f[s->n] = (uint8_t)leaf->threepower;
  }
  s->n++;
}

/* LSD radix sort of n survivors by r (k bits), 16-bit digits, carrying number and f. */
static int sort_by_r(uint32_t k, int64_t n, uint64_t* r, uint64_t* number, uint8_t* f) {
  uint64_t* r2 = (uint64_t*)malloc((size_t)n * 8 + 8);
  uint64_t* n2 = (uint64_t*)malloc((size_t)n * 8 + 8);
  uint8_t* f2 = (uint8_t*)malloc((size_t)n + 1);
  int64_t* cnt = (int64_t*)malloc(sizeof(int64_t) * 65537);
  if (!r2 || !n2 || !f2 || !cnt) {
    free(r2); free(n2); free(f2); free(cnt);
    return -1;
  }
  for (uint32_t shift = 0; shift < k; shift += 16) {
    memset(cnt, 0, sizeof(int64_t) * 65537);
    for (int64_t i = 0; i < n; ++i) cnt[((r[i] >> shift) & 0xFFFF) + 1]++;
    for (int i = 0; i < 65536; ++i) cnt[i + 1] += cnt[i];
    for (int64_t i = 0; i < n; ++i) {
      int64_t d = cnt[(r[i] >> shift) & 0xFFFF]++;
      r2[d] = r[i];
      n2[d] = number[i];
      f2[d] = f[i];
    }
    memcpy(r, r2, (size_t)n * 8);
    memcpy(number, n2, (size_t)n * 8);
    memcpy(f, f2, (size_t)n);
  }
  free(r2); free(n2); free(f2); free(cnt);
  return 0;
}

int64_t or_sieve_sorted(uint32_t k, uint64_t* r, uint64_t* number, uint8_t* f, int64_t cap) {
  if (k < 2 || k > 40) return -1;
  or_node root = {3, 0, 9, 0, 2, 2, 1, 0};
  if (!r || !number || !f) return s1_walk(&root, k, OR_SEM_CXX, NULL, NULL);
  sorted_ctx s = {r, number, f, cap, 0};
  int64_t n = s1_walk(&root, k, OR_SEM_CXX, sorted_leaf, &s);
  if (n < 0 || n > cap) return n;
  return sort_by_r(k, n, r, number, f) ? -1 : n;
}

/* One thread of or_sieve_windows: subtrees below frontier nodes taken in turn, leaves
 * appended to a buffer of its own that grows as needed. */
typedef struct {
  const or_node* nodes;
  int64_t nn;
  atomic_int_fast64_t* next;
  uint32_t k, w, ntops;
  const uint64_t* tops;
  uint64_t *r, *num;
  uint8_t* f;
  int64_t n, cap;
  int err;
} win_job;

static void win_leaf(void* vctx, const or_node* leaf) {
  win_job* j = (win_job*)vctx;
  if (j->err) return;
  if (j->n == j->cap) {
    const int64_t c = j->cap ? 2 * j->cap : (1 << 16);
    uint64_t* r = (uint64_t*)realloc(j->r, (size_t)c * 8);
    if (r) j->r = r;
    uint64_t* num = (uint64_t*)realloc(j->num, (size_t)c * 8);
    if (num) j->num = num;
    uint8_t* f = (uint8_t*)realloc(j->f, (size_t)c);
    if (f) j->f = f;
    if (!r || !num || !f) {
      j->err = 1;
      return;
    }
    j->cap = c;
  }
  j->r[j->n] = leaf->base_lo;
  j->num[j->n] = leaf->number_lo;
  j->f[j->n] = (uint8_t)leaf->threepower;
  j->n++;
}

static void* win_worker(void* v) {
  win_job* j = (win_job*)v;
  for (;;) {
    const int64_t i = atomic_fetch_add(j->next, 1);
    if (i >= j->nn) break;
    if (s1_walk_win(&j->nodes[i], j->k, OR_SEM_CXX, win_leaf, j, j->w, j->ntops, j->tops) < 0)
      j->err = 1;
  }
  return NULL;
}

/* The survivors of or_sieve_sorted(k) whose residue r has r >> w equal to one of tops[]:
 * the same walk, below depth w only along the paths of those windows.  For chunks of a
 * range narrower than 2^k, whose residues mod 2^k form such windows, these are every
 * survivor the chunks use, at the cost of the tree to depth w.  The subtrees below depth
 * min(w, 14) are walked on `threads` threads.  Returns the number of survivors; they are
 * written (sorted by r) only if that is at most cap, so a caller with too small a buffer
 * calls again with a larger one. */
int64_t or_sieve_windows(uint32_t k, uint32_t w, uint32_t ntops, const uint64_t* tops,
                         int threads, uint64_t* r, uint64_t* number, uint8_t* f,
                         int64_t cap) {
  if (k < 2 || k > 40 || w < 2 || w > k || !ntops || !tops) return -1;
  for (uint32_t i = 0; i < ntops; ++i)
    if ((tops[i] >> (k - w)) != 0) return -1;
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  or_node root = {3, 0, 9, 0, 2, 2, 1, 0};
  const uint32_t split = w < 14 ? w : 14;
  const int64_t nn = or_search1(&root, split, OR_SEM_CXX, NULL, 0);
  if (nn < 0) return -1;
  or_node* nodes = (or_node*)malloc(sizeof(or_node) * (size_t)(nn > 0 ? nn : 1));
  win_job* jobs = (win_job*)calloc((size_t)threads, sizeof(win_job));
  pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  atomic_int_fast64_t next;
  atomic_init(&next, 0);
  int rc = (nodes && jobs && th) ? 0 : -1;
  if (rc == 0) or_search1(&root, split, OR_SEM_CXX, nodes, nn);
  int started = 0;
  for (int t = 0; rc == 0 && t < threads; ++t) {
    jobs[t] = (win_job){nodes, nn, &next, k, w, ntops, tops, NULL, NULL, NULL, 0, 0, 0};
    if (pthread_create(&th[t], NULL, win_worker, &jobs[t]) != 0) rc = -1;
    else started++;
  }
  int64_t n = 0;
  for (int t = 0; t < started; ++t) {
    pthread_join(th[t], NULL);
    if (jobs[t].err) rc = -1;
    n += jobs[t].n;
  }
  if (rc == 0 && r && number && f && n <= cap) {
    int64_t o = 0;
    for (int t = 0; t < started; ++t) {
      memcpy(r + o, jobs[t].r, (size_t)jobs[t].n * 8);
      memcpy(number + o, jobs[t].num, (size_t)jobs[t].n * 8);
      memcpy(f + o, jobs[t].f, (size_t)jobs[t].n);
      o += jobs[t].n;
    }
    if (sort_by_r(k, n, r, number, f)) rc = -1;
  }
  for (int t = 0; t < threads && jobs; ++t) {
    free(jobs[t].r);
    free(jobs[t].num);
    free(jobs[t].f);
  }
  free(nodes);
  free(jobs);
  free(th);
  return rc ? -1 : n;
}

/* ------------------------------------------------------------------------------------
 * precompute: collatz_cudav7.cu:413-494 (C++), main.rs:101-107,125-203 (Rust).
 * ---------------------------------------------------------------------------------- */

/* collatz_cudav7.cu:415-421.  In C++, `(x) & 0b11 == 0` parses as `x & (0b11 == 0)`,
 * which is `x & 0`: the test is never true.  main.rs:104 binds `&` first and tests the
 * two bits.  Both versions are kept, selected by semantics. */
int or_doubleeven(uint32_t upordown, uint32_t i, uint32_t B, uint32_t semantics) {
  uint32_t a = (uint32_t)__builtin_ctz((upordown >> i) + 1u);
  if (i + a + 2 >= B) return 0;
  if (semantics != OR_SEM_RUST) return 0;
  return ((upordown >> (i + a)) & 3u) == 0;
}

/* `x > y` as the precompute evaluates it: signed in collatz_cudav7.cu:449,461 and main.rs,
 * unsigned (int + u32 promotes to u32) in collatz_openclv2.cpp:319,331. */
static inline int gt_sem(int32_t x, int32_t y, int uns) {
  return uns ? ((uint32_t)x > (uint32_t)y) : (x > y);
}

/* The B-step dip of one residue n0 of number (collatz_cudav7.cu:436-467): B steps of the
 * shifted map on n0 give the parity word; then, step by step, cur = 306 (steps) - 485 (odd
 * steps) and its running maximum with the path-merge terms.  Only the low B bits of number
 * decide the first B steps, so n0 = number mod 2^B. */
int32_t or_la_dip(uint64_t n0, uint32_t B, uint32_t semantics) {
  const int uns = semantics == OR_SEM_OPENCL;
  uint32_t upordown = 0;
  uint64_t n = n0;
  for (uint32_t i = 0; i < B; ++i) {
    if (n % 2 == 0) {
      n = n + n / 2;
      upordown |= 1u << i;
    } else {
      n = n / 2 + 1;
    }
  }
  int32_t cur = 0, mx = 0;
  uint32_t ne = 0;
  for (uint32_t i = 0; i < B; ++i) {
    if (upordown & (1u << i)) {
      if (gt_sem(cur + (int32_t)OR_NUMERATOR, mx, uns) &&
          or_doubleeven(upordown, i, B, semantics))
        mx = cur + (int32_t)OR_NUMERATOR;
      cur -= (int32_t)(OR_DENOMINATOR - OR_NUMERATOR);
      ne = 0;
    } else {
      cur += (int32_t)OR_NUMERATOR;
      ne += 1;
      if (ne % 2 == 0 && gt_sem(cur + (int32_t)(OR_DENOMINATOR - OR_NUMERATOR), mx, uns))
        mx = cur + (int32_t)(OR_DENOMINATOR - OR_NUMERATOR);
    }
    if (cur > mx) mx = cur;
  }
  return mx;
}

int or_precompute_bv(const or_knobs* kn, uint64_t* bv) {
  if (!kn || !bv || kn->B < 6 || kn->B > 30 || kn->A > kn->N) return -1;
  const uint32_t B = kn->B, mbi = kn->N - kn->A, tpo = or_tpo(mbi);
  const int32_t first = (int32_t)(int16_t)(OR_DENOMINATOR * tpo - OR_NUMERATOR * mbi);
  const uint64_t words = 1ull << (B - 6);
  const uint32_t mask = (B >= 32) ? 0xFFFFFFFFu : ((1u << B) - 1u);
  uint32_t magic[64];
  for (uint32_t a = 0; a < kn->num_bv && a < 64; ++a) magic[a] = or_magic(tpo + a);
  memset(bv, 0, (size_t)kn->num_bv * words * 8);
  for (uint64_t n0 = 0; n0 < (1ull << B); ++n0) {
    const int32_t mx = or_la_dip(n0, B, kn->semantics);
    for (uint32_t a = 0; a < kn->num_bv; ++a) {
      if (mx <= first + (int32_t)(a * OR_DENOMINATOR)) {
        uint32_t idx = ((uint32_t)n0 * magic[a]) & mask;
        bv[a * words + idx / 64] |= 1ull << (idx % 64);
      }
    }
  }
  return 0;
}

static int mod9_keep(uint32_t r) { return r == 0 || r == 1 || r == 3 || r == 6 || r == 7; }

/* collatz_cudav7.cu:482-494: row r, bit i set iff (r + i * 2^MBI) mod 9 in {0,1,3,6,7}. */
int or_mod9_rows(uint32_t mbi, uint32_t A, uint64_t* rows) {
  if (!rows || A > 20 || mbi > 120) return -1;
  const uint64_t words = (A >= 6) ? (1ull << (A - 6)) : 1;
  const uint32_t step = (uint32_t)(((u128)1 << mbi) % 9);
  memset(rows, 0, 9 * words * 8);
  for (uint32_t a = 0; a < 9; ++a) {
    uint32_t red = a;
    for (uint64_t i = 0; i < (1ull << A); ++i) {
      if (mod9_keep(red)) rows[a * words + i / 64] |= 1ull << (i % 64);
      red += step;
      if (red >= 9) red -= 9;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------------------
 * The iteration schedule (collatz_cudav7.cu:280-366 for the kernel, :684-715 for
 * search3).  One check = odd run, even run, odd run, even run, then hash ^= low 32 bits
 * and the test number < base.
 *   odd run : a = ctz(number | 2^8) <= 8,  number = (number >> a) * 3^a
 *   even run: number -= 1; b = ctz(number | 2^31) <= 31; number = (number >> b) + 1
 * If number reaches 2^cutoff after an odd run it is reported and replaced by 2.
 * Values are kept exactly in three u64 limbs (the cutoff is at most 187, so an odd run,
 * which multiplies by at most (3/2)^8 < 2^5, cannot leave 192 bits).
 * ---------------------------------------------------------------------------------- */
typedef struct {
  uint64_t w[3];
} big3;

static inline int big_cmp(const big3* x, const big3* y) {
  for (int i = 2; i >= 0; --i) {
    if (x->w[i] != y->w[i]) return x->w[i] < y->w[i] ? -1 : 1;
  }
  return 0;
}

static inline void big_shr(big3* x, uint32_t s) {
  if (s == 0) return;
  x->w[0] = (x->w[0] >> s) | (x->w[1] << (64 - s));
  x->w[1] = (x->w[1] >> s) | (x->w[2] << (64 - s));
  x->w[2] >>= s;
}

static inline void big_mul_small(big3* x, uint64_t m) {
  u128 carry = 0;
  for (int i = 0; i < 3; ++i) {
    u128 t = (u128)x->w[i] * m + carry;
    x->w[i] = (uint64_t)t;
    carry = t >> 64;
  }
}

static inline void big_sub1(big3* x) {
  for (int i = 0; i < 3; ++i) {
    if (x->w[i]-- != 0) break;
  }
}

static inline void big_add1(big3* x) {
  for (int i = 0; i < 3; ++i) {
    if (++x->w[i] != 0) break;
  }
}

static inline int big_ge_pow2(const big3* x, uint32_t c) {
  uint32_t wi = c / 64, bi = c % 64;
  for (uint32_t i = 2; i > wi; --i)
    if (x->w[i]) return 1;
  return (x->w[wi] >> bi) != 0;
}

static const uint64_t POW3_SMALL[9] = {1, 3, 9, 27, 81, 243, 729, 2187, 6561};

void or_traj(const uint64_t base[2], const uint64_t num0[3], uint32_t cutoff, int second_gt,
             or_traj_out* out) {
  big3 v = {{num0[0], num0[1], num0[2]}};
  big3 mx = v;
  big3 b = {{base[0], base[1], 0}};
  uint32_t hash = 0, ovf = 0;
  uint64_t steps = 0;
  if (cutoff > 187) cutoff = 187;
  for (;;) {
    for (int half = 0; half < 2; ++half) {
      uint32_t a = (uint32_t)__builtin_ctzll(v.w[0] | 0x100u);
      big_shr(&v, a);
      big_mul_small(&v, POW3_SMALL[a]);
      steps += a;
      if (big_cmp(&v, &mx) > 0) mx = v;
      int over = big_ge_pow2(&v, cutoff);
      if (over && second_gt && half == 1) {
        /* collatz_cudav7.cu:702 uses '>' here: exactly 2^cutoff does not count. */
        big3 p = {{0, 0, 0}};
        p.w[cutoff / 64] = 1ull << (cutoff % 64);
        over = big_cmp(&v, &p) > 0;
      }
      if (over) {
        ovf++;
        v.w[0] = 2;
        v.w[1] = 0;
        v.w[2] = 0;
      }
      big_sub1(&v);
      uint32_t bb = (uint32_t)__builtin_ctzll(v.w[0] | 0x80000000u);
      big_shr(&v, bb);
      big_add1(&v);
      steps += bb;
    }
    hash ^= (uint32_t)v.w[0];
    if (big_cmp(&v, &b) < 0) break;
  }
  out->hash = hash;
  out->overflow = ovf;
  out->steps = steps;
  out->maxnum[0] = mx.w[0];
  out->maxnum[1] = mx.w[1];
  out->maxnum[2] = mx.w[2];
  out->final[0] = v.w[0];
  out->final[1] = v.w[1];
  out->final[2] = v.w[2];
}

/* ------------------------------------------------------------------------------------
 * main.rs's search3 (main.rs:321-341) with its 16-step table (precompute, main.rs:111-123):
 * entry[n0] for n0 < 2^16 holds t | S^16(n0) << 5, t the odd steps among 16 steps of the
 * shifted map S (m even: 3m/2; m odd: m/2 + 1, rounding down).  A block is
 * number = (number >> 16) * 3^t + S^16(number mod 2^16).  After each block the tally grows
 * by one; the candidate stops when number < base, or when number >= 2^cutoff (main.rs:336
 * tests (number >> 64) >= 2^56, i.e. cutoff 120, and prints the start).  Kept exactly in
 * three u64 limbs: main.rs's u128 would wrap for a block that starts in [2^118.6, 2^120).
 * ---------------------------------------------------------------------------------- */
static uint32_t g_ms16[1u << 16];
static pthread_once_t g_ms16_once = PTHREAD_ONCE_INIT;

static void ms16_init(void) {
  for (uint32_t n0 = 0; n0 < (1u << 16); ++n0) {
    uint32_t t = 0;
    uint64_t n = n0;
    for (int i = 0; i < 16; ++i) {
      if ((n & 1) == 0) {
        n = n + n / 2;
        t++;
      } else {
        n = n / 2 + 1;
      }
    }
    g_ms16[n0] = t | ((uint32_t)n << 5);
  }
}

static inline void big_add_small(big3* x, uint64_t a) {
  u128 s = (u128)x->w[0] + a;
  x->w[0] = (uint64_t)s;
  if ((s >> 64) == 0) return;
  for (int i = 1; i < 3; ++i)
    if (++x->w[i] != 0) break;
}

int or_rust_search3(const uint64_t base[2], const uint64_t num0[3], uint32_t cutoff,
                    uint64_t* tally) {
  pthread_once(&g_ms16_once, ms16_init);
  if (cutoff > 180) cutoff = 180;
  big3 v = {{num0[0], num0[1], num0[2]}};
  const big3 b = {{base[0], base[1], 0}};
  uint64_t t = 0;
  for (;;) {
    const uint32_t e = g_ms16[v.w[0] & 0xFFFFu];
    uint64_t p3 = 1;
    for (uint32_t i = 0; i < (e & 31u); ++i) p3 *= 3;
    big_shr(&v, 16);
    big_mul_small(&v, p3);
    big_add_small(&v, e >> 5);
    t++;
    if (big_cmp(&v, &b) < 0) {
      *tally = t;
      return 0;
    }
    if (big_ge_pow2(&v, cutoff)) {
      *tally = t;
      return 1;
    }
  }
}

/* ------------------------------------------------------------------------------------
 * Reference mode: search2 (collatz_cudav7.cu:654-682, main.rs:280-319) and whole case
 * lines (collatz_cudav7.cu:95-183 with launchkernel :565-649).
 * ---------------------------------------------------------------------------------- */
static inline int bit_at(const uint64_t* words, uint64_t i) { return (words[i / 64] >> (i % 64)) & 1; }

typedef struct {
  const or_knobs* kn;
  const uint64_t* bv;
  const uint64_t* mod9;
  uint32_t hash;
  uint64_t cand;
  uint64_t ovf;
  int count_only;
  uint64_t numcases;     /* main.rs semantics: 16-step blocks of search3, summed */
  uint64_t* ovf_starts;  /* optional: the start n of each overflow report, two words each */
  uint64_t ovf_cap;
} leaf_ctx;

static void note_overflow(leaf_ctx* c, u128 n) {
  if (c->ovf_starts && c->ovf < c->ovf_cap) {
    c->ovf_starts[2 * c->ovf] = (uint64_t)n;
    c->ovf_starts[2 * c->ovf + 1] = (uint64_t)(n >> 64);
  }
  c->ovf++;
}

static void leaf_candidates(leaf_ctx* c, const or_node* leaf) {
  const or_knobs* kn = c->kn;
  const uint32_t mbi = kn->N - kn->A, tpo = or_tpo(mbi), f = leaf->threepower;
  const uint32_t bidx = f - tpo;
  const uint32_t mask = (1u << kn->B) - 1u;
  const uint64_t bvwords = 1ull << (kn->B - 6);
  const uint64_t m9words = (kn->A >= 6) ? (1ull << (kn->A - 6)) : 1;
  const u128 base = mk128(leaf->base_lo, leaf->base_hi);
  const u128 number = mk128(leaf->number_lo, leaf->number_hi);
  const uint32_t red = (uint32_t)(base % 9);
  const int use_bv = bidx < kn->num_bv;
  const uint32_t idx = use_bv ? ((or_magic(f) * (uint32_t)number) & mask) : 0;
  const u128 p3 = pow3_128(f);
  for (uint64_t a = 0; a < (1ull << kn->A); ++a) {
    if (!bit_at(c->mod9 + red * m9words, a)) continue;
    if (use_bv && !bit_at(c->bv + bidx * bvwords, (idx + a) & mask)) continue;
    if (kn->semantics == OR_SEM_RUST) {
      /* main.rs:309-313 */
      uint64_t nl = (uint64_t)number;
      if (leaf->doubleeven == 2 && (nl + a) % 2 == 1) continue;
      if (leaf->doubleeven == 1) {
        if (f % 2 == 0 && (nl + a) % 4 == 1) continue;
        if (f % 2 == 1 && (nl + 3 * a) % 4 == 1) continue;
      }
    }
    c->cand++;
    if (c->count_only) continue;
    u128 n = base + ((u128)a << mbi);
    u128 v = number + (u128)a * p3;
    uint64_t b2[2] = {(uint64_t)n, (uint64_t)(n >> 64)};
    uint64_t v3[3] = {(uint64_t)v, (uint64_t)(v >> 64), 0};
    if (kn->semantics == OR_SEM_RUST) {
      /* main.rs:314 -> search3 (:321-341): 16-step blocks, no hash. */
      uint64_t tally = 0;
      if (or_rust_search3(b2, v3, kn->cutoff, &tally)) note_overflow(c, n);
      c->numcases += tally;
      continue;
    }
    or_traj_out o;
    /* Leaves without a bit vector run on the CPU in the reference (search2/search3,
     * collatz_cudav7.cu:517-519), whose second overflow test is '>'. */
    or_traj(b2, v3, kn->cutoff, !use_bv, &o);
    c->hash ^= o.hash;
    for (uint32_t i = 0; i < o.overflow; ++i) note_overflow(c, n);
  }
}

uint32_t or_leaf_hash(const or_knobs* kn, const uint64_t* bv, const uint64_t* mod9,
                      const or_node* leaf, uint64_t* ncand) {
  leaf_ctx c = {kn, bv, mod9, 0, 0, 0, 0, 0, NULL, 0};
  leaf_candidates(&c, leaf);
  if (ncand) *ncand = c.cand;
  return c.hash;
}

typedef struct {
  leaf_ctx lc;
  or_case_out* out;
  uint32_t tpo;
} case_ctx;

static void case_leaf(void* vctx, const or_node* leaf) {
  case_ctx* c = (case_ctx*)vctx;
  uint32_t b = leaf->threepower - c->tpo;
  c->out->bin_leaves[b < 16 ? b : 16]++;
  leaf_candidates(&c->lc, leaf);
}

int or_case_run_starts(const or_knobs* kn, const uint64_t* bv, const uint64_t* mod9,
                       const or_node* root, or_case_out* out, uint64_t* ovf_starts,
                       uint64_t ovf_cap) {
  if (!kn || !bv || !mod9 || !root || !out || kn->A > kn->N) return -1;
  memset(out, 0, sizeof(*out));
  case_ctx c = {{kn, bv, mod9, 0, 0, 0, 0, 0, ovf_starts, ovf_cap}, out, or_tpo(kn->N - kn->A)};
  int64_t n = s1_walk(root, kn->N - kn->A, kn->semantics, case_leaf, &c);
  if (n < 0) return -1;
  out->leaves = (uint64_t)n;
  out->complete_hash = c.lc.hash;
  out->candidates = c.lc.cand;
  out->overflow_count = c.lc.ovf;
  out->numcases = c.lc.numcases;
  /* One launch per full batch of 2^16 leaves plus one flush per non-empty bin
   * (collatz_cudav7.cu:532-533, :177-179). */
  for (int b = 0; b < 16; ++b) out->launches += (out->bin_leaves[b] + 65535) / 65536;
  return 0;
}

int or_case_run(const or_knobs* kn, const uint64_t* bv, const uint64_t* mod9,
                const or_node* root, or_case_out* out) {
  return or_case_run_starts(kn, bv, mod9, root, out, NULL, 0);
}

/* The same case run on `threads` threads: the subtree is cut at a frontier a few levels
 * below the root (search1's own nodes there), the frontier nodes are shared out in
 * interleaved order, each thread runs its nodes as case lines of their own, and the
 * results are merged (sums, XOR of hashes, per-bin leaves; the overflow starts of each
 * thread in turn).  The same numbers as or_case_run_starts; only the order of the listed
 * starts differs. */
typedef struct {
  const or_knobs* kn;
  const uint64_t* bv;
  const uint64_t* mod9;
  const or_node* nodes;
  int64_t nnodes;
  int tid, nthreads;
  or_case_out out;
  uint64_t* starts;
  uint64_t cap;
  int rc;
} case_job;

static void* case_worker(void* vj) {
  case_job* j = (case_job*)vj;
  memset(&j->out, 0, sizeof(j->out));
  j->rc = 0;
  for (int64_t i = j->tid; i < j->nnodes; i += j->nthreads) {
    or_case_out o;
    const uint64_t used = j->out.overflow_count;
    const uint64_t room = used < j->cap ? j->cap - used : 0;
    if (or_case_run_starts(j->kn, j->bv, j->mod9, &j->nodes[i], &o,
                           room ? j->starts + 2 * used : NULL, room) != 0) {
      j->rc = -1;
      return NULL;
    }
    j->out.leaves += o.leaves;
    j->out.candidates += o.candidates;
    j->out.overflow_count += o.overflow_count;
    j->out.numcases += o.numcases;
    j->out.complete_hash ^= o.complete_hash;
    for (int b = 0; b < 17; ++b) j->out.bin_leaves[b] += o.bin_leaves[b];
  }
  return NULL;
}

int or_case_run_mt(const or_knobs* kn, const uint64_t* bv, const uint64_t* mod9,
                   const or_node* root, or_case_out* out, uint64_t* ovf_starts,
                   uint64_t ovf_cap, int threads) {
  if (!kn || !root || !out || kn->A > kn->N) return -1;
  const uint32_t depth = kn->N - kn->A;
  if (threads <= 1 || depth < root->sigfig + 12)
    return or_case_run_starts(kn, bv, mod9, root, out, ovf_starts, ovf_cap);
  if (threads > 256) threads = 256;
  const uint32_t split = root->sigfig + 10;
  const int64_t nn = or_search1(root, split, kn->semantics, NULL, 0);
  if (nn < 0) return -1;
  or_node* nodes = (or_node*)malloc(sizeof(or_node) * (size_t)(nn > 0 ? nn : 1));
  case_job* jobs = (case_job*)calloc((size_t)threads, sizeof(case_job));
  pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  uint64_t** bufs = (uint64_t**)calloc((size_t)threads, sizeof(uint64_t*));
  int rc = (nodes && jobs && th && bufs) ? 0 : -1;
  if (rc == 0) or_search1(root, split, kn->semantics, nodes, nn);
  for (int t = 0; rc == 0 && t < threads; ++t) {
    bufs[t] = ovf_starts ? (uint64_t*)malloc(16 * (size_t)(ovf_cap ? ovf_cap : 1)) : NULL;
    jobs[t] = (case_job){kn, bv, mod9, nodes, nn, t, threads, {0}, bufs[t],
                         ovf_starts ? ovf_cap : 0, 0};
    if (pthread_create(&th[t], NULL, case_worker, &jobs[t]) != 0) {
      rc = -1;
      threads = t;
    }
  }
  memset(out, 0, sizeof(*out));
  uint64_t listed = 0;
  for (int t = 0; t < threads; ++t) {
    pthread_join(th[t], NULL);
    if (jobs[t].rc) rc = -1;
    const or_case_out* o = &jobs[t].out;
    out->leaves += o->leaves;
    out->candidates += o->candidates;
    out->overflow_count += o->overflow_count;
    out->numcases += o->numcases;
    out->complete_hash ^= o->complete_hash;
    for (int b = 0; b < 17; ++b) out->bin_leaves[b] += o->bin_leaves[b];
    if (ovf_starts && bufs[t]) {
      const uint64_t k = o->overflow_count < ovf_cap ? o->overflow_count : ovf_cap;
      for (uint64_t i = 0; i < k && listed < ovf_cap; ++i, ++listed) {
        ovf_starts[2 * listed] = bufs[t][2 * i];
        ovf_starts[2 * listed + 1] = bufs[t][2 * i + 1];
      }
    }
    free(bufs[t]);
  }
  for (int b = 0; b < 16; ++b) out->launches += (out->bin_leaves[b] + 65535) / 65536;
  free(nodes);
  free(jobs);
  free(th);
  free(bufs);
  return rc;
}

int64_t or_rust_count(const or_knobs* kn, const uint64_t* bv, const uint64_t* mod9,
                      const or_node* root) {
  case_ctx c;
  or_case_out scratch;
  memset(&scratch, 0, sizeof(scratch));
  leaf_ctx lc = {kn, bv, mod9, 0, 0, 0, 1, 0, NULL, 0};
  c.lc = lc;
  c.out = &scratch;
  c.tpo = or_tpo(kn->N - kn->A);
  int64_t n = s1_walk(root, kn->N - kn->A, kn->semantics, case_leaf, &c);
  if (n < 0) return -1;
  return (int64_t)c.lc.cand;
}

/* ------------------------------------------------------------------------------------
 * Range mode.  n = q*2^k + r with r a sieve survivor; start value
 * number = T^k(r) + 1 + q*3^f (Lemma 2.3, paper p.4; collatz_cudav7.cu:262-278),
 * mod 9 filter on n (collatz_cudav7.cu:482-494), then the schedule above.
 * ---------------------------------------------------------------------------------- */
void or_chunk_merge(or_chunk* into, const or_chunk* from) {
  if (from->count > 0) {
    int take = into->count == 0;
    if (!take) {
      big3 a = {{from->max_exc[0], from->max_exc[1], from->max_exc[2]}};
      big3 b = {{into->max_exc[0], into->max_exc[1], into->max_exc[2]}};
      int c = big_cmp(&a, &b);
      if (c > 0) take = 1;
      if (c == 0 && mk128(from->max_start[0], from->max_start[1]) <
                        mk128(into->max_start[0], into->max_start[1]))
        take = 1;
    }
    if (take) {
      memcpy(into->max_exc, from->max_exc, sizeof(into->max_exc));
      memcpy(into->max_start, from->max_start, sizeof(into->max_start));
    }
  }
  into->n_checked += from->n_checked;
  into->count += from->count;
  into->stop_sum += from->stop_sum;
  into->stop_xor ^= from->stop_xor;
  into->overflow_count += from->overflow_count;
  into->hash ^= from->hash;
  into->flags |= from->flags;
}

static int64_t lower_bound_u64(const uint64_t* a, int64_t n, uint64_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

typedef struct {
  uint32_t k, C, cutoff;
  int mod9;
  uint32_t la_bits, la_sem;
  int64_t nres;
  const uint64_t* r;
  const uint64_t* number;
  const uint8_t* f;
  u128 lo, hi;
  uint64_t nchunks;
  or_chunk* chunks;
  atomic_ullong next;
  u128 p3[64];
} range_ctx;

static void range_chunk(range_ctx* c, uint64_t ci, or_chunk* out) {
  memset(out, 0, sizeof(*out));
  const u128 j = (c->lo >> c->C) + ci;
  u128 c_lo = j << c->C;
  u128 c_hi = c_lo + ((u128)1 << c->C);
  if (c_lo < c->lo) c_lo = c->lo;
  if (c_hi > c->hi) c_hi = c->hi;
  out->chunk_index = (uint64_t)j;
  out->n_checked = (uint64_t)(c_hi - c_lo);
  const u128 kspan = (u128)1 << c->k;
  big3 best = {{0, 0, 0}};
  u128 best_start = 0;
  for (u128 q = c_lo >> c->k; q <= (c_hi - 1) >> c->k; ++q) {
    const u128 qbase = q << c->k;
    const uint64_t rlo = c_lo > qbase ? (uint64_t)(c_lo - qbase) : 0;
    const u128 rhi128 = c_hi - qbase;
    const uint64_t rhi = rhi128 > kspan ? (uint64_t)kspan : (uint64_t)rhi128;
    const int64_t i0 = lower_bound_u64(c->r, c->nres, rlo);
    const int64_t i1 = lower_bound_u64(c->r, c->nres, rhi);
    const uint64_t qlo = (uint64_t)q, qhi = (uint64_t)(q >> 64);
    for (int64_t i = i0; i < i1; ++i) {
      const u128 n = qbase + c->r[i];
      if (c->mod9 && !mod9_keep((uint32_t)(n % 9))) continue;
      /* number = number_r + q * 3^f, up to 192 bits */
      const u128 p = c->p3[c->f[i]];
      const uint64_t plo = (uint64_t)p, phi = (uint64_t)(p >> 64);
      big3 v = {{0, 0, 0}};
      u128 t0 = (u128)qlo * plo;
      u128 t1 = (u128)qlo * phi + (t0 >> 64);
      u128 t2 = (u128)qhi * plo + (uint64_t)t1;
      v.w[0] = (uint64_t)t0;
      v.w[1] = (uint64_t)t2;
      v.w[2] = (uint64_t)(t1 >> 64) + (uint64_t)(t2 >> 64) + qhi * phi;
      u128 s = (u128)v.w[0] + c->number[i];
      v.w[0] = (uint64_t)s;
      s = (u128)v.w[1] + (uint64_t)(s >> 64);
      v.w[1] = (uint64_t)s;
      v.w[2] += (uint64_t)(s >> 64);
      /* Step 2 at depth k (collatz_cudav7.cu:219-235 with BV_i of :423-480): keep the
       * candidate iff the dip of number mod 2^B is at most 485 f - 306 k. */
      if (c->la_bits) {
        const int32_t dip = or_la_dip(v.w[0] & ((1ull << c->la_bits) - 1), c->la_bits, c->la_sem);
        if (dip > (int32_t)(OR_DENOMINATOR * c->f[i]) - (int32_t)(OR_NUMERATOR * c->k)) continue;
      }
      const uint64_t b2[2] = {(uint64_t)n, (uint64_t)(n >> 64)};
      or_traj_out o;
      or_traj(b2, v.w, c->cutoff, 0, &o);
      const uint64_t J = c->k + o.steps;
      out->count++;
      out->hash ^= o.hash;
      out->stop_sum += J;
      out->stop_xor ^= J;
      out->overflow_count += o.overflow;
      big3 m = {{o.maxnum[0], o.maxnum[1], o.maxnum[2]}};
      int cmp = big_cmp(&m, &best);
      if (out->count == 1 || cmp > 0 || (cmp == 0 && n < best_start)) {
        best = m;
        best_start = n;
      }
    }
  }
  if (out->count) {
    big_sub1(&best); /* T-value = number - 1 */
    memcpy(out->max_exc, best.w, sizeof(best.w));
    out->max_start[0] = (uint64_t)best_start;
    out->max_start[1] = (uint64_t)(best_start >> 64);
    if (out->overflow_count) out->flags |= 1u;
    big_add1(&best);
    if (big_ge_pow2(&best, OR_WIDE_LOG2)) out->flags |= 2u;
  }
}

static void* range_worker(void* vctx) {
  range_ctx* c = (range_ctx*)vctx;
  for (;;) {
    uint64_t ci = atomic_fetch_add(&c->next, 1);
    if (ci >= c->nchunks) break;
    range_chunk(c, ci, &c->chunks[ci]);
  }
  return NULL;
}

int or_verify_range(uint32_t k, int64_t nres, const uint64_t* r, const uint64_t* number,
                    const uint8_t* f, int mod9, const uint64_t lo[2], const uint64_t hi[2],
                    uint32_t C, uint32_t cutoff, int threads, or_chunk* chunks,
                    or_chunk* total) {
  return or_verify_range_la(k, nres, r, number, f, mod9, 0, 0, lo, hi, C, cutoff, threads,
                            chunks, total);
}

int or_verify_range_la(uint32_t k, int64_t nres, const uint64_t* r, const uint64_t* number,
                       const uint8_t* f, int mod9, uint32_t la_bits, uint32_t la_sem,
                       const uint64_t lo[2], const uint64_t hi[2], uint32_t C,
                       uint32_t cutoff, int threads, or_chunk* chunks, or_chunk* total) {
  if (k < 2 || k > 40 || C > 100 || !chunks || !total || nres < 0) return -1;
  if (la_bits && (la_bits < 1 || la_bits > 30 || la_sem == OR_SEM_RUST)) return -1;
  const u128 L = mk128(lo[0], lo[1]), H = mk128(hi[0], hi[1]);
  if (L >= H) return -1;
  range_ctx* c = (range_ctx*)calloc(1, sizeof(range_ctx));
  if (!c) return -1;
  c->k = k;
  c->C = C;
  c->cutoff = cutoff;
  c->mod9 = mod9;
  c->la_bits = la_bits;
  c->la_sem = la_sem;
  c->nres = nres;
  c->r = r;
  c->number = number;
  c->f = f;
  c->lo = L;
  c->hi = H;
  c->nchunks = (uint64_t)(((H - 1) >> C) - (L >> C) + 1);
  c->chunks = chunks;
  atomic_init(&c->next, 0);
  for (int i = 0; i < 64; ++i) c->p3[i] = pow3_128((uint32_t)i);
  if (threads <= 1) {
    range_worker(c);
  } else {
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, range_worker, c);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    free(th);
  }
  memset(total, 0, sizeof(*total));
  for (uint64_t i = 0; i < c->nchunks; ++i) or_chunk_merge(total, &chunks[i]);
  total->chunk_index = (uint64_t)(L >> C);
  free(c);
  return 0;
}
