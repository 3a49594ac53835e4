// collatz_run: command-line front end with the reference program's process protocol.
//
//   echo "3 2 2 9 1 0" | collatz_run [--N 45] [--A 10] [--cutoff 100] [--semantics cuda]
// reads one case line "base sigfig threepower number doubleeven numevens" from stdin
// (collatz_cudav7.cu:95-101), runs the whole reference method on the GPU, prints every start
// whose trajectory reached 2^cutoff as hex on stdout (collatz_cudav7.cu:298) and the summary
// line on stderr (collatz_cudav7.cu:181-182; the OpenCL program's two-line form with
// --semantics opencl, collatz_openclv2.cpp:279-280).  Spot checks (collatz_cudav7.cu:619-635)
// re-run one tile of 256 leaves in every P (P the largest power of two not above --spot-freq
// F; the default F = 2,560,000 is the reference's one launch of 2^16 leaves in 10000, and
// gives P = 2^21) by a second kernel and compare; "Spot checks" counts them, and each tile
// whose rerun differs prints "Spot check failed" as the reference does.
//
//   echo "3 2 2 9" | collatz_run --semantics rust [--N 40] [--A 6] [--cutoff 120]
// is main.rs's protocol (main.rs:85-98): the first four numbers on stdin (doubleeven and
// numevens are 0 there, main.rs:97), "Done with precompute" and "numcases = X" on stderr, and
// in decimal on stdout every start whose trajectory reached 2^cutoff after a 16-step block
// (main.rs:336-338).
//
//   collatz_run [--N 45] [--A 10] ... --cases FILE
// runs every case line of FILE in turn (the reference's users start one process per line),
// with one "line: launches L hash H" line each on stderr and a closing line with the XOR of
// the complete hashes.
//
//   collatz_run --range LO HI [--k 32] [--chunk 24] [--cutoff 187] [--form plain|persistent]
//               [--lookahead B [--semantics cuda|opencl]] [--out FILE]
// verifies [LO, HI) in range mode and prints one line per chunk and a total line; with
// --lookahead the sieve carries the Step 2 filter of B steps; --out also writes the records
// as a versioned result file (format below).
#include <algorithm>
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/collatz.h"

using u128 = unsigned __int128;

static u128 parse_u128(const char* s) {
  u128 v = 0;
  for (; *s; ++s) {
    if (*s < '0' || *s > '9') {
      std::fprintf(stderr, "not a decimal number: %s\n", s);
      std::exit(2);
    }
    v = v * 10 + (unsigned)(*s - '0');
  }
  return v;
}

static std::string to_dec(u128 v) {
  if (!v) return "0";
  std::string s;
  while (v) {
    s.insert(s.begin(), (char)('0' + (int)(v % 10)));
    v /= 10;
  }
  return s;
}

// ---- versioned per-chunk result file ("collatz-chunks 1") --------------------------------
// Text, one item per line: a header with the knobs of the run, one line per chunk record
// (the fields of collatz_chunk in header order; counts in decimal, the max excursion and
// its start as 0x-prefixed lowercase hex, the hash as eight uppercase hex digits), the
// merged total, the listed overflow starts in ascending order, and a last line
// "end <fnv>" whose FNV-1a 32-bit hash covers every byte before it.  collatz_amd/chunkfile.py reads and writes the same
// bytes.
static std::string hex_words(const uint64_t* w, int n) {
  int top = n - 1;
  while (top > 0 && w[top] == 0) --top;
  char buf[80];
  int len = std::snprintf(buf, sizeof(buf), "0x%llx", (unsigned long long)w[top]);
  for (int i = top - 1; i >= 0; --i)
    len += std::snprintf(buf + len, sizeof(buf) - len, "%016llx", (unsigned long long)w[i]);
  return std::string(buf, len);
}

static std::string chunk_fields(const collatz_chunk& c) {
  char buf[320];
  std::snprintf(buf, sizeof(buf), "%" PRIu64 " %" PRIu64 " %" PRIu64 " %" PRIu64 " %" PRIu64
                " %s %s %" PRIu64 " %08X %u",
                c.chunk_index, c.n_checked, c.count, c.stop_sum, c.stop_xor,
                hex_words(c.max_exc, 3).c_str(), hex_words(c.max_start, 2).c_str(),
                c.overflow_count, c.hash, c.flags);
  return buf;
}

static uint32_t fnv1a(const std::string& s) {
  uint32_t h = 2166136261u;
  for (unsigned char ch : s) {
    h ^= ch;
    h *= 16777619u;
  }
  return h;
}

static bool write_chunk_file(const char* path, uint32_t k, uint32_t lookahead, uint32_t sem,
                             uint32_t chunk, uint32_t cutoff, u128 lo, u128 hi,
                             const std::vector<collatz_chunk>& chunks, const collatz_chunk& total,
                             const uint64_t* ovf, uint64_t novf) {
  std::string b = "collatz-chunks 1\n";
  char buf[160];
  const uint64_t l2[2] = {(uint64_t)lo, (uint64_t)(lo >> 64)};
  const uint64_t h2[2] = {(uint64_t)hi, (uint64_t)(hi >> 64)};
  std::snprintf(buf, sizeof(buf), "k %u\nmod9 1\nlookahead %u %s\nchunk_log2 %u\ncutoff_log2 %u\n",
                k, lookahead, sem == COLLATZ_SEM_OPENCL ? "opencl" : "cuda", chunk, cutoff);
  b += buf;
  b += "range " + hex_words(l2, 2) + " " + hex_words(h2, 2) + "\n";
  std::snprintf(buf, sizeof(buf), "chunks %zu\n", chunks.size());
  b += buf;
  b += "fields chunk_index n_checked count stop_sum stop_xor max_exc max_start overflow_count "
       "hash flags\n";
  for (const collatz_chunk& c : chunks) b += "chunk " + chunk_fields(c) + "\n";
  b += "total " + chunk_fields(total) + "\n";
  std::snprintf(buf, sizeof(buf), "overflow %" PRIu64 "\n", novf);
  b += buf;
  // The device lists overflow events in the order its waves report them; the file lists
  // them in ascending order, so that it does not depend on scheduling.
  std::vector<u128> starts(novf);
  for (uint64_t i = 0; i < novf; ++i) starts[i] = ((u128)ovf[2 * i + 1] << 64) | ovf[2 * i];
  std::sort(starts.begin(), starts.end());
  for (const u128 v : starts) {
    const uint64_t w[2] = {(uint64_t)v, (uint64_t)(v >> 64)};
    b += "start " + hex_words(w, 2) + "\n";
  }
  std::snprintf(buf, sizeof(buf), "end %08x\n", fnv1a(b));
  b += buf;
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  const bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
  return std::fclose(f) == 0 && ok;
}

static std::string hex_start(uint64_t lo, uint64_t hi) {
  // "%X%08X%08X" of the start's 32-bit limbs 2, 1, 0 (collatz_cudav7.cu:298); for starts
  // below 2^96 the same text as the CPU form "%lX%016lX" (:691).
  char buf[64];
  if (hi >> 32) std::snprintf(buf, sizeof(buf), "%X%08X%08X%08X", (unsigned)(hi >> 32), (unsigned)hi,
                              (unsigned)(lo >> 32), (unsigned)lo);
  else std::snprintf(buf, sizeof(buf), "%X%08X%08X", (unsigned)hi, (unsigned)(lo >> 32), (unsigned)lo);
  return buf;
}

static int usage() {
  std::fprintf(stderr,
               "usage: collatz_run [--N n] [--A a] [--B b] [--cutoff c] [--semantics cuda|opencl|rust]"
               " [--spot-freq F] < case line\n"
               "       collatz_run [same knobs] --cases FILE\n"
               "       collatz_run --range LO HI [--k k] [--chunk c] [--cutoff c] [--limbs 2|3]"
               " [--form plain|persistent] [--lookahead B] [--out FILE]\n");
  return 2;
}

int main(int argc, char** argv) {
  uint32_t N = 45, A = 10, B = 24, cutoff = 0, k = 32, chunk = 24, limbs = 2, form = 0;
  uint32_t lookahead = 0, spot_freq = 2560000;
  bool a_given = false;
  const char* out_path = nullptr;
  const char* cases_path = nullptr;
  uint32_t sem = COLLATZ_SEM_CUDA;
  bool range = false;
  u128 lo = 0, hi = 0;
  for (int i = 1; i < argc; ++i) {
    const char* a = argv[i];
    auto next = [&]() -> const char* {
      if (i + 1 >= argc) std::exit(usage());
      return argv[++i];
    };
    if (!std::strcmp(a, "--N")) N = std::atoi(next());
    else if (!std::strcmp(a, "--A")) A = std::atoi(next()), a_given = true;
    else if (!std::strcmp(a, "--B")) B = std::atoi(next());
    else if (!std::strcmp(a, "--cutoff")) cutoff = std::atoi(next());
    else if (!std::strcmp(a, "--k")) k = std::atoi(next());
    else if (!std::strcmp(a, "--chunk")) chunk = std::atoi(next());
    else if (!std::strcmp(a, "--limbs")) limbs = std::atoi(next());
    else if (!std::strcmp(a, "--form")) form = !std::strcmp(next(), "persistent") ? 1 : 0;
    else if (!std::strcmp(a, "--lookahead")) lookahead = std::atoi(next());
    else if (!std::strcmp(a, "--out")) out_path = next();
    else if (!std::strcmp(a, "--cases")) cases_path = next();
    else if (!std::strcmp(a, "--spot-freq")) spot_freq = (uint32_t)std::strtoul(next(), nullptr, 10);
    else if (!std::strcmp(a, "--semantics")) {
      const char* v = next();
      sem = !std::strcmp(v, "rust") ? COLLATZ_SEM_RUST
            : !std::strcmp(v, "opencl") ? COLLATZ_SEM_OPENCL : COLLATZ_SEM_CUDA;
    } else if (!std::strcmp(a, "--range")) {
      range = true;
      lo = parse_u128(next());
      hi = parse_u128(next());
    } else {
      return usage();
    }
  }
  if (range) {
    collatz_sieve_params sp = {k, 1, -1, lookahead, sem, 0};
    collatz_sieve* sv = nullptr;
    if (collatz_sieve_build(&sp, &sv) != COLLATZ_OK) {
      std::fprintf(stderr, "sieve: %s\n", collatz_last_error());
      return 1;
    }
    const uint64_t nch = (uint64_t)(((hi - 1) >> chunk) - (lo >> chunk) + 1);
    std::vector<collatz_chunk> chunks(nch);
    std::vector<uint64_t> ovf(2 * 4096);
    collatz_range_result res;
    std::memset(&res, 0, sizeof(res));
    res.chunks = chunks.data();
    res.chunks_cap = nch;
    res.overflow_starts = ovf.data();
    res.overflow_cap = 4096;
    collatz_verify_opts o = {chunk, limbs, cutoff ? cutoff : 187u, form};
    const uint64_t l2[2] = {(uint64_t)lo, (uint64_t)(lo >> 64)};
    const uint64_t h2[2] = {(uint64_t)hi, (uint64_t)(hi >> 64)};
    const auto t0 = std::chrono::steady_clock::now();
    int rc = collatz_verify_range(sv, l2, h2, &o, nullptr, &res);
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != COLLATZ_OK && rc != COLLATZ_ECAPACITY) {
      std::fprintf(stderr, "verify: %s\n", collatz_last_error());
      return 1;
    }
    for (uint64_t i = 0; i < nch; ++i) {
      const collatz_chunk& c = chunks[i];
      const u128 ms = ((u128)c.max_start[1] << 64) | c.max_start[0];
      std::printf("chunk %" PRIu64 " n %" PRIu64 " count %" PRIu64 " stop_sum %" PRIu64
                  " stop_xor %" PRIu64 " hash %08X max_start %s flags %u\n",
                  c.chunk_index, c.n_checked, c.count, c.stop_sum, c.stop_xor, c.hash,
                  to_dec(ms).c_str(), c.flags);
    }
    const uint64_t nlisted = std::min<uint64_t>(res.overflow_written, 4096);
    for (uint64_t i = 0; i < nlisted; ++i)
      std::printf("overflow %s\n", hex_start(ovf[2 * i], ovf[2 * i + 1]).c_str());
    if (out_path &&
        !write_chunk_file(out_path, k, lookahead, sem, chunk, cutoff ? cutoff : 187u, lo, hi,
                          chunks, res.total, ovf.data(), nlisted)) {
      std::fprintf(stderr, "cannot write %s\n", out_path);
      return 1;
    }
    std::fprintf(stderr,
                 "Verified %s integers, %" PRIu64 " iterated, hash %08X, %.3f s (%.3e n/s).\n",
                 to_dec(hi - lo).c_str(), res.total.count, res.total.hash, s,
                 (double)(hi - lo) / s);
    collatz_sieve_free(sv);
    return 0;
  }
  const bool rust = sem == COLLATZ_SEM_RUST;
  // main.rs's knobs: LASTBITS 6, eight bit vectors and its overflow test at 2^120
  // (main.rs:20, :31, :336); MAXBITS stays --N (main.rs's 72 is beyond one run here).
  if (rust && !a_given) A = 6;
  collatz_ref_params p = {N, A, B, rust ? 8u : 16u, sem, cutoff ? cutoff : (rust ? 120u : 100u),
                          -1, spot_freq};
  std::vector<uint64_t> ovf(2 * (1 << 16));
  // One case line "base sigfig threepower number doubleeven numevens" through the whole
  // reference program; false (with a message) if the line does not parse or the run fails.
  // Under main.rs semantics the line is main.rs's input: its first four numbers, with
  // doubleeven and numevens 0 whatever follows (main.rs:88-97).
  auto run_line = [&](const std::string& line, collatz_ref_result& r) -> bool {
    std::istringstream iss(line);
    uint64_t t[6] = {0};
    for (int i = 0; i < (rust ? 4 : 6); ++i) {
      if (!(iss >> t[i])) {
        std::fprintf(stderr, "collatz_run: not a case line: %s\n", line.c_str());
        return false;
      }
    }
    if (rust) t[4] = t[5] = 0;
    collatz_node node;
    std::memset(&node, 0, sizeof(node));
    node.base[0] = t[0];
    node.sigfig = (uint32_t)t[1];
    node.threepower = (uint32_t)t[2];
    node.number[0] = t[3];
    node.doubleeven = (uint32_t)t[4];
    node.numevens = (uint32_t)t[5];
    std::memset(&r, 0, sizeof(r));
    r.overflow_starts = ovf.data();
    r.overflow_cap = 1 << 16;
    const int rc = collatz_reference_run(&p, &node, nullptr, &r);
    // A spot check that found a difference is reported as the reference reports it, and the
    // run's outputs are printed as usual (collatz_cudav7.cu:631-633).
    const bool spot_failed = rc == COLLATZ_EDEVICE && r.spot_check_failures > 0;
    if (rc != COLLATZ_OK && rc != COLLATZ_ECAPACITY && !spot_failed) {
      std::fprintf(stderr, "collatz_run: %s\n", collatz_last_error());
      return false;
    }
    for (uint64_t i = 0; i < r.spot_check_failures; ++i) std::fprintf(stderr, "Spot check failed\n");
    // Overflow starts: hex as the CUDA and OpenCL programs print them (collatz_cudav7.cu:298),
    // decimal as main.rs does (main.rs:337).
    for (uint64_t i = 0; i < std::min<uint64_t>(r.overflow_written, 1 << 16); ++i) {
      if (rust) std::printf("%s\n", to_dec(((u128)ovf[2 * i + 1] << 64) | ovf[2 * i]).c_str());
      else std::printf("%s\n", hex_start(ovf[2 * i], ovf[2 * i + 1]).c_str());
    }
    return true;
  };
  if (cases_path) {
    // The case-file driver: every line of the file in turn (what the reference's users do
    // with one process per line), one summary line each on stderr, then the XOR of the
    // complete hashes (for lines that partition a tree, the whole tree's hash).
    std::FILE* f = std::fopen(cases_path, "r");
    if (!f) {
      std::fprintf(stderr, "collatz_run: cannot open %s\n", cases_path);
      return 1;
    }
    char buf[512];
    uint32_t all = 0;
    uint64_t lines = 0, launches = 0, cands = 0, numcases = 0, spots = 0;
    const auto t0 = std::chrono::steady_clock::now();
    while (std::fgets(buf, sizeof(buf), f)) {
      std::string line(buf);
      while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
      if (line.find_first_not_of(" \t") == std::string::npos) continue;
      collatz_ref_result r;
      if (!run_line(line, r)) {
        std::fclose(f);
        return 1;
      }
      if (rust)
        std::fprintf(stderr, "%s: numcases %" PRIu64 "\n", line.c_str(), r.numcases);
      else
        std::fprintf(stderr, "%s: launches %" PRIu64 " hash %X\n", line.c_str(), r.launches,
                     r.complete_hash);
      all ^= r.complete_hash;
      ++lines;
      launches += r.launches;
      cands += r.candidates;
      numcases += r.numcases;
      spots += r.spot_checks;
    }
    std::fclose(f);
    std::fflush(stdout);
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rust)
      std::fprintf(stderr,
                   "%" PRIu64 " case lines, %" PRIu64 " candidates, numcases = %" PRIu64
                   ", spot checks %" PRIu64 ", %.3f s.\n", lines, cands, numcases, spots, s);
    else
      std::fprintf(stderr,
                   "%" PRIu64 " case lines, %" PRIu64 " launches, %" PRIu64 " candidates, "
                   "XOR of complete hashes = %X, spot checks %" PRIu64 ", %.3f s.\n", lines,
                   launches, cands, all, spots, s);
    return 0;
  }
  std::string line;
  if (rust) {
    // main.rs reads the whole of stdin and splits it on whitespace (main.rs:88-94).
    std::ostringstream in;
    in << std::cin.rdbuf();
    line = in.str();
    // main.rs prints this once its tables are built (main.rs:85-86); here the run itself
    // builds them on the device.
    std::fprintf(stderr, "Done with precompute\n");
  } else {
    std::getline(std::cin, line);
  }
  collatz_ref_result r;
  if (!run_line(line, r)) return 1;
  std::fflush(stdout);
  if (rust) {
    std::fprintf(stderr, "numcases = %" PRIu64 "\n", r.numcases);
  } else if (sem == COLLATZ_SEM_OPENCL) {
    std::fprintf(stderr, "Number of kernel launches = %" PRIu64 " \n", r.launches);
    std::fprintf(stderr, "Spot checks = %u. Complete hash = %X.\n", (uint32_t)r.spot_checks,
                 r.complete_hash);
  } else {
    std::fprintf(stderr, "Number of kernel launches = %" PRIu64 ". ", r.launches);
    std::fprintf(stderr, "Spot checks = %u. Complete hash = %X.\n", (uint32_t)r.spot_checks,
                 r.complete_hash);
  }
  return 0;
}
