// valu_peak: the int32 VALU issue ceiling of this GPU, measured.
//
// Two kinds of kernels, each at full occupancy (32 waves per CU) over every CU:
//   * one instruction kind (the ones the range kernels issue most: add, add with carry,
//     32-bit multiply low/high, 32x32->64 multiply-add, funnel shift, find-first-bit,
//     compare + select, shift, xor), written in inline asm so that the count is exact,
//     in eight independent chains per lane (latency hidden by the chains and the waves);
//   * the reference schedule's check (two odd and two even runs on three 32-bit words,
//     the step sum and the hash, schedule.hpp) in a register loop with no memory traffic:
//     the ceiling of the algorithm's own instruction mix.
// Each kernel runs three times; the fastest counts.  Lane 0 of the first wave reads the
// shader clock counter (s_memtime) and the constant-rate clock (s_memrealtime) at its
// start and end, which gives the shader clock during the run.  Prints one JSON object.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "schedule.hpp"

namespace {

constexpr int kBlock = 256;
constexpr int kChains = 8;

struct ClockSample {
  unsigned long long t0, t1, r0, r1;
};

__device__ __forceinline__ void clock_begin(ClockSample* cs) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    cs->t0 = __builtin_amdgcn_s_memtime();
    cs->r0 = __builtin_amdgcn_s_memrealtime();
  }
}

__device__ __forceinline__ void clock_end(ClockSample* cs) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    cs->t1 = __builtin_amdgcn_s_memtime();
    cs->r1 = __builtin_amdgcn_s_memrealtime();
  }
}

enum Op {
  kAdd, kAddCarry, kMulLo, kMulHi, kMad64, kAlignbit, kFfbl, kCmpSel, kShift, kXor,
  kMul24, kMulHi24, kMad24, kAdd3, kLshlOr, kOr, kBfe, kNumOps
};
const char* kOpNames[kNumOps] = {"v_add_u32", "v_add_co_u32+v_addc_co_u32", "v_mul_lo_u32",
                                 "v_mul_hi_u32", "v_mad_u64_u32", "v_alignbit_b32",
                                 "v_ffbl_b32", "v_cmp_gt_u32+v_cndmask_b32", "v_lshrrev_b32",
                                 "v_xor_b32", "v_mul_u32_u24", "v_mul_hi_u32_u24",
                                 "v_mad_u32_u24", "v_add3_u32", "v_lshl_or_b32", "v_or_b32",
                                 "v_bfe_u32"};

// Eight chains of `iters` x 4 instructions of kind OP per lane (a pair counts as two).
template <int OP>
__global__ __launch_bounds__(kBlock) void k_op(uint32_t iters, uint32_t seed, uint32_t* out,
                                               ClockSample* cs) {
  clock_begin(cs);
  uint32_t x[kChains];
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
#pragma unroll
  for (int c = 0; c < kChains; ++c) x[c] = t * 2654435761u + c * 40503u + seed;
  uint32_t y = seed | 1u, z = t | 0x10001u;
  for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if constexpr (OP == kAdd) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) asm volatile("v_add_u32_e32 %0, %0, %1" : "+v"(x[c]) : "v"(y));
      } else if constexpr (OP == kAddCarry) {
#pragma unroll
        for (int c = 0; c < kChains; c += 2) {
          uint64_t cy;
          asm volatile(
              "v_add_co_u32_e64 %0, %2, %0, %3\n\t"
              "v_addc_co_u32_e64 %1, %2, %1, %3, %2"
              : "+v"(x[c]), "+v"(x[c + 1]), "=&s"(cy)
              : "v"(y));
        }
      } else if constexpr (OP == kMulLo) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(x[c]) : "v"(z));
      } else if constexpr (OP == kMulHi) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(x[c]) : "v"(z));
      } else if constexpr (OP == kMad64) {
#pragma unroll
        for (int c = 0; c < kChains; c += 2) {
          uint64_t acc = ((uint64_t)x[c + 1] << 32) | x[c];
          uint64_t cy;
          asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(acc), "=&s"(cy) : "v"(y), "v"(z));
          x[c] = (uint32_t)acc;
          x[c + 1] = (uint32_t)(acc >> 32);
        }
      } else if constexpr (OP == kAlignbit) {
#pragma unroll
        for (int c = 0; c < kChains; ++c)
          asm volatile("v_alignbit_b32 %0, %0, %1, %2" : "+v"(x[c]) : "v"(y), "v"(z));
      } else if constexpr (OP == kFfbl) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) asm volatile("v_ffbl_b32_e32 %0, %0" : "+v"(x[c]));
      } else if constexpr (OP == kCmpSel) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) {
          uint64_t m;
          asm volatile(
              "v_cmp_gt_u32_e64 %1, %0, %2\n\t"
              "v_cndmask_b32_e64 %0, %0, %2, %1"
              : "+v"(x[c]), "=&s"(m)
              : "v"(y));
        }
      } else if constexpr (OP == kShift) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) asm volatile("v_lshrrev_b32_e32 %0, %1, %0" : "+v"(x[c]) : "v"(y));
      } else if constexpr (OP == kXor) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) asm volatile("v_xor_b32_e32 %0, %0, %1" : "+v"(x[c]) : "v"(y));
      } else if constexpr (OP == kMul24) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) asm volatile("v_mul_u32_u24_e32 %0, %0, %1" : "+v"(x[c]) : "v"(z));
      } else if constexpr (OP == kMulHi24) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) asm volatile("v_mul_hi_u32_u24_e32 %0, %0, %1" : "+v"(x[c]) : "v"(z));
      } else if constexpr (OP == kMad24) {
#pragma unroll
        for (int c = 0; c < kChains; ++c)
          asm volatile("v_mad_u32_u24 %0, %0, %1, %2" : "+v"(x[c]) : "v"(z), "v"(y));
      } else if constexpr (OP == kAdd3) {
#pragma unroll
        for (int c = 0; c < kChains; ++c)
          asm volatile("v_add3_u32 %0, %0, %1, %2" : "+v"(x[c]) : "v"(y), "v"(z));
      } else if constexpr (OP == kLshlOr) {
#pragma unroll
        for (int c = 0; c < kChains; ++c)
          asm volatile("v_lshl_or_b32 %0, %0, %1, %2" : "+v"(x[c]) : "v"(y), "v"(z));
      } else if constexpr (OP == kOr) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) asm volatile("v_or_b32_e32 %0, %0, %1" : "+v"(x[c]) : "v"(y));
      } else {
#pragma unroll
        for (int c = 0; c < kChains; ++c)
          asm volatile("v_bfe_u32 %0, %0, %1, %2" : "+v"(x[c]) : "v"(y), "v"(z));
      }
    }
  }
  uint32_t h = 0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) h ^= x[c];
  out[t] = h;
  clock_end(cs);
}

// VALU instructions per lane and loop iteration of k_op<OP> (the compare + select pair
// and the add-with-carry pair count as two; the multiply-add chains are four of eight).
constexpr double kInstrPerIter[kNumOps] = {32, 32, 32, 32, 16, 32, 32, 64, 32, 32,
                                           32, 32, 32, 32, 32, 32, 32};

// The schedule's check on three 32-bit words in registers: odd run, even run, odd run,
// even run, the step sum and the hash (schedule.hpp), `iters` times per lane.
__global__ __launch_bounds__(kBlock) void k_checks(uint32_t iters, uint32_t seed, uint32_t* out,
                                                   ClockSample* cs) {
  __shared__ cz::Pow3Lds p3;
  cz::stage_pow3(p3);
  __syncthreads();
  clock_begin(cs);
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  cz::Num<3> v;
  v.w[0] = t * 2654435761u + seed;
  v.w[1] = t ^ 0x9E3779B9u;
  v.w[2] = 1u + (t & 7u);
  uint32_t steps = 0, hash = 0;
  for (uint32_t i = 0; i < iters; ++i) {
    const uint32_t a1 = cz::odd_run<3>(v, p3);
    const uint32_t b1 = cz::even_run<3>(v);
    const uint32_t a2 = cz::odd_run<3>(v, p3);
    const uint32_t b2 = cz::even_run<3>(v);
    steps += a1 + b1 + a2 + b2;
    hash ^= v.w[0];
    // Restart from a fresh large value when the trajectory has fallen to the 2, 3 cycle,
    // so every check runs on a multi-word value (uniform in the count of instructions).
    if ((v.w[1] | v.w[2]) == 0u) {
      v.w[1] = hash | 1u;
      v.w[2] = 1u;
    }
  }
  out[t] = steps ^ hash ^ v.w[1];
  clock_end(cs);
}

#define CHECK(x)                                                                   \
  do {                                                                             \
    hipError_t e = (x);                                                            \
    if (e != hipSuccess) {                                                         \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));                  \
      std::exit(1);                                                                \
    }                                                                              \
  } while (0)

struct Timed {
  double ms;
  double sclk_mhz;
};

template <typename F>
Timed time_kernel(F launch, ClockSample* dcs, double rtc_khz) {
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  Timed best{1e30, 0};
  for (int rep = 0; rep < 3; ++rep) {
    CHECK(hipEventRecord(a, nullptr));
    launch();
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(b, nullptr));
    CHECK(hipEventSynchronize(b));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, a, b));
    ClockSample cs;
    CHECK(hipMemcpy(&cs, dcs, sizeof(cs), hipMemcpyDeviceToHost));
    const double rt_s = (double)(cs.r1 - cs.r0) / (rtc_khz * 1e3);
    const double mhz = rt_s > 0 ? (double)(cs.t1 - cs.t0) / rt_s / 1e6 : 0;
    if (ms < best.ms) best = Timed{ms, mhz};
  }
  CHECK(hipEventDestroy(a));
  CHECK(hipEventDestroy(b));
  return best;
}

}  // namespace

int main(int argc, char** argv) {
  const uint32_t iters = argc > 1 ? (uint32_t)std::strtoul(argv[1], nullptr, 10) : 16384;
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  int rtc_khz = 0;
  CHECK(hipDeviceGetAttribute(&rtc_khz, hipDeviceAttributeWallClockRate, 0));
  const int cus = prop.multiProcessorCount;
  const uint32_t blocks = (uint32_t)cus * 8;  // 8 blocks of 4 waves: 32 waves per CU
  const double lanes = (double)blocks * kBlock;
  uint32_t* dout = nullptr;
  ClockSample* dcs = nullptr;
  CHECK(hipMalloc(&dout, (size_t)blocks * kBlock * 4));
  CHECK(hipMalloc(&dcs, sizeof(ClockSample)));
  std::string js = "{\"device\": \"" + std::string(prop.gcnArchName) + "\", \"cus\": " +
                   std::to_string(cus) + ", \"max_clock_mhz\": " +
                   std::to_string(prop.clockRate / 1000) + ", \"rtc_khz\": " +
                   std::to_string(rtc_khz) + ", \"waves_per_cu\": 32, \"ops\": {";
  auto report = [&](const char* name, double instr_per_lane, Timed tm, bool first) {
    const double lane_ops = instr_per_lane * lanes;
    const double tops = lane_ops / (tm.ms * 1e-3) / 1e12;
    // Wave instructions per CU per shader clock: 1.0 is one wave64 instruction per cycle
    // per CU (four SIMDs, each taking a wave64 instruction every four cycles).
    const double per_cu_cycle = lane_ops / 64 / cus / (tm.ms * 1e-3 * tm.sclk_mhz * 1e6);
    char buf[256];
    std::snprintf(buf, sizeof(buf),
                  "%s\"%s\": {\"tops\": %.3f, \"ms\": %.3f, \"sclk_mhz\": %.0f, "
                  "\"wave_instr_per_cu_cycle\": %.3f}",
                  first ? "" : ", ", name, tops, tm.ms, tm.sclk_mhz, per_cu_cycle);
    js += buf;
  };
  const dim3 grid(blocks), block(kBlock);
#define RUN_OP(OP)                                                                           \
  {                                                                                          \
    Timed tm = time_kernel([&] { hipLaunchKernelGGL(k_op<OP>, grid, block, 0, nullptr, iters, \
                                                    7u, dout, dcs); }, dcs, rtc_khz);        \
    report(kOpNames[OP], kInstrPerIter[OP] * iters, tm, OP == 0);                            \
  }
  RUN_OP(kAdd) RUN_OP(kAddCarry) RUN_OP(kMulLo) RUN_OP(kMulHi) RUN_OP(kMad64)
  RUN_OP(kAlignbit) RUN_OP(kFfbl) RUN_OP(kCmpSel) RUN_OP(kShift) RUN_OP(kXor)
  RUN_OP(kMul24) RUN_OP(kMulHi24) RUN_OP(kMad24) RUN_OP(kAdd3) RUN_OP(kLshlOr) RUN_OP(kOr)
  RUN_OP(kBfe)
#undef RUN_OP
  js += "}";
  // The schedule's check loop: checks per second, and the algorithmic lane-ops of the
  // check at three words (13 L + 7 = 46) per second.
  const uint32_t citers = iters / 4;
  Timed tm = time_kernel([&] { hipLaunchKernelGGL(k_checks, grid, block, 0, nullptr, citers,
                                                  11u, dout, dcs); }, dcs, rtc_khz);
  const double checks = (double)citers * lanes;
  char buf[320];
  std::snprintf(buf, sizeof(buf),
                ", \"checks_loop\": {\"words\": 3, \"checks_per_s\": %.4e, "
                "\"algorithmic_tops\": %.3f, \"ms\": %.3f, \"sclk_mhz\": %.0f}}",
                checks / (tm.ms * 1e-3), checks * 46 / (tm.ms * 1e-3) / 1e12, tm.ms,
                tm.sclk_mhz);
  js += buf;
  std::printf("%s\n", js.c_str());
  CHECK(hipFree(dout));
  CHECK(hipFree(dcs));
  return 0;
}
