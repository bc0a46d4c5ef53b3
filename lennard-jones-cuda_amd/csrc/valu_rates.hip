// valu_rates.hip -- issue cost of the vector instructions of the pair loops,
// measured on the device (a standalone program, not part of libljhip.so).
//
// Each kernel runs 8 independent chains per lane of one instruction kind
// (or a mix) for kIter iterations, 8 four-wave workgroups per CU (8 waves per
// SIMD, as the force kernels run), and the time per wave instruction per SIMD
// is reported: for a mix, the time of one iteration of the mix.  Comparing a
// mix against the sum of its parts shows which instructions issue beside the
// FMA pipe.  The clocks are settled first by ~0.5 s of FMA load.
//
// Output: one JSON line per kernel.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kIter = 4096;

#define CHECK(x)                                                                   \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                 \
      std::exit(1);                                                                \
    }                                                                              \
  } while (0)

enum Op { FMA, MUL, RCP, CVT, SUB, FMA8_RCP2, FMA8_CVT4, FMA8_SUB4, kOps };
const char* kName[kOps] = {"v_fma_f32",       "v_mul_f32",       "v_rcp_f32",
                           "v_cvt_f32_i32",   "v_sub_u32",       "8 fma + 2 rcp",
                           "8 fma + 4 cvt",   "8 fma + 4 sub"};
// instructions per iteration of each kernel (8 chains)
const int kPerIter[kOps] = {8, 8, 8, 8, 8, 10, 12, 12};

__device__ __forceinline__ unsigned cvt_bits(unsigned u) { return __float_as_uint((float)(int)u); }

template <int OP>
__global__ __launch_bounds__(256) void k_chain(float* out, float a, float b)
{
  float x[8];
  unsigned u[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    x[c] = 1e-3f * (float)threadIdx.x + (float)c;
    u[c] = threadIdx.x + (unsigned)c;
  }
  for (int i = 0; i < kIter; ++i) {
    if constexpr (OP == FMA) {
#pragma unroll
      for (int c = 0; c < 8; ++c) x[c] = __builtin_fmaf(x[c], a, b);
    } else if constexpr (OP == MUL) {
#pragma unroll
      for (int c = 0; c < 8; ++c) x[c] *= a;
    } else if constexpr (OP == RCP) {
#pragma unroll
      for (int c = 0; c < 8; ++c) x[c] = __builtin_amdgcn_rcpf(x[c]);
    } else if constexpr (OP == CVT) {
#pragma unroll
      for (int c = 0; c < 8; ++c) u[c] = cvt_bits(u[c]);
    } else if constexpr (OP == SUB) {
#pragma unroll
      for (int c = 0; c < 8; ++c) u[c] = u[(c + 1) & 7] - u[c];
    } else if constexpr (OP == FMA8_RCP2) {
#pragma unroll
      for (int c = 0; c < 6; ++c) x[c] = __builtin_fmaf(x[c], a, b);
      x[6] = __builtin_amdgcn_rcpf(x[6]);
      x[7] = __builtin_amdgcn_rcpf(x[7]);
      x[0] = __builtin_fmaf(x[0], a, b);
      x[1] = __builtin_fmaf(x[1], a, b);
    } else if constexpr (OP == FMA8_CVT4) {
#pragma unroll
      for (int c = 0; c < 8; ++c) x[c] = __builtin_fmaf(x[c], a, b);
#pragma unroll
      for (int c = 0; c < 4; ++c) u[c] = cvt_bits(u[c]);
    } else {
#pragma unroll
      for (int c = 0; c < 8; ++c) x[c] = __builtin_fmaf(x[c], a, b);
#pragma unroll
      for (int c = 0; c < 4; ++c) u[c] = u[(c + 1) & 3] - u[c];
    }
  }
  float r = 0.f;
#pragma unroll
  for (int c = 0; c < 8; ++c) r += x[c] + (float)u[c];
  out[blockIdx.x * 256 + threadIdx.x] = r;
}

template <int OP>
float time_ms(float* d, int blocks, int reps)
{
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_chain<OP>, dim3(blocks), dim3(256), 0, 0, d, 0.999f, 1e-3f);
  CHECK(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_chain<OP>, dim3(blocks), dim3(256), 0, 0, d, 0.999f, 1e-3f);
  CHECK(hipEventRecord(e1, 0));
  CHECK(hipEventSynchronize(e1));
  CHECK(hipGetLastError());
  float ms = 0.f;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  return ms / reps;
}

template <int OP>
void report(float* d, int blocks, int simds)
{
  const float ms = time_ms<OP>(d, blocks, 10);
  // iterations per SIMD: waves per SIMD x kIter
  const double iters = (double)blocks * 4 / simds * kIter;
  const double ns_iter = ms * 1e6 / iters;
  std::printf("{\"kernel\": \"%s\", \"ms\": %.4f, \"ns_per_iteration_per_simd\": %.4f, "
              "\"instructions_per_iteration\": %d, \"ns_per_instruction\": %.4f}\n",
              kName[OP], ms, ns_iter, kPerIter[OP], ns_iter / kPerIter[OP]);
}

}  // namespace

int main()
{
  int dev = 0, cus = 0;
  CHECK(hipGetDevice(&dev));
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = cus * 8;  // eight four-wave workgroups per CU
  float* d = nullptr;
  CHECK(hipMalloc(&d, sizeof(float) * (size_t)blocks * 256));
  for (int r = 0; r < 100; ++r) time_ms<FMA>(d, blocks, 1);  // settle the clocks
  std::printf("{\"cus\": %d, \"waves_per_simd\": 8, \"iterations\": %d}\n", cus, kIter);
  report<FMA>(d, blocks, cus * 4);
  report<MUL>(d, blocks, cus * 4);
  report<RCP>(d, blocks, cus * 4);
  report<CVT>(d, blocks, cus * 4);
  report<SUB>(d, blocks, cus * 4);
  report<FMA8_RCP2>(d, blocks, cus * 4);
  report<FMA8_CVT4>(d, blocks, cus * 4);
  report<FMA8_SUB4>(d, blocks, cus * 4);
  CHECK(hipFree(d));
  return 0;
}
