// Exhaustive check of the fast correctly rounded square root and reciprocal
// of the contact loop (sqrt_rn_mid, rcp_rn_mid) against the compiler's
// correctly rounded sqrtf and 1.0f / x, on every float of their ranges.
#include "fc_internal.hpp"
#include "fc_device.hpp"

namespace fc {
namespace {

constexpr uint32_t kSqrtLo = 31u << 23;   // 2^-96
constexpr uint32_t kSqrtHi = 223u << 23;  // 2^96
constexpr uint32_t kRcpLo = 79u << 23;    // 2^-48
constexpr uint32_t kRcpHi = 175u << 23;   // 2^48

// out: [0] sqrt mismatches, [1] reciprocal mismatches (all-ones significands
// excluded), [2] reciprocal mismatches at all-ones significands;
// firstBad: [0] smallest mismatching sqrt argument, [1] smallest
// mismatching reciprocal argument (both as bits, 0xffffffff if none).
__global__ __launch_bounds__(256) void k_check_pair_math(unsigned long long* out,
                                                         uint32_t* firstBad) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t first = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long badSqrt = 0, badRcp = 0, badOnes = 0;
  uint32_t minSqrt = 0xffffffffu, minRcp = 0xffffffffu;
  for (uint32_t i = kSqrtLo + first; i <= kSqrtHi; i += stride) {
    const float x = __uint_as_float(i);
    if (__float_as_uint(sqrt_rn_mid(x)) != __float_as_uint(sqrtf(x))) {
      ++badSqrt;
      minSqrt = i < minSqrt ? i : minSqrt;
    }
  }
  for (uint32_t i = kRcpLo + first; i <= kRcpHi; i += stride) {
    const float b = __uint_as_float(i);
    const bool bad = __float_as_uint(rcp_rn_mid(b)) != __float_as_uint(1.0f / b);
    if ((i & 0x7fffffu) == 0x7fffffu) {
      badOnes += bad ? 1u : 0u;
    } else if (bad) {
      ++badRcp;
      minRcp = i < minRcp ? i : minRcp;
    }
  }
  if (badSqrt) atomicAdd(&out[0], badSqrt);
  if (badRcp) atomicAdd(&out[1], badRcp);
  if (badOnes) atomicAdd(&out[2], badOnes);
  if (minSqrt != 0xffffffffu) atomicMin(&firstBad[0], minSqrt);
  if (minRcp != 0xffffffffu) atomicMin(&firstBad[1], minRcp);
}

}  // namespace
}  // namespace fc

extern "C" fc_status fc_selftest_pair_math(fc_context* ctx, uint64_t out[5]) {
  using namespace fc;
  if (!ctx || !out) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  unsigned long long* dCounts = nullptr;
  uint32_t* dFirst = nullptr;
  FC_HIP_TRY(hipMalloc(&dCounts, 3 * sizeof(unsigned long long)));
  hipError_t e = hipMalloc(&dFirst, 2 * sizeof(uint32_t));
  const uint32_t none[2] = {0xffffffffu, 0xffffffffu};
  unsigned long long counts[3] = {0, 0, 0};
  uint32_t firstBad[2] = {0xffffffffu, 0xffffffffu};
  if (e == hipSuccess)
    e = hipMemcpyAsync(dCounts, counts, sizeof(counts), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dFirst, none, sizeof(none), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_check_pair_math, dim3(4096), dim3(256), 0, ctx->stream, dCounts, dFirst);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(counts, dCounts, sizeof(counts), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(firstBad, dFirst, sizeof(firstBad), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(dCounts);
  if (dFirst) hipFree(dFirst);
  if (e != hipSuccess) return hip_error(e, "pair math self-test");
  out[0] = counts[0];
  out[1] = counts[1];
  out[2] = counts[2];
  out[3] = firstBad[0];
  out[4] = firstBad[1];
  return FC_SUCCESS;
}
