// Device helpers shared by the radix sort and the kernels that produce its
// digit histogram (radix_sort.hip, dem_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "fc_internal.hpp"

namespace fc {

// Block-wide exclusive scan of the values held by threads 0..255 (one per
// digit).  Every thread of the block must call it (it synchronises); threads
// from 256 on contribute nothing and get an unspecified result.
// s_scratch holds 4 words.
__device__ __forceinline__ uint32_t scan256_exclusive(uint32_t v, uint32_t* s_scratch) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(inc, off);
    if (lane >= (uint32_t)off) inc += t;
  }
  if (lane == 63u && wave < 4u) s_scratch[wave] = inc;
  __syncthreads();
  uint32_t base = 0;
  for (uint32_t w = 0; w < wave && w < 4u; ++w) base += s_scratch[w];
  __syncthreads();
  return base + inc - v;
}

// Adds one key's digits to a block histogram in LDS.  When every active lane
// of the wave holds the same digit (typical for the high digits of spatially
// ordered keys) one lane adds the whole count, so same-address LDS atomics
// do not serialise.
__device__ __forceinline__ void histogram_add_key(uint32_t* s_hist, int passes, uint32_t key,
                                                  bool valid) {
  const uint64_t active = __ballot(valid);
  for (int p = 0; p < passes; ++p) {
    const uint32_t d = (key >> (p * RS_RADIX_BITS)) & (RS_RADIX - 1);
    const uint32_t first = __shfl(d, (int)(__ffsll((long long)active) - 1));
    const uint64_t same = __ballot(valid && d == first);
    if (same == active) {
      if (valid && (active & ((1ull << (threadIdx.x & 63u)) - 1ull)) == 0ull)
        atomicAdd(&s_hist[p * RS_RADIX + d], (uint32_t)__popcll(active));
    } else if (valid) {
      atomicAdd(&s_hist[p * RS_RADIX + d], 1u);
    }
  }
}

// Ends a histogram launch over n keys.  s_hist holds this block's counts
// ([passes][256], complete and synchronised).  Non-zero bins are added to
// copy (blockIdx % RS_HIST_COPIES) of the global histogram, so that many
// blocks adding the same digit do not all contend for one address.  The last
// block to take a ticket sums the copies into exclusive bucket offsets per
// pass, zeroes every copy for the next sort, resets the ticket and bumps the
// sort epoch.
//
// No release fence is used: a device-scope fence writes back the whole L2
// of the XCD, which in a kernel that just streamed its particle data costs
// far more than the histogram.  Instead the last block re-reads the copies
// until every pass sums to n, which proves that every block's adds have
// landed (the wait is bounded; hitting the bound raises DEVERR_HIST_STALLED).
// Every thread of the block (at least 256) must call it.
__device__ __forceinline__ void histogram_flush_and_finish(const uint32_t* s_hist, int passes,
                                                           uint32_t n, uint32_t* gHist,
                                                           uint32_t* offsets, uint32_t* counters,
                                                           uint32_t* devFlags,
                                                           uint32_t* s_scratch) {
  __shared__ uint32_t s_last;
  __shared__ uint32_t s_sum[RS_MAX_PASSES];
  uint32_t* copy = gHist + (size_t)(blockIdx.x % RS_HIST_COPIES) * RS_MAX_PASSES * RS_RADIX;
  for (int i = threadIdx.x; i < passes * RS_RADIX; i += blockDim.x) {
    const uint32_t c = s_hist[i];
    if (c) atomicAdd(&copy[i], c);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t ticket = atomicAdd(&counters[RSC_HIST_DONE], 1u);
    s_last = (ticket == gridDim.x - 1u) ? 1u : 0u;
  }
  __syncthreads();
  if (!s_last) return;

  const uint32_t d = threadIdx.x;
  uint32_t total[RS_MAX_PASSES];
  for (uint32_t spins = 0;; ++spins) {
    if (threadIdx.x < RS_MAX_PASSES) s_sum[threadIdx.x] = 0u;
    __syncthreads();
    for (int p = 0; p < passes; ++p) {
      total[p] = 0;
      if (d < (uint32_t)RS_RADIX) {
        for (int c = 0; c < RS_HIST_COPIES; ++c)
          total[p] += __hip_atomic_load(gHist + ((size_t)c * RS_MAX_PASSES + p) * RS_RADIX + d,
                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (total[p]) atomicAdd(&s_sum[p], total[p]);
      }
    }
    __syncthreads();
    bool complete = true;
    for (int p = 0; p < passes; ++p) complete = complete && s_sum[p] == n;
    __syncthreads();
    if (complete) break;
    if (spins > (1u << 20)) {
      if (threadIdx.x == 0) atomicOr(devFlags, DEVERR_HIST_STALLED);
      break;
    }
    __builtin_amdgcn_s_sleep(2);
  }
  for (int p = 0; p < passes; ++p) {
    if (d < (uint32_t)RS_RADIX)
      for (int c = 0; c < RS_HIST_COPIES; ++c)
        __hip_atomic_store(gHist + ((size_t)c * RS_MAX_PASSES + p) * RS_RADIX + d, 0u,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t excl = scan256_exclusive(d < (uint32_t)RS_RADIX ? total[p] : 0u, s_scratch);
    if (d < (uint32_t)RS_RADIX) offsets[p * RS_RADIX + d] = excl;
  }
  if (threadIdx.x == 0) {
    counters[RSC_HIST_DONE] = 0u;
    counters[RSC_EPOCH] = counters[RSC_EPOCH] + 1u;
  }
}

}  // namespace fc
