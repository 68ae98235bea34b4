// Device helpers shared by the radix sort and the kernels that produce its
// digit histogram (radix_sort.hip, dem_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "fc_internal.hpp"

namespace fc {

// Exclusive prefix sum of v over threads 0..kN-1 (kN a multiple of 64, at
// most the block size); every thread of the block must call it.  Threads
// from kN on contribute nothing and get an unspecified result.  s_scratch
// holds kN / 64 words.
template <int kN>
__device__ __forceinline__ uint32_t scan_exclusive(uint32_t v, uint32_t* s_scratch) {
  constexpr uint32_t kW = (uint32_t)kN / 64u;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(inc, off);
    if (lane >= (uint32_t)off) inc += t;
  }
  if (lane == 63u && wave < kW) s_scratch[wave] = inc;
  __syncthreads();
  uint32_t base = 0;
  for (uint32_t w = 0; w < wave && w < kW; ++w) base += s_scratch[w];
  __syncthreads();
  return base + inc - v;
}

// Block-wide exclusive scan of the values held by threads 0..255 (one per
// digit).  Every thread of the block must call it (it synchronises); threads
// from 256 on contribute nothing and get an unspecified result.
// s_scratch holds 4 words.
__device__ __forceinline__ uint32_t scan256_exclusive(uint32_t v, uint32_t* s_scratch) {
  return scan_exclusive<256>(v, s_scratch);
}

// Histogram dataflow of the sort (no tickets, no fences, no last block):
//   H[p] = gHist + p * RS_HIST_COPIES * RS_RADIX holds the digit-p histogram
//   of the keys in RS_HIST_COPIES privatised copies (block b adds to copy
//   b % RS_HIST_COPIES so that blocks adding the same digit do not all
//   contend for one address).
//   - The kernel that produces the keys counts digit 0 into H[0]; its block 0
//     zeroes H[1..] (none is in use outside the scatter passes) and bumps
//     the sort epoch (no other block of that kernel reads either).
//   - Scatter pass p sums H[p] in every block and counts digit p+1 of its
//     keys into H[p+1]; block 0 of pass 1 zeroes H[0] for the next sort.
// This holds whatever the pass count of consecutive sorts.
// Every consumer runs in a later launch than the producer of what it reads,
// so plain (non-returning) atomics suffice.

__device__ __forceinline__ uint32_t* hist_buffer(uint32_t* gHist, int pass) {
  return gHist + (size_t)pass * RS_HIST_COPIES * RS_RADIX;
}

// Adds digit d of one key per lane to a 256-bin LDS histogram.  When every
// active lane of the wave holds the same digit (typical for the high digits
// of spatially ordered keys) one lane adds the whole count, so same-address
// LDS atomics do not serialise.  Must be called by whole waves.
__device__ __forceinline__ void histogram_add_digit(uint32_t* s_hist, uint32_t d, bool valid) {
  const uint64_t active = __ballot(valid);
  const uint32_t first = __shfl(d, (int)(__ffsll((long long)active) - 1));
  const uint64_t same = __ballot(valid && d == first);
  if (same == active) {
    if (valid && (active & ((1ull << (threadIdx.x & 63u)) - 1ull)) == 0ull)
      atomicAdd(&s_hist[d], (uint32_t)__popcll(active));
  } else if (valid) {
    atomicAdd(&s_hist[d], 1u);
  }
}

// Adds a complete, synchronised 256-bin LDS histogram to H[pass].
__device__ __forceinline__ void histogram_flush(const uint32_t* s_hist, uint32_t* gHist,
                                                int pass) {
  uint32_t* copy = hist_buffer(gHist, pass) + (size_t)(blockIdx.x % RS_HIST_COPIES) * RS_RADIX;
  for (int i = threadIdx.x; i < RS_RADIX; i += blockDim.x) {
    const uint32_t c = s_hist[i];
    if (c) atomicAdd(&copy[i], c);
  }
}

// Block 0 of a key-producing kernel: zero H[1..] and bump the epoch for the
// coming sort.
__device__ __forceinline__ void histogram_producer_prologue(uint32_t* gHist, int passes,
                                                            uint32_t* counters) {
  (void)passes;
  if (blockIdx.x != 0) return;
  uint32_t* h = hist_buffer(gHist, 1);
  for (int i = threadIdx.x; i < (RS_MAX_PASSES - 1) * RS_HIST_COPIES * RS_RADIX; i += blockDim.x)
    h[i] = 0u;
  if (threadIdx.x == 0) counters[RSC_EPOCH] = counters[RSC_EPOCH] + 1u;
}

// Total of digit d (thread d < 256) over the copies of H[pass].
__device__ __forceinline__ uint32_t histogram_total(const uint32_t* gHist, int pass, uint32_t d) {
  const uint32_t* h = hist_buffer(const_cast<uint32_t*>(gHist), pass);
  uint32_t t = 0;
#pragma unroll
  for (int c = 0; c < RS_HIST_COPIES; ++c) t += h[(size_t)c * RS_RADIX + d];
  return t;
}

}  // namespace fc
