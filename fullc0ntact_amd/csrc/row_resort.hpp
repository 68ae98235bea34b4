// Per-row counts of the next step's keys, for the row resort of a
// multi-step run (row_resort.hip).  A grid row is the gridSize.x cells
// that share (y, z); its hash range is [row << shift, (row + 1) << shift)
// with shift = log2(gridSize.x), so the spheres of a row are one contiguous
// range of the sorted order.
#pragma once

#include "fc_internal.hpp"

namespace fc {

constexpr uint32_t kRowGroupShift = 8;  // rows per group of the two-level scan: 256

// Set in a next key stored for the row resort when the sphere changes cell
// (hashes are below 2^30 wherever the row resort runs).
constexpr uint32_t kMovedBit = 0x80000000u;

// The per-row counters come in kRowCopies copies, [row][copy]: the counts
// of the sphere at sorted slot k go to copy (k >> 8) % kRowCopies, so the
// many waves that count into one row at the same moment spread their
// atomics over several addresses.  The changer pass takes the same copy
// for slot k, and the scan gives every (row, copy) its own bucket range.
constexpr uint32_t kRowCopies = 8;

__device__ __forceinline__ uint32_t row_copy(uint32_t slot) {
  return (slot >> 8) & (kRowCopies - 1u);
}

struct RowHists {
  uint32_t* rowHist = nullptr;  // [rows][copies] spheres whose next key lies in the row
  uint32_t* chgHist = nullptr;  // [rows][copies] of those, the ones whose current row differs
  uint32_t shift = 0;           // log2(gridSize.x)
};

// Lanes a, a+1, ..., 63 of a wave (none when a >= 64).
__device__ __forceinline__ uint64_t lanes_from(uint32_t a) {
  return a >= 64u ? 0ull : (~0ull << a);
}

__device__ __forceinline__ uint32_t lowest_lane(uint64_t m) {
  return (uint32_t)__ffsll((unsigned long long)m) - 1u;
}

// Adds this wave's next keys to the row counts; `slot` is the lane's sorted
// slot (the lanes of a wave hold 64 consecutive slots of one 256-slot
// tile) and `oldKey` the key of the cell the sphere is in now.  A run of
// lanes with the same row adds its count with one atomic.  Every lane of
// the wave must call it.
__device__ __forceinline__ void row_histograms(const RowHists& rh, uint32_t slot, uint32_t key,
                                               uint32_t oldKey, bool valid) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t r = key >> rh.shift;
  const uint32_t at = r * kRowCopies + row_copy(slot);
  const uint32_t rPrev = __shfl_up(r, 1);
  const bool head = valid && (lane == 0u || r != rPrev);
  const uint64_t bounds = __ballot(head) | ~__ballot(valid);
  if (head) {
    const uint64_t later = bounds & lanes_from(lane + 1u);
    atomicAdd(&rh.rowHist[at], (later ? lowest_lane(later) : 64u) - lane);
  }
  if (valid && r != (oldKey >> rh.shift)) atomicAdd(&rh.chgHist[at], 1u);
}

}  // namespace fc
