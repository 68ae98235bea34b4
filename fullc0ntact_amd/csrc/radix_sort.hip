// LSD onesweep radix sort of (key, value) uint32 pairs for gfx950.
//
// Replaces thrust::sort_by_key (particleSystem.cu:265-270), whose radix path
// is stable; with values starting as the identity the output order is the
// lexicographic (hash, original index) order.
//
// Structure (one histogram launch + one launch per 8-bit digit):
//   k_rs_histogram  every block counts all digits of its share of the keys
//                   into LDS with wave-aggregated adds, writes a partial
//                   histogram; the last block to finish (ticket counter)
//                   sums the partials, writes exclusive bucket offsets per
//                   pass and bumps the sort epoch.
//   k_rs_scatter    one 4096-key tile per block, tiles taken in launch order
//                   from a ticket counter.  Each wave ranks its 1024 keys
//                   with ballot-based digit matching into a wave-private LDS
//                   histogram, the block publishes its per-digit counts and
//                   resolves its global prefix by decoupled look-back over
//                   the preceding tiles, then stages the tile digit-sorted in
//                   LDS and writes it out so consecutive lanes write
//                   consecutive addresses.
//
// No memset is needed between sorts: the ticket counters reset themselves
// (the block holding the last ticket stores 0) and look-back words carry the
// epoch of the pass that wrote them, so words from earlier sorts read as
// "not ready".  That keeps the whole sort capturable in a hipGraph.
//
// Forward progress: a block only waits for tiles with smaller tickets, which
// were handed to blocks that had already started.  The wait is additionally
// bounded; hitting the bound raises DEVERR_SORT_STALLED instead of hanging.
#include "fc_internal.hpp"

namespace fc {

namespace {

constexpr int kWaves = RS_THREADS / 64;
constexpr int kWaveKeys = 64 * RS_ITEMS;
constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagPrefix = 2ull << 62;
constexpr uint64_t kFlagMask = 3ull << 62;
constexpr uint32_t kEpochMask = (1u << 30) - 1u;
constexpr uint32_t kSpinLimit = 1u << 22;

__device__ __forceinline__ uint64_t pack_status(uint64_t flag, uint32_t epoch, uint32_t value) {
  return flag | ((uint64_t)(epoch & kEpochMask) << 32) | (uint64_t)value;
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Lanes (among `active`) whose 8-bit digit equals this lane's digit.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t active) {
  uint64_t peers = active;
#pragma unroll
  for (int b = 0; b < RS_RADIX_BITS; ++b) {
    const uint32_t bit = (d >> b) & 1u;
    const uint64_t m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

// Block-wide exclusive scan of one value per thread (256 threads, 4 waves).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_waveTotals) {
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t t = __shfl_up(inc, off);
    if (lane >= (uint32_t)off) inc += t;
  }
  if (lane == 63) s_waveTotals[wave] = inc;
  __syncthreads();
  uint32_t base = 0;
  for (uint32_t w = 0; w < wave; ++w) base += s_waveTotals[w];
  __syncthreads();
  return base + inc - v;
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_histogram(
    const uint32_t* __restrict__ keys, uint32_t n, int passes, int bits,
    uint32_t* __restrict__ partial, uint32_t* __restrict__ offsets, uint32_t* counters,
    uint32_t* devFlags) {
  __shared__ uint32_t s_hist[RS_MAX_PASSES][RS_RADIX];
  __shared__ uint32_t s_waveTotals[kWaves];
  __shared__ uint32_t s_last;
  for (int i = threadIdx.x; i < RS_MAX_PASSES * RS_RADIX; i += RS_THREADS)
    (&s_hist[0][0])[i] = 0u;
  __syncthreads();

  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  const uint32_t gwave = blockIdx.x * kWaves + wave;
  const uint32_t nwaves = gridDim.x * kWaves;
  const uint32_t highMask = bits >= 32 ? 0u : ~((1u << bits) - 1u);
  uint32_t rangeErr = 0u;
  for (uint32_t slice = gwave; (size_t)slice * 64u < n; slice += nwaves) {
    const uint32_t idx = slice * 64u + lane;
    const bool valid = idx < n;
    const uint32_t key = valid ? keys[idx] : 0u;
    rangeErr |= key & highMask;
    const uint64_t active = __ballot(valid);
    for (int p = 0; p < passes; ++p) {
      const uint32_t d = (key >> (p * RS_RADIX_BITS)) & (RS_RADIX - 1);
      const uint64_t peers = match_digit(d, active);
      const uint64_t lower = peers & ((1ull << lane) - 1ull);
      if (valid && lower == 0ull) atomicAdd(&s_hist[p][d], (uint32_t)__popcll(peers));
    }
  }
  if (rangeErr) atomicOr(devFlags, DEVERR_SORT_KEY_RANGE);
  __syncthreads();

  uint32_t* mine = partial + (size_t)blockIdx.x * RS_MAX_PASSES * RS_RADIX;
  for (int i = threadIdx.x; i < passes * RS_RADIX; i += RS_THREADS) mine[i] = (&s_hist[0][0])[i];
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t ticket = atomicAdd(&counters[RSC_HIST_DONE], 1u);
    s_last = (ticket == gridDim.x - 1u) ? 1u : 0u;
  }
  __syncthreads();
  if (!s_last) return;

  // Last block: every other block's partial is visible after its fence and
  // ticket; sum them and publish exclusive bucket offsets per pass.
  __threadfence();
  const uint32_t d = threadIdx.x;
  for (int p = 0; p < passes; ++p) {
    uint32_t total = 0;
    for (uint32_t b = 0; b < gridDim.x; ++b)
      total += __hip_atomic_load(&partial[((size_t)b * RS_MAX_PASSES + p) * RS_RADIX + d],
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t excl = block_exclusive_scan(total, s_waveTotals);
    offsets[p * RS_RADIX + d] = excl;
  }
  if (threadIdx.x == 0) {
    counters[RSC_HIST_DONE] = 0u;
    counters[RSC_EPOCH] = counters[RSC_EPOCH] + 1u;
  }
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(
    const uint32_t* __restrict__ keysIn, const uint32_t* __restrict__ valsIn,
    uint32_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut, uint32_t n, int pass,
    const uint32_t* __restrict__ offsets, uint64_t* status, uint32_t* counters,
    uint32_t* devFlags) {
  __shared__ uint32_t s_keys[RS_TILE];
  __shared__ uint32_t s_vals[RS_TILE];
  __shared__ uint32_t s_waveHist[kWaves][RS_RADIX];
  __shared__ uint32_t s_tileStart[RS_RADIX];
  __shared__ uint32_t s_digitBase[RS_RADIX];
  __shared__ uint32_t s_waveTotals[kWaves];
  __shared__ uint32_t s_tile;

  const uint32_t numTiles = (n + RS_TILE - 1) / RS_TILE;
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  const int shift = pass * RS_RADIX_BITS;

  if (threadIdx.x == 0) {
    uint32_t t = atomicAdd(&counters[RSC_TILE0 + pass], 1u);
    if (t == numTiles - 1u) atomicExch(&counters[RSC_TILE0 + pass], 0u);
    s_tile = t;
  }
  for (int i = threadIdx.x; i < kWaves * RS_RADIX; i += RS_THREADS) (&s_waveHist[0][0])[i] = 0u;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t epoch = (counters[RSC_EPOCH] * 4u + (uint32_t)pass) & kEpochMask;
  const uint32_t tileBase = tile * RS_TILE;
  const uint32_t waveBase = tileBase + wave * kWaveKeys;

  uint32_t key[RS_ITEMS], val[RS_ITEMS], rank[RS_ITEMS];
#pragma unroll
  for (int i = 0; i < RS_ITEMS; ++i) {
    const uint32_t idx = waveBase + i * 64u + lane;
    const bool valid = idx < n;
    key[i] = valid ? keysIn[idx] : 0xffffffffu;
    val[i] = valid ? (valsIn ? valsIn[idx] : idx) : 0u;
  }
  // Rank within the wave, slice by slice in index order (stable).
#pragma unroll
  for (int i = 0; i < RS_ITEMS; ++i) {
    const uint32_t idx = waveBase + i * 64u + lane;
    const bool valid = idx < n;
    const uint64_t active = __ballot(valid);
    const uint32_t d = (key[i] >> shift) & (RS_RADIX - 1);
    const uint64_t peers = match_digit(d, active);
    const uint32_t lower = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
    const uint32_t before = s_waveHist[wave][d];
    rank[i] = before + lower;
    if (valid && lower == 0u) s_waveHist[wave][d] = before + (uint32_t)__popcll(peers);
  }
  __syncthreads();

  // Per digit: wave offsets inside the tile and the tile total.
  const uint32_t d = threadIdx.x;
  uint32_t running = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    uint32_t c = s_waveHist[w][d];
    s_waveHist[w][d] = running;
    running += c;
  }
  const uint32_t tileCount = running;
  uint64_t* myStatus = status + (size_t)tile * RS_RADIX + d;
  if (tile == 0u) {
    __hip_atomic_store(myStatus, pack_status(kFlagPrefix, epoch, tileCount), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  } else {
    __hip_atomic_store(myStatus, pack_status(kFlagAgg, epoch, tileCount), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  const uint32_t tileStart = block_exclusive_scan(tileCount, s_waveTotals);
  s_tileStart[d] = tileStart;

  // Decoupled look-back for this digit.
  uint32_t prefix = 0;
  if (tile > 0u) {
    int t = (int)tile - 1;
    uint32_t spins = 0;
    while (t >= 0) {
      const uint64_t w = __hip_atomic_load(status + (size_t)t * RS_RADIX + d, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
      const uint64_t flag = w & kFlagMask;
      const uint32_t wEpoch = (uint32_t)(w >> 32) & kEpochMask;
      if (flag == 0ull || wEpoch != epoch) {
        if (++spins > kSpinLimit) {
          atomicOr(devFlags, DEVERR_SORT_STALLED);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      prefix += (uint32_t)w;
      if (flag == kFlagPrefix) break;
      --t;
    }
    __hip_atomic_store(myStatus, pack_status(kFlagPrefix, epoch, prefix + tileCount),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  s_digitBase[d] = offsets[pass * RS_RADIX + d] + prefix - tileStart;
  __syncthreads();

  // Stage the tile digit-sorted in LDS.
#pragma unroll
  for (int i = 0; i < RS_ITEMS; ++i) {
    const uint32_t idx = waveBase + i * 64u + lane;
    if (idx < n) {
      const uint32_t dg = (key[i] >> shift) & (RS_RADIX - 1);
      const uint32_t lp = s_tileStart[dg] + s_waveHist[wave][dg] + rank[i];
      if (lp < (uint32_t)RS_TILE) {
        s_keys[lp] = key[i];
        s_vals[lp] = val[i];
      }
    }
  }
  __syncthreads();

  const uint32_t validCount = (n - tileBase) < (uint32_t)RS_TILE ? (n - tileBase) : (uint32_t)RS_TILE;
#pragma unroll
  for (int i = 0; i < RS_ITEMS; ++i) {
    const uint32_t s = i * RS_THREADS + threadIdx.x;
    if (s < validCount) {
      const uint32_t k = s_keys[s];
      const uint32_t dg = (k >> shift) & (RS_RADIX - 1);
      const uint32_t g = s_digitBase[dg] + s;
      if (g < n) {  // always true unless a look-back stalled
        keysOut[g] = k;
        valsOut[g] = s_vals[s];
      }
    }
  }
}

}  // namespace

int key_bits_for_cells(uint32_t numCells) {
  if (numCells <= 1u) return 0;
  uint32_t m = numCells - 1u;
  int b = 0;
  while (m) { ++b; m >>= 1; }
  return b;
}

hipError_t radix_sort_pairs(hipStream_t stream, SortScratch& s, uint32_t* devFlags,
                            const uint32_t* keysIn, const uint32_t* valsIn, uint32_t* keysOut,
                            uint32_t* valsOut, uint32_t n, int bits, StageTimer* timer) {
  if (n == 0) return hipSuccess;
  int passes = (bits + RS_RADIX_BITS - 1) / RS_RADIX_BITS;
  if (passes < 1) passes = 1;  // still produces the (identity) values
  // A single pass cannot run in place; a second pass over an all-zero digit
  // is a stable copy, so in-place one-digit sorts take two passes.
  const bool inPlace = keysIn == keysOut || (valsIn != nullptr && valsIn == valsOut);
  if (passes == 1 && inPlace) passes = 2;
  if (passes > RS_MAX_PASSES) return hipErrorInvalidValue;
  const uint32_t numTiles = (n + RS_TILE - 1) / RS_TILE;
  if (s.capacity < n || s.statusCapacity < (size_t)numTiles * RS_RADIX) return hipErrorInvalidValue;

  uint32_t histBlocks = (n + RS_TILE - 1) / RS_TILE;
  if (histBlocks > (uint32_t)RS_MAX_HIST_BLOCKS) histBlocks = RS_MAX_HIST_BLOCKS;
  if (timer) timer->begin(stream, "sort_histogram");
  hipLaunchKernelGGL(k_rs_histogram, dim3(histBlocks), dim3(RS_THREADS), 0, stream, keysIn, n,
                     passes, bits, s.partial, s.offsets, s.counters, devFlags);
  hipError_t e = hipGetLastError();
  if (timer) timer->end(stream);
  if (e != hipSuccess) return e;

  // Ping-pong so that the last pass lands in keysOut/valsOut and no pass
  // reads and writes the same buffer.
  const uint32_t* kin = keysIn;
  const uint32_t* vin = valsIn;
  for (int p = 0; p < passes; ++p) {
    uint32_t* kout;
    uint32_t* vout;
    if (p == passes - 1) {
      kout = keysOut;
      vout = valsOut;
    } else if (((passes - 1 - p) & 1) == 1) {
      kout = s.keysA;
      vout = s.valsA;
    } else {
      kout = s.keysB;
      vout = s.valsB;
    }
    static const char* names[RS_MAX_PASSES] = {"sort_pass0", "sort_pass1", "sort_pass2",
                                               "sort_pass3"};
    if (timer) timer->begin(stream, names[p]);
    hipLaunchKernelGGL(k_rs_scatter, dim3(numTiles), dim3(RS_THREADS), 0, stream, kin, vin, kout,
                       vout, n, p, s.offsets, s.status, s.counters, devFlags);
    e = hipGetLastError();
    if (timer) timer->end(stream);
    if (e != hipSuccess) return e;
    kin = kout;
    vin = vout;
  }
  return hipSuccess;
}

}  // namespace fc
