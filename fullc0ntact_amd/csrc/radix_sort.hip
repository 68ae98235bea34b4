// LSD onesweep radix sort of (key, value) uint32 pairs for gfx950.
//
// Replaces thrust::sort_by_key (particleSystem.cu:265-270), whose radix path
// is stable; with values starting as the identity the output order is the
// lexicographic (hash, original index) order.
//
// Structure (one histogram launch + one launch per 8-bit digit):
//   digit histograms flow from the kernel that writes the keys (digit 0)
//                   and from each pass (digit p+1) to the next pass, see
//                   radix_common.hpp; k_rs_histogram produces digit 0 when
//                   the keys come from the caller.
//   k_rs_scatter    one 4096-key tile per block, tiles taken in launch order
//                   from a ticket counter.  Each wave ranks its 1024 keys
//                   with ballot-based digit matching into a wave-private LDS
//                   histogram, the block publishes its per-digit counts and
//                   resolves its global prefix with a two-level look-back
//                   (below), then stages the tile digit-sorted in LDS and
//                   writes it out so consecutive lanes write consecutive
//                   addresses.
//
// Two-level look-back: tiles form groups of kGroup consecutive tickets.  A
// tile reads the counts of its group's earlier tiles all at once; the last
// tile of each group publishes the group's count and, after a decoupled
// look-back over the earlier groups' words (window kGroupWindow), the
// group's inclusive prefix.  Every tile finds its group's exclusive prefix
// by the same group-level look-back.  All tiles of a pass start together,
// so a tile-level chain would make the last tile walk back through every
// tile while inclusive prefixes propagate one look-back at a time; the group
// level shortens that chain kGroup-fold for a few more loads per tile.
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
#include "radix_common.hpp"

namespace fc {

namespace {

constexpr int kWaves = RS_THREADS / 64;
// Sorts of at least this many keys use 8192-key tiles (k_rs_scatter).
constexpr uint32_t kWideTileKeys = 1u << 22;
constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagPrefix = 2ull << 62;
constexpr uint64_t kFlagMask = 3ull << 62;
constexpr uint32_t kEpochMask = (1u << 30) - 1u;
constexpr uint32_t kSpinLimit = 1u << 22;
constexpr int kGroup = RS_LOOKBACK_GROUP;  // tiles per look-back group
// Tile counts up to which a pass is launched without the ticket counter:
// one 1024-thread block per compute unit of an MI355X (256) is resident at
// once.
#ifndef FC_RS_UNTICKETED_TILES
#define FC_RS_UNTICKETED_TILES 256
#endif
constexpr uint32_t kUnticketedTiles = FC_RS_UNTICKETED_TILES;
#ifndef FC_RS_GROUP_WINDOW
#define FC_RS_GROUP_WINDOW 4
#endif
constexpr int kGroupWindow = FC_RS_GROUP_WINDOW;  // group words per round trip

__device__ __forceinline__ uint64_t pack_status(uint64_t flag, uint32_t epoch, uint32_t value) {
  return flag | ((uint64_t)(epoch & kEpochMask) << 32) | (uint64_t)value;
}

// Lanes (among `active`) whose kBits-bit digit equals this lane's digit.
template <int kBits = RS_RADIX_BITS>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t active) {
  uint64_t peers = active;
#pragma unroll
  for (int b = 0; b < kBits; ++b) {
    const uint32_t bit = (d >> b) & 1u;
    const uint64_t m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

// Standalone digit-0 histogram (used when the keys come from the caller).
__global__ __launch_bounds__(RS_THREADS) void k_rs_histogram(
    const uint32_t* __restrict__ keys, uint32_t n, int passes, int bits,
    uint32_t* __restrict__ gHist, uint32_t* counters, uint32_t* devFlags) {
  __shared__ uint32_t s_hist[RS_RADIX];
  for (int i = threadIdx.x; i < RS_RADIX; i += RS_THREADS) s_hist[i] = 0u;
  histogram_producer_prologue(gHist, passes, counters);
  __syncthreads();
  const uint32_t highMask = bits >= 32 ? 0u : ~((1u << bits) - 1u);
  uint32_t rangeErr = 0u;
  const uint32_t numTiles = (n + RS_TILE - 1) / RS_TILE;
  for (uint32_t tile = blockIdx.x; tile < numTiles; tile += gridDim.x) {
    uint32_t key[RS_ITEMS];
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
      const uint32_t idx = tile * RS_TILE + i * RS_THREADS + threadIdx.x;
      key[i] = idx < n ? keys[idx] : 0u;
    }
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
      const uint32_t idx = tile * RS_TILE + i * RS_THREADS + threadIdx.x;
      if (idx < n) rangeErr |= key[i] & highMask;
      histogram_add_digit(s_hist, key[i] & (RS_RADIX - 1), idx < n);
    }
  }
  if (rangeErr) atomicOr(devFlags, DEVERR_SORT_KEY_RANGE);
  __syncthreads();
  histogram_flush(s_hist, gHist, 0);
}

// One LSD pass over one tile of RS_THREADS x kItems keys per block (16
// waves x kItems keys each): 4096-key tiles, or 8192-key tiles for large
// sorts (radix_sort_pairs picks).
// With ga.pos (the last pass of a sort that gathers the particle state) the
// output loop also copies the state of each output's value to its final
// position: the writes follow the keys' (digit-run) order.  Issuing these
// loads with the keys instead and storing from the loading thread measured
// 50 us per pass instead of 29 at 1M keys.
template <int kItems>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(
    const uint32_t* __restrict__ keysIn, const uint32_t* __restrict__ valsIn,
    uint32_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut, uint32_t n, int pass,
    int passes, uint32_t* gHist, uint64_t* status, uint64_t* groupStatus, uint32_t* counters,
    uint32_t* devFlags, const uint32_t* __restrict__ clearHash, uint32_t* __restrict__ cellStart,
    uint32_t* __restrict__ cellEnd, SortGather ga, int ticketed) {
  constexpr uint32_t kTile = (uint32_t)RS_THREADS * kItems;
  constexpr uint32_t kTileWaveKeys = 64u * kItems;
  __shared__ uint32_t s_keys[kTile];
  __shared__ uint32_t s_vals[kTile];
  __shared__ uint32_t s_waveHist[kWaves][RS_RADIX];
  __shared__ uint32_t s_tileStart[RS_RADIX];
  __shared__ uint32_t s_digitBase[RS_RADIX];
  __shared__ uint32_t s_nextHist[RS_RADIX];
  __shared__ uint32_t s_scratch[RS_RADIX / 64];
  __shared__ uint32_t s_tile;

  const uint32_t numTiles = (n + kTile - 1) / kTile;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = threadIdx.x >> 6;
  const int shift = pass * RS_RADIX_BITS;

  // Tiles in launch order: from a ticket counter, or -- when every block of
  // the launch can be resident at once -- the block index itself (a block
  // then only waits for blocks that are running).
  const uint32_t epochBase = counters[RSC_EPOCH];
  if (ticketed && threadIdx.x == 0) {
    uint32_t t = atomicAdd(&counters[RSC_TILE0 + pass], 1u);
    if (t == numTiles - 1u) atomicExch(&counters[RSC_TILE0 + pass], 0u);
    s_tile = t;
  }
  for (int i = threadIdx.x; i < kWaves * RS_RADIX; i += RS_THREADS) (&s_waveHist[0][0])[i] = 0u;
  if (threadIdx.x < (uint32_t)RS_RADIX) s_nextHist[threadIdx.x] = 0u;
  if (blockIdx.x == 0 && pass == 1) {  // H[0] was consumed by pass 0
    uint32_t* h = hist_buffer(gHist, 0);
    for (int i = threadIdx.x; i < RS_HIST_COPIES * RS_RADIX; i += RS_THREADS) h[i] = 0u;
  }
  __syncthreads();
  const uint32_t tile = ticketed ? s_tile : blockIdx.x;
  const uint32_t epoch = (epochBase * 4u + (uint32_t)pass) & kEpochMask;
  const uint32_t tileBase = tile * kTile;
  const uint32_t waveBase = tileBase + wave * kTileWaveKeys;
  if (clearHash != nullptr) {
    // Reset the previous update's cells: cell h was written iff it starts a
    // run of the previous sorted hashes.  clearHash is only rewritten by the
    // last pass, which runs after this one.
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const uint32_t idx = tileBase + i * RS_THREADS + threadIdx.x;
      if (idx < n) {
        const uint32_t h = clearHash[idx];
        if (idx == 0 || clearHash[idx - 1] != h) {
          cellStart[h] = CELL_EMPTY;
          cellEnd[h] = CELL_EMPTY;
        }
      }
    }
  }

  uint32_t key[kItems], val[kItems], rank[kItems];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint32_t idx = waveBase + i * 64u + lane;
    const bool valid = idx < n;
    key[i] = valid ? keysIn[idx] : 0xffffffffu;
    val[i] = valid ? (valsIn ? valsIn[idx] : idx) : 0u;
  }
  // Global total of this thread's digit (threads 0..255), requested while
  // the keys are in flight.
  const uint32_t digitTotal =
      threadIdx.x < (uint32_t)RS_RADIX ? histogram_total(gHist, pass, threadIdx.x) : 0u;
  // Rank within the wave, slice by slice in index order (stable), and count
  // the next pass's digit.
  const bool countNext = pass + 1 < passes;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint32_t idx = waveBase + i * 64u + lane;
    const bool valid = idx < n;
    if (countNext)
      histogram_add_digit(s_nextHist, (key[i] >> (shift + RS_RADIX_BITS)) & (RS_RADIX - 1), valid);
    const uint64_t active = __ballot(valid);
    const uint32_t d = (key[i] >> shift) & (RS_RADIX - 1);
    const uint64_t peers = match_digit(d, active);
    const uint32_t lower = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
    const uint32_t before = s_waveHist[wave][d];
    rank[i] = before + lower;
    if (valid && lower == 0u) s_waveHist[wave][d] = before + (uint32_t)__popcll(peers);
  }
  __syncthreads();

  // Threads 0..255 own one digit each: wave offsets inside the tile, the
  // tile total, publication and look-back.
  const uint32_t d = threadIdx.x;
  const bool digitThread = d < (uint32_t)RS_RADIX;
  uint32_t tileCount = 0;
  uint64_t* myStatus = status + (size_t)tile * RS_RADIX + (digitThread ? d : 0u);
  if (digitThread) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      uint32_t c = s_waveHist[w][d];
      s_waveHist[w][d] = tileCount;
      tileCount += c;
    }
    __hip_atomic_store(myStatus, pack_status(kFlagAgg, epoch, tileCount), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  const uint32_t tileStart = scan256_exclusive(digitThread ? tileCount : 0u, s_scratch);
  const uint32_t digitGlobal = scan256_exclusive(digitTotal, s_scratch);
  if (digitThread) {
    s_tileStart[d] = tileStart;
    bool stalled = false;
    // Counts of the group's earlier tiles, all requested at once.
    const uint32_t g = tile / (uint32_t)kGroup, first = g * (uint32_t)kGroup;
    const uint32_t nIn = tile - first;
    uint64_t w[kGroup - 1];
#pragma unroll
    for (int i = 0; i < kGroup - 1; ++i)
      w[i] = (uint32_t)i < nIn ? __hip_atomic_load(status + (size_t)(first + i) * RS_RADIX + d,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                               : 0ull;
    uint32_t inGroup = 0;
#pragma unroll
    for (int i = 0; i < kGroup - 1; ++i) {
      if ((uint32_t)i < nIn) {
        uint32_t spins = 0;
        while (!stalled && ((w[i] & kFlagMask) == 0ull ||
                            ((uint32_t)(w[i] >> 32) & kEpochMask) != epoch)) {
          if (++spins > kSpinLimit) {
            stalled = true;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          w[i] = __hip_atomic_load(status + (size_t)(first + i) * RS_RADIX + d, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        inGroup += (uint32_t)w[i];
      }
    }
    uint64_t* gWord = groupStatus + (size_t)g * RS_RADIX + d;
    const bool leader = nIn == (uint32_t)kGroup - 1u;
    if (leader)
      __hip_atomic_store(gWord, pack_status(g == 0u ? kFlagPrefix : kFlagAgg, epoch, inGroup + tileCount),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // Decoupled look-back over the earlier groups, kGroupWindow words per
    // round trip, consumed newest to oldest until an inclusive prefix.
    uint32_t before = 0;
    if (g > 0u) {
      int t = (int)g - 1;
      uint32_t spins = 0;
      while (!stalled) {
        uint64_t v[kGroupWindow];
#pragma unroll
        for (int i = 0; i < kGroupWindow; ++i)
          v[i] = (t - i >= 0) ? __hip_atomic_load(groupStatus + (size_t)(t - i) * RS_RADIX + d,
                                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                              : 0ull;
        int used = 0;
        bool done = false;
#pragma unroll
        for (int i = 0; i < kGroupWindow; ++i) {
          if (done || used != i) break;
          if (t - i < 0) { done = true; break; }  // unreachable: group 0 holds a prefix
          const uint64_t flag = v[i] & kFlagMask;
          const uint32_t vEpoch = (uint32_t)(v[i] >> 32) & kEpochMask;
          if (flag == 0ull || vEpoch != epoch) break;  // not published yet
          before += (uint32_t)v[i];
          ++used;
          if (flag == kFlagPrefix) done = true;
        }
        if (done) break;
        t -= used;
        if (used == 0) {
          if (++spins > kSpinLimit) {
            stalled = true;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      if (leader)
        __hip_atomic_store(gWord, pack_status(kFlagPrefix, epoch, before + inGroup + tileCount),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (stalled) atomicOr(devFlags, DEVERR_SORT_STALLED);
    const uint32_t prefix = before + inGroup;
    s_digitBase[d] = digitGlobal + prefix - tileStart;
  }
  __syncthreads();

  // Stage the tile digit-sorted in LDS.
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint32_t idx = waveBase + i * 64u + lane;
    if (idx < n) {
      const uint32_t dg = (key[i] >> shift) & (RS_RADIX - 1);
      const uint32_t lp = s_tileStart[dg] + s_waveHist[wave][dg] + rank[i];
      if (lp < (uint32_t)kTile) {
        s_keys[lp] = key[i];
        s_vals[lp] = val[i];
      }
    }
  }
  __syncthreads();

  const uint32_t validCount = (n - tileBase) < (uint32_t)kTile ? (n - tileBase) : (uint32_t)kTile;
  // Output: every item's key, value and destination first, then (the last
  // pass of a gathering sort) all of the items' state loads together, then
  // the stores.  The state arrays may alias nothing here, but the compiler
  // cannot know it: loads placed after a store wait for the store's round
  // trip, which serialised the gather.
  uint32_t ok[kItems], ks[kItems], vs[kItems], gs[kItems];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint32_t s = i * RS_THREADS + threadIdx.x;
    ok[i] = 0u;
    ks[i] = vs[i] = gs[i] = 0u;
    if (s < validCount) {
      ks[i] = s_keys[s];
      const uint32_t dg = (ks[i] >> shift) & (RS_RADIX - 1);
      gs[i] = s_digitBase[dg] + s;
      vs[i] = s_vals[s];
      ok[i] = gs[i] < n ? 1u : 0u;  // always true unless a look-back stalled
    }
  }
  if (ga.pos) {
    float4 pv[kItems], vv[kItems];
    uint32_t id[kItems];
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const bool live = ok[i] && vs[i] < n;  // v < n: values index the particle arrays
      const uint32_t v = live ? vs[i] : 0u;
      pv[i] = ga.pos[v];
      vv[i] = ga.vel[v];
      id[i] = ga.idOfSlot ? ga.idOfSlot[v] : 0u;
    }
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      if (!ok[i]) continue;
      const uint32_t g = gs[i];
      keysOut[g] = ks[i];
      valsOut[g] = vs[i];
      if (vs[i] < n) {
        ga.sortedPos[g] = pv[i];
        ga.sortedVel[g] = make_float4(vv[i].x, vv[i].y, vv[i].z, __uint_as_float(ks[i]));
        if (ga.idOfSlot) ga.indexOut[g] = id[i];
        if (ga.sortedQ) ga.sortedQ[g] = quantize_pos(ga.frame, pv[i].x, pv[i].y, pv[i].z);
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      if (!ok[i]) continue;
      keysOut[gs[i]] = ks[i];
      valsOut[gs[i]] = vs[i];
    }
  }
  if (countNext) histogram_flush(s_nextHist, gHist, pass + 1);
}

// The sort of n <= RS_TILE keys by counting, over many blocks: a key's place
// in the stable order is the number of (key, index) pairs below its own,
// so every block stages all n pairs in LDS and counts for 16 of them, a
// wave per key (lane s compares against the pairs j = s mod 64, so that
// a wave reads 64 consecutive words), with no exchange between blocks.  C packs (key << 12) | index, 32 bits wide when the keys' width
// leaves 12 bits for the index, else 64.  The block also resets the cells
// of its 16 entries of the previous sorted hashes (clearHash), which no
// later write of this launch touches, and block 0 zeroes H[0] for the next
// sort.  The sorted pairs go to keysStage/valsStage; k_small_emit writes
// the outputs in a second launch, so that inputs and outputs may alias.
constexpr uint32_t kRankLanes = 64;                  // lanes per key: one wave
constexpr uint32_t kRankThreads = 1024;
constexpr uint32_t kRankKeys = kRankThreads / kRankLanes;  // keys per block
static_assert(RS_TILE % kRankLanes == 0 && (1u << 12) >= (uint32_t)RS_TILE, "12-bit index");

template <typename C>
__global__ __launch_bounds__(kRankThreads) void k_small_rank(
    const uint32_t* __restrict__ keysIn, const uint32_t* __restrict__ valsIn, uint32_t n,
    const uint32_t* __restrict__ clearHash, uint32_t* __restrict__ cellStart,
    uint32_t* __restrict__ cellEnd, uint32_t* __restrict__ keysStage,
    uint32_t* __restrict__ valsStage, uint32_t* gHist) {
  __shared__ C s_pair[RS_TILE];
  const uint32_t tid = threadIdx.x;
  if (blockIdx.x == 0) {
    uint32_t* h0 = hist_buffer(gHist, 0);
    for (uint32_t i = tid; i < RS_HIST_COPIES * RS_RADIX; i += kRankThreads) h0[i] = 0u;
  }
  if (clearHash != nullptr && tid < kRankKeys) {
    const uint32_t idx = blockIdx.x * kRankKeys + tid;
    if (idx < n) {
      const uint32_t h = clearHash[idx];
      if (idx == 0 || clearHash[idx - 1] != h) {
        cellStart[h] = CELL_EMPTY;
        cellEnd[h] = CELL_EMPTY;
      }
    }
  }
  const uint32_t padded = (n + kRankLanes - 1u) & ~(kRankLanes - 1u);
  for (uint32_t j = tid; j < padded; j += kRankThreads)
    s_pair[j] = j < n ? (((C)keysIn[j] << 12) | (C)j) : ~(C)0;
  const uint32_t seg = tid & (kRankLanes - 1u);
  const uint32_t i = blockIdx.x * kRankKeys + tid / kRankLanes;
  const uint32_t val = (seg == 0u && i < n) ? (valsIn ? valsIn[i] : i) : 0u;
  __syncthreads();
  const C mine = i < n ? s_pair[i] : ~(C)0;
  uint32_t below = 0;
#pragma unroll 8
  for (uint32_t j = seg; j < padded; j += kRankLanes) below += s_pair[j] < mine ? 1u : 0u;
#pragma unroll
  for (uint32_t d = kRankLanes / 2; d > 0; d >>= 1) below += __shfl_xor(below, d, kRankLanes);
  if (seg == 0u && i < n) {
    keysStage[below] = (uint32_t)(mine >> 12);
    valsStage[below] = val;
  }
}

// The outputs of the sort by counting, one thread per sorted position:
// keys and values, the gathered state, and the cell bounds (as
// k_cell_bounds writes them).
__global__ __launch_bounds__(64) void k_small_emit(
    const uint32_t* __restrict__ keysStage, const uint32_t* __restrict__ valsStage,
    uint32_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut, uint32_t n,
    uint32_t* __restrict__ cellStart, uint32_t* __restrict__ cellEnd, uint32_t numCells,
    SortGather ga) {
  const uint32_t g = blockIdx.x * 64u + threadIdx.x;
  if (g >= n) return;
  const uint32_t k = keysStage[g], v = valsStage[g];
  const uint32_t kp = g > 0 ? keysStage[g - 1] : 0u;
  float4 pv = make_float4(0.f, 0.f, 0.f, 0.f), vv = pv;
  uint32_t id = 0;
  if (ga.pos && v < n) {
    pv = ga.pos[v];
    vv = ga.vel[v];
    if (ga.idOfSlot) id = ga.idOfSlot[v];
  }
  keysOut[g] = k;
  valsOut[g] = v;
  if (ga.pos && v < n) {
    ga.sortedPos[g] = pv;
    ga.sortedVel[g] = make_float4(vv.x, vv.y, vv.z, __uint_as_float(k));
    if (ga.idOfSlot) ga.indexOut[g] = id;
    if (ga.sortedQ) ga.sortedQ[g] = quantize_pos(ga.frame, pv.x, pv.y, pv.z);
  }
  if (g == 0 || k != kp) {
    if (k < numCells) cellStart[k] = g;
    if (g > 0 && kp < numCells) cellEnd[kp] = g;
  }
  if (g == n - 1 && k < numCells) cellEnd[k] = g + 1;
}

}  // namespace

hipError_t launch_sort_small(hipStream_t stream, SortScratch& s, const uint32_t* keysIn,
                             const uint32_t* valsIn, uint32_t* keysOut, uint32_t* valsOut,
                             uint32_t n, int passes, const uint32_t* clearHash,
                             uint32_t* cellStart, uint32_t* cellEnd, uint32_t numCells,
                             const SortGather& gather) {
  if (n == 0) return hipSuccess;
  if (n > (uint32_t)RS_TILE || passes < 1 || passes > RS_MAX_PASSES) return hipErrorInvalidValue;
  // The keys are cells, below numCells: (key << 12) | index fits 32 bits
  // for keys of up to 20 bits.
  int bits = key_bits_for_cells(numCells);
  if (bits > passes * RS_RADIX_BITS) bits = passes * RS_RADIX_BITS;
  const dim3 rankGrid((n + kRankKeys - 1) / kRankKeys);
  if (bits <= 20)
    hipLaunchKernelGGL(k_small_rank<uint32_t>, rankGrid, dim3(kRankThreads), 0, stream, keysIn,
                       valsIn, n, clearHash, cellStart, cellEnd, s.keysA, s.valsA, s.histogram);
  else
    hipLaunchKernelGGL(k_small_rank<uint64_t>, rankGrid, dim3(kRankThreads), 0, stream, keysIn,
                       valsIn, n, clearHash, cellStart, cellEnd, s.keysA, s.valsA, s.histogram);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_small_emit, dim3((n + 63) / 64), dim3(64), 0, stream, s.keysA, s.valsA,
                     keysOut, valsOut, n, cellStart, cellEnd, numCells, gather);
  return hipGetLastError();
}

int key_bits_for_cells(uint32_t numCells) {
  if (numCells <= 1u) return 0;
  uint32_t m = numCells - 1u;
  int b = 0;
  while (m) { ++b; m >>= 1; }
  return b;
}

int radix_passes(int bits, bool inPlace) {
  (void)inPlace;
  // At least two passes: a single pass could not run in place, and the
  // histogram buffers are recycled by the pass after the one that read them.
  // A pass over an all-zero digit is a stable copy.
  int passes = (bits + RS_RADIX_BITS - 1) / RS_RADIX_BITS;
  return passes < 2 ? 2 : passes;
}

hipError_t radix_sort_pairs(hipStream_t stream, SortScratch& s, uint32_t* devFlags,
                            const uint32_t* keysIn, const uint32_t* valsIn, uint32_t* keysOut,
                            uint32_t* valsOut, uint32_t n, int bits, StageTimer* timer,
                            bool histogramReady, int minPasses, const uint32_t* clearHash,
                            uint32_t* cellStart, uint32_t* cellEnd, const SortGather* gather) {
  if (n == 0) return hipSuccess;
  const bool inPlace = keysIn == keysOut || (valsIn != nullptr && valsIn == valsOut);
  int passes = radix_passes(bits, inPlace);
  if (passes < minPasses) passes = minPasses;
  if (clearHash != nullptr && passes < 2) return hipErrorInvalidValue;
  if (passes > RS_MAX_PASSES) return hipErrorInvalidValue;
  // Large sorts take 8192-key tiles: half the tiles and look-back chain,
  // longer digit runs per tile.  The batched ensemble's step (four passes
  // over 11.9M keys) measured 1.656 against 1.735 ms with them; at 1M keys
  // a pass took 21.9 against 18.5 us.
  const bool wide = n >= kWideTileKeys;
  const uint32_t tileKeys = wide ? 2u * (uint32_t)RS_TILE : (uint32_t)RS_TILE;
  const uint32_t numTiles = (n + tileKeys - 1) / tileKeys;
  if (s.capacity < n || s.statusCapacity < (size_t)numTiles * RS_RADIX ||
      s.groupStatusCapacity < (size_t)lookback_groups(numTiles) * RS_RADIX)
    return hipErrorInvalidValue;
  hipError_t e = hipSuccess;
  if (!histogramReady) {
    uint32_t histBlocks = numTiles;
    if (histBlocks > (uint32_t)RS_MAX_HIST_BLOCKS) histBlocks = RS_MAX_HIST_BLOCKS;
    if (timer) timer->begin(stream, "sort_histogram");
    hipLaunchKernelGGL(k_rs_histogram, dim3(histBlocks), dim3(RS_THREADS), 0, stream, keysIn, n,
                       passes, bits, s.histogram, s.counters, devFlags);
    e = hipGetLastError();
    if (timer) timer->end(stream);
    if (e != hipSuccess) return e;
  }

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
    const SortGather ga = (gather && p == passes - 1) ? *gather : SortGather{};
    const int ticketed = numTiles > kUnticketedTiles ? 1 : 0;
    if (wide)
      hipLaunchKernelGGL(k_rs_scatter<2 * RS_ITEMS>, dim3(numTiles), dim3(RS_THREADS), 0, stream,
                         kin, vin, kout, vout, n, p, passes, s.histogram, s.status,
                         s.groupStatus, s.counters, devFlags, p == 0 ? clearHash : nullptr,
                         cellStart, cellEnd, ga, ticketed);
    else
      hipLaunchKernelGGL(k_rs_scatter<RS_ITEMS>, dim3(numTiles), dim3(RS_THREADS), 0, stream,
                         kin, vin, kout, vout, n, p, passes, s.histogram, s.status,
                         s.groupStatus, s.counters, devFlags, p == 0 ? clearHash : nullptr,
                         cellStart, cellEnd, ga, ticketed);
    e = hipGetLastError();
    if (timer) timer->end(stream);
    if (e != hipSuccess) return e;
    kin = kout;
    vin = vout;
  }
  return hipSuccess;
}

}  // namespace fc
