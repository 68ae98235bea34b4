// Row resort: the sort of a multi-step run's steady-state steps.
//
// Within a run the fused collide stores each sphere's next hash at its
// sorted slot, so the keys arrive in the previous step's (cell, id) order
// and only the spheres that changed cell are out of place (8-9 % per step
// on a settled bed).  The result must be the order of the reference's
// stable sort by hash of the ids (particleSystem.cu:265-270 sorts
// (hash, index) pairs whose index is the identity), that is (cell, id).
// A grid row (the gridSize.x cells that share y and z) is a contiguous
// range of that order, so each row is sorted on its own:
//
//   k_row_scan        row starts of the next order (the current ones are
//                     kept), the changers' bucket starts, and the list of
//                     the occupied rows, from the per-row counts the key
//                     producer accumulated (256-row groups pass their
//                     totals on to the later groups);
//   k_row_changers    buckets the spheres whose row changes by their new
//                     row; clears the cells of rows that empty;
//   k_row_sort        a fixed grid over the occupied rows.  A row's members
//                     are the spheres of its current range whose next key
//                     stays in it plus its incomers.  Those that keep their
//                     cell are already in (cell, id) order; the others (the
//                     movers, few) are ranked among themselves, and each
//                     member's place is its rank in its own list plus the
//                     number of the other list below it.  Rows beyond LDS,
//                     or with many movers, are ranked by counting.  The
//                     members are written at the row's start with the
//                     gathered state and the new cell tables.
//
// This replaces the radix passes and the cell-bounds kernel of those
// steps; the first sort of a run (keys in id order) stays a radix sort.
#include "row_resort.hpp"

namespace fc {
namespace {

constexpr int kScanThreads = 256;
// The row sort: one row per 128-thread block (9 KB of LDS, 82 VGPRs: 10
// blocks per CU) over a fixed grid of 8192 blocks, so that the ~3800
// occupied rows of a 1M-sphere bed are in flight together (each row is a
// chain of dependent loads; capping the kernel at 64 or 72 VGPRs for more
// blocks per CU made it slower).
constexpr int kSortThreads = 128;
constexpr uint32_t kRowCap = 768;      // members of a row held in LDS
constexpr uint32_t kMoverCap = 256;    // movers ranked by counting
constexpr uint32_t kPerThread = kRowCap / kSortThreads;
constexpr uint32_t kMoversPerThread = kMoverCap / kSortThreads;
constexpr uint32_t kSortBlocks = 8192; // grid of the row sort
static_assert(kRowCap % kSortThreads == 0 && kMoverCap % kSortThreads == 0, "whole per-thread shares");
constexpr uint64_t kNoMember = ~0ull;
static_assert(kScanThreads == (1 << kRowGroupShift), "one row per scan thread");

// Exclusive scan of one value per thread over a kThreads-thread block;
// returns the thread's prefix and writes the block total to *total.  s_wave
// holds kThreads/64 words; the caller separates two uses of one s_wave by a
// barrier.
template <int kThreads = kScanThreads>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave,
                                                         uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d);
    if (lane >= (uint32_t)d) x += y;
  }
  if (lane == 63u) s_wave[wave] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (uint32_t w = 0; w < kThreads / 64; ++w) {
    const uint32_t t = s_wave[w];
    if (w < wave) before += t;
    all += t;
  }
  *total = all;
  return before + x - v;
}

// Sum of one value per thread over the block (every thread gets it).
template <int kThreads = kScanThreads>
__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* s_wave) {
  uint32_t total;
  block_exclusive_scan<kThreads>(v, s_wave, &total);
  return total;
}

// The kRowCopies counters of row r (16-byte aligned).
__device__ __forceinline__ void load_copies(const uint32_t* a, uint32_t r, uint32_t (&c)[kRowCopies]) {
  const uint4* p = reinterpret_cast<const uint4*>(a + (size_t)r * kRowCopies);
  const uint4 lo = p[0], hi = p[1];
  c[0] = lo.x; c[1] = lo.y; c[2] = lo.z; c[3] = lo.w;
  c[4] = hi.x; c[5] = hi.y; c[6] = hi.z; c[7] = hi.w;
}

__device__ __forceinline__ void zero_copies(uint32_t* a, uint32_t r) {
  uint4* p = reinterpret_cast<uint4*>(a + (size_t)r * kRowCopies);
  p[0] = make_uint4(0u, 0u, 0u, 0u);
  p[1] = make_uint4(0u, 0u, 0u, 0u);
}
static_assert(kRowCopies == 8, "two 16-byte words of counters per row");

constexpr uint32_t kScanSpinLimit = 1u << 22;

// One block per 256-row group, one row per thread.  The block publishes
// its group's totals (spheres, changers, occupied rows), each tagged with
// this step's epoch, then adds the totals of the earlier groups as they
// appear (every group publishes before it waits, and waits only on earlier
// ones; the wait is bounded and reported).  Each row then takes its start
// in the next order (base + prefix within the group) and each (row, copy)
// its changers' bucket start; every occupied row gets its entry in row
// order (next start and count, current start and count, the changers'
// first bucket start and number), so that the row sort reads one entry
// per row; the counters are zeroed for the next producer.
__global__ __launch_bounds__(kScanThreads) void k_row_scan(RowTables T, uint32_t* devFlags) {
  __shared__ uint32_t s_wave[3][kScanThreads / 64];
  const uint32_t g = blockIdx.x;
  const uint32_t epoch = *T.epoch + 1u;  // the row sort advances *T.epoch after this kernel
  const uint32_t r = (g << kRowGroupShift) + threadIdx.x;
  const bool in = r < T.rows;
  uint32_t x[kRowCopies], y[kRowCopies];
  uint32_t h = 0, c = 0, curStart = 0, curCount = 0;
  if (in) {
    // the row's current start and count: only this thread rewrites them
    curStart = T.startNext[r];
    curCount = T.count[r];
    load_copies(T.rowHist, r, x);
    load_copies(T.chgHist, r, y);
#pragma unroll
    for (uint32_t q = 0; q < kRowCopies; ++q) {
      h += x[q];
      c += y[q];
    }
  }
  uint32_t th, tc, to;
  uint32_t ph = block_exclusive_scan(h, s_wave[0], &th);
  uint32_t pc = block_exclusive_scan(c, s_wave[1], &tc);
  uint32_t po = block_exclusive_scan(h != 0u ? 1u : 0u, s_wave[2], &to);
  uint64_t* pub = T.groupWords + (size_t)g * 3;
  if (threadIdx.x == 0) {
    const uint64_t tag = (uint64_t)epoch << 32;
    __hip_atomic_store(pub + 0, tag | th, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(pub + 1, tag | tc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(pub + 2, tag | to, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // the earlier groups' totals
  uint32_t bh = 0, bc = 0, bo = 0;
  bool stalled = false;
  for (uint32_t i = threadIdx.x; i < g; i += kScanThreads) {
    const uint64_t* w = T.groupWords + (size_t)i * 3;
    uint64_t v[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      v[j] = __hip_atomic_load(w + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint32_t spins = 0;
      while ((uint32_t)(v[j] >> 32) != epoch) {
        if (++spins > kScanSpinLimit) {
          stalled = true;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        v[j] = __hip_atomic_load(w + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    bh += (uint32_t)v[0];
    bc += (uint32_t)v[1];
    bo += (uint32_t)v[2];
  }
  if (stalled) atomicOr(devFlags, DEVERR_SORT_STALLED);
  __syncthreads();  // s_wave's first use is done
  bh = block_sum(bh, s_wave[0]);
  bc = block_sum(bc, s_wave[1]);
  bo = block_sum(bo, s_wave[2]);
  ph += bh;
  pc += bc;
  po += bo;
  if (in) {
    T.startNext[r] = ph;
    T.count[r] = h;
    if (h != 0u) {
      T.rowList[2 * (size_t)po] = make_uint4(r, ph, h, curStart);
      T.rowList[2 * (size_t)po + 1] = make_uint4(curCount, pc, c, 0u);
    }
    uint32_t st[kRowCopies];
#pragma unroll
    for (uint32_t q = 0; q < kRowCopies; ++q) {
      st[q] = pc;
      pc += y[q];
    }
    uint4* cs = reinterpret_cast<uint4*>(T.chgStart + (size_t)r * kRowCopies);
    cs[0] = make_uint4(st[0], st[1], st[2], st[3]);
    cs[1] = make_uint4(st[4], st[5], st[6], st[7]);
    zero_copies(T.rowHist, r);
    zero_copies(T.chgHist, r);
    zero_copies(T.chgCursor, r);
  }
  if (g == T.groups - 1u && threadIdx.x == kScanThreads - 1) {
    // the last thread's prefixes plus its own counts are the totals
    T.startNext[T.rows] = ph + h;  // past the last row: the sphere count
    T.chgStart[(size_t)T.rows * kRowCopies] = pc;
    *T.listCount = po + (h != 0u ? 1u : 0u);
  }
}

// Start of a run: zero the counters (a run that stopped early may have left
// counts behind) and move the epoch past any tag a stopped scan published.
__global__ __launch_bounds__(256) void k_row_reset(RowTables T) {
  for (size_t i = (size_t)blockIdx.x * 256u + threadIdx.x; i < T.zeroedWords;
       i += (size_t)gridDim.x * 256u)
    T.zeroed[i] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) *T.epoch += 2u;
}

// First position of a sorted hash array whose hash lies in row r or beyond.
__device__ __forceinline__ uint32_t row_start_in(const uint32_t* hash, uint32_t n, uint32_t rows,
                                                 uint32_t shift, uint32_t r) {
  if (r >= rows) return n;
  const uint32_t first = r << shift;
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (hash[mid] < first)
      lo = mid + 1u;
    else
      hi = mid;
  }
  return lo;
}

// Row starts and counts of a sorted hash array (binary searches per row).
__global__ __launch_bounds__(256) void k_row_starts(RowTables T, const uint32_t* __restrict__ hash,
                                                    uint32_t n) {
  const uint32_t r = blockIdx.x * 256u + threadIdx.x;
  if (r > T.rows) return;
  const uint32_t lo = row_start_in(hash, n, T.rows, T.shift, r);
  T.startNext[r] = lo;
  if (r < T.rows) T.count[r] = row_start_in(hash, n, T.rows, T.shift, r + 1u) - lo;
}

// One lane per current sorted entry k (keys[k] is the next key with
// kMovedBit when the cell changes, hash[k] the current one), each wave over
// kChgTiles consecutive 64-slot tiles with every tile's loads, then every
// tile's bucket positions, in flight together.  A sphere whose row changes
// is listed in its new row's bucket with its slot, next key and id.  A
// current cell none of whose spheres stays is cleared in the tables (the
// row sort, which runs next, writes every cell of the next order, this one
// too if spheres move in): the tile's ballots tell a cell's first lane
// whether any of the cell's lanes stays; only a cell that goes on past the
// tile reads on.
constexpr uint32_t kChgTiles = 2;  // 1: 12.0 us, 2: 11.6, 4: 12.8, 8: 15.5 on the 1M bed
__global__ __launch_bounds__(256) void k_row_changers(RowTables T, const uint32_t* __restrict__ keys,
                                                      const uint32_t* __restrict__ idOfSlot,
                                                      const uint32_t* __restrict__ hash,
                                                      uint32_t* __restrict__ cellStart,
                                                      uint32_t* __restrict__ cellEnd, uint32_t n,
                                                      uint32_t numCells, uint32_t* devFlags) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t tile0 = (blockIdx.x * 4u + (threadIdx.x >> 6)) * kChgTiles;
  uint32_t raw[kChgTiles], old[kChgTiles], before[kChgTiles], after[kChgTiles];
#pragma unroll
  for (uint32_t i = 0; i < kChgTiles; ++i) {
    const uint32_t k = (tile0 + i) * 64u + lane;
    const bool valid = k < n;
    raw[i] = valid ? keys[k] : 0u;
    old[i] = valid ? hash[k] : 0u;
    before[i] = (lane == 0u && valid && k > 0) ? hash[k - 1] : 0u;
    after[i] = (lane == 63u && valid && k + 1 < n) ? hash[k + 1] : 0u;
  }
  bool chg[kChgTiles];
  uint32_t bucket[kChgTiles], id[kChgTiles], lo[kChgTiles], hi[kChgTiles];
#pragma unroll
  for (uint32_t i = 0; i < kChgTiles; ++i) {
    const uint32_t k = (tile0 + i) * 64u + lane;
    const bool valid = k < n;
    uint32_t oldPrev = __shfl_up(old[i], 1), oldNext = __shfl_down(old[i], 1);
    if (lane == 0u) oldPrev = before[i];
    if (lane == 63u) oldNext = after[i];
    const bool first = valid && (k == 0 || oldPrev != old[i]);
    const bool last = valid && (k + 1 == n || oldNext != old[i]);
    const uint64_t lastMask = __ballot(last);
    const uint64_t staysMask = __ballot(valid && (raw[i] & kMovedBit) == 0u);
    if (first) {
      const uint64_t ends = lastMask & lanes_from(lane);
      const uint32_t t = ends ? lowest_lane(ends) : 63u;
      bool stays = (staysMask & lanes_from(lane) & ~lanes_from(t + 1u)) != 0ull;
      // the cell goes on past the tile's last lane
      for (uint32_t k2 = k + (64u - lane); !stays && ends == 0ull && k2 < n && hash[k2] == old[i];
           ++k2)
        stays = (keys[k2] & kMovedBit) == 0u;
      if (!stays && old[i] < numCells) {
        cellStart[old[i]] = CELL_EMPTY;
        cellEnd[old[i]] = CELL_EMPTY;
      }
    }
    const uint32_t r = (raw[i] & ~kMovedBit) >> T.shift;
    chg[i] = valid && r != (old[i] >> T.shift);
    bucket[i] = 0u;
    id[i] = lo[i] = hi[i] = 0u;
    if (chg[i]) {
      if (r >= T.rows) {
        atomicOr(devFlags, DEVERR_ROW_COUNT);
        chg[i] = false;
      } else {
        bucket[i] = r * kRowCopies + row_copy(k);
        id[i] = idOfSlot[k];
        lo[i] = T.chgStart[bucket[i]];
        hi[i] = T.chgStart[bucket[i] + 1];
      }
    }
  }
  // one atomic per bucket and tile: the lowest lane of each bucket's lanes
  // (its leader) adds their count, every lane takes the leader's result
  // plus its rank among them; the leaders of both tiles add in one go
  uint32_t leader[kChgTiles], rank[kChgTiles], cnt[kChgTiles], at[kChgTiles];
#pragma unroll
  for (uint32_t i = 0; i < kChgTiles; ++i) {
    leader[i] = lane;
    rank[i] = cnt[i] = 0u;
    uint64_t pending = __ballot(chg[i]);
    while (pending != 0ull) {
      const uint32_t l = lowest_lane(pending);
      const uint32_t bl = __builtin_amdgcn_readlane(bucket[i], l);
      const uint64_t peers = __ballot(chg[i] && bucket[i] == bl);
      if ((peers >> lane) & 1ull) {
        leader[i] = l;
        rank[i] = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
      }
      if (lane == l) cnt[i] = (uint32_t)__popcll(peers);
      pending &= ~peers;
    }
  }
#pragma unroll
  for (uint32_t i = 0; i < kChgTiles; ++i)
    at[i] = (chg[i] && leader[i] == lane) ? atomicAdd(&T.chgCursor[bucket[i]], cnt[i]) : 0u;
#pragma unroll
  for (uint32_t i = 0; i < kChgTiles; ++i) at[i] = __shfl(at[i], (int)leader[i]) + rank[i];
#pragma unroll
  for (uint32_t i = 0; i < kChgTiles; ++i) {
    if (!chg[i]) continue;
    const uint32_t k = (tile0 + i) * 64u + lane;
    const uint32_t pos = lo[i] + at[i];
    if (pos < hi[i])
      T.chgList[pos] = make_uint4(k, raw[i] & ~kMovedBit, id[i], 0u);
    else
      atomicOr(devFlags, DEVERR_ROW_COUNT);
  }
}

// Writes a member at position q: the sorted hash and id, the gathered
// state, and the cell-table entries where the cell changes.
__device__ __forceinline__ void emit_member(const RowOutputs& O, uint32_t q, uint32_t key,
                                            uint32_t id, uint32_t slot, bool firstOfCell,
                                            bool lastOfCell) {
  O.hash[q] = key;
  O.index[q] = id;
  const float4 p = O.pos[slot];
  const float4 v = O.vel[slot];
  O.sortedPos[q] = p;
  O.sortedVel[q] = make_float4(v.x, v.y, v.z, __uint_as_float(key));
  O.sortedQ[q] = quantize_pos(O.frame, p.x, p.y, p.z);
  if (firstOfCell) O.cellStart[key] = q;
  if (lastOfCell) O.cellEnd[key] = q + 1u;
}

// Writes the row's m members, in order in s_key/s_slot, from position o0:
// each thread loads the state of up to kEmitBatch members before it stores
// any, so the gathers' latencies overlap.
constexpr uint32_t kEmitBatch = 4;
__device__ __forceinline__ void emit_row(const RowOutputs& O, const uint64_t* s_key,
                                         const uint32_t* s_slot, uint32_t o0, uint32_t m,
                                         uint32_t rowBase) {
  for (uint32_t base = 0; base < m; base += kSortThreads * kEmitBatch) {
    float4 p[kEmitBatch], v[kEmitBatch];
#pragma unroll
    for (uint32_t t = 0; t < kEmitBatch; ++t) {
      const uint32_t i = base + t * kSortThreads + threadIdx.x;
      if (i < m) {
        const uint32_t slot = s_slot[i];
        p[t] = O.pos[slot];
        v[t] = O.vel[slot];
      }
    }
#pragma unroll
    for (uint32_t t = 0; t < kEmitBatch; ++t) {
      const uint32_t i = base + t * kSortThreads + threadIdx.x;
      if (i >= m) continue;
      const uint64_t ck = s_key[i];
      const uint32_t x = (uint32_t)(ck >> 32);
      const uint32_t key = rowBase | x;
      const uint32_t q = o0 + i;
      O.hash[q] = key;
      O.index[q] = (uint32_t)ck;
      O.sortedPos[q] = p[t];
      O.sortedVel[q] = make_float4(v[t].x, v[t].y, v[t].z, __uint_as_float(key));
      O.sortedQ[q] = quantize_pos(O.frame, p[t].x, p[t].y, p[t].z);
      if (i == 0u || (uint32_t)(s_key[i - 1] >> 32) != x) O.cellStart[key] = q;
      if (i + 1u == m || (uint32_t)(s_key[i + 1] >> 32) != x) O.cellEnd[key] = q + 1u;
    }
  }
}

__device__ __forceinline__ uint64_t member_key(uint32_t key, uint32_t shift, uint32_t id) {
  return ((uint64_t)(key & ((1u << shift) - 1u)) << 32) | id;
}

// Candidate u of row r: entry a0 + u of the current order for u < A (a
// member if its next key stays in the row), else incomer u - A.  Returns
// the member's (cell in row, id) key, or kNoMember.
__device__ __forceinline__ uint64_t candidate(const RowTables& T, const uint32_t* keys,
                                              const uint32_t* idOfSlot, uint32_t r, uint32_t a0,
                                              uint32_t A, uint32_t c0, uint32_t u,
                                              uint32_t* slotOut) {
  if (u < A) {
    const uint32_t k = a0 + u;
    const uint32_t key = keys[k] & ~kMovedBit;
    *slotOut = k;
    if ((key >> T.shift) != r) return kNoMember;
    return member_key(key, T.shift, idOfSlot[k]);
  }
  const uint4 c = T.chgList[c0 + (u - A)];
  *slotOut = c.x;
  if ((c.y >> T.shift) != r) return kNoMember;
  return member_key(c.y, T.shift, c.z);
}

// Number of entries of the ascending s_key[lo, lo + len) below x.
__device__ __forceinline__ uint32_t count_below(const uint64_t* s_key, uint32_t lo, uint32_t len,
                                                uint64_t x) {
  uint32_t a = 0, b = len;
  while (a < b) {
    const uint32_t mid = (a + b) >> 1;
    if (s_key[lo + mid] < x)
      a = mid + 1u;
    else
      b = mid;
  }
  return a;
}

// The members of row r (more than LDS holds, or too many movers to rank):
// each member's place is the number of members below it, counted over the
// candidates staged through LDS in kRowCap chunks.  O(m x candidates).
__device__ void rank_by_counting(const RowTables& T, const uint32_t* keys,
                                 const uint32_t* idOfSlot, const RowOutputs& O, uint32_t r,
                                 uint32_t o0, uint32_t m, uint32_t a0, uint32_t A, uint32_t c0,
                                 uint32_t C, uint64_t* s_key, uint32_t* s_wave,
                                 uint32_t* devFlags) {
  const uint32_t tid = threadIdx.x;
  const uint32_t rowBase = r << T.shift;
  const uint32_t U = A + C;
  uint32_t members = 0;
  for (uint32_t tb = 0; tb < U; tb += kSortThreads) {
    const uint32_t u = tb + tid;
    uint32_t slot = 0;
    const uint64_t mine = u < U ? candidate(T, keys, idOfSlot, r, a0, A, c0, u, &slot) : kNoMember;
    uint32_t below = 0, sameBelow = 0, sameAbove = 0;
    for (uint32_t w = 0; w < U; w += kRowCap) {
      __syncthreads();
      for (uint32_t i = tid; i < kRowCap; i += kSortThreads) {
        uint32_t sl;
        s_key[i] =
            w + i < U ? candidate(T, keys, idOfSlot, r, a0, A, c0, w + i, &sl) : kNoMember;
      }
      __syncthreads();
      if (mine != kNoMember) {
        const uint32_t xm = (uint32_t)(mine >> 32);
        for (uint32_t i = 0; i < kRowCap; ++i) {
          const uint64_t o = s_key[i];
          if (o == kNoMember) continue;
          below += o < mine ? 1u : 0u;
          const bool same = (uint32_t)(o >> 32) == xm;
          sameBelow += (same && o < mine) ? 1u : 0u;
          sameAbove += (same && o > mine) ? 1u : 0u;
        }
      }
    }
    if (mine != kNoMember) {
      ++members;
      if (below < m)
        emit_member(O, o0 + below, rowBase | (uint32_t)(mine >> 32), (uint32_t)mine, slot,
                    sameBelow == 0u, sameAbove == 0u);
      else
        atomicOr(devFlags, DEVERR_ROW_COUNT);
    }
  }
  if (block_sum<kSortThreads>(members, s_wave) != m && tid == 0)
    atomicOr(devFlags, DEVERR_ROW_COUNT);
}

// The next keys carry kMovedBit; the sorted hashes are written to O.hash
// while other blocks still read their rows, so the current hashes are not
// read here.  A row's loads form a short chain: its entry (read before the
// list length, as the launch has at most one block per row), then its
// current range's keys and ids together with its incomers' records, then,
// after the merge in LDS, the members' state.
__global__ __launch_bounds__(kSortThreads) void k_row_sort(
    RowTables T, const uint32_t* __restrict__ keys, const uint32_t* __restrict__ idOfSlot,
    RowOutputs O, uint32_t* devFlags) {
  // pool: the stayers in order from the front, the movers from the back;
  // after the merge the row's members in order from the front
  __shared__ uint64_t s_key[kRowCap];
  __shared__ uint32_t s_slot[kRowCap];
  __shared__ uint32_t s_wave[kSortThreads / 64];
  __shared__ uint32_t s_count[kPerThread][kSortThreads / 64];  // stayers per chunk and wave
  __shared__ uint32_t s_movers;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  uint4 e0 = T.rowList[2 * (size_t)blockIdx.x], e1 = T.rowList[2 * (size_t)blockIdx.x + 1];
  const uint32_t listed = *T.listCount;
  if (blockIdx.x == 0 && tid == 0) *T.epoch += 1u;  // the scan of this step is done
  for (uint32_t j = blockIdx.x; j < listed; j += gridDim.x) {
    if (j != blockIdx.x) {
      e0 = T.rowList[2 * (size_t)j];
      e1 = T.rowList[2 * (size_t)j + 1];
    }
    const uint32_t r = e0.x, o0 = e0.y, m = e0.z, a0 = e0.w;
    const uint32_t A = e1.x, c0 = e1.y, C = e1.z;
    const uint32_t rowBase = r << T.shift;
    __syncthreads();  // the previous row's LDS reads are done
    if (m > kRowCap) {
      rank_by_counting(T, keys, idOfSlot, O, r, o0, m, a0, A, c0, C, s_key, s_wave, devFlags);
      continue;
    }
    // gather, kRowCap candidates at a time (one group for every row of at
    // most kRowCap current spheres and incomers): candidate u = g0 + 128 t
    // + tid, loads first; a stayer's place is the number of stayers before
    // it in candidate order (wave ballots per 128-candidate chunk, counts
    // exchanged in LDS); movers fill the pool from the back
    if (tid == 0) s_movers = 0u;
    const uint32_t U = A + C;
    uint32_t stayers = 0;  // the same in every thread
    for (uint32_t g0 = 0; g0 < U; g0 += kRowCap) {
      uint32_t raw[kPerThread], idv[kPerThread], slot[kPerThread];
#pragma unroll
      for (uint32_t t = 0; t < kPerThread; ++t) {
        const uint32_t u = g0 + t * kSortThreads + tid;
        raw[t] = 0u;
        idv[t] = 0u;
        slot[t] = 0u;
        if (u < A) {
          slot[t] = a0 + u;
          raw[t] = keys[a0 + u];
          idv[t] = idOfSlot[a0 + u];
        } else if (u < U) {
          const uint4 c = T.chgList[c0 + (u - A)];
          slot[t] = c.x;
          raw[t] = c.y | kMovedBit;  // a sphere that changes row changes cell
          idv[t] = c.z;
        }
      }
      uint64_t stays[kPerThread];
#pragma unroll
      for (uint32_t t = 0; t < kPerThread; ++t) {
        const uint32_t u = g0 + t * kSortThreads + tid;
        const bool member = u < U && ((raw[t] & ~kMovedBit) >> T.shift) == r;
        stays[t] = __ballot(member && (raw[t] & kMovedBit) == 0u);
        if (lane == 0u) s_count[t][wave] = (uint32_t)__popcll(stays[t]);
      }
      __syncthreads();
      uint32_t before = stayers;  // stayers of the earlier chunks
#pragma unroll
      for (uint32_t t = 0; t < kPerThread; ++t) {
        const uint32_t u = g0 + t * kSortThreads + tid;
        const uint32_t w0 = s_count[t][0], w1 = s_count[t][1];
        const uint32_t key = raw[t] & ~kMovedBit;
        if (u < U && (key >> T.shift) == r) {
          const uint32_t below = __builtin_amdgcn_mbcnt_hi(
              (uint32_t)(stays[t] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)stays[t], 0u));
          const uint32_t pos = (raw[t] & kMovedBit) == 0u
                                   ? before + (wave != 0u ? w0 : 0u) + below
                                   : kRowCap - 1u - atomicAdd(&s_movers, 1u);
          if (pos < kRowCap) {  // (beyond it the counts disagree: checked below)
            s_key[pos] = member_key(key, T.shift, idv[t]);
            s_slot[pos] = slot[t];
          }
        }
        before += w0 + w1;
      }
      stayers = before;
      __syncthreads();  // the pool is written; s_count may be reused
    }
    if (U == 0u) {  // a listed row has members: the counts disagree
      if (tid == 0) atomicOr(devFlags, DEVERR_ROW_COUNT);
      continue;
    }
    const uint32_t movers = s_movers;
    if (stayers + movers != m) {  // the counts and the members disagree: write nothing
      if (tid == 0) atomicOr(devFlags, DEVERR_ROW_COUNT);
      continue;
    }
    const uint32_t mBase = kRowCap - movers;  // movers at [mBase, kRowCap)
    if (movers <= kMoverCap) {
      // rank the movers among themselves and put them in order
      uint64_t mk[kMoversPerThread];
      uint32_t ms[kMoversPerThread], mr[kMoversPerThread];
#pragma unroll
      for (uint32_t t = 0; t < kMoversPerThread; ++t) {
        const uint32_t i = tid + t * kSortThreads;
        mk[t] = kNoMember;
        ms[t] = 0u;
        mr[t] = 0u;
        if (i < movers) {
          mk[t] = s_key[mBase + i];
          ms[t] = s_slot[mBase + i];
          uint32_t below = 0;
          for (uint32_t q = 0; q < movers; ++q) below += s_key[mBase + q] < mk[t] ? 1u : 0u;
          mr[t] = below;
        }
      }
      __syncthreads();
#pragma unroll
      for (uint32_t t = 0; t < kMoversPerThread; ++t)
        if (tid + t * kSortThreads < movers) {
          s_key[mBase + mr[t]] = mk[t];
          s_slot[mBase + mr[t]] = ms[t];
        }
      __syncthreads();
      // each member's place: its rank in its list + the other list's below it
      uint64_t sk[kPerThread];
      uint32_t ss[kPerThread], sr[kPerThread];
#pragma unroll
      for (uint32_t t = 0; t < kPerThread; ++t) {
        const uint32_t i = tid + t * kSortThreads;
        sk[t] = kNoMember;
        ss[t] = 0u;
        sr[t] = 0u;
        if (i < stayers) {
          sk[t] = s_key[i];
          ss[t] = s_slot[i];
          sr[t] = i + count_below(s_key, mBase, movers, sk[t]);
        }
      }
#pragma unroll
      for (uint32_t t = 0; t < kMoversPerThread; ++t) {
        const uint32_t i = tid + t * kSortThreads;
        if (i < movers) {
          mk[t] = s_key[mBase + i];
          ms[t] = s_slot[mBase + i];
          mr[t] = i + count_below(s_key, 0u, stayers, mk[t]);
        }
      }
      __syncthreads();
#pragma unroll
      for (uint32_t t = 0; t < kPerThread; ++t)
        if (tid + t * kSortThreads < stayers) {
          s_key[sr[t]] = sk[t];
          s_slot[sr[t]] = ss[t];
        }
#pragma unroll
      for (uint32_t t = 0; t < kMoversPerThread; ++t)
        if (tid + t * kSortThreads < movers) {
          s_key[mr[t]] = mk[t];
          s_slot[mr[t]] = ms[t];
        }
    } else {
      // many movers (a transient): count instead
      rank_by_counting(T, keys, idOfSlot, O, r, o0, m, a0, A, c0, C, s_key, s_wave, devFlags);
      continue;
    }
    __syncthreads();
    emit_row(O, s_key, s_slot, o0, m, rowBase);
  }
}

}  // namespace

bool row_resort_grid_ok(const fc_sim_params& P) {
  const uint64_t rows = (uint64_t)P.gridSize.y * P.gridSize.z;
  return rows >= 1 && rows <= kMaxRows && P.gridSize.x <= (1u << 20);
}

RowHists row_hists(const RowTables& T) {
  RowHists h;
  h.rowHist = T.rowHist;
  h.chgHist = T.chgHist;
  h.shift = T.shift;
  return h;
}

fc_status ensure_row_tables(RowTables& T, const fc_sim_params& P, uint32_t n) {
  uint32_t shift = 0;
  while ((1u << shift) < P.gridSize.x) ++shift;
  const uint32_t rows = P.gridSize.y * P.gridSize.z;
  const uint32_t groups = (rows + (1u << kRowGroupShift) - 1u) >> kRowGroupShift;
  if (T.block != nullptr && T.rows == rows && T.shift == shift && T.capacity >= n)
    return FC_SUCCESS;
  free_row_tables(T);
  // count, startNext [rows + 1]; chgStart [rows x copies + 1]; rowList
  // [2 rows, 16-byte]; groupWords [3 groups, 64-bit]; listCount, epoch [2];
  // rowHist, chgHist, chgCursor [rows x copies] (zeroed); chgList [n, 16-byte]
  // every array up to the counters starts on a 64-word (256-byte) boundary,
  // chgList on a 32-byte one
  const size_t cells = (size_t)rows * kRowCopies;
  auto up = [](size_t x) { return (x + 63) & ~(size_t)63; };
  const size_t zeroWords = 3 * cells;
  const size_t words = 2 * up(rows + 1) + up(cells + 1) + up(8 * (size_t)rows) +
                       up(6 * (size_t)groups) + up(2) + zeroWords + 4 * (size_t)(n ? n : 1);
  FC_HIP_TRY(hipMalloc(&T.block, words * sizeof(uint32_t)));
  uint32_t* w = T.block;
  T.count = w;
  w += up(rows + 1);
  T.startNext = w;
  w += up(rows + 1);
  T.chgStart = w;
  w += up(cells + 1);
  T.rowList = reinterpret_cast<uint4*>(w);
  w += up(8 * (size_t)rows);
  T.groupWords = reinterpret_cast<uint64_t*>(w);
  w += up(6 * (size_t)groups);
  T.listCount = w;
  T.epoch = w + 1;
  w += up(2);
  T.zeroed = w;
  T.zeroedWords = zeroWords;
  T.rowHist = w;
  w += cells;
  T.chgHist = w;
  w += cells;
  T.chgCursor = w;
  w += cells;
  T.chgList = reinterpret_cast<uint4*>(w);
  T.rows = rows;
  T.shift = shift;
  T.groups = groups;
  T.capacity = n;
  FC_HIP_TRY(hipMemset(T.block, 0, words * sizeof(uint32_t)));
  return FC_SUCCESS;
}

void free_row_tables(RowTables& T) {
  if (T.block) hipFree(T.block);
  T = RowTables{};
}

hipError_t launch_row_clear(hipStream_t s, const RowTables& T) {
  const uint32_t blocks = (uint32_t)((T.zeroedWords + 255u) / 256u);
  hipLaunchKernelGGL(k_row_reset, dim3(blocks < 4096u ? blocks : 4096u), dim3(256), 0, s, T);
  return hipGetLastError();
}

hipError_t launch_row_scan(hipStream_t s, const RowTables& T, uint32_t* devFlags) {
  hipLaunchKernelGGL(k_row_scan, dim3(T.groups), dim3(kScanThreads), 0, s, T, devFlags);
  return hipGetLastError();
}

hipError_t launch_row_starts(hipStream_t s, const RowTables& T, const uint32_t* sortedHash,
                             uint32_t n) {
  hipLaunchKernelGGL(k_row_starts, dim3((T.rows + 1u + 255u) / 256u), dim3(256), 0, s, T,
                     sortedHash, n);
  return hipGetLastError();
}

hipError_t launch_row_changers(hipStream_t s, const RowTables& T, const uint32_t* keys,
                               const uint32_t* idOfSlot, const uint32_t* hash,
                               uint32_t* cellStart, uint32_t* cellEnd, uint32_t n,
                               uint32_t numCells, uint32_t* devFlags) {
  if (n == 0) return hipSuccess;
  const uint32_t perBlock = 256u * kChgTiles;
  hipLaunchKernelGGL(k_row_changers, dim3((n + perBlock - 1u) / perBlock), dim3(256), 0, s, T, keys,
                     idOfSlot, hash, cellStart, cellEnd, n, numCells, devFlags);
  return hipGetLastError();
}

hipError_t launch_row_sort(hipStream_t s, const RowTables& T, const uint32_t* keys,
                           const uint32_t* idOfSlot, const RowOutputs& O, uint32_t* devFlags) {
  const uint32_t blocks = T.rows < kSortBlocks ? T.rows : kSortBlocks;
  hipLaunchKernelGGL(k_row_sort, dim3(blocks), dim3(kSortThreads), 0, s, T, keys, idOfSlot, O,
                     devFlags);
  return hipGetLastError();
}

}  // namespace fc
