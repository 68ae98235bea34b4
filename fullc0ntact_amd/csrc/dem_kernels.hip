// Particle-step kernels for gfx950: integrate, calcHash, reorder /
// find-cell-start and the 27-cell collide.
//
// Layout in HBM: positions and velocities are float4[N] arrays (x, y, z, w),
// read and written as one 16-byte vector per particle per lane, so a wave64
// touches 1 KiB of consecutive memory per access.  Hash, index and cell
// tables are uint32 arrays.  Parameters arrive by value as kernel arguments.
#include "fc_internal.hpp"
#include "fc_device.hpp"
#include "radix_common.hpp"
#include "distmap.hpp"
#include "collide.hpp"
#include "row_resort.hpp"

#include <cmath>
#include <cstring>

namespace fc {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(size_t n, int block) {
  size_t g = (n + (size_t)block - 1) / (size_t)block;
  return (unsigned)(g == 0 ? 1 : g);
}

// Fills p (blockIdx.y 0) and p2 (blockIdx.y 1) with value.
__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* __restrict__ p0,
                                                     uint32_t* __restrict__ p2, uint32_t value,
                                                     size_t count) {
  uint32_t* __restrict__ p = blockIdx.y == 0 ? p0 : p2;
  size_t tid = (size_t)blockIdx.x * kBlock + threadIdx.x;
  size_t stride = (size_t)gridDim.x * kBlock;
  // 16-byte stores over the aligned body, scalar stores for the ragged ends.
  size_t head = ((16 - ((uintptr_t)p & 15)) & 15) / 4;
  if (head > count) head = count;
  if (tid < head) p[tid] = value;
  uint32_t* body = p + head;
  size_t nvec = (count - head) / 4;
  uint4 v = make_uint4(value, value, value, value);
  for (size_t i = tid; i < nvec; i += stride) reinterpret_cast<uint4*>(body)[i] = v;
  size_t tail = head + nvec * 4;
  if (tid < count - tail) p[tail + tid] = value;
}

// Plain 16-byte-per-lane copy (bandwidth probe).  Launched with one lane per
// element: on MI355X a capped grid that strides over 1 GiB (2048 or 8192
// blocks) copies at 4.5-4.8 TB/s, one element per lane at 6.2 TB/s.
__global__ __launch_bounds__(kBlock) void k_copy16(uint4* __restrict__ dst,
                                                   const uint4* __restrict__ src, size_t count) {
  size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) dst[i] = src[i];
}

// integrate_functor (particles_kernel.cu:39-75) over zip(pos, vel).
__global__ __launch_bounds__(kBlock) void k_integrate(fc_sim_params P, float4* __restrict__ pos,
                                                      float4* __restrict__ vel, float dt,
                                                      uint32_t n) {
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float4 p4 = pos[i];
  float4 v4 = vel[i];
  v3f p = xyz(p4), v = xyz(v4);
  integrate_box(P, P.particleRadius, dt, p, v);
  pos[i] = make_float4(p.x, p.y, p.z, p4.w);
  vel[i] = make_float4(v.x, v.y, v.z, v4.w);
}

// calcHashD (particles_kernel.cu:97-115).
__global__ __launch_bounds__(kBlock) void k_calc_hash(fc_sim_params P, uint32_t* __restrict__ hash,
                                                      uint32_t* __restrict__ index,
                                                      const float4* __restrict__ pos, uint32_t n) {
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float4 p = pos[i];
  hash[i] = hash_of(P, p.x, p.y, p.z);
  index[i] = i;
}

// The state the row-range collide reads, built from the reference-style
// sorted arrays of fc_collide / fc_eval_forces, which carry neither: each
// sorted sphere's velocity with the bits of its cell hash in w (the hash of
// its position, as calcHash computes it, so the hash it was sorted by when
// the arrays come from reorderDataAndFindCellStart of those hashes), and its
// 16-bit coordinates in `frame`.
__global__ __launch_bounds__(kBlock) void k_rows_state(fc_sim_params P, QuantFrame frame,
                                                       const float4* __restrict__ sortedPos,
                                                       const float4* __restrict__ sortedVel,
                                                       uint32_t n, float4* __restrict__ velOut,
                                                       uint2* __restrict__ qOut) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 p = sortedPos[i], v = sortedVel[i];
  velOut[i] = make_float4(v.x, v.y, v.z, __uint_as_float(hash_of(P, p.x, p.y, p.z)));
  qOut[i] = quantize_pos(frame, p.x, p.y, p.z);
}

// integrate + calcHash in one pass (the update order runs them back to back
// on the same arrays, so fusing them changes no value), with two more jobs
// folded into the same read of the particles:
//  * when prevSortedHash is given, the cells occupied after the previous
//    update are reset to EMPTY: cell h was written iff h starts a run in the
//    sorted hash array, so exactly those entries are cleared (instead of the
//    reference's memset of the whole table, particleSystem.cu:198);
//  * the digit histogram of the new keys for the radix sort's passes.
// Each thread handles kIhPer particles strided by the block size.
constexpr int kIhPer = 4;

__global__ __launch_bounds__(kBlock) void k_integrate_hash(
    fc_sim_params P, BoundaryArg B, float4* __restrict__ pos, float4* __restrict__ vel, float dt,
    uint32_t* __restrict__ keysOut, const uint32_t* __restrict__ prevSortedHash,
    uint32_t* __restrict__ cellStart, uint32_t* __restrict__ cellEnd, uint32_t n, int passes,
    uint32_t* gHist, uint32_t* counters) {
  __shared__ uint32_t s_hist[RS_RADIX];
  for (int i = threadIdx.x; i < RS_RADIX; i += kBlock) s_hist[i] = 0u;
  histogram_producer_prologue(gHist, passes, counters);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kIhPer; ++u) {
    const uint32_t i = blockIdx.x * (kBlock * kIhPer) + u * kBlock + threadIdx.x;
    uint32_t key = 0;
    if (i < n) {
    if (prevSortedHash != nullptr) {
      const uint32_t h = prevSortedHash[i];
      if (i == 0 || prevSortedHash[i - 1] != h) {
        cellStart[h] = CELL_EMPTY;
        cellEnd[h] = CELL_EMPTY;
      }
    }
    const float4 p4 = pos[i];
    const float4 v4 = vel[i];
    v3f p = xyz(p4), v = xyz(v4);
    integrate_with(P, B, p4.w, dt, p, v);
    pos[i] = make_float4(p.x, p.y, p.z, p4.w);
    vel[i] = make_float4(v.x, v.y, v.z, v4.w);
    key = hash_of(P, p.x, p.y, p.z);
    keysOut[i] = key;
    }
    histogram_add_digit(s_hist, key & (RS_RADIX - 1), i < n);
  }
  __syncthreads();
  histogram_flush(s_hist, gHist, 0);
}

// reorderDataAndFindCellStartD (particles_kernel.cu:119-182).  The
// neighbour's hash comes from the previous lane through a cross-lane
// shuffle; lane 0 of each wave reads it from memory.
// vals are the sorted values: particle ids, or -- with idOfSlot, in the
// sorted-slot layout kept between the fused steps of a run -- the slots of
// oldPos/oldVel that hold each particle, whose ids idOfSlot gives; then the
// ids are written to indexOut.
__global__ __launch_bounds__(kBlock) void k_reorder(
    uint32_t* __restrict__ cellStart, uint32_t* __restrict__ cellEnd,
    float4* __restrict__ sortedPos, float4* __restrict__ sortedVel,
    const uint32_t* __restrict__ hash, const uint32_t* __restrict__ vals,
    const uint32_t* __restrict__ idOfSlot, uint32_t* __restrict__ indexOut,
    const float4* __restrict__ oldPos, const float4* __restrict__ oldVel, uint32_t n,
    uint32_t numCells) {
  uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  uint32_t h = (k < n) ? hash[k] : 0u;
  uint32_t hprev = __shfl_up(h, 1);
  if (k >= n) return;
  if ((threadIdx.x & 63u) == 0u && k > 0) hprev = hash[k - 1];
  // Hashes produced by calcHash are < numCells; the guards only keep
  // caller-supplied out-of-range hashes from writing outside the tables.
  if (k == 0 || h != hprev) {
    if (h < numCells) cellStart[h] = k;
    if (k > 0 && hprev < numCells) cellEnd[hprev] = k;
  }
  if (k == n - 1 && h < numCells) cellEnd[h] = k + 1;
  const uint32_t src = vals[k];
  sortedPos[k] = oldPos[src];
  sortedVel[k] = oldVel[src];
  if (idOfSlot) indexOut[k] = idOfSlot[src];
}

// The cell-table half of k_reorder (the gather done by the sort's last
// pass).
__global__ __launch_bounds__(kBlock) void k_cell_bounds(uint32_t* __restrict__ cellStart,
                                                        uint32_t* __restrict__ cellEnd,
                                                        const uint32_t* __restrict__ hash,
                                                        uint32_t n, uint32_t numCells) {
  uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  uint32_t h = (k < n) ? hash[k] : 0u;
  uint32_t hprev = __shfl_up(h, 1);
  if (k >= n) return;
  if ((threadIdx.x & 63u) == 0u && k > 0) hprev = hash[k - 1];
  if (k == 0 || h != hprev) {
    if (h < numCells) cellStart[h] = k;
    if (k > 0 && hprev < numCells) cellEnd[hprev] = k;
  }
  if (k == n - 1 && h < numCells) cellEnd[h] = k + 1;
}

// Sum of the pair forces of cells [s, e) on particle k, in index order,
// with a one-ahead prefetch of the next neighbour's position and velocity.
__device__ __forceinline__ v3f cell_sum(const fc_sim_params& P, v3f pos, v3f vel, uint32_t k,
                                        uint32_t s, uint32_t e, const float4* __restrict__ sp,
                                        const float4* __restrict__ sv, float rA, bool poly) {
  v3f cf = make_v3(0.0f, 0.0f, 0.0f);
  if (s >= e) return cf;
  float4 pn = sp[s], vn = sv[s];
  for (uint32_t j = s; j < e; ++j) {
    const float4 pc = pn, vc = vn;
    if (j + 1 < e) {
      pn = sp[j + 1];
      vn = sv[j + 1];
    }
    if (j != k) {
      v3f f = collide_spheres(pos, xyz(pc), vel, xyz(vc), rA, poly ? pc.w : rA, P.spring,
                              P.damping, P.shear, P.attraction);
      cf = make_v3(cf.x + f.x, cf.y + f.y, cf.z + f.z);
    }
  }
  return cf;
}

// collideD / collideCell (particles_kernel.cu:223-293): one lane per sorted
// particle; for z, y, x in {-1,0,1} (that nesting is the summation order),
// each neighbour cell's forces are summed into a per-cell partial that is
// then added to the total, then the collider sphere is added, and
// vel + force is scattered to the particle's original slot.  The 9
// cell-table entries of one z-plane of
// the neighbourhood (start and end) are loaded together before any
// neighbour is visited, so each plane costs one table latency instead of a
// chain of 9-18 dependent loads.  Empty cells contribute nothing: their
// partial sum is +0 and the running total is never -0, so skipping the add
// leaves every bit unchanged.
// forceOut != nullptr is evalForces_launcher (particledem_um.cu:538-589):
// no cursor sphere, and the force also stored as float3 at the original
// index.
__global__ __launch_bounds__(kBlock) void k_collide_v2(
    fc_sim_params P, float4* __restrict__ newVel, const float4* __restrict__ sortedPos,
    const float4* __restrict__ sortedVel, const uint32_t* __restrict__ index,
    const uint32_t* __restrict__ cellStart, const uint32_t* __restrict__ cellEnd, uint32_t n,
    int poly, uint32_t* devFlags, float* __restrict__ forceOut) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  const float4 p4 = sortedPos[k];
  const v3f pos = xyz(p4);
  const v3f vel = xyz(sortedVel[k]);
  const int gx = cell_coord(pos.x, P.worldOrigin.x, P.cellSize.x);
  const int gy = cell_coord(pos.y, P.worldOrigin.y, P.cellSize.y);
  const int gz = cell_coord(pos.z, P.worldOrigin.z, P.cellSize.z);
  const float r = poly ? p4.w : P.particleRadius;
  if (poly && !(r <= P.particleRadius)) atomicOr(devFlags, DEVERR_RADIUS);
  v3f force = make_v3(0.0f, 0.0f, 0.0f);
#pragma unroll 1
  for (int dz = -1; dz <= 1; ++dz) {
    uint32_t st[9], en[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      const uint32_t h = grid_hash(gx + (c % 3) - 1, gy + (c / 3) - 1, gz + dz, P.gridSize);
      st[c] = cellStart[h];
      en[c] = cellEnd[h];
    }
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      if (st[c] != CELL_EMPTY) {
        const uint32_t e = en[c] > n ? n : en[c];
        v3f cf = cell_sum(P, pos, vel, k, st[c], e, sortedPos, sortedVel, r, poly != 0);
        force = make_v3(force.x + cf.x, force.y + cf.y, force.z + cf.z);
      }
    }
  }
  const uint32_t orig = index[k];
  if (forceOut == nullptr) {
    force = add_collider(P, pos, vel, r, force);
  } else {  // evalForces: no cursor sphere, the force stored as float3
    forceOut[3 * (size_t)orig] = force.x;
    forceOut[3 * (size_t)orig + 1] = force.y;
    forceOut[3 * (size_t)orig + 2] = force.z;
  }
  newVel[orig] = make_float4(vel.x + force.x, vel.y + force.y, vel.z + force.z, 0.0f);
}

// The collide stage.  newVel[index[k]] = (vel + force, 0), and with posOut
// also posOut[index[k]] = sortedPos[k] (the last step of a multi-step run
// returns the positions to id order).
// With kFuse the kernel also performs the next update's first stage for the
// same particle, which the update order would run next: integrate (from the
// collided velocity and the unchanged position), the cell hash and the
// radix sort's digit histogram of the hashes.  The integrated state is
// written at the sorted slot k (pos[k], newVel[k], coalesced), the hash at
// the particle's id (keysOut[id], so that the stable sort still orders ties
// by id), slotOf[id] = k (the value the sort carries) and idOfSlot[k] = id
// (for the next reorder).  Every value is the one the separate stages would
// compute.
// kPoly takes each sphere's radius from pos.w (a radius above
// particleRadius raises DEVERR_RADIUS: the neighbour search and the contact
// pre-filter are sized by particleRadius).
// kRows selects the row-range neighbour walk (collide_one_rows), which
// needs the sorted hashes in sortedVel.w and the fixed-point coordinates
// sortedQ with their filter bound qThreshold.
// forceOut != nullptr (not fused) is evalForces_launcher
// (particledem_um.cu:538-589): no cursor sphere, and the force also stored
// as float3 at the original index.
// slotKeys (fused): the next key is stored at the sorted slot k (coalesced)
// instead of the key at the id and the slot map, for the row resort
// (row_resort.hip), with kMovedBit when the cell changes; the per-row
// counts it needs are accumulated here (rows) instead of the radix digit
// histogram.
// At least kCollideWavesPerSimd waves per SIMD: the register budget this
// sets (96 VGPRs, a few values spilled to scratch) beats 4 waves at 104
// VGPRs; the fused launch on the settled 1M-sphere bed measured 91.0 us
// against 94.3 us.
constexpr int kCollideWavesPerSimd = 5;
template <bool kFuse, bool kPoly, bool kRows>
__global__ __launch_bounds__(kBlock, kCollideWavesPerSimd) void k_collide_v3(
    fc_sim_params P, BoundaryArg B, float contactD2, float4* __restrict__ newVel,
    const float4* __restrict__ sortedPos, const float4* __restrict__ sortedVel,
    const uint32_t* __restrict__ index, const uint32_t* __restrict__ cellStart,
    const uint32_t* __restrict__ cellEnd, uint32_t n, uint32_t xcdGroup, float4* __restrict__ pos,
    float dt, uint32_t* __restrict__ keysOut, int passes, uint32_t* gHist, uint32_t* counters,
    uint32_t* devFlags, uint32_t* __restrict__ slotOf, uint32_t* __restrict__ idOfSlot,
    v3f invCell, const uint2* __restrict__ sortedQ, uint32_t qThreshold,
    float* __restrict__ forceOut, int slotKeys, RowHists rows) {
  __shared__ uint32_t s_list[kListCap + 1][kBlock];
  __shared__ uint32_t s_hist[kFuse ? RS_RADIX : 1];
  const uint32_t k = xcd_grouped_block(blockIdx.x, gridDim.x, xcdGroup) * kBlock + threadIdx.x;
  const bool valid = k < n;
  // rows.rowHist: the next sort is a row resort, which takes per-row counts
  // instead of the radix digit histogram (the same for every block)
  const bool digitHist = kFuse && rows.rowHist == nullptr;
  if (digitHist) {
    for (int i = threadIdx.x; i < RS_RADIX; i += kBlock) s_hist[i] = 0u;
    histogram_producer_prologue(gHist, passes, counters);
  }
  uint32_t key = 0, oldKey = 0;
  if (valid) {
    const float4 p4 = sortedPos[k];
    const float rA = kPoly ? p4.w : P.particleRadius;
    if (kPoly && !(rA <= P.particleRadius)) atomicOr(devFlags, DEVERR_RADIUS);
    const float4 v4 = sortedVel[k];
    oldKey = __float_as_uint(v4.w);  // the gather keeps the sorted hash in w
    const v3f pos0 = xyz(p4), vel0 = xyz(v4);
    v3f force =
        kRows ? collide_one_rows<kPoly, kBlock>(P, invCell, contactD2, k, pos0, vel0, rA,
                                                sortedPos, sortedVel, sortedQ, qThreshold,
                                                cellStart, cellEnd, n, s_list)
              : collide_one_v3<kPoly, kBlock>(P, contactD2, k, pos0, vel0, rA, sortedPos,
                                              sortedVel, cellStart, cellEnd, n, s_list);
    const uint32_t orig = index[k];
    if (kFuse || forceOut == nullptr) {
      force = add_collider(P, pos0, vel0, rA, force);
    } else {  // evalForces: no cursor sphere, the force stored as float3
      forceOut[3 * (size_t)orig] = force.x;
      forceOut[3 * (size_t)orig + 1] = force.y;
      forceOut[3 * (size_t)orig + 2] = force.z;
    }
    const v3f nv = make_v3(vel0.x + force.x, vel0.y + force.y, vel0.z + force.z);
    if (!kFuse) {
      newVel[orig] = make_float4(nv.x, nv.y, nv.z, 0.0f);
      if (pos) pos[orig] = p4;
    } else {
      v3f p = xyz(p4), v = nv;
      integrate_with(P, B, p4.w, dt, p, v);
      pos[k] = make_float4(p.x, p.y, p.z, p4.w);
      newVel[k] = make_float4(v.x, v.y, v.z, 0.0f);
      key = hash_of_q(P, cell_quotients(P, invCell, p));
      if (slotKeys) {
        // at the slot, with whether the sphere changed cell, for the row resort
        keysOut[k] = key != oldKey ? (key | kMovedBit) : key;
      } else {
        keysOut[orig] = key;
        slotOf[orig] = k;
      }
      idOfSlot[k] = orig;
    }
  }
  if (digitHist) {
    __syncthreads();
    histogram_add_digit(s_hist, key & (RS_RADIX - 1), valid);
    __syncthreads();
    histogram_flush(s_hist, gHist, 0);
  } else if (kFuse) {
    row_histograms(rows, k, key, oldKey, valid);
  }
}

// The collide of small systems (few spheres: the per-lane walk of
// k_collide_v3 is then bound by the latency of its chain of dependent
// loads, not by throughput), with each sphere's nine neighbour rows on nine
// lanes of a wave (collide_row_cells: seven spheres per wave, lane 63
// idle).  The row-0 lane gathers the other rows' cell partials by lane
// shuffles and adds the 27 in the reference order (z, then y, then x), then
// the cursor sphere; the rest is k_collide_v3's tail (the scatter, or with
// kFuse the next update's integrate, hash and histogram), with the same
// arguments and results.
constexpr int kSplitSpheresPerWave = 7;

template <bool kFuse, bool kPoly>
__global__ __launch_bounds__(kBlock) void k_collide_split(
    fc_sim_params P, BoundaryArg B, float contactD2, float4* __restrict__ newVel,
    const float4* __restrict__ sortedPos, const float4* __restrict__ sortedVel,
    const uint32_t* __restrict__ index, const uint32_t* __restrict__ cellStart,
    const uint32_t* __restrict__ cellEnd, uint32_t n, float4* __restrict__ pos, float dt,
    uint32_t* __restrict__ keysOut, int passes, uint32_t* gHist, uint32_t* counters,
    uint32_t* devFlags, uint32_t* __restrict__ slotOf, uint32_t* __restrict__ idOfSlot,
    v3f invCell, float* __restrict__ forceOut) {
  __shared__ uint32_t s_hist[kFuse ? RS_RADIX : 1];
  const int lane = (int)(threadIdx.x & 63u);
  const uint32_t wave = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int sl = lane / 9, r = lane - 9 * (lane / 9);
  const uint32_t k = wave * kSplitSpheresPerWave + (uint32_t)sl;
  const bool valid = lane < 9 * kSplitSpheresPerWave && k < n;
  if (kFuse) {
    for (int i = threadIdx.x; i < RS_RADIX; i += kBlock) s_hist[i] = 0u;
    histogram_producer_prologue(gHist, passes, counters);
  }
  v3f cf[3];
  float4 p4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  v3f vel0 = make_v3(0.0f, 0.0f, 0.0f);
  float rA = P.particleRadius;
  if (valid) {
    p4 = sortedPos[k];
    vel0 = xyz(sortedVel[k]);
    rA = kPoly ? p4.w : P.particleRadius;
    collide_row_cells<kPoly>(P, invCell, contactD2, k, xyz(p4), vel0, rA, r, sortedPos,
                             sortedVel, cellStart, cellEnd, n, cf);
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) cf[c] = make_v3(0.0f, 0.0f, 0.0f);
  }
  // the row-0 lane of each sphere adds the 27 cell partials in order
  const int base = 9 * sl;
  v3f force = make_v3(0.0f, 0.0f, 0.0f);
#pragma unroll
  for (int rr = 0; rr < 9; ++rr) {
    const int src = base + rr < 64 ? base + rr : 63;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float x = __shfl(cf[c].x, src), y = __shfl(cf[c].y, src), z = __shfl(cf[c].z, src);
      force = make_v3(force.x + x, force.y + y, force.z + z);
    }
  }
  uint32_t key = 0;
  const bool owner = valid && r == 0;
  if (owner) {
    if (kPoly && !(rA <= P.particleRadius)) atomicOr(devFlags, DEVERR_RADIUS);
    const v3f pos0 = xyz(p4);
    const uint32_t orig = index[k];
    if (kFuse || forceOut == nullptr) {
      force = add_collider(P, pos0, vel0, rA, force);
    } else {  // evalForces: no cursor sphere, the force stored as float3
      forceOut[3 * (size_t)orig] = force.x;
      forceOut[3 * (size_t)orig + 1] = force.y;
      forceOut[3 * (size_t)orig + 2] = force.z;
    }
    const v3f nv = make_v3(vel0.x + force.x, vel0.y + force.y, vel0.z + force.z);
    if (!kFuse) {
      newVel[orig] = make_float4(nv.x, nv.y, nv.z, 0.0f);
      if (pos) pos[orig] = p4;
    } else {
      v3f p = pos0, v = nv;
      integrate_with(P, B, p4.w, dt, p, v);
      pos[k] = make_float4(p.x, p.y, p.z, p4.w);
      newVel[k] = make_float4(v.x, v.y, v.z, 0.0f);
      key = hash_of_q(P, cell_quotients(P, invCell, p));
      keysOut[orig] = key;
      slotOf[orig] = k;
      idOfSlot[k] = orig;
    }
  }
  if (kFuse) {
    __syncthreads();
    histogram_add_digit(s_hist, key & (RS_RADIX - 1), owner);
    __syncthreads();
    histogram_flush(s_hist, gHist, 0);
  }
}

}  // namespace

hipError_t launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t value, size_t count,
                           uint32_t* p2) {
  if (count == 0) return hipSuccess;
  size_t vec = count / 4 + 2;
  unsigned grid = grid_for(vec, kBlock);
  if (grid > 2048u) grid = 2048u;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid, p2 ? 2 : 1), dim3(kBlock), 0, s, p, p2, value, count);
  return hipGetLastError();
}

hipError_t launch_copy16(hipStream_t s, void* dst, const void* src, size_t bytes) {
  const size_t count = bytes / 16;
  if (count == 0) return hipSuccess;
  const size_t blocks = (count + kBlock - 1) / kBlock;
  const unsigned grid = blocks < (1u << 30) ? (unsigned)blocks : (1u << 30);
  hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(kBlock), 0, s, (uint4*)dst, (const uint4*)src,
                     count);
  return hipGetLastError();
}

hipError_t launch_integrate(hipStream_t s, const fc_sim_params& P, float4* pos, float4* vel,
                            float dt, uint32_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_integrate, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, P, pos, vel, dt,
                     n);
  return hipGetLastError();
}

hipError_t launch_rows_state(hipStream_t s, const fc_sim_params& P, const QuantFrame& frame,
                             const float4* sortedPos, const float4* sortedVel, uint32_t n,
                             float4* velOut, uint2* qOut) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_rows_state, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, P, frame,
                     sortedPos, sortedVel, n, velOut, qOut);
  return hipGetLastError();
}

hipError_t launch_calc_hash(hipStream_t s, const fc_sim_params& P, uint32_t* hash,
                            uint32_t* index, const float4* pos, uint32_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_calc_hash, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, P, hash, index,
                     pos, n);
  return hipGetLastError();
}

hipError_t launch_integrate_hash(hipStream_t s, const fc_sim_params& P, const BoundaryArg& B,
                                 float4* pos, float4* vel, float dt, uint32_t* keysOut,
                                 const uint32_t* prevSortedHash, uint32_t* cellStart,
                                 uint32_t* cellEnd, uint32_t n, SortScratch& sort, int passes,
                                 uint32_t* devFlags) {
  (void)devFlags;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_integrate_hash, dim3(grid_for(n, kBlock * kIhPer)), dim3(kBlock), 0, s, P, B,
                     pos, vel, dt, keysOut, prevSortedHash, cellStart, cellEnd, n, passes,
                     sort.histogram, sort.counters);
  return hipGetLastError();
}

hipError_t launch_reorder(hipStream_t s, uint32_t* cellStart, uint32_t* cellEnd, float4* sortedPos,
                          float4* sortedVel, const uint32_t* hash, const uint32_t* index,
                          const float4* oldPos, const float4* oldVel, uint32_t n,
                          uint32_t numCells, const uint32_t* idOfSlot, uint32_t* indexOut) {
  if (n == 0) return hipSuccess;
  if (idOfSlot && !indexOut) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_reorder, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, cellStart, cellEnd,
                     sortedPos, sortedVel, hash, index, idOfSlot, indexOut, oldPos, oldVel, n,
                     numCells);
  return hipGetLastError();
}

hipError_t launch_cell_bounds(hipStream_t s, uint32_t* cellStart, uint32_t* cellEnd,
                              const uint32_t* hash, uint32_t n, uint32_t numCells) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cell_bounds, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, cellStart,
                     cellEnd, hash, n, numCells);
  return hipGetLastError();
}

// Smallest float t with sqrtf(t) >= R: every d2 >= t has sqrtf(d2) >= R
// (sqrtf is correctly rounded and monotone), so d2 < t is a lossless
// pre-test for the contact condition sqrtf(d2) < R.
float contact_d2_threshold(float R) {
  if (!(R > 0.0f)) return 0.0f;
  if (std::isinf(R)) return INFINITY;
  float t = R * R;
  while (t > 0.0f && std::sqrt(std::nextafter(t, 0.0f)) >= R) t = std::nextafter(t, 0.0f);
  while (std::sqrt(t) < R) t = std::nextafter(t, INFINITY);
  return t;
}

// The row-range walk needs the sorted hash in sortedVel.w and the
// fixed-point coordinates with kSortedPosPad entries of padding
// (GatheredState: a system's arrays gathered by its sort, or the ones
// k_rows_state builds), 27 distinct neighbour cells (every grid dimension
// >= 4 here) and cell byte offsets below 2^32.
static bool rows_walk_ok(const fc_sim_params& P) {
  return P.gridSize.x >= 4 && P.gridSize.y >= 4 && P.gridSize.z >= 4 &&
         P.numCells <= (1u << 30);
}

// Exact reciprocal of c when c is a positive normal power of two whose
// reciprocal is normal too, else 0.
static float pow2_reciprocal(float c) {
  uint32_t b;
  std::memcpy(&b, &c, 4);
  const uint32_t ex = (b >> 23) & 0xffu;
  return ((b & 0x807fffffu) == 0u && ex >= 1u && ex <= 253u) ? 1.0f / c : 0.0f;
}

// Reciprocal cell sizes for cell_quotients: all three or none.
v3f inv_cell(const fc_sim_params& P) {
  const float ix = pow2_reciprocal(P.cellSize.x), iy = pow2_reciprocal(P.cellSize.y),
              iz = pow2_reciprocal(P.cellSize.z);
  if (ix == 0.0f || iy == 0.0f || iz == 0.0f) return v3f{0.0f, 0.0f, 0.0f};
  return v3f{ix, iy, iz};
}

// Bound on the squared fixed-point distance (in units of 1/scale) of every
// pair closer than R, for the row-range collide's candidate filter: each
// coordinate lies within 0.51 units of (p - lo) * scale after the float
// rounding of that product and rint, so a pair's difference vector is
// within 1.02 * sqrt(3) < 1.8 units of the scaled true one; the bound adds
// 2.5 units, which also covers the rounding of the reference's own distance
// (~1e-7 relative).  Every pair is a candidate when the bound leaves the
// int16 range of the coordinate differences.
uint32_t quant_threshold(float R, float scale) {
  const double b = (double)R * (double)scale + 2.5;
  if (!(b < 32000.0)) return 0xffffffffu;
  return (uint32_t)std::ceil(b * b);
}

// Systems up to this many spheres take the row-split collide (k_collide_split).
#ifndef FC_SPLIT_MAX_SPHERES
#define FC_SPLIT_MAX_SPHERES 16384
#endif
constexpr uint32_t kSplitMaxSpheres = FC_SPLIT_MAX_SPHERES;

template <bool kFuse, bool kPoly>
static void launch_v3(hipStream_t s, const GatheredState* gs, dim3 g, dim3 b,
                      const fc_sim_params& P,
                      const BoundaryArg& B, float d2, float4* newVel, const float4* sortedPos,
                      const float4* sortedVel, const uint32_t* index, const uint32_t* cellStart,
                      const uint32_t* cellEnd, uint32_t n, float4* pos, float dt,
                      uint32_t* keysOut, int passes, uint32_t* gHist, uint32_t* counters,
                      uint32_t* devFlags, uint32_t* slotOf, uint32_t* idOfSlot,
                      float* forceOut = nullptr, bool slotKeys = false,
                      const RowHists* rows = nullptr) {
  // (the split kernel reads the reference layout only: no gathered arrays)
  if (rows_walk_ok(P) && n <= kSplitMaxSpheres) {
    const uint32_t waves = (n + kSplitSpheresPerWave - 1) / kSplitSpheresPerWave;
    hipLaunchKernelGGL((k_collide_split<kFuse, kPoly>), dim3((waves + 3) / 4), b, 0, s, P, B, d2,
                       newVel, sortedPos, sortedVel, index, cellStart, cellEnd, n, pos, dt,
                       keysOut, passes, gHist, counters, devFlags, slotOf, idOfSlot, inv_cell(P),
                       forceOut);
  } else if (gs != nullptr && rows_walk_ok(P))
    hipLaunchKernelGGL((k_collide_v3<kFuse, kPoly, true>), g, b, 0, s, P, B, d2, newVel,
                       sortedPos, sortedVel, index, cellStart, cellEnd, n, kCollideXcdGroup, pos,
                       dt, keysOut, passes, gHist, counters, devFlags, slotOf, idOfSlot,
                       inv_cell(P), gs->sortedQ,
                       quant_threshold(P.particleRadius + P.particleRadius, gs->frame.scale),
                       forceOut, slotKeys ? 1 : 0, rows ? *rows : RowHists{});
  else
    hipLaunchKernelGGL((k_collide_v3<kFuse, kPoly, false>), g, b, 0, s, P, B, d2, newVel,
                       sortedPos, sortedVel, index, cellStart, cellEnd, n, kCollideXcdGroup, pos,
                       dt, keysOut, passes, gHist, counters, devFlags, slotOf, idOfSlot,
                       inv_cell(P), nullptr, 0u, forceOut, 0, RowHists{});
}

hipError_t launch_collide(hipStream_t s, const fc_sim_params& P, bool poly, uint32_t* devFlags,
                          float4* newVel, const float4* sortedPos, const float4* sortedVel,
                          const uint32_t* index, const uint32_t* cellStart,
                          const uint32_t* cellEnd, uint32_t n, float4* posOut,
                          const GatheredState* gathered, float* forceOut) {
  if (n == 0) return hipSuccess;
  dim3 g(grid_for(n, kBlock)), b(kBlock);
  const float d2 = contact_d2_threshold(P.particleRadius + P.particleRadius);
  // The contact-list kernel packs a sorted index into 27 bits and assumes
  // the three cells of a neighbour row are distinct (G.x >= 4).
  if (n >= (1u << 27) || P.gridSize.x < 4) {
    if (posOut) return hipErrorInvalidValue;  // only the multi-step run asks for positions
    hipLaunchKernelGGL(k_collide_v2, g, b, 0, s, P, newVel, sortedPos, sortedVel, index,
                       cellStart, cellEnd, n, poly ? 1 : 0, devFlags, forceOut);
    return hipGetLastError();
  }
  if (poly)
    launch_v3<false, true>(s, gathered, g, b, P, BoundaryArg{}, d2, newVel, sortedPos, sortedVel,
                           index, cellStart, cellEnd, n, posOut, 0.0f, nullptr, 0, nullptr,
                           nullptr, devFlags, nullptr, nullptr, forceOut);
  else
    launch_v3<false, false>(s, gathered, g, b, P, BoundaryArg{}, d2, newVel, sortedPos, sortedVel,
                            index, cellStart, cellEnd, n, posOut, 0.0f, nullptr, 0, nullptr,
                            nullptr, devFlags, nullptr, nullptr, forceOut);
  return hipGetLastError();
}

// Whether a run's fused collide stores the next key at the sorted slot for
// a row resort (the row-range kernel on a grid whose row tables fit; the
// row-split kernel of small systems keeps the key at the id and the slot
// map, for the radix sort).
bool collide_slot_keys(const fc_sim_params& P, uint32_t n) {
  return n > kSplitMaxSpheres && rows_walk_ok(P) && row_resort_grid_ok(P);
}

bool collide_rows_state_pays(const fc_sim_params& P, uint32_t n) {
  return rows_walk_ok(P) && n > kSplitMaxSpheres && n < (1u << 27);
}

bool collide_can_fuse(const fc_sim_params& P, uint32_t n) {
  return n > 0 && n < (1u << 27) && P.gridSize.x >= 4;
}

hipError_t launch_collide_fused(hipStream_t s, const fc_sim_params& P, const BoundaryArg& B,
                                float4* pos, float4* vel,
                                const float4* sortedPos, const float4* sortedVel,
                                const uint32_t* index, const uint32_t* cellStart,
                                const uint32_t* cellEnd, uint32_t n, float dt, uint32_t* keysOut,
                                SortScratch& sort, int passes, uint32_t* devFlags,
                                uint32_t* slotOf, uint32_t* idOfSlot,
                                const GatheredState* gathered, bool slotKeys,
                                const RowHists* rows) {
  if (!collide_can_fuse(P, n)) return hipErrorInvalidValue;
  // slot keys go with the row counts, for the row resort
  if (slotKeys != (rows != nullptr)) return hipErrorInvalidValue;
  if (slotKeys && !(gathered != nullptr && collide_slot_keys(P, n))) return hipErrorInvalidValue;
  const float d2 = contact_d2_threshold(P.particleRadius + P.particleRadius);
  const dim3 g(grid_for(n, kBlock)), b(kBlock);
  if (B.poly)
    launch_v3<true, true>(s, gathered, g, b, P, B, d2, vel, sortedPos, sortedVel, index, cellStart,
                          cellEnd, n, pos, dt, keysOut, passes, sort.histogram, sort.counters,
                          devFlags, slotOf, idOfSlot, nullptr, slotKeys, rows);
  else
    launch_v3<true, false>(s, gathered, g, b, P, B, d2, vel, sortedPos, sortedVel, index, cellStart,
                           cellEnd, n, pos, dt, keysOut, passes, sort.histogram, sort.counters,
                           devFlags, slotOf, idOfSlot, nullptr, slotKeys, rows);
  return hipGetLastError();
}

}  // namespace fc
