// Core of the contact-list collide (one lane per sorted sphere), shared by
// the single-system kernels (dem_kernels.hip) and the batched-scene kernels
// (batch.hip), and the XCD-aware block mapping.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "fc_device.hpp"
#include "fc_internal.hpp"

namespace fc {

// Logical block index for XCD-aware scheduling.  The hardware hands
// consecutive block ids to the 8 XCDs round-robin; remapping so that each
// XCD receives runs of `group` consecutive logical blocks keeps spatially
// adjacent sorted particles (which share neighbours) in one XCD's L2.
__device__ __forceinline__ uint32_t xcd_grouped_block(uint32_t b, uint32_t nb, uint32_t group) {
  constexpr uint32_t kXcds = 8;
  if (group == 0u) {
    // Contiguous chunks: XCD x (= b % 8) takes logical blocks
    // [x*q + min(x, r), ...) in its own dispatch order, q = nb/8, r = nb%8,
    // so each XCD sweeps one slab of the sorted particles and its L2 keeps
    // the neighbouring z-planes the slab's collide re-reads.
    const uint32_t x = b % kXcds, local = b / kXcds, q = nb / kXcds, r = nb % kXcds;
    return x * q + (x < r ? x : r) + local;
  }
  const uint32_t span = kXcds * group;
  const uint32_t full = nb / span * span;  // blocks in complete spans
  if (group <= 1u || b >= full) return b;
  const uint32_t base = b / span * span, r = b - base;
  return base + (r % kXcds) * group + r / kXcds;
}

// XCD mapping of the collide's blocks (0 = contiguous chunks per XCD, else
// runs of this many consecutive blocks per XCD).  On the settled 1M-sphere
// bed, an earlier kernel measured 118 us per launch with runs of 16, 121 us
// with runs of 4 and 136 us with contiguous chunks; the row-range kernel
// measured 97.2 us with runs of 16, 97.4 / 98.4 / 97.8 us with runs of
// 8 / 32 / 64 and 103.1 us with the hardware's round-robin order.
#ifndef FC_COLLIDE_XCD_GROUP
#define FC_COLLIDE_XCD_GROUP 16
#endif
constexpr uint32_t kCollideXcdGroup = FC_COLLIDE_XCD_GROUP;

// Capacity of the per-lane contact list of k_collide_v3.
#ifndef FC_LIST_CAP
#define FC_LIST_CAP 28
#endif
constexpr int kListCap = FC_LIST_CAP;
// Candidates per iteration of collide_one_rows' phase A and contacts per
// iteration of its phase B (loads issued together).  The kernel is bound by
// the latency of its dependent neighbour loads more than by issue: on the
// settled 1M-sphere bed, batches of (2, 1) measured 136 us per launch,
// (4, 2) 122 us and (4, 4) 113 us; (8, 4) needed 97 VGPRs and gained
// nothing while phase A loaded one candidate per lane request.  Raising
// occupancy instead (list capacity 20 or 16 with 6-8 waves per SIMD)
// spilled registers and was slower.  With two candidates per 16-byte load
// (the fused kernel of a later build): (4, 4) 98.3 us, (8, 4) 95.5; with
// the cheaper squared distance of quant_d2 as well, (8, 4) 93.6-94.4,
// (12, 4) 95.0-95.1, (16, 4) 109.7 (registers spilled), (8, 2) 96.5-97.0,
// (8, 6) 93.7-93.9.
#ifndef FC_ROWS_BATCH
#define FC_ROWS_BATCH 8
#endif
#ifndef FC_PAIR_BATCH
#define FC_PAIR_BATCH 4
#endif
constexpr int kRowsBatch = FC_ROWS_BATCH;
constexpr int kPairBatch = FC_PAIR_BATCH;
static_assert(kRowsBatch >= 1 && (uint32_t)kRowsBatch - 1u <= kSortedPosPad, "row padding");
// Candidates whose loads are issued together in k_collide_v3.
constexpr int kBatch = 4;

// The cursor sphere of collideD (particles_kernel.cu:286-288): its pair
// force (at rest, no attraction) added to the 27 cells' sum.
__device__ __forceinline__ v3f add_collider(const fc_sim_params& P, v3f pos, v3f vel, float rA,
                                            v3f force) {
  const v3f cpos = make_v3(P.colliderPos.x, P.colliderPos.y, P.colliderPos.z);
  const v3f f = collide_spheres(pos, cpos, vel, make_v3(0.0f, 0.0f, 0.0f), rA, P.colliderRadius,
                                P.spring, P.damping, P.shear, 0.0f);
  return make_v3(force.x + f.x, force.y + f.y, force.z + f.z);
}

// One lane per sorted particle, two phases.
//  A: for each of the 9 neighbour rows, the candidates of its 3 cells are
//     walked as one flattened sequence; a candidate whose squared distance
//     is >= contactD2 (the smallest float whose square root reaches the
//     contact distance, so sqrtf(d2) < R implies d2 < contactD2) cannot
//     touch and is dropped.  Survivors are appended to a per-lane list in
//     LDS with their cell number (0..26).
//  B: the list is walked flat: each entry gets the reference pair force and
//     the per-cell partial is added to the total whenever the cell number
//     changes.  Dropped candidates contributed +0 to a partial that is never
//     -0, and an all-dropped cell's +0 partial to a total that is never -0,
//     so the result is bit-identical to the 27-cell loop, while the wave
//     spends the expensive pair arithmetic only on (near-)contacts.
// A lane whose list would overflow falls back to the direct 27-cell loop.
// Returns the sum of the 27 cells' partials (the caller adds the cursor
// sphere, add_collider, and the velocity).
// kPoly: radii come from pos.w (rA for this sphere); contactD2 is then the
// threshold of the largest contact distance 2 * particleRadius, which
// bounds every rA + rB.
template <bool kPoly, int BS>
__device__ __forceinline__ v3f collide_one_v3(const fc_sim_params& P, float contactD2,
                                              uint32_t k, v3f pos, v3f vel, float rA,
                                              const float4* __restrict__ sortedPos,
                                              const float4* __restrict__ sortedVel,
                                              const uint32_t* __restrict__ cellStart,
                                              const uint32_t* __restrict__ cellEnd, uint32_t n,
                                              uint32_t (*s_list)[BS]) {
  const int gx = cell_coord(pos.x, P.worldOrigin.x, P.cellSize.x);
  const int gy = cell_coord(pos.y, P.worldOrigin.y, P.cellSize.y);
  const int gz = cell_coord(pos.z, P.worldOrigin.z, P.cellSize.z);

  // ---- phase A ----
  // All 27 table entries are requested before any is used (one latency).
  uint32_t st[27], en[27];
#pragma unroll
  for (int c = 0; c < 27; ++c) {
    const uint32_t h =
        grid_hash(gx + (c % 3) - 1, gy + (c / 3) % 3 - 1, gz + c / 9 - 1, P.gridSize);
    st[c] = cellStart[h];
    en[c] = cellEnd[h];
  }
  uint32_t cnt = 0;
  bool overflow = false;
#pragma unroll
  for (int row = 0; row < 9; ++row) {
    const uint32_t s0 = st[row * 3], s1 = st[row * 3 + 1], s2 = st[row * 3 + 2];
    const uint32_t e0 = en[row * 3], e1 = en[row * 3 + 1], e2 = en[row * 3 + 2];
    const uint32_t l0 = s0 == CELL_EMPTY ? 0u : (e0 > n ? n : e0) - s0;
    const uint32_t l1 = s1 == CELL_EMPTY ? 0u : (e1 > n ? n : e1) - s1;
    const uint32_t l2 = s2 == CELL_EMPTY ? 0u : (e2 > n ? n : e2) - s2;
    const uint32_t total = l0 + l1 + l2;
    const uint32_t cellBase = (uint32_t)(row * 3);
    for (uint32_t t0 = 0; t0 < total; t0 += kBatch) {
      float4 q[kBatch];
      uint32_t jj[kBatch], cc[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {  // kBatch independent loads in flight
        const uint32_t t = t0 + u;
        uint32_t j, cell;
        if (t < l0) { j = s0 + t; cell = cellBase; }
        else if (t < l0 + l1) { j = s1 + (t - l0); cell = cellBase + 1; }
        else { j = s2 + (t - l0 - l1); cell = cellBase + 2; }
        jj[u] = j;
        cc[u] = cell;
        if (t < total) q[u] = sortedPos[j];
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        if (t0 + u < total) {
          const float dx = q[u].x - pos.x, dy = q[u].y - pos.y, dzz = q[u].z - pos.z;
          float d2 = dx * dx;
          d2 = d2 + dy * dy;
          d2 = d2 + dzz * dzz;
          if (d2 < contactD2 && jj[u] != k) {
            if (cnt < (uint32_t)kListCap) s_list[cnt][threadIdx.x] = jj[u] | (cc[u] << 27);
            else overflow = true;
            ++cnt;
          }
        }
      }
    }
  }

  v3f force = make_v3(0.0f, 0.0f, 0.0f);
  if (!overflow) {
    // ---- phase B ----
    v3f cf = make_v3(0.0f, 0.0f, 0.0f);
    uint32_t curCell = 0xffffffffu;
    for (uint32_t e0 = 0; e0 < cnt; e0 += kBatch) {
      float4 qp[kBatch], qv[kBatch];
      uint32_t ww[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        if (e0 + u < cnt) {
          ww[u] = s_list[e0 + u][threadIdx.x];
          const uint32_t j = ww[u] & 0x07ffffffu;
          qp[u] = sortedPos[j];
          qv[u] = sortedVel[j];
        }
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        if (e0 + u < cnt) {
          const uint32_t cell = ww[u] >> 27;
          if (cell != curCell) {
            force = make_v3(force.x + cf.x, force.y + cf.y, force.z + cf.z);
            cf = make_v3(0.0f, 0.0f, 0.0f);
            curCell = cell;
          }
          v3f f = collide_spheres(pos, xyz(qp[u]), vel, xyz(qv[u]), rA, kPoly ? qp[u].w : rA,
                                  P.spring, P.damping, P.shear, P.attraction);
          cf = make_v3(cf.x + f.x, cf.y + f.y, cf.z + f.z);
        }
      }
    }
    force = make_v3(force.x + cf.x, force.y + cf.y, force.z + cf.z);
  } else {
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const uint32_t h = grid_hash(gx + dx, gy + dy, gz + dz, P.gridSize);
          const uint32_t start = cellStart[h];
          if (start == CELL_EMPTY) continue;
          const uint32_t end = cellEnd[h] > n ? n : cellEnd[h];
          v3f cf = make_v3(0.0f, 0.0f, 0.0f);
          for (uint32_t j = start; j < end; ++j)
            if (j != k) {
              const float4 pj = sortedPos[j];
              v3f f = collide_spheres(pos, xyz(pj), vel, xyz(sortedVel[j]), rA,
                                      kPoly ? pj.w : rA, P.spring, P.damping, P.shear,
                                      P.attraction);
              cf = make_v3(cf.x + f.x, cf.y + f.y, cf.z + f.z);
            }
          force = make_v3(force.x + cf.x, force.y + cf.y, force.z + cf.z);
        }
  }
  return force;
}

// Load through a 32-bit byte offset from a uniform base pointer, so that the
// address is the base (scalar registers) plus one 32-bit offset register.
template <typename T>
__device__ __forceinline__ T load_off(const T* __restrict__ base, uint32_t byteOff) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byteOff);
}

// Squared distance of two packed fixed-point coordinate triples
// (quantize_pos): the x and y differences in 16-bit lanes and one packed
// dot product, added to the square of the z difference (exact: the z
// coordinates are the low halves, so the difference fits 17 bits and its
// square 32), all modulo 2^32.  Exact whenever every difference fits int16
// (the sum of three squares of at most 2^15 stays below 2^32); a larger
// difference wraps and may read short, which only keeps a far pair as a
// candidate.
__device__ __forceinline__ uint32_t quant_d2(uint2 a, uint2 b) {
  typedef short s16x2 __attribute__((ext_vector_type(2)));
  const s16x2 dxy = __builtin_bit_cast(s16x2, a.x) - __builtin_bit_cast(s16x2, b.x);
  const int dz = (int)a.y - (int)b.y;
  return (uint32_t)__builtin_amdgcn_sdot2(dxy, dxy, __mul24(dz, dz), false);
}

// Row-range variant of collide_one_v3 for grids with every dimension >= 4,
// numCells <= 2^30 and n < 2^28, on the arrays of a sort that gathered the
// state (SortGather): sortedVel[j].w holds the bits of sphere j's sorted
// hash, and sortedPos has kRowsBatch - 1 readable entries past the last
// sphere.
//
// The three cells of a neighbour row (dx = -1, 0, 1) have consecutive
// hashes unless the row wraps in x, so their particles are one contiguous
// range [lo, lo + len) of the sorted order, visited in the cell order of
// the 27-cell loop.  (A wrapping row is still one range when one of its two
// hash runs is empty; a lane with a row of two non-empty runs takes the
// direct loop.)  Cells out of contact reach are cut from the ranges first.
//
// Phase A walks the three rows of each z-plane as one flattened sequence
// and keeps the sorted indices of candidates that pass a pre-filter in a
// per-lane list in LDS.  The pre-filter reads 8 bytes per candidate, the
// 16-bit fixed-point coordinates sortedQ, and keeps every candidate whose
// exact integer squared distance to this sphere's is below qThreshold, a
// bound that every pair in contact (by the reference's rounded distance)
// stays under: rounding moves a fixed-point distance by less than the
// bound's margin and clamping into the frame only shortens it.  The extra
// near-misses get the exact contact test in phase B and add +0; so does a
// pair with a NaN sphere, whatever its coordinates read.
// Phase B walks the list and detects a change of cell by a change of the
// neighbour's sorted hash (all 27 cells are distinct when every grid
// dimension is >= 3), adding each cell's partial to the total in cell
// order.  As in collide_one_v3 every dropped or non-touching candidate
// contributed exactly +0, so the result is bit-identical to the 27-cell
// loop.  Returns the 27 cells' sum, as collide_one_v3 does.
template <bool kPoly, int BS>
__device__ __forceinline__ v3f collide_one_rows(const fc_sim_params& P, v3f invCell,
                                                float contactD2, uint32_t k, v3f pos, v3f vel,
                                                float rA, const float4* __restrict__ sortedPos,
                                                const float4* __restrict__ sortedVel,
                                                const uint2* __restrict__ sortedQ,
                                                uint32_t qThreshold,
                                                const uint32_t* __restrict__ cellStart,
                                                const uint32_t* __restrict__ cellEnd, uint32_t n,
                                                uint32_t (*s_list)[BS]) {
  const v3f q = cell_quotients(P, invCell, pos);
  const int gx = cell_of_quotient(q.x), gy = cell_of_quotient(q.y), gz = cell_of_quotient(q.z);
  const uint32_t Gx = P.gridSize.x, Gy = P.gridSize.y;
  const uint32_t mx = Gx - 1u, my = Gy - 1u, mz = P.gridSize.z - 1u;
  const uint32_t xc = (uint32_t)gx & mx;
  const uint32_t x0 = ((uint32_t)gx - 1u) & mx, x2 = ((uint32_t)gx + 1u) & mx;

  // ---- row ranges ----
  uint32_t s[27], e[27];
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const uint32_t yy = ((uint32_t)gy + (uint32_t)(r % 3) - 1u) & my;
    const uint32_t zz = ((uint32_t)gz + (uint32_t)(r / 3) - 1u) & mz;
    const uint32_t rowBase = (zz * Gy + yy) * Gx;
    const uint32_t o0 = (rowBase + x0) << 2, o1 = (rowBase + xc) << 2, o2 = (rowBase + x2) << 2;
    s[3 * r] = load_off(cellStart, o0);
    s[3 * r + 1] = load_off(cellStart, o1);
    s[3 * r + 2] = load_off(cellStart, o2);
    e[3 * r] = load_off(cellEnd, o0);
    e[3 * r + 1] = load_off(cellEnd, o1);
    e[3 * r + 2] = load_off(cellEnd, o2);
  }

  // ---- reachable cells ----
  // Lower bounds on the squared distance from this sphere to the boxes of
  // its lower and upper neighbour cells along each axis: the fraction of its
  // cell quotient below / above it, shrunk by a margin (2^-20 (|q| + 1)
  // cells) that covers the rounding of both spheres' quotients, times the
  // cell size.  A sphere hashed to a neighbour cell lies in that box or, when
  // the index wrapped, at least G - 2 >= 2 cells further along the axis, so
  // no sphere of a cell whose bound reaches the pre-filter threshold (with a
  // 2^-16 relative margin for the rounding of d2) can touch this one, and the
  // cell is cut from its row's range.  The cut cells are always at the ends
  // of a row, which stays one range.  A NaN bound keeps the cell.
  const float thrA = fmaxf(contactD2 * (1.0f + 0x1p-18f), contactD2 + 0x1p-140f);
  const float thrCut = thrA * (1.0f + 0x1p-16f);
  v3f lo2, hi2;
  {
    const float fx = q.x - floorf(q.x), fy = q.y - floorf(q.y), fz = q.z - floorf(q.z);
    const float ux = fmaf(fabsf(q.x), 0x1p-20f, 0x1p-20f);
    const float uy = fmaf(fabsf(q.y), 0x1p-20f, 0x1p-20f);
    const float uz = fmaf(fabsf(q.z), 0x1p-20f, 0x1p-20f);
    const float lx = fmaxf(fx - ux, 0.0f) * P.cellSize.x, hx = fmaxf(1.0f - fx - ux, 0.0f) * P.cellSize.x;
    const float ly = fmaxf(fy - uy, 0.0f) * P.cellSize.y, hy = fmaxf(1.0f - fy - uy, 0.0f) * P.cellSize.y;
    const float lz = fmaxf(fz - uz, 0.0f) * P.cellSize.z, hz = fmaxf(1.0f - fz - uz, 0.0f) * P.cellSize.z;
    lo2 = make_v3(lx * lx, ly * ly, lz * lz);
    hi2 = make_v3(hx * hx, hy * hy, hz * hz);
  }

  const bool wrapLo = xc == 0u, wrapHi = xc == mx;
  uint32_t lo[9], len[9];
  bool split = false;
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const int dy = r % 3 - 1, dz = r / 3 - 1;
    const float gyz = (dy < 0 ? lo2.y : (dy > 0 ? hi2.y : 0.0f)) +
                      (dz < 0 ? lo2.z : (dz > 0 ? hi2.z : 0.0f));
    const bool n0 = s[3 * r] != CELL_EMPTY && !(gyz + lo2.x >= thrCut);
    const bool n1 = s[3 * r + 1] != CELL_EMPTY && !(gyz >= thrCut);
    const bool n2 = s[3 * r + 2] != CELL_EMPTY && !(gyz + hi2.x >= thrCut);
    const uint32_t e0 = e[3 * r] > n ? n : e[3 * r];
    const uint32_t e1 = e[3 * r + 1] > n ? n : e[3 * r + 1];
    const uint32_t e2 = e[3 * r + 2] > n ? n : e[3 * r + 2];
    // An empty row starts at 0, so that the padding slots of a plane whose
    // last row is empty stay in the array.
    const uint32_t a = n0 ? s[3 * r] : (n1 ? s[3 * r + 1] : (n2 ? s[3 * r + 2] : 0u));
    const uint32_t b = n2 ? e2 : (n1 ? e1 : e0);
    lo[r] = a;
    len[r] = (n0 || n1 || n2) && b > a ? b - a : 0u;
    split = split || (wrapLo && n0 && (n1 || n2)) || (wrapHi && n2 && (n0 || n1));
  }

  // ---- phase A: candidates within the contact pre-filter ----
  // The three rows of a z-plane are walked as one flattened sequence of
  // slots in which each row's length is rounded up to even, so that slots
  // 2p and 2p + 1 are adjacent spheres of one row and one 16-byte load
  // brings both candidates' coordinates; kRowsBatch slots per iteration
  // with their loads in flight together.  Slot t of the plane is sphere
  // t + q(t), q stepping at the (even) row boundaries; the list keeps the
  // candidates' sorted indices.  The extra slot of an odd row (the sphere
  // after the row) and the slots past the plane's end are masked off; the
  // latter read at most kRowsBatch - 1 entries past the last row's range
  // (inside the array's padding).  Only the centre plane holds the sphere
  // itself.
  typedef uint32_t u32x4_a8 __attribute__((ext_vector_type(4), aligned(8)));
  static_assert(kRowsBatch % 2 == 0, "whole slot pairs per iteration");
  const uint2 me = sortedQ[k];
  uint32_t cnt = 0;
#pragma unroll
  for (int pl = 0; pl < 3; ++pl) {
    const uint32_t l0 = len[3 * pl], l1 = len[3 * pl + 1], l2 = len[3 * pl + 2];
    const uint32_t b1 = (l0 + 1u) & ~1u, b2 = b1 + ((l1 + 1u) & ~1u);
    const uint32_t tot = b2 + ((l2 + 1u) & ~1u);
    const uint32_t e1 = b1 + l1, e2 = b2 + l2;  // ends of the real slots (row 0: l0)
    const uint32_t q0 = lo[3 * pl], q1 = lo[3 * pl + 1] - b1,
                   q2 = lo[3 * pl + 2] - b2;  // modulo 2^32
#pragma unroll 1
    for (uint32_t t = 0; t < tot; t += kRowsBatch) {
      uint32_t idx[kRowsBatch / 2], end[kRowsBatch / 2];
      u32x4_a8 qq[kRowsBatch / 2];
#pragma unroll
      for (int h = 0; h < kRowsBatch / 2; ++h) {
        const uint32_t th = t + 2u * (uint32_t)h;
        const bool r0 = th < b1, r1 = th < b2;
        idx[h] = th + (r0 ? q0 : (r1 ? q1 : q2));
        end[h] = r0 ? l0 : (r1 ? e1 : e2);
        qq[h] = *reinterpret_cast<const u32x4_a8*>(reinterpret_cast<const char*>(sortedQ) +
                                                   (idx[h] << 3));
      }
#pragma unroll
      for (int u = 0; u < kRowsBatch; ++u) {
        const int h = u >> 1;
        const uint32_t j = idx[h] + (uint32_t)(u & 1);
        const uint2 c = (u & 1) ? make_uint2(qq[h].z, qq[h].w) : make_uint2(qq[h].x, qq[h].y);
        const uint32_t d2 = quant_d2(c, me);
        s_list[cnt < (uint32_t)kListCap ? cnt : (uint32_t)kListCap][threadIdx.x] = j;
        // (the first slot of the first pair is always a real one)
        const bool in = (u == 0 || t + (uint32_t)u < end[h]) && (pl != 1 || j != k);
        cnt += (d2 < qThreshold && in) ? 1u : 0u;
      }
    }
  }

  v3f force = make_v3(0.0f, 0.0f, 0.0f);
  if (!split && cnt <= (uint32_t)kListCap) {
    // ---- phase B ----
    // kPairBatch list entries per iteration (entries past cnt load the
    // sphere itself and are masked off).
    v3f cf = make_v3(0.0f, 0.0f, 0.0f);
    uint32_t curHash = 0xffffffffu;  // no cell has this hash (numCells <= 2^30)
#pragma unroll 1
    for (uint32_t i = 0; i < cnt; i += kPairBatch) {
      float4 qp[kPairBatch], qv[kPairBatch];
#pragma unroll
      for (int u = 0; u < kPairBatch; ++u) {
        const uint32_t iu = i + (uint32_t)u;
        uint32_t j = s_list[iu < (uint32_t)kListCap ? iu : (uint32_t)kListCap][threadIdx.x];
        j = (u == 0 || iu < cnt) ? j : k;  // a stale entry may lie in the padding
        const uint32_t off = j << 4;
        qp[u] = load_off(sortedPos, off);
        qv[u] = load_off(sortedVel, off);  // .w holds the sorted hash
      }
#pragma unroll
      for (int u = 0; u < kPairBatch; ++u) {
        if (u == 0 || i + (uint32_t)u < cnt) {
          const uint32_t hu = __float_as_uint(qv[u].w);
          if (hu != curHash) {
            force = make_v3(force.x + cf.x, force.y + cf.y, force.z + cf.z);
            cf = make_v3(0.0f, 0.0f, 0.0f);
            curHash = hu;
          }
          v3f f = collide_spheres_fast(pos, xyz(qp[u]), vel, xyz(qv[u]), rA,
                                       kPoly ? qp[u].w : rA, P.spring, P.damping, P.shear,
                                       P.attraction);
          cf = make_v3(cf.x + f.x, cf.y + f.y, cf.z + f.z);
        }
      }
    }
    force = make_v3(force.x + cf.x, force.y + cf.y, force.z + cf.z);
  } else if (!split) {
    // The list overflowed (a densely packed bed: up to ~60 contacts): walk
    // the nine row ranges as one flattened sequence, kPairBatch candidates
    // per iteration with their loads in flight together, each candidate
    // through the exact pair force (non-touching ones add +0), the cell
    // partial added when the sorted hash changes -- the list path's terms
    // in the list path's order.
    uint32_t B[9], Q[9];  // slots before row r; slot t of row r is sphere t + Q[r]
    uint32_t total = 0;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      B[r] = total;
      Q[r] = lo[r] - total;  // modulo 2^32
      total += len[r];
    }
    v3f cf = make_v3(0.0f, 0.0f, 0.0f);
    uint32_t curHash = 0xffffffffu;
#pragma unroll 1
    for (uint32_t t = 0; t < total; t += kPairBatch) {
      uint32_t jj[kPairBatch];
      float4 qp[kPairBatch], qv[kPairBatch];
#pragma unroll
      for (int u = 0; u < kPairBatch; ++u) {
        const uint32_t tu = t + (uint32_t)u < total ? t + (uint32_t)u : total - 1u;
        uint32_t qr = Q[0];
#pragma unroll
        for (int r = 1; r < 9; ++r) qr = tu >= B[r] ? Q[r] : qr;
        jj[u] = tu + qr;
        qp[u] = sortedPos[jj[u]];
        qv[u] = sortedVel[jj[u]];  // .w holds the sorted hash
      }
#pragma unroll
      for (int u = 0; u < kPairBatch; ++u) {
        if (t + (uint32_t)u < total && jj[u] != k) {
          const uint32_t hu = __float_as_uint(qv[u].w);
          if (hu != curHash) {
            force = make_v3(force.x + cf.x, force.y + cf.y, force.z + cf.z);
            cf = make_v3(0.0f, 0.0f, 0.0f);
            curHash = hu;
          }
          v3f f = collide_spheres_fast(pos, xyz(qp[u]), vel, xyz(qv[u]), rA,
                                       kPoly ? qp[u].w : rA, P.spring, P.damping, P.shear,
                                       P.attraction);
          cf = make_v3(cf.x + f.x, cf.y + f.y, cf.z + f.z);
        }
      }
    }
    force = make_v3(force.x + cf.x, force.y + cf.y, force.z + cf.z);
  } else {
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const uint32_t h = grid_hash(gx + dx, gy + dy, gz + dz, P.gridSize);
          const uint32_t start = cellStart[h];
          if (start == CELL_EMPTY) continue;
          const uint32_t end = cellEnd[h] > n ? n : cellEnd[h];
          v3f cf = make_v3(0.0f, 0.0f, 0.0f);
          for (uint32_t jj = start; jj < end; ++jj)
            if (jj != k) {
              const float4 pj = sortedPos[jj];
              v3f f = collide_spheres(pos, xyz(pj), vel, xyz(sortedVel[jj]), rA,
                                      kPoly ? pj.w : rA, P.spring, P.damping, P.shear,
                                      P.attraction);
              cf = make_v3(cf.x + f.x, cf.y + f.y, cf.z + f.z);
            }
          force = make_v3(force.x + cf.x, force.y + cf.y, force.z + cf.z);
        }
  }
  return force;
}

// One neighbour row of collide_one_rows for the row-split collide of small
// systems, where nine lanes share a sphere (row r = 3 (dz + 1) + (dy + 1)):
// the partial force sums of the row's three cells (dx = -1, 0, 1), each in
// index order from +0, with every candidate through the exact pair force
// (non-touching ones add +0).  Cells out of contact reach are cut as in
// collide_one_rows (their partials stay +0, which is what the reference
// adds for them); the cells are walked one by one, so a row that wraps in x
// needs no special case.  Requires every grid dimension >= 4.
template <bool kPoly>
__device__ __forceinline__ void collide_row_cells(const fc_sim_params& P, v3f invCell,
                                                  float contactD2, uint32_t k, v3f pos, v3f vel,
                                                  float rA, int r,
                                                  const float4* __restrict__ sortedPos,
                                                  const float4* __restrict__ sortedVel,
                                                  const uint32_t* __restrict__ cellStart,
                                                  const uint32_t* __restrict__ cellEnd,
                                                  uint32_t n, v3f (&cf)[3]) {
  const v3f q = cell_quotients(P, invCell, pos);
  const int gx = cell_of_quotient(q.x), gy = cell_of_quotient(q.y), gz = cell_of_quotient(q.z);
  const uint32_t Gx = P.gridSize.x, Gy = P.gridSize.y;
  const uint32_t mx = Gx - 1u, my = Gy - 1u, mz = P.gridSize.z - 1u;
  const int dy = r % 3 - 1, dz = r / 3 - 1;
  const uint32_t yy = ((uint32_t)gy + (uint32_t)dy) & my;
  const uint32_t zz = ((uint32_t)gz + (uint32_t)dz) & mz;
  const uint32_t rowBase = (zz * Gy + yy) * Gx;
  uint32_t st[3], en[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const uint32_t xx = ((uint32_t)gx + (uint32_t)(c - 1)) & mx;
    st[c] = cellStart[rowBase + xx];
    en[c] = cellEnd[rowBase + xx];
  }
  // lower bounds on the squared distance to the row's cells (collide_one_rows)
  const float thrA = fmaxf(contactD2 * (1.0f + 0x1p-18f), contactD2 + 0x1p-140f);
  const float thrCut = thrA * (1.0f + 0x1p-16f);
  const float fx = q.x - floorf(q.x), fy = q.y - floorf(q.y), fz = q.z - floorf(q.z);
  const float ux = fmaf(fabsf(q.x), 0x1p-20f, 0x1p-20f);
  const float uy = fmaf(fabsf(q.y), 0x1p-20f, 0x1p-20f);
  const float uz = fmaf(fabsf(q.z), 0x1p-20f, 0x1p-20f);
  const float lx = fmaxf(fx - ux, 0.0f) * P.cellSize.x, hx = fmaxf(1.0f - fx - ux, 0.0f) * P.cellSize.x;
  const float ly = fmaxf(fy - uy, 0.0f) * P.cellSize.y, hy = fmaxf(1.0f - fy - uy, 0.0f) * P.cellSize.y;
  const float lz = fmaxf(fz - uz, 0.0f) * P.cellSize.z, hz = fmaxf(1.0f - fz - uz, 0.0f) * P.cellSize.z;
  const float gyz = (dy < 0 ? ly * ly : (dy > 0 ? hy * hy : 0.0f)) +
                    (dz < 0 ? lz * lz : (dz > 0 ? hz * hz : 0.0f));
  const bool keep[3] = {!(gyz + lx * lx >= thrCut), !(gyz >= thrCut), !(gyz + hx * hx >= thrCut)};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    cf[c] = make_v3(0.0f, 0.0f, 0.0f);
    if (st[c] == CELL_EMPTY || !keep[c]) continue;
    const uint32_t e = en[c] > n ? n : en[c];
#pragma unroll 1
    for (uint32_t j0 = st[c]; j0 < e; j0 += 2) {
      const uint32_t j1 = j0 + 1u < e ? j0 + 1u : j0;
      const float4 p0 = sortedPos[j0], p1 = sortedPos[j1];
      const float4 v0 = sortedVel[j0], v1 = sortedVel[j1];
      if (j0 != k) {
        const v3f f = collide_spheres_fast(pos, xyz(p0), vel, xyz(v0), rA, kPoly ? p0.w : rA,
                                           P.spring, P.damping, P.shear, P.attraction);
        cf[c] = make_v3(cf[c].x + f.x, cf[c].y + f.y, cf[c].z + f.z);
      }
      if (j0 + 1u < e && j1 != k) {
        const v3f f = collide_spheres_fast(pos, xyz(p1), vel, xyz(v1), rA, kPoly ? p1.w : rA,
                                           P.spring, P.damping, P.shear, P.attraction);
        cf[c] = make_v3(cf[c].x + f.x, cf[c].y + f.y, cf[c].z + f.z);
      }
    }
  }
}

}  // namespace fc
