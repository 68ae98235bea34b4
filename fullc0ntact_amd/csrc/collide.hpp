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
// bed, runs of 16 measured 118 us per launch, runs of 4 121 us, contiguous
// chunks 136 us.
#ifndef FC_COLLIDE_XCD_GROUP
#define FC_COLLIDE_XCD_GROUP 16
#endif
constexpr uint32_t kCollideXcdGroup = FC_COLLIDE_XCD_GROUP;

// Capacity of the per-lane contact list of k_collide_v3.
constexpr int kListCap = 28;
// Candidates whose loads are issued together in k_collide_v3.
constexpr int kBatch = 4;

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
  v3f cpos = make_v3(P.colliderPos.x, P.colliderPos.y, P.colliderPos.z);
  v3f f = collide_spheres(pos, cpos, vel, make_v3(0.0f, 0.0f, 0.0f), rA, P.colliderRadius,
                          P.spring, P.damping, P.shear, 0.0f);
  force = make_v3(force.x + f.x, force.y + f.y, force.z + f.z);
  return make_v3(vel.x + force.x, vel.y + force.y, vel.z + force.z);
}

}  // namespace fc
