// Distance map in HBM and its device-side query (shared by distmap.hip and
// the particle kernels that apply the distance-map boundary).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "fc_device.hpp"

namespace fc {

// A DistanceMap<float> (grid/distancemap.h:36-143) on a tensor-product
// lattice.  Layout in HBM: one float4 per lattice vertex holding the closest
// point and the signed distance (xyz, w) so that one 16-byte load serves the
// trilinear interpolation of both; one float4 per vertex holding the normal;
// the lattice coordinates per axis (the reference stores 12 bytes of
// coordinates per vertex).  Vertex (ix,iy,iz) has index
// (iz*(cy+1) + iy)*(cx+1) + ix (vertexIndices, distancemap.cuh:281-302).
struct DistMapView {
  int cx, cy, cz;
  float cellSize;
  float bmin[3];
  float bmax[3];
  const float* xs;
  const float* ys;
  const float* zs;
  const float4* cpd;
  const float4* nrm;
  // Optional: per block of 8^3 cells, the smallest and largest distance of
  // its 9^3 vertices (distmap_far), with the blocks per x and y row.
  const float2* blockBounds = nullptr;
  int bx = 0, by = 0;
};

constexpr int kMapBlockShift = 3;  // 8 cells per block edge

__device__ __forceinline__ int dm_clamp_cell(int c, int cells) {
  // q on the upper face gives c == cells, whose +1 corner lies outside the
  // lattice; the last cell is used instead.
  return c >= cells ? cells - 1 : (c < 0 ? 0 : c);
}

// queryMap (distancemap.cuh:154-191) with trilinearInterpolateDistance and
// trilinearInterpolateCP (:227-279), keeping the reference's arithmetic:
// distance stages are evaluated in binary64 (the 1.0 - x_d literals promote
// them) with one float product per stage and rounded to float per stage; the
// closest point uses the double weights narrowed to float.  Returns false
// (distance 5.0) outside the closed bounding box.
__device__ __forceinline__ bool distmap_query(const DistMapView& m, v3f q, float& dist, v3f& cp,
                                              v3f* nrm) {
  if (!(m.bmin[0] <= q.x && q.x <= m.bmax[0] && m.bmin[1] <= q.y && q.y <= m.bmax[1] &&
        m.bmin[2] <= q.z && q.z <= m.bmax[2])) {
    dist = 5.0f;
    return false;
  }
  const float inv = (float)(1.0 / (double)m.cellSize);
  const int x = dm_clamp_cell((int)(fabsf(q.x - m.bmin[0]) * inv), m.cx);
  const int y = dm_clamp_cell((int)(fabsf(q.y - m.bmin[1]) * inv), m.cy);
  const int z = dm_clamp_cell((int)(fabsf(q.z - m.bmin[2]) * inv), m.cz);
  const int64_t ix = m.cx + 1, islice = (int64_t)(m.cx + 1) * (m.cy + 1);
  const int64_t base = (int64_t)z * islice + (int64_t)y * ix + x;
  const float4 v0 = m.cpd[base], v1 = m.cpd[base + 1];
  const float4 v3 = m.cpd[base + ix], v2 = m.cpd[base + ix + 1];
  const float4 v4 = m.cpd[base + islice], v5 = m.cpd[base + islice + 1];
  const float4 v7 = m.cpd[base + ix + islice], v6 = m.cpd[base + ix + islice + 1];
  const float xd = (q.x - m.xs[x]) / (m.xs[x + 1] - m.xs[x]);
  const float yd = (q.y - m.ys[y]) / (m.ys[y + 1] - m.ys[y]);
  const float zd = (q.z - m.zs[z]) / (m.zs[z + 1] - m.zs[z]);
  const double ox = 1.0 - (double)xd, oy = 1.0 - (double)yd, oz = 1.0 - (double)zd;
  const float c00 = (float)((double)v0.w * ox + (double)(v1.w * xd));
  const float c10 = (float)((double)v3.w * ox + (double)(v2.w * xd));
  const float c01 = (float)((double)v4.w * ox + (double)(v5.w * xd));
  const float c11 = (float)((double)v7.w * ox + (double)(v6.w * xd));
  const float c0 = (float)((double)c00 * oy + (double)(c10 * yd));
  const float c1 = (float)((double)c01 * oy + (double)(c11 * yd));
  dist = (float)((double)c0 * oz + (double)(c1 * zd));
  const float wx = (float)ox, wy = (float)oy, wz = (float)oz;
  const float p00x = v0.x * wx + v1.x * xd, p00y = v0.y * wx + v1.y * xd, p00z = v0.z * wx + v1.z * xd;
  const float p10x = v3.x * wx + v2.x * xd, p10y = v3.y * wx + v2.y * xd, p10z = v3.z * wx + v2.z * xd;
  const float p01x = v4.x * wx + v5.x * xd, p01y = v4.y * wx + v5.y * xd, p01z = v4.z * wx + v5.z * xd;
  const float p11x = v7.x * wx + v6.x * xd, p11y = v7.y * wx + v6.y * xd, p11z = v7.z * wx + v6.z * xd;
  const float p0x = p00x * wy + p10x * yd, p0y = p00y * wy + p10y * yd, p0z = p00z * wy + p10z * yd;
  const float p1x = p01x * wy + p11x * yd, p1y = p01y * wy + p11y * yd, p1z = p01z * wy + p11z * yd;
  cp = make_v3(p0x * wz + p1x * zd, p0y * wz + p1y * zd, p0z * wz + p1z * zd);
  if (nrm) {
    const float4 n0 = m.nrm[base];
    *nrm = make_v3(n0.x, n0.y, n0.z);
  }
  return true;
}

// normalize(float3) with the header's host definition rsqrtf = 1/sqrtf
// (cutil_math.cuh:51-54, 397-401).
__device__ __forceinline__ v3f normalize_ref(v3f v) {
  const float inv = 1.0f / sqrtf(dot3(v, v));
  return make_v3(v.x * inv, v.y * inv, v.z * inv);
}

// Whether the map distance at q is certainly at least r, so that queryMap
// there would make no contact (it tests dist < r).  The query takes the
// cell of q and mixes its eight corner distances with weights from q's
// place in the cell: a convex mix, never below the smallest corner, except
// that rounding can put q a few 10^-5 cells outside its cell and the mix
// then extrapolates by that much times the corners' spread (on a lattice
// whose coordinates are the uniform ones; the caller supplies blockBounds
// only for such maps).  So a block whose smallest vertex distance m
// satisfies m >= r + 10^-3 (max - min) + 10^-6 |m| is clear of contact for
// every q in it.  False outside the bounding box (the query's own 5.0
// sentinel applies there) and whenever the bounds are absent or NaN.
__device__ __forceinline__ bool distmap_far(const DistMapView& m, v3f q, float r) {
  if (m.blockBounds == nullptr) return false;
  if (!(m.bmin[0] <= q.x && q.x <= m.bmax[0] && m.bmin[1] <= q.y && q.y <= m.bmax[1] &&
        m.bmin[2] <= q.z && q.z <= m.bmax[2]))
    return false;
  const float inv = (float)(1.0 / (double)m.cellSize);
  const int x = dm_clamp_cell((int)(fabsf(q.x - m.bmin[0]) * inv), m.cx);
  const int y = dm_clamp_cell((int)(fabsf(q.y - m.bmin[1]) * inv), m.cy);
  const int z = dm_clamp_cell((int)(fabsf(q.z - m.bmin[2]) * inv), m.cz);
  const float2 b = m.blockBounds[((size_t)(z >> kMapBlockShift) * m.by + (y >> kMapBlockShift)) *
                                     m.bx +
                                 (x >> kMapBlockShift)];
  return b.x >= r + 1e-3f * (b.y - b.x) + 1e-6f * fabsf(b.x);
}

// d_integrateSystem (particledem_um.cu:604-638) with d_checkBoundaryDistMap
// and checkDistMap (boundaries.cuh:38-89): semi-implicit Euler, push-out of
// the distance-map boundary, a cylinder of radius 1 around the z axis, and
// the planes z = -1 and z = 10.  r is the sphere's radius.
__device__ __forceinline__ void integrate_distmap(const fc_sim_params& P, const DistMapView& m,
                                                  float r, float dt, v3f& pos, v3f& vel) {
  vel = make_v3(vel.x + P.gravity.x * dt, vel.y + P.gravity.y * dt, vel.z + P.gravity.z * dt);
  vel = make_v3(vel.x * P.globalDamping, vel.y * P.globalDamping, vel.z * P.globalDamping);
  pos = make_v3(pos.x + vel.x * dt, pos.y + vel.y * dt, pos.z + vel.z * dt);
  const float bd = P.boundaryDamping;
  const v3f pc = make_v3(pos.x, pos.y, 0.0f);  // taken before the push-out
  const float len = sqrtf(dot3(pc, pc));
  float d = 0.0f;
  v3f cp = make_v3(0.0f, 0.0f, 0.0f);
  const bool near = !distmap_far(m, pos, r);
  if (near) distmap_query(m, pos, d, cp, nullptr);
  if (near && d < r) {
    const v3f n = normalize_ref(make_v3(pos.x - cp.x, pos.y - cp.y, pos.z - cp.z));
    const float pen = r - d;
    const float s2 = (float)(0.1 * (double)r);
    pos = make_v3((pos.x + n.x * pen) + n.x * s2, (pos.y + n.y * pen) + n.y * s2,
                  (pos.z + n.z * pen) + n.z * s2);
    vel.x *= 0.5f * bd;
    vel.y *= 0.5f * bd;
    vel.z *= 0.7f * bd;
  }
  if (len > 1.0f) {
    const float pen = len - 1.0f;
    const v3f n = normalize_ref(pc);
    const float s = pen + r;
    pos = make_v3(pos.x - n.x * s, pos.y - n.y * s, pos.z - n.z * s);
    vel.x *= bd;
    vel.y *= bd;
  }
  if (pos.z > 10.0f - r) { pos.z = 10.0f - r; vel.z *= bd; }
  if (pos.z < -1.0f + r) { pos.z = -1.0f + r; vel.z *= bd; }
}

// Boundary of a particle system: the box walls of integrate_functor, or
// the distance-map boundary when hasMap; poly takes each sphere's radius
// from pos.w instead of the parameter radius.
struct BoundaryArg {
  int hasMap;
  int poly;
  DistMapView map;
};

__device__ __forceinline__ void integrate_with(const fc_sim_params& P, const BoundaryArg& B,
                                               float w, float dt, v3f& pos, v3f& vel) {
  const float r = B.poly ? w : P.particleRadius;
  if (B.hasMap)
    integrate_distmap(P, B.map, r, dt, pos, vel);
  else
    integrate_box(P, r, dt, pos, vel);
}

}  // namespace fc
