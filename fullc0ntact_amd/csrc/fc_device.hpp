// Device-side arithmetic of the particle step, shared by the HIP kernels.
//
// Every function keeps the exact IEEE operation order of the reference
// kernels so that, compiled without FMA contraction, the GPU reproduces the
// CPU restatement bit for bit.  Vector arithmetic is spelled out per
// component on a plain struct instead of HIP's vector operators, so the
// order of every add and multiply is visible here.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include "fc_dem.h"

namespace fc {

struct v3f {
  float x, y, z;
};

__device__ __forceinline__ v3f make_v3(float x, float y, float z) { return v3f{x, y, z}; }
__device__ __forceinline__ v3f xyz(const float4& a) { return v3f{a.x, a.y, a.z}; }

// dot(a,b) = a.x*b.x + a.y*b.y + a.z*b.z, evaluated left to right
// (cutil_math.cuh:379-382).
__device__ __forceinline__ float dot3(v3f a, v3f b) {
  float t = a.x * b.x;
  t = t + a.y * b.y;
  return t + a.z * b.z;
}

// Cell coordinate along one axis: (int)floor((p - origin) / cellSize)
// (calcGridPos, particles_kernel.cu:78-85).  The float quotient is
// converted with saturation to [-2^31, 2^31-1] and NaN maps to 0, which
// makes out-of-range and non-finite positions deterministic (the reference
// leaves them undefined).
__device__ __forceinline__ int cell_coord(float p, float origin, float cell) {
  float q = floorf((p - origin) / cell);
  if (!(q == q)) return 0;
  if (q >= 2147483648.0f) return 2147483647;
  if (q < -2147483648.0f) return (int)0x80000000;
  return (int)q;
}

// Linear cell id with power-of-two wrap (calcGridHash,
// particles_kernel.cu:88-94).  The reference multiplies with __umul24; the
// operands are below 2^24 for every grid with x*y < 2^24, where the 32-bit
// product is identical.
__device__ __forceinline__ uint32_t grid_hash(int gx, int gy, int gz, const fc_uint3& G) {
  uint32_t x = (uint32_t)gx & (G.x - 1u);
  uint32_t y = (uint32_t)gy & (G.y - 1u);
  uint32_t z = (uint32_t)gz & (G.z - 1u);
  return (z * G.y) * G.x + y * G.x + x;
}

__device__ __forceinline__ uint32_t hash_of(const fc_sim_params& P, float px, float py, float pz) {
  int gx = cell_coord(px, P.worldOrigin.x, P.cellSize.x);
  int gy = cell_coord(py, P.worldOrigin.y, P.cellSize.y);
  int gz = cell_coord(pz, P.worldOrigin.z, P.cellSize.z);
  return grid_hash(gx, gy, gz, P.gridSize);
}

// cell_coord split into its quotient and the saturating floor, for kernels
// that also use the quotient.
__device__ __forceinline__ int cell_of_quotient(float q) {
  q = floorf(q);
  if (!(q == q)) return 0;
  if (q >= 2147483648.0f) return 2147483647;
  if (q < -2147483648.0f) return (int)0x80000000;
  return (int)q;
}

// (p - origin) / cellSize per axis.  invCell holds the exact reciprocals of
// the cell sizes when all three are powers of two (then x / c and x * (1/c)
// are the same real number and round identically, so the multiply is the
// division), else zeros; the choice is uniform across the launch.
__device__ __forceinline__ v3f cell_quotients(const fc_sim_params& P, v3f invCell, v3f p) {
  const v3f d = make_v3(p.x - P.worldOrigin.x, p.y - P.worldOrigin.y, p.z - P.worldOrigin.z);
  if (invCell.x != 0.0f) return make_v3(d.x * invCell.x, d.y * invCell.y, d.z * invCell.z);
  return make_v3(d.x / P.cellSize.x, d.y / P.cellSize.y, d.z / P.cellSize.z);
}

__device__ __forceinline__ uint32_t hash_of_q(const fc_sim_params& P, v3f q) {
  return grid_hash(cell_of_quotient(q.x), cell_of_quotient(q.y), cell_of_quotient(q.z),
                   P.gridSize);
}

// Pair force of sphere B on sphere A (collideSpheres,
// particles_kernel.cu:185-218):
//   d = pB - pA; dist = sqrt(d.d); R = rA + rB
//   dist < R:  n = d * (1/dist)               (float3 / float, cutil_math.cuh:345-349)
//              u = vB - vA;  ut = u - (u.n) n
//              f = (-k (R - dist)) n + c u + s ut + a d
//   else 0.
// contact_force is the touching branch, given dist and inv = 1/dist.
__device__ __forceinline__ v3f contact_force(v3f rel, float dist, float inv, float collideDist,
                                             v3f vA, v3f vB, float spring, float damping,
                                             float shear, float attraction) {
  v3f nrm = make_v3(rel.x * inv, rel.y * inv, rel.z * inv);
  v3f rv = make_v3(vB.x - vA.x, vB.y - vA.y, vB.z - vA.z);
  float rvn = dot3(rv, nrm);
  v3f tv = make_v3(rv.x - nrm.x * rvn, rv.y - nrm.y * rvn, rv.z - nrm.z * rvn);
  float s = -spring * (collideDist - dist);
  v3f f = make_v3(nrm.x * s, nrm.y * s, nrm.z * s);
  f = make_v3(f.x + rv.x * damping, f.y + rv.y * damping, f.z + rv.z * damping);
  f = make_v3(f.x + tv.x * shear, f.y + tv.y * shear, f.z + tv.z * shear);
  // With attraction == 0 the term is +-0 (rel is finite here), which can
  // only turn a -0 component into +0.  Every caller adds the result to a
  // sum that starts at +0 and therefore never holds -0, where x + (-0)
  // and x + (+0) are equal, so the term is skipped exactly.
  if (__builtin_expect(attraction != 0.0f, 0))
    f = make_v3(f.x + rel.x * attraction, f.y + rel.y * attraction, f.z + rel.z * attraction);
  return f;
}

__device__ __forceinline__ v3f collide_spheres(v3f pA, v3f pB, v3f vA, v3f vB, float rA, float rB,
                                               float spring, float damping, float shear,
                                               float attraction) {
  v3f rel = make_v3(pB.x - pA.x, pB.y - pA.y, pB.z - pA.z);
  float dist = sqrtf(dot3(rel, rel));
  float collideDist = rA + rB;
  v3f f = make_v3(0.0f, 0.0f, 0.0f);
  if (dist < collideDist)
    f = contact_force(rel, dist, 1.0f / dist, collideDist, vA, vB, spring, damping, shear,
                      attraction);
  return f;
}

// Correctly rounded square root for x in [2^-96, 2^96].  v_sqrt_f32 is
// within 1 ulp of the root, so the rounded root is s or one of its
// neighbours sDn, sUp.  x <= sDn s (the sign of one FMA residual, which is
// exact in sign) puts the root at or below the geometric mean of sDn and s;
// between that mean and the midpoint of sDn and s lies no root of a float
// (their squares differ by a quarter of ulp(s)^2, less than the spacing of
// the floats there), so the root rounds to sDn.  x > sUp s rounds up in the
// same way.  This is also the compiler's own expansion of sqrtf, without
// its rescaling of tiny and pass-through of zero and infinite arguments,
// which this range never needs.
__device__ __forceinline__ float sqrt_rn_mid(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sDn = __uint_as_float(__float_as_uint(s) - 1u);
  const float sUp = __uint_as_float(__float_as_uint(s) + 1u);
  float r = fmaf(-sDn, s, x) <= 0.0f ? sDn : s;
  r = fmaf(-sUp, s, x) > 0.0f ? sUp : r;
  return r;
}

// Correctly rounded 1 / b for b in [2^-48, 2^48] whose significand is not
// all ones: v_rcp_f32 is within 1 ulp of 1/b, and one Newton step with an
// FMA residual, y + y (1 - b y), rounds it correctly (Markstein, IBM J. Res.
// Dev. 34(1):111-119, 1990, whose one exception is the all-ones
// significand).  fc_selftest_pair_math checks sqrt_rn_mid and rcp_rn_mid
// against the correctly rounded operations on every argument of their
// ranges.
__device__ __forceinline__ float rcp_rn_mid(float b) {
  const float y = __builtin_amdgcn_rcpf(b);
  const float e = fmaf(-b, y, 1.0f);
  return fmaf(e, y, y);
}

__device__ __forceinline__ bool rcp_rn_mid_ok(float b) {
  return b >= 0x1p-48f && b <= 0x1p48f && (__float_as_uint(b) & 0x7fffffu) != 0x7fffffu;
}

// collide_spheres, bit for bit, with the square root and the reciprocal
// taken through sqrt_rn_mid / rcp_rn_mid when every active lane's argument
// is in their range.  The choice is a wave-uniform branch, so the general
// sequences cost nothing while no lane needs them.
__device__ __forceinline__ v3f collide_spheres_fast(v3f pA, v3f pB, v3f vA, v3f vB, float rA,
                                                    float rB, float spring, float damping,
                                                    float shear, float attraction) {
  v3f rel = make_v3(pB.x - pA.x, pB.y - pA.y, pB.z - pA.z);
  const float d2 = dot3(rel, rel);
  const bool mid = d2 >= 0x1p-96f && d2 <= 0x1p96f;
  float dist;
  if (__builtin_amdgcn_ballot_w64(!mid) == 0u)
    dist = sqrt_rn_mid(d2);
  else
    dist = sqrtf(d2);
  float collideDist = rA + rB;
  v3f f = make_v3(0.0f, 0.0f, 0.0f);
  if (dist < collideDist) {
    float inv;
    if (__builtin_amdgcn_ballot_w64(!rcp_rn_mid_ok(dist)) == 0u)
      inv = rcp_rn_mid(dist);
    else
      inv = 1.0f / dist;
    f = contact_force(rel, dist, inv, collideDist, vA, vB, spring, damping, shear, attraction);
  }
  return f;
}

// integrate_functor body (particles_kernel.cu:46-74): semi-implicit Euler
// and the box walls at +-1 (not derived from worldOrigin), tested in the
// order x+, x-, y+, z+, z-, y-.  r is the sphere's radius (the parameter
// radius, or pos.w for polydisperse systems).
__device__ __forceinline__ void integrate_box(const fc_sim_params& P, float r, float dt, v3f& pos,
                                              v3f& vel) {
  vel = make_v3(vel.x + P.gravity.x * dt, vel.y + P.gravity.y * dt, vel.z + P.gravity.z * dt);
  vel = make_v3(vel.x * P.globalDamping, vel.y * P.globalDamping, vel.z * P.globalDamping);
  pos = make_v3(pos.x + vel.x * dt, pos.y + vel.y * dt, pos.z + vel.z * dt);
  const float hi = 1.0f - r;
  const float lo = -1.0f + r;
  const float bd = P.boundaryDamping;
  if (pos.x > hi) { pos.x = hi; vel.x *= bd; }
  if (pos.x < lo) { pos.x = lo; vel.x *= bd; }
  if (pos.y > hi) { pos.y = hi; vel.y *= bd; }
  if (pos.z > hi) { pos.z = hi; vel.z *= bd; }
  if (pos.z < lo) { pos.z = lo; vel.z *= bd; }
  if (pos.y < lo) { pos.y = lo; vel.y *= bd; }
}

}  // namespace fc
