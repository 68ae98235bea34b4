// Distance-map boundary on gfx950: upload of a reference-layout map, map
// queries (plain and rigidly transformed), and the exact point-mesh builder.
#include "fc_internal.hpp"
#include "distmap.hpp"

#include <cfloat>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iterator>
#include <new>
#include <string>
#include <vector>

namespace fc {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(size_t n, int block) {
  size_t g = (n + (size_t)block - 1) / (size_t)block;
  return (unsigned)(g == 0 ? 1 : g);
}

// Batched queryMap over float[3] points (the reference's vector3 arrays).
// queryFBM (distancemap.cuh:195-225): 1 where the interpolated distance of
// a point inside the box is <= 0, else 0 (0 outside the box).
__global__ __launch_bounds__(kBlock) void k_distmap_query_fbm(DistMapView m,
                                                              const float* __restrict__ pts,
                                                              uint32_t n,
                                                              int32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const v3f q = make_v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  float d = 0.0f;
  v3f c = make_v3(0.0f, 0.0f, 0.0f);
  const bool inside = distmap_query(m, q, d, c, nullptr);
  out[i] = (inside && d <= 0.0f) ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void k_distmap_query(DistMapView m,
                                                          const float* __restrict__ pts,
                                                          uint32_t n, float* __restrict__ dist,
                                                          float* __restrict__ cp,
                                                          float* __restrict__ nrm) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const v3f q = make_v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  float d = 0.0f;
  v3f c = make_v3(0.0f, 0.0f, 0.0f), nn = make_v3(0.0f, 0.0f, 0.0f);
  distmap_query(m, q, d, c, &nn);
  dist[i] = d;
  if (cp) {
    cp[3 * i] = c.x;
    cp[3 * i + 1] = c.y;
    cp[3 * i + 2] = c.z;
  }
  if (nrm) {
    nrm[3 * i] = nn.x;
    nrm[3 * i + 1] = nn.y;
    nrm[3 * i + 2] = nn.z;
  }
}

// Matrix3x3<T> * Vector3<T> (math/matrix3x3.h:313-321): row major,
// accumulated from zero.
__device__ __forceinline__ v3f mat3_mul(const float* M, v3f v) {
  float r[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float s = 0.0f;
    s += M[i * 3 + 0] * v.x;
    s += M[i * 3 + 1] * v.y;
    s += M[i * 3 + 2] * v.z;
    r[i] = s;
  }
  return make_v3(r[0], r[1], r[2]);
}

// eval_distmap_kernel (difi.cu:427-474).
// eval_distmap for one vertex p of body 1 (difi.cu:427-474): to world space
// by body 1's pose, into body 0's model frame, queryMap there, then the
// world closest point, the unit normal from the vertex to it and the
// midpoint contact.  Returns false outside body 0's box (distance 1000).
__device__ __forceinline__ bool eval_vertex(const DistMapView& m, const fc_trans_info& T, v3f p,
                                            float& d, v3f& contact, v3f& normal) {
  v3f qw = mat3_mul(T.m2w1, p);
  qw = make_v3(qw.x + T.origin1.x, qw.y + T.origin1.y, qw.z + T.origin1.z);
  const v3f q = mat3_mul(T.w2m0, make_v3(qw.x - T.origin0.x, qw.y - T.origin0.y, qw.z - T.origin0.z));
  if (!(m.bmin[0] <= q.x && q.x <= m.bmax[0] && m.bmin[1] <= q.y && q.y <= m.bmax[1] &&
        m.bmin[2] <= q.z && q.z <= m.bmax[2])) {
    d = 1000.0f;
    return false;
  }
  d = 1000.0f;
  v3f cp = make_v3(0.0f, 0.0f, 0.0f);
  distmap_query(m, q, d, cp, nullptr);
  v3f c0 = mat3_mul(T.m2w0, cp);
  c0 = make_v3(c0.x + T.origin0.x, c0.y + T.origin0.y, c0.z + T.origin0.z);
  // Vector3::normalize (math/vector3.h:159-167): scale by (float)(1.0/|n|).
  const v3f nn = make_v3(c0.x - qw.x, c0.y - qw.y, c0.z - qw.z);
  const float mag = sqrtf(dot3(nn, nn));
  const float invMag = (float)(1.0 / (double)mag);
  normal = make_v3(nn.x * invMag, nn.y * invMag, nn.z * invMag);
  contact = make_v3((c0.x + qw.x) * 0.5f, (c0.y + qw.y) * 0.5f, (c0.z + qw.z) * 0.5f);
  return true;
}

__global__ __launch_bounds__(kBlock) void k_distmap_eval(DistMapView m, fc_trans_info T,
                                                         const float* __restrict__ v, uint32_t n,
                                                         float* __restrict__ cps,
                                                         float* __restrict__ normals,
                                                         float* __restrict__ distance) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float d;
  v3f c, nn;
  if (eval_vertex(m, T, make_v3(v[3 * i], v[3 * i + 1], v[3 * i + 2]), d, c, nn)) {
    normals[3 * i] = nn.x;
    normals[3 * i + 1] = nn.y;
    normals[3 * i + 2] = nn.z;
    cps[3 * i] = c.x;
    cps[3 * i + 1] = c.y;
    cps[3 * i + 2] = c.z;
  }
  distance[i] = d;
}

// ColliderMeshMesh<gpu>::collide (collidermeshmesh.cuh:89-172) keeps the
// vertices with eval distance < threshold as contacts, in vertex order.
// Pass 1 counts them per block; pass 2 writes each at its block's offset
// plus its rank in the block (ballots), so the list is in vertex order.
__global__ __launch_bounds__(kBlock) void k_mesh_contacts_count(DistMapView m, fc_trans_info T,
                                                                const float* __restrict__ v,
                                                                uint32_t n, float threshold,
                                                                uint32_t* __restrict__ blockCounts) {
  __shared__ uint32_t s_w[kBlock / 64];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  bool hit = false;
  if (i < n) {
    float d;
    v3f c, nn;
    eval_vertex(m, T, make_v3(v[3 * i], v[3 * i + 1], v[3 * i + 2]), d, c, nn);
    hit = d < threshold;
  }
  const uint64_t b = __ballot(hit);
  if ((threadIdx.x & 63u) == 0u) s_w[threadIdx.x >> 6] = (uint32_t)__popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += s_w[w];
    blockCounts[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kBlock) void k_mesh_contacts_write(
    DistMapView m, fc_trans_info T, const float* __restrict__ v, uint32_t n, float threshold,
    const uint32_t* __restrict__ blockOffsets, uint32_t capacity, float* __restrict__ cps,
    float* __restrict__ normals, uint32_t* __restrict__ ids) {
  __shared__ uint32_t s_w[kBlock / 64];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  bool hit = false;
  float d = 1000.0f;
  v3f c = make_v3(0.0f, 0.0f, 0.0f), nn = c;
  if (i < n) {
    eval_vertex(m, T, make_v3(v[3 * i], v[3 * i + 1], v[3 * i + 2]), d, c, nn);
    hit = d < threshold;
  }
  const uint64_t b = __ballot(hit);
  if (lane == 0u) s_w[wave] = (uint32_t)__popcll(b);
  __syncthreads();
  uint32_t rank = blockOffsets[blockIdx.x] + (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
  for (uint32_t w = 0; w < wave; ++w) rank += s_w[w];
  if (hit && rank < capacity) {
    cps[3 * (size_t)rank] = c.x;
    cps[3 * (size_t)rank + 1] = c.y;
    cps[3 * (size_t)rank + 2] = c.z;
    normals[3 * (size_t)rank] = nn.x;
    normals[3 * (size_t)rank + 1] = nn.y;
    normals[3 * (size_t)rank + 2] = nn.z;
    ids[rank] = i;
  }
}

// integrateSystem with the distance-map boundary on caller arrays.
__global__ __launch_bounds__(kBlock) void k_integrate_distmap(fc_sim_params P, DistMapView m,
                                                              float4* __restrict__ pos,
                                                              float4* __restrict__ vel, float dt,
                                                              uint32_t n) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 p4 = pos[i], v4 = vel[i];
  v3f p = xyz(p4), v = xyz(v4);
  integrate_distmap(P, m, P.particleRadius, dt, p, v);
  pos[i] = make_float4(p.x, p.y, p.z, p4.w);
  vel[i] = make_float4(v.x, v.y, v.z, v4.w);
}

// ---------------------------------------------------------------------------
// Builder: exact signed distance of every lattice vertex to a closed mesh.
// ---------------------------------------------------------------------------

// Squared distance from q to triangle (B, V1, V2) and the closest point, by
// the region method of D. Eberly, "Distance Between Point and Triangle in
// 3D" (Geometric Tools), in binary64.  (The reference's variant,
// distance/distancetriangle.cpp:55-190, departs from it in regions 4 and 6.)
__device__ __forceinline__ double point_triangle_dist2(const double q[3], const double* B,
                                                       const double* V1, const double* V2,
                                                       double cp[3]) {
  double E0[3], E1[3], D[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    E0[k] = V1[k] - B[k];
    E1[k] = V2[k] - B[k];
    D[k] = B[k] - q[k];
  }
  const double a = E0[0] * E0[0] + E0[1] * E0[1] + E0[2] * E0[2];
  const double b = E0[0] * E1[0] + E0[1] * E1[1] + E0[2] * E1[2];
  const double c = E1[0] * E1[0] + E1[1] * E1[1] + E1[2] * E1[2];
  const double d = E0[0] * D[0] + E0[1] * D[1] + E0[2] * D[2];
  const double e = E1[0] * D[0] + E1[1] * D[1] + E1[2] * D[2];
  const double det = a * c - b * b;
  double s = b * e - c * d, t = b * d - a * e;
  if (s + t <= det) {
    if (s < 0.0) {
      if (t < 0.0) {  // region 4
        if (d < 0.0) {
          t = 0.0;
          s = (-d >= a) ? 1.0 : -d / a;
        } else {
          s = 0.0;
          t = (e >= 0.0) ? 0.0 : ((-e >= c) ? 1.0 : -e / c);
        }
      } else {  // region 3
        s = 0.0;
        t = (e >= 0.0) ? 0.0 : ((-e >= c) ? 1.0 : -e / c);
      }
    } else if (t < 0.0) {  // region 5
      t = 0.0;
      s = (d >= 0.0) ? 0.0 : ((-d >= a) ? 1.0 : -d / a);
    } else {  // region 0
      const double inv = 1.0 / det;
      s *= inv;
      t *= inv;
    }
  } else {
    if (s < 0.0) {  // region 2
      const double tmp0 = b + d, tmp1 = c + e;
      if (tmp1 > tmp0) {
        const double numer = tmp1 - tmp0, denom = a - 2.0 * b + c;
        s = (numer >= denom) ? 1.0 : numer / denom;
        t = 1.0 - s;
      } else {
        s = 0.0;
        t = (tmp1 <= 0.0) ? 1.0 : ((e >= 0.0) ? 0.0 : -e / c);
      }
    } else if (t < 0.0) {  // region 6
      const double tmp0 = b + e, tmp1 = a + d;
      if (tmp1 > tmp0) {
        const double numer = tmp1 - tmp0, denom = a - 2.0 * b + c;
        t = (numer >= denom) ? 1.0 : numer / denom;
        s = 1.0 - t;
      } else {
        t = 0.0;
        s = (tmp1 <= 0.0) ? 1.0 : ((d >= 0.0) ? 0.0 : -d / a);
      }
    } else {  // region 1
      const double numer = (c + e) - (b + d);
      if (numer <= 0.0) {
        s = 0.0;
      } else {
        const double denom = a - 2.0 * b + c;
        s = (numer >= denom) ? 1.0 : numer / denom;
      }
      t = 1.0 - s;
    }
  }
  double r2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    cp[k] = (B[k] + s * E0[k]) + t * E1[k];
    const double diff = cp[k] - q[k];
    r2 += diff * diff;
  }
  return r2;
}

// Solid angle of a triangle seen from q (Van Oosterom & Strackee 1983).
__device__ __forceinline__ double solid_angle(const double q[3], const double* A, const double* B,
                                              const double* C) {
  double a[3], b[3], c[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    a[k] = A[k] - q[k];
    b[k] = B[k] - q[k];
    c[k] = C[k] - q[k];
  }
  const double la = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  const double lb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  const double lc = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  const double det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                     a[2] * (b[0] * c[1] - b[1] * c[0]);
  const double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  const double bc = b[0] * c[0] + b[1] * c[1] + b[2] * c[2];
  const double ca = c[0] * a[0] + c[1] * a[1] + c[2] * a[2];
  const double den = la * lb * lc + ab * lc + bc * la + ca * lb;
  return 2.0 * atan2(det, den);
}

constexpr int kTriTile = 128;

// One lane per query point (lattice vertex or free point); triangles are
// streamed through LDS in tiles that every lane of the block reads
// (broadcast reads, no bank conflicts).  Minimum over triangles in order
// (first minimum wins), sign from the generalized winding number.
template <bool kLattice>
__global__ __launch_bounds__(kBlock) void k_mesh_sdf(
    const double* __restrict__ verts, uint32_t nv, const uint32_t* __restrict__ tris, uint32_t nt,
    uint64_t npts, const double* __restrict__ pts, int cx, int cy, double bx, double by,
    double bz, double cs, float4* __restrict__ cpd, float4* __restrict__ nrm,
    double* __restrict__ sdfOut, double* __restrict__ cpOut) {
  __shared__ double s_tri[kTriTile][9];
  const uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool valid = v < npts;
  double q[3] = {0.0, 0.0, 0.0};
  if (valid) {
    if (kLattice) {
      const uint64_t ix = (uint64_t)cx + 1, islice = ix * ((uint64_t)cy + 1);
      const uint64_t iz = v / islice, rem = v - iz * islice, iy = rem / ix, ixx = rem - iy * ix;
      q[0] = bx + (double)ixx * cs;
      q[1] = by + (double)iy * cs;
      q[2] = bz + (double)iz * cs;
    } else {
      q[0] = pts[3 * v];
      q[1] = pts[3 * v + 1];
      q[2] = pts[3 * v + 2];
    }
  }
  double best = DBL_MAX, bcp[3] = {0.0, 0.0, 0.0}, w = 0.0;
  for (uint32_t t0 = 0; t0 < nt; t0 += kTriTile) {
    __syncthreads();
    for (int k = threadIdx.x; k < kTriTile * 3; k += kBlock) {
      const uint32_t t = t0 + k / 3;
      if (t < nt) {
        uint32_t vi = tris[3 * t + k % 3];
        if (vi >= nv) vi = nv - 1;  // out-of-range indices must not read past the mesh
        s_tri[k / 3][3 * (k % 3)] = verts[3 * vi];
        s_tri[k / 3][3 * (k % 3) + 1] = verts[3 * vi + 1];
        s_tri[k / 3][3 * (k % 3) + 2] = verts[3 * vi + 2];
      }
    }
    __syncthreads();
    if (!valid) continue;
    const uint32_t cnt = (nt - t0) < (uint32_t)kTriTile ? (nt - t0) : (uint32_t)kTriTile;
    for (uint32_t t = 0; t < cnt; ++t) {
      const double* T = s_tri[t];
      double c3[3];
      const double d2 = point_triangle_dist2(q, T, T + 3, T + 6, c3);
      if (d2 < best) {
        best = d2;
        bcp[0] = c3[0];
        bcp[1] = c3[1];
        bcp[2] = c3[2];
      }
      w += solid_angle(q, T, T + 3, T + 6);
    }
  }
  if (!valid) return;
  w /= 4.0 * 3.14159265358979323846;
  const double dist = sqrt(best);
  const double sd = (w > 0.5) ? -dist : dist;
  if (kLattice) {
    cpd[v] = make_float4((float)bcp[0], (float)bcp[1], (float)bcp[2], (float)sd);
    nrm[v] = make_float4((float)(q[0] - bcp[0]), (float)(q[1] - bcp[1]), (float)(q[2] - bcp[2]),
                         0.0f);
  } else {
    sdfOut[v] = sd;
    if (cpOut) {
      cpOut[3 * v] = bcp[0];
      cpOut[3 * v + 1] = bcp[1];
      cpOut[3 * v + 2] = bcp[2];
    }
  }
}

}  // namespace

hipError_t launch_integrate_distmap(hipStream_t s, const fc_sim_params& P, const DistMapView& m,
                                    float4* pos, float4* vel, float dt, uint32_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_integrate_distmap, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, P, m,
                     pos, vel, dt, n);
  return hipGetLastError();
}

hipError_t launch_mesh_sdf_points(hipStream_t s, const double* verts, uint32_t nv,
                                  const uint32_t* tris,
                                  uint32_t nt, const double* pts, uint64_t np, double* sdf,
                                  double* cp) {
  if (np == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mesh_sdf<false>, dim3(grid_for(np, kBlock)), dim3(kBlock), 0, s, verts, nv,
                     tris, nt, np, pts, 0, 0, 0.0, 0.0, 0.0, 0.0, nullptr, nullptr, sdf, cp);
  return hipGetLastError();
}

}  // namespace fc

using namespace fc;

struct fc_distmap {
  fc_context* ctx = nullptr;
  int cells[3] = {0, 0, 0};
  float cellSize = 0.0f;
  float bmin[3] = {0, 0, 0};
  float bmax[3] = {0, 0, 0};
  float* xs = nullptr;
  float* ys = nullptr;
  float* zs = nullptr;
  float4* cpd = nullptr;
  float4* nrm = nullptr;
  size_t V = 0;
  std::vector<int32_t> fbm;  // stateFBM_ per vertex (host; empty = all 0)
};

namespace fc {

namespace {
// Smallest and largest vertex distance of each block of 8^3 cells (its
// 9^3 vertices); a NaN distance makes the block's bounds NaN.
__global__ __launch_bounds__(kBlock) void k_distmap_block_bounds(DistMapView m,
                                                                 float2* __restrict__ out,
                                                                 int bz) {
  const int nb = m.bx * m.by * bz;
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= nb) return;
  const int ix = c % m.bx, iy = (c / m.bx) % m.by, iz = c / (m.bx * m.by);
  const int x0 = ix << kMapBlockShift, y0 = iy << kMapBlockShift, z0 = iz << kMapBlockShift;
  const int x1 = min(x0 + (1 << kMapBlockShift), m.cx), y1 = min(y0 + (1 << kMapBlockShift), m.cy),
            z1 = min(z0 + (1 << kMapBlockShift), m.cz);
  const int64_t sx = m.cx + 1, sxy = (int64_t)(m.cx + 1) * (m.cy + 1);
  float lo = INFINITY, hi = -INFINITY;
  bool nan = false;
  for (int z = z0; z <= z1; ++z)
    for (int y = y0; y <= y1; ++y)
      for (int x = x0; x <= x1; ++x) {
        const float d = m.cpd[z * sxy + y * sx + x].w;
        nan = nan || d != d;
        lo = fminf(lo, d);
        hi = fmaxf(hi, d);
      }
  out[c] = nan ? make_float2(NAN, NAN) : make_float2(lo, hi);
}
}  // namespace

fc_status distmap_block_bounds(const fc_distmap* m, hipStream_t s, float2** out, int* bx,
                               int* by) {
  *out = nullptr;
  *bx = *by = 0;
  // only on the uniform lattice bmin + i cellSize (see distmap_far)
  const float* coords[3] = {m->xs, m->ys, m->zs};
  for (int a = 0; a < 3; ++a) {
    std::vector<float> c((size_t)m->cells[a] + 1);
    FC_HIP_TRY(hipMemcpy(c.data(), coords[a], c.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < c.size(); ++i) {
      const double want = (double)m->bmin[a] + (double)i * (double)m->cellSize;
      if (!(std::fabs((double)c[i] - want) <= 1e-5 * (double)m->cellSize)) return FC_SUCCESS;
    }
  }
  const int nbx = (m->cells[0] + (1 << kMapBlockShift) - 1) >> kMapBlockShift;
  const int nby = (m->cells[1] + (1 << kMapBlockShift) - 1) >> kMapBlockShift;
  const int nbz = (m->cells[2] + (1 << kMapBlockShift) - 1) >> kMapBlockShift;
  const size_t nb = (size_t)nbx * nby * nbz;
  float2* d = nullptr;
  FC_HIP_TRY(hipMalloc(&d, nb * sizeof(float2)));
  DistMapView v = distmap_view(m);
  v.bx = nbx;
  v.by = nby;
  hipLaunchKernelGGL(k_distmap_block_bounds, dim3((unsigned)((nb + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, s, v, d, nbz);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(d);
    return hip_error(e, "distance-map block bounds");
  }
  *out = d;
  *bx = nbx;
  *by = nby;
  return FC_SUCCESS;
}

DistMapView distmap_view(const fc_distmap* m) {
  DistMapView v;
  v.cx = m->cells[0];
  v.cy = m->cells[1];
  v.cz = m->cells[2];
  v.cellSize = m->cellSize;
  for (int a = 0; a < 3; ++a) {
    v.bmin[a] = m->bmin[a];
    v.bmax[a] = m->bmax[a];
  }
  v.xs = m->xs;
  v.ys = m->ys;
  v.zs = m->zs;
  v.cpd = m->cpd;
  v.nrm = m->nrm;
  return v;
}

}  // namespace fc

namespace {

fc_status alloc_map(fc_context* ctx, const int32_t cells[3], fc_distmap** out) {
  for (int a = 0; a < 3; ++a)
    if (cells[a] < 1) return set_error(FC_ERROR_INVALID_VALUE, "cells must be >= 1 per axis");
  const size_t V = (size_t)(cells[0] + 1) * (cells[1] + 1) * (cells[2] + 1);
  if (V > ((size_t)1 << 33)) return set_error(FC_ERROR_CAPACITY, "distance map too large");
  fc_distmap* m = new fc_distmap();
  m->ctx = ctx;
  for (int a = 0; a < 3; ++a) m->cells[a] = cells[a];
  m->V = V;
  hipError_t e = hipMalloc(&m->xs, (cells[0] + 1) * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->ys, (cells[1] + 1) * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->zs, (cells[2] + 1) * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->cpd, V * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc(&m->nrm, V * sizeof(float4));
  if (e != hipSuccess) {
    fc_distmap_destroy(m);
    return hip_error(e, "distance map allocation");
  }
  *out = m;
  return FC_SUCCESS;
}

// Text map file of RigidBody::buildDistanceMapFromFile /
// storeDistanceMapToFile (rigidbody.cpp:1214-1376): "dim0 dim1", "cx cy cz",
// "cellSize", then (cx+1)(cy+1)(cz+1) lines each of vertex coordinates,
// closest points, normals (3 values), distances and FBM states (1 value).
// Like the reference's stream reads, each record's numbers are read across
// whitespace and line breaks and the rest of its last line is skipped.
struct MapFileReader {
  std::string text;
  size_t at = 0;
  bool ok = true;

  void skip_space() {
    while (at < text.size() && (text[at] == ' ' || text[at] == '\t' || text[at] == '\n' ||
                                text[at] == '\r' || text[at] == '\v' || text[at] == '\f'))
      ++at;
  }
  void end_line() {
    while (at < text.size() && text[at] != '\n') ++at;
    if (at < text.size()) ++at;
  }
  double real() {
    skip_space();
    if (at >= text.size()) { ok = false; return 0.0; }
    double v = 0.0;
    const char* b = text.data() + at;
    const char* e = text.data() + text.size();
    if (*b == '+') ++b;  // from_chars takes no leading plus; stream extraction does
    auto r = std::from_chars(b, e, v);
    if (r.ec != std::errc() && r.ec != std::errc::result_out_of_range) { ok = false; return 0.0; }
    at = (size_t)(r.ptr - text.data());
    return v;
  }
  long long integer() {
    skip_space();
    if (at >= text.size()) { ok = false; return 0; }
    long long v = 0;
    const char* b = text.data() + at;
    const char* e = text.data() + text.size();
    if (*b == '+') ++b;
    auto r = std::from_chars(b, e, v);
    if (r.ec != std::errc()) { ok = false; return 0; }
    at = (size_t)(r.ptr - text.data());
    return v;
  }
};

struct MapFileData {
  int32_t dim[2] = {0, 0};
  int32_t cells[3] = {0, 0, 0};
  double cellSize = 0.0;
  std::vector<double> coords, cps, normals, dist;  // 3, 3, 3, 1 per vertex
  std::vector<int32_t> fbm;
};

fc_status read_map_file(const char* path, bool headerOnly, MapFileData& d) {
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) return set_error(FC_ERROR_INVALID_VALUE, std::string("cannot open ") + path);
  MapFileReader r;
  try {
    r.text.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  } catch (const std::bad_alloc&) {
    return set_error(FC_ERROR_OUT_OF_MEMORY, "map file: text buffer");
  }
  d.dim[0] = (int32_t)r.integer();
  d.dim[1] = (int32_t)r.integer();
  r.end_line();
  for (int a = 0; a < 3; ++a) d.cells[a] = (int32_t)r.integer();
  r.end_line();
  d.cellSize = r.real();
  r.end_line();
  if (!r.ok) return set_error(FC_ERROR_INVALID_VALUE, "map file: bad header");
  for (int a = 0; a < 3; ++a)
    if (d.cells[a] < 1 || d.cells[a] > 4096)
      return set_error(FC_ERROR_INVALID_VALUE, "map file: cell counts outside [1, 4096]");
  if (headerOnly) return FC_SUCCESS;
  const size_t V = (size_t)(d.cells[0] + 1) * (d.cells[1] + 1) * (d.cells[2] + 1);
  if (V > (size_t(1) << 27)) return set_error(FC_ERROR_CAPACITY, "map file: more than 2^27 vertices");
  try {
    d.coords.resize(3 * V);
    d.cps.resize(3 * V);
    d.normals.resize(3 * V);
    d.dist.resize(V);
    d.fbm.resize(V);
  } catch (const std::bad_alloc&) {
    return set_error(FC_ERROR_OUT_OF_MEMORY, "map file: host arrays");
  }
  std::vector<double>* tri[3] = {&d.coords, &d.cps, &d.normals};
  for (auto* t : tri) {
    for (size_t v = 0; v < V && r.ok; ++v) {
      (*t)[3 * v] = r.real();
      (*t)[3 * v + 1] = r.real();
      (*t)[3 * v + 2] = r.real();
      r.end_line();
    }
  }
  for (size_t v = 0; v < V && r.ok; ++v) {
    d.dist[v] = r.real();
    r.end_line();
  }
  for (size_t v = 0; v < V && r.ok; ++v) {
    d.fbm[v] = (int32_t)r.integer();
    r.end_line();
  }
  if (!r.ok) return set_error(FC_ERROR_INVALID_VALUE, "map file: truncated or malformed data");
  return FC_SUCCESS;
}

}  // namespace

extern "C" {

fc_status fc_distmap_file_read(const char* path, int32_t cells[3], double* cellSize,
                               int32_t dim[2], double* vertexCoords, double* contactPoints,
                               double* normals, double* distance, int32_t* stateFBM) {
  if (!path || !cells || !cellSize) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  const bool headerOnly =
      !vertexCoords && !contactPoints && !normals && !distance && !stateFBM;
  MapFileData d;
  FC_TRY(read_map_file(path, headerOnly, d));
  for (int a = 0; a < 3; ++a) cells[a] = d.cells[a];
  *cellSize = d.cellSize;
  if (dim) {
    dim[0] = d.dim[0];
    dim[1] = d.dim[1];
  }
  if (headerOnly) return FC_SUCCESS;
  const size_t V = d.dist.size();
  if (vertexCoords) std::copy(d.coords.begin(), d.coords.end(), vertexCoords);
  if (contactPoints) std::copy(d.cps.begin(), d.cps.end(), contactPoints);
  if (normals) std::copy(d.normals.begin(), d.normals.end(), normals);
  if (distance) std::copy(d.dist.begin(), d.dist.end(), distance);
  if (stateFBM) std::copy(d.fbm.begin(), d.fbm.begin() + V, stateFBM);
  return FC_SUCCESS;
}

fc_status fc_distmap_load(fc_context* ctx, const char* path, const float bmin[3],
                          const float bmax[3], fc_distmap** out) {
  if (!ctx || !path || !out) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if ((bmin == nullptr) != (bmax == nullptr))
    return set_error(FC_ERROR_INVALID_VALUE, "give both box corners or neither");
  *out = nullptr;
  MapFileData d;
  FC_TRY(read_map_file(path, false, d));
  const size_t V = d.dist.size();
  // binary64 -> binary32 as the conversion constructor does
  // (DistanceMap<float,cpu>(DistanceMap<Real,cpu>*), distancemap.cpp:129-178)
  std::vector<float> c, p, nn, ds;
  try {
    c.resize(3 * V);
    p.resize(3 * V);
    nn.resize(3 * V);
    ds.resize(V);
  } catch (const std::bad_alloc&) {
    return set_error(FC_ERROR_OUT_OF_MEMORY, "map file: float arrays");
  }
  for (size_t i = 0; i < 3 * V; ++i) {
    c[i] = (float)d.coords[i];
    p[i] = (float)d.cps[i];
    nn[i] = (float)d.normals[i];
  }
  for (size_t v = 0; v < V; ++v) ds[v] = (float)d.dist[v];
  float lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = bmin ? bmin[a] : c[a];
    hi[a] = bmax ? bmax[a] : c[3 * (V - 1) + a];
  }
  fc_distmap* m = nullptr;
  FC_TRY(fc_distmap_create(ctx, d.cells, (float)d.cellSize, lo, hi, c.data(), nn.data(), p.data(),
                           ds.data(), &m));
  m->fbm = std::move(d.fbm);
  *out = m;
  return FC_SUCCESS;
}

fc_status fc_distmap_save(fc_context* ctx, const fc_distmap* m, const char* path) {
  if (!ctx || !m || !path) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  const size_t V = m->V;
  std::vector<float> c(3 * V), p(3 * V), nn(3 * V), ds(V);
  FC_TRY(fc_distmap_download(ctx, m, c.data(), nn.data(), p.data(), ds.data()));
  std::FILE* f = std::fopen(path, "w");
  if (!f) return set_error(FC_ERROR_INVALID_VALUE, std::string("cannot write ") + path);
  // %.9g: nine significant digits give every binary32 value back exactly
  // (the reference writes six).
  std::fprintf(f, "%d %d\n", m->cells[0] + 1, (m->cells[1] + 1) * (m->cells[2] + 1));
  std::fprintf(f, "%d %d %d\n", m->cells[0], m->cells[1], m->cells[2]);
  std::fprintf(f, "%.9g\n", (double)m->cellSize);
  const float* tri[3] = {c.data(), p.data(), nn.data()};
  for (const float* t : tri)
    for (size_t v = 0; v < V; ++v)
      std::fprintf(f, "%.9g %.9g %.9g\n", (double)t[3 * v], (double)t[3 * v + 1],
                   (double)t[3 * v + 2]);
  for (size_t v = 0; v < V; ++v) std::fprintf(f, "%.9g\n", (double)ds[v]);
  for (size_t v = 0; v < V; ++v) std::fprintf(f, "%d\n", m->fbm.empty() ? 0 : (int)m->fbm[v]);
  const bool bad = std::ferror(f) != 0;
  if (std::fclose(f) != 0 || bad) return set_error(FC_ERROR_INVALID_VALUE, "map file write failed");
  return FC_SUCCESS;
}

fc_status fc_distmap_get_fbm(const fc_distmap* m, int32_t* stateFBM) {
  if (!m || !stateFBM) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  for (size_t v = 0; v < m->V; ++v) stateFBM[v] = m->fbm.empty() ? 0 : m->fbm[v];
  return FC_SUCCESS;
}

fc_status fc_distmap_set_fbm(fc_distmap* m, const int32_t* stateFBM) {
  if (!m) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (!stateFBM) {
    m->fbm.clear();
    return FC_SUCCESS;
  }
  m->fbm.assign(stateFBM, stateFBM + m->V);
  return FC_SUCCESS;
}

fc_status fc_distmap_query_fbm(fc_context* ctx, const fc_distmap* m, const float* points,
                               uint32_t n, int32_t* out) {
  if (!ctx || !m) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (n == 0) return FC_SUCCESS;
  if (!points || !out) return set_error(FC_ERROR_INVALID_VALUE, "NULL array");
  hipLaunchKernelGGL(k_distmap_query_fbm, dim3(grid_for(n, kBlock)), dim3(kBlock), 0,
                     ctx->stream, distmap_view(m), points, n, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_error(e, "distmap queryFBM");
  return FC_SUCCESS;
}

fc_status fc_distmap_create(fc_context* ctx, const int32_t cells[3], float cellSize,
                            const float bmin[3], const float bmax[3], const float* vertexCoords,
                            const float* normals, const float* contactPoints,
                            const float* distance, fc_distmap** out) {
  if (!ctx || !cells || !bmin || !bmax || !vertexCoords || !normals || !contactPoints ||
      !distance || !out)
    return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  *out = nullptr;
  fc_distmap* m = nullptr;
  FC_TRY(alloc_map(ctx, cells, &m));
  const int cx = cells[0], cy = cells[1], cz = cells[2];
  const size_t ix = cx + 1, islice = ix * (cy + 1), V = m->V;
  std::vector<float> xs(cx + 1), ys(cy + 1), zs(cz + 1);
  for (int i = 0; i <= cx; ++i) xs[i] = vertexCoords[3 * i];
  for (int i = 0; i <= cy; ++i) ys[i] = vertexCoords[3 * (i * ix) + 1];
  for (int i = 0; i <= cz; ++i) zs[i] = vertexCoords[3 * (i * islice) + 2];
  std::vector<float4> cpd(V), nrm(V);
  for (size_t v = 0; v < V; ++v) {
    const size_t z = v / islice, y = (v % islice) / ix, x = v % ix;
    if (vertexCoords[3 * v] != xs[x] || vertexCoords[3 * v + 1] != ys[y] ||
        vertexCoords[3 * v + 2] != zs[z]) {
      fc_distmap_destroy(m);
      return set_error(FC_ERROR_INVALID_VALUE,
                       "vertex coordinates are not a tensor-product lattice");
    }
    cpd[v] = make_float4(contactPoints[3 * v], contactPoints[3 * v + 1],
                         contactPoints[3 * v + 2], distance[v]);
    nrm[v] = make_float4(normals[3 * v], normals[3 * v + 1], normals[3 * v + 2], 0.0f);
  }
  m->cellSize = cellSize;
  for (int a = 0; a < 3; ++a) {
    m->bmin[a] = bmin[a];
    m->bmax[a] = bmax[a];
  }
  hipStream_t st = ctx->stream;
  hipError_t e = hipMemcpyAsync(m->xs, xs.data(), xs.size() * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(m->ys, ys.data(), ys.size() * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(m->zs, zs.data(), zs.size() * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(m->cpd, cpd.data(), V * 16, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(m->nrm, nrm.data(), V * 16, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    fc_distmap_destroy(m);
    return hip_error(e, "distance map upload");
  }
  *out = m;
  return FC_SUCCESS;
}

fc_status fc_distmap_build(fc_context* ctx, const double* vertices, uint32_t numVertices,
                           const uint32_t* triangles, uint32_t numTriangles,
                           const int32_t cells[3], float cellSize, const float bmin[3],
                           fc_distmap** out) {
  if (!ctx || !vertices || !triangles || !cells || !bmin || !out)
    return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (numTriangles == 0 || numVertices == 0)
    return set_error(FC_ERROR_INVALID_VALUE, "empty mesh");
  if (!(cellSize > 0.0f)) return set_error(FC_ERROR_INVALID_VALUE, "cellSize must be positive");
  for (uint64_t i = 0; i < 3ull * numTriangles; ++i)
    if (triangles[i] >= numVertices)
      return set_error(FC_ERROR_INVALID_VALUE, "triangle index out of range");
  *out = nullptr;
  fc_distmap* m = nullptr;
  FC_TRY(alloc_map(ctx, cells, &m));
  const double cs = cellSize, b[3] = {bmin[0], bmin[1], bmin[2]};
  std::vector<float> ax[3];
  for (int a = 0; a < 3; ++a) {
    ax[a].resize(cells[a] + 1);
    for (int i = 0; i <= cells[a]; ++i) ax[a][i] = (float)(b[a] + (double)i * cs);
    m->bmin[a] = bmin[a];
    m->bmax[a] = (float)(b[a] + (double)cells[a] * cs);
  }
  m->cellSize = cellSize;
  hipStream_t st = ctx->stream;
  double* dv = nullptr;
  uint32_t* dt = nullptr;
  hipError_t e = hipMalloc(&dv, (size_t)numVertices * 3 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dt, (size_t)numTriangles * 3 * sizeof(uint32_t));
  if (e == hipSuccess)
    e = hipMemcpyAsync(dv, vertices, (size_t)numVertices * 24, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dt, triangles, (size_t)numTriangles * 12, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(m->xs, ax[0].data(), ax[0].size() * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(m->ys, ax[1].data(), ax[1].size() * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(m->zs, ax[2].data(), ax[2].size() * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mesh_sdf<true>, dim3(grid_for(m->V, kBlock)), dim3(kBlock), 0, st, dv,
                       numVertices, dt, numTriangles, (uint64_t)m->V, nullptr, cells[0], cells[1], b[0], b[1],
                       b[2], cs, m->cpd, m->nrm, nullptr, nullptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  hipFree(dv);
  hipFree(dt);
  if (e != hipSuccess) {
    fc_distmap_destroy(m);
    return hip_error(e, "distance map build");
  }
  *out = m;
  return FC_SUCCESS;
}

fc_status fc_distmap_destroy(fc_distmap* m) {
  if (!m) return FC_SUCCESS;
  // No use of m->ctx: the context may already be gone; hipFree waits for
  // the device.
  hipFree(m->xs);
  hipFree(m->ys);
  hipFree(m->zs);
  hipFree(m->cpd);
  hipFree(m->nrm);
  delete m;
  return FC_SUCCESS;
}

fc_status fc_distmap_info(const fc_distmap* m, int32_t cells[3], float* cellSize, float bmin[3],
                          float bmax[3]) {
  if (!m) return set_error(FC_ERROR_INVALID_VALUE, "map is NULL");
  for (int a = 0; a < 3; ++a) {
    if (cells) cells[a] = m->cells[a];
    if (bmin) bmin[a] = m->bmin[a];
    if (bmax) bmax[a] = m->bmax[a];
  }
  if (cellSize) *cellSize = m->cellSize;
  return FC_SUCCESS;
}

fc_status fc_distmap_download(fc_context* ctx, const fc_distmap* m, float* vertexCoords,
                              float* normals, float* contactPoints, float* distance) {
  if (!ctx || !m) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  const size_t V = m->V, ix = m->cells[0] + 1, islice = ix * (m->cells[1] + 1);
  std::vector<float4> cpd(V), nrm(V);
  std::vector<float> ax[3];
  const float* src[3] = {m->xs, m->ys, m->zs};
  hipStream_t st = ctx->stream;
  hipError_t e = hipSuccess;
  for (int a = 0; a < 3 && e == hipSuccess; ++a) {
    ax[a].resize(m->cells[a] + 1);
    e = hipMemcpyAsync(ax[a].data(), src[a], ax[a].size() * 4, hipMemcpyDeviceToHost, st);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(cpd.data(), m->cpd, V * 16, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(nrm.data(), m->nrm, V * 16, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_error(e, "distance map download");
  for (size_t v = 0; v < V; ++v) {
    const size_t z = v / islice, y = (v % islice) / ix, x = v % ix;
    if (vertexCoords) {
      vertexCoords[3 * v] = ax[0][x];
      vertexCoords[3 * v + 1] = ax[1][y];
      vertexCoords[3 * v + 2] = ax[2][z];
    }
    if (contactPoints) {
      contactPoints[3 * v] = cpd[v].x;
      contactPoints[3 * v + 1] = cpd[v].y;
      contactPoints[3 * v + 2] = cpd[v].z;
    }
    if (distance) distance[v] = cpd[v].w;
    if (normals) {
      normals[3 * v] = nrm[v].x;
      normals[3 * v + 1] = nrm[v].y;
      normals[3 * v + 2] = nrm[v].z;
    }
  }
  return FC_SUCCESS;
}

fc_status fc_distmap_query(fc_context* ctx, const fc_distmap* m, const float* points, uint32_t n,
                           float* dist, float* cp, float* normal) {
  if (!ctx || !m) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (n == 0) return FC_SUCCESS;
  if (!points || !dist) return set_error(FC_ERROR_INVALID_VALUE, "NULL array");
  hipLaunchKernelGGL(k_distmap_query, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, ctx->stream,
                     distmap_view(m), points, n, dist, cp, normal);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_error(e, "distmap query");
  return FC_SUCCESS;
}

fc_status fc_distmap_eval(fc_context* ctx, const fc_distmap* m, const float* v, uint32_t n,
                          const fc_trans_info* info, float* cps, float* normals,
                          float* distance) {
  if (!ctx || !m || !info) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (n == 0) return FC_SUCCESS;
  if (!v || !cps || !normals || !distance) return set_error(FC_ERROR_INVALID_VALUE, "NULL array");
  hipLaunchKernelGGL(k_distmap_eval, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, ctx->stream,
                     distmap_view(m), *info, v, n, cps, normals, distance);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_error(e, "eval_distmap");
  return FC_SUCCESS;
}

fc_status fc_mesh_mesh_contacts(fc_context* ctx, const fc_distmap* m, const float* v, uint32_t n,
                                const fc_trans_info* info, float threshold, float* cps,
                                float* normals, uint32_t* vertexIds, uint32_t capacity,
                                uint32_t* numContacts) {
  if (!ctx || !m || !info || !numContacts) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  *numContacts = 0;
  if (n == 0) return FC_SUCCESS;
  if (!v || (capacity > 0 && (!cps || !normals || !vertexIds)))
    return set_error(FC_ERROR_INVALID_VALUE, "NULL array");
  const uint32_t blocks = grid_for(n, kBlock);
  uint32_t* dCounts = nullptr;
  hipError_t e = hipMalloc(&dCounts, (size_t)blocks * sizeof(uint32_t));
  if (e != hipSuccess) return hip_error(e, "contact counts");
  hipStream_t st = ctx->stream;
  const DistMapView view = distmap_view(m);
  hipLaunchKernelGGL(k_mesh_contacts_count, dim3(blocks), dim3(kBlock), 0, st, view, *info, v, n,
                     threshold, dCounts);
  e = hipGetLastError();
  std::vector<uint32_t> counts(blocks);
  if (e == hipSuccess)
    e = hipMemcpyAsync(counts.data(), dCounts, (size_t)blocks * 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  uint64_t total = 0;
  for (uint32_t b = 0; b < blocks; ++b) {
    const uint32_t c = counts[b];
    counts[b] = (uint32_t)total;
    total += c;
  }
  if (e == hipSuccess && total > 0 && capacity > 0) {
    e = hipMemcpyAsync(dCounts, counts.data(), (size_t)blocks * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_mesh_contacts_write, dim3(blocks), dim3(kBlock), 0, st, view, *info, v,
                         n, threshold, dCounts, capacity, cps, normals, vertexIds);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
  }
  hipFree(dCounts);
  if (e != hipSuccess) return hip_error(e, "mesh-mesh contacts");
  *numContacts = (uint32_t)total;
  return FC_SUCCESS;
}

fc_status fc_integrate_system_distmap(fc_context* ctx, const fc_distmap* m, float* pos, float* vel,
                                      float deltaTime, uint32_t numParticles) {
  if (!ctx || !m) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (numParticles == 0) return FC_SUCCESS;
  if (!pos || !vel) return set_error(FC_ERROR_INVALID_VALUE, "NULL array");
  hipError_t e = launch_integrate_distmap(ctx->stream, ctx->params, distmap_view(m), (float4*)pos,
                                          (float4*)vel, deltaTime, numParticles);
  if (e != hipSuccess) return hip_error(e, "integrate (distance map)");
  return FC_SUCCESS;
}

fc_status fc_mesh_signed_distance(fc_context* ctx, const double* vertices, uint32_t numVertices,
                                  const uint32_t* triangles, uint32_t numTriangles,
                                  const double* points, uint32_t numPoints, double* sdf,
                                  double* closest) {
  if (!ctx || !vertices || !triangles || !points || !sdf)
    return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (numPoints == 0) return FC_SUCCESS;
  if (numTriangles == 0) return set_error(FC_ERROR_INVALID_VALUE, "empty mesh");
  if (numVertices == 0) return set_error(FC_ERROR_INVALID_VALUE, "empty mesh");
  hipError_t e = launch_mesh_sdf_points(ctx->stream, vertices, numVertices, triangles,
                                        numTriangles, points, numPoints, sdf, closest);
  if (e != hipSuccess) return hip_error(e, "mesh signed distance");
  return FC_SUCCESS;
}

}  // extern "C"
