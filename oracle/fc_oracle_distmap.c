/*
 * fc_oracle_distmap.c - CPU restatement of FullC0ntact's distance-map
 * boundary: map query, distance-map integrate, transformed batch query, and
 * the exact point-mesh distance used to build maps.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker); see fc_oracle.h.  Built with
 * -ffp-contract=off and without fast-math.
 */
#include "fc_oracle.h"

#include <math.h>
#include <float.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* DistanceMap<float,gpu>::queryMap (cuda_addon/distancemap.cuh:154-191),    */
/* vertexIndices (:281-302), trilinearInterpolateDistance (:227-252),        */
/* trilinearInterpolateCP (:254-279).                                        */
/* ------------------------------------------------------------------------ */
static inline int clamp_cell(int c, int cells) {
  /* q on the upper face gives c == cells, whose +1 corner lies outside the
   * lattice; the last cell is used instead. */
  if (c >= cells) c = cells - 1;
  if (c < 0) c = 0;
  return c;
}

int orc_distmap_query(const orc_distmap* m, const float q[3], float* dist, float cp[3],
                      float nrm[3]) {
  /* AABB3::isPointInside (shapes/aabb3.h:173-181): closed box. */
  if (!(m->bmin[0] <= q[0] && q[0] <= m->bmax[0] && m->bmin[1] <= q[1] && q[1] <= m->bmax[1] &&
        m->bmin[2] <= q[2] && q[2] <= m->bmax[2])) {
    *dist = 5.0f;
    return 0;
  }
  /* float invCellSize = 1.0/cellSize_ (double quotient rounded to float). */
  const float inv = (float)(1.0 / (double)m->cellSize);
  const int cx = m->cells[0], cy = m->cells[1], cz = m->cells[2];
  const int x = clamp_cell((int)(fabsf(q[0] - m->bmin[0]) * inv), cx);
  const int y = clamp_cell((int)(fabsf(q[1] - m->bmin[1]) * inv), cy);
  const int z = clamp_cell((int)(fabsf(q[2] - m->bmin[2]) * inv), cz);
  const long ix = cx + 1, islice = (long)(cx + 1) * (cy + 1);
  const long base = (long)z * islice + (long)y * ix + x;
  const long i0 = base, i1 = base + 1, i2 = base + ix + 1, i3 = base + ix;
  const long i4 = base + islice, i5 = base + islice + 1, i6 = base + ix + islice + 1,
             i7 = base + ix + islice;
  /* Corner coordinates: vertexCoords_[i0].x, [i1].x, [i2].y, [i0].y, [i4].z,
   * [i0].z of a tensor-product lattice. */
  const float xd = (q[0] - m->xs[x]) / (m->xs[x + 1] - m->xs[x]);
  const float yd = (q[1] - m->ys[y]) / (m->ys[y + 1] - m->ys[y]);
  const float zd = (q[2] - m->zs[z]) / (m->zs[z + 1] - m->zs[z]);
  const float* D = m->dist;
  /* T c00 = distance_[i0] * (1.0 - x_d) + distance_[i1] * x_d: the first
   * product is double, the second float, the sum double, the result float. */
  const float c00 = (float)((double)D[i0] * (1.0 - (double)xd) + (double)(D[i1] * xd));
  const float c10 = (float)((double)D[i3] * (1.0 - (double)xd) + (double)(D[i2] * xd));
  const float c01 = (float)((double)D[i4] * (1.0 - (double)xd) + (double)(D[i5] * xd));
  const float c11 = (float)((double)D[i7] * (1.0 - (double)xd) + (double)(D[i6] * xd));
  const float c0 = (float)((double)c00 * (1.0 - (double)yd) + (double)(c10 * yd));
  const float c1 = (float)((double)c01 * (1.0 - (double)yd) + (double)(c11 * yd));
  *dist = (float)((double)c0 * (1.0 - (double)zd) + (double)(c1 * zd));
  /* Vector3<float>::operator*(T) narrows the double weight (1.0 - x_d) to
   * float; everything else is float. */
  const float wx = (float)(1.0 - (double)xd), wy = (float)(1.0 - (double)yd),
              wz = (float)(1.0 - (double)zd);
  const float* C = m->cp;
  for (int a = 0; a < 3; ++a) {
    const float p00 = C[3 * i0 + a] * wx + C[3 * i1 + a] * xd;
    const float p10 = C[3 * i3 + a] * wx + C[3 * i2 + a] * xd;
    const float p01 = C[3 * i4 + a] * wx + C[3 * i5 + a] * xd;
    const float p11 = C[3 * i7 + a] * wx + C[3 * i6 + a] * xd;
    const float p0 = p00 * wy + p10 * yd;
    const float p1 = p01 * wy + p11 * yd;
    cp[a] = p0 * wz + p1 * zd;
    if (nrm) nrm[a] = m->nrm[3 * i0 + a];
  }
  return 1;
}

void orc_distmap_query_batch(const orc_distmap* m, const float* q, uint32_t n, float* dist,
                             float* cp, float* nrm) {
  long i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < (long)n; ++i) {
    float c[3] = {0, 0, 0}, nn[3] = {0, 0, 0};
    orc_distmap_query(m, q + 3 * i, dist + i, c, nn);
    memcpy(cp + 3 * i, c, sizeof c);
    memcpy(nrm + 3 * i, nn, sizeof nn);
  }
}

/* normalize(float3) with the header's rsqrtf(x) = 1/sqrtf(x)
 * (cutil_math.cuh:51-54, 397-401). */
static inline void normalize3(const float v[3], float out[3]) {
  const float inv = 1.0f / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  out[0] = v[0] * inv;
  out[1] = v[1] * inv;
  out[2] = v[2] * inv;
}

/* d_integrateSystem + d_checkBoundaryDistMap + checkDistMap
 * (particledem_um.cu:604-638, boundaries.cuh:38-89): integrate, push out of
 * the distance-map boundary, a cylinder of radius 1 around z, and the planes
 * z = -1 and z = 10.  vel.w is preserved (the reference stores pos.w there,
 * particledem_um.cu:636). */
void orc_integrate_distmap(float* pos4, float* vel4, uint32_t n, float dt, const orc_params* P,
                           const orc_distmap* m) {
  orc_integrate_distmap_ex(pos4, vel4, n, dt, P, m, 0);
}

/* With poly each sphere's radius is pos.w. */
void orc_integrate_distmap_ex(float* pos4, float* vel4, uint32_t n, float dt, const orc_params* P,
                              const orc_distmap* m, int poly) {
  const float bd = P->boundaryDamping;
  long i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < (long)n; ++i) {
    float* pp = pos4 + 4 * i;
    float* vv = vel4 + 4 * i;
    const float r = poly ? pp[3] : P->particleRadius;
    float pos[3] = {pp[0], pp[1], pp[2]}, vel[3] = {vv[0], vv[1], vv[2]};
    const float g[3] = {P->gravity.x, P->gravity.y, P->gravity.z};
    for (int a = 0; a < 3; ++a) vel[a] = vel[a] + g[a] * dt;
    for (int a = 0; a < 3; ++a) vel[a] = vel[a] * P->globalDamping;
    for (int a = 0; a < 3; ++a) pos[a] = pos[a] + vel[a] * dt;
    /* d_checkBoundaryDistMap: the cylinder test uses the position before the
     * map push-out. */
    const float pc[3] = {pos[0], pos[1], 0.0f};
    const float len = sqrtf(pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]);
    /* checkDistMap */
    float d = 0.0f, cpv[3] = {0, 0, 0};
    orc_distmap_query(m, pos, &d, cpv, NULL);
    if (d < r) {
      const float n0[3] = {pos[0] - cpv[0], pos[1] - cpv[1], pos[2] - cpv[2]};
      float nn[3];
      normalize3(n0, nn);
      const float pen = r - d;
      const float s2 = (float)(0.1 * (double)r);
      for (int a = 0; a < 3; ++a) pos[a] = (pos[a] + nn[a] * pen) + nn[a] * s2;
      vel[0] *= 0.5f * bd;
      vel[1] *= 0.5f * bd;
      vel[2] *= 0.7f * bd;
    }
    if (len > 1.0f) {
      const float pen = len - 1.0f;
      float nn[3];
      normalize3(pc, nn);
      const float s = pen + r;
      for (int a = 0; a < 3; ++a) pos[a] = pos[a] - nn[a] * s;
      vel[0] *= bd;
      vel[1] *= bd;
    }
    if (pos[2] > 10.0f - r) { pos[2] = 10.0f - r; vel[2] *= bd; }
    if (pos[2] < -1.0f + r) { pos[2] = -1.0f + r; vel[2] *= bd; }
    pp[0] = pos[0]; pp[1] = pos[1]; pp[2] = pos[2];
    vv[0] = vel[0]; vv[1] = vel[1]; vv[2] = vel[2];
  }
}

/* Matrix3x3<T>::operator*(Vector3) (math/matrix3x3.h:313-321), row major,
 * accumulated from zero. */
static inline void mat3_mul(const float M[9], const float v[3], float out[3]) {
  for (int i = 0; i < 3; ++i) {
    float s = 0.0f;
    for (int j = 0; j < 3; ++j) s += M[i * 3 + j] * v[j];
    out[i] = s;
  }
}

/* eval_distmap_kernel (cuda_addon/difi.cu:427-474): vertices of body 1 into
 * the model frame of body 0, map query, closest point back to world space,
 * normal = normalize(cp_world - q_world) (Vector3::normalize,
 * math/vector3.h:159-167), contact point = midpoint.  Points outside the box
 * get distance 1000 and leave cps/normals untouched. */
void orc_distmap_eval(const orc_distmap* m, const float* v, uint32_t n, const orc_transinfo* t,
                      float* cps, float* normals, float* distance) {
  long i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < (long)n; ++i) {
    float qw[3], tmp[3], q[3];
    mat3_mul(t->m2w1, v + 3 * i, tmp);
    for (int a = 0; a < 3; ++a) qw[a] = tmp[a] + t->origin1[a];
    for (int a = 0; a < 3; ++a) tmp[a] = qw[a] - t->origin0[a];
    mat3_mul(t->w2m0, tmp, q);
    if (!(m->bmin[0] <= q[0] && q[0] <= m->bmax[0] && m->bmin[1] <= q[1] && q[1] <= m->bmax[1] &&
          m->bmin[2] <= q[2] && q[2] <= m->bmax[2])) {
      distance[i] = 1000.0f;
      continue;
    }
    float d = 1000.0f, cp[3] = {0, 0, 0};
    orc_distmap_query(m, q, &d, cp, NULL);
    float c0[3];
    mat3_mul(t->m2w0, cp, tmp);
    for (int a = 0; a < 3; ++a) c0[a] = tmp[a] + t->origin0[a];
    float nn[3] = {c0[0] - qw[0], c0[1] - qw[1], c0[2] - qw[2]};
    const float mag = sqrtf(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
    const float invMag = (float)(1.0 / (double)mag);
    for (int a = 0; a < 3; ++a) normals[3 * i + a] = nn[a] * invMag;
    for (int a = 0; a < 3; ++a) cps[3 * i + a] = (c0[a] + qw[a]) * 0.5f;
    distance[i] = d;
  }
}

/* ------------------------------------------------------------------------ */
/* Exact point-triangle distance: D. Eberly, "Distance Between Point and     */
/* Triangle in 3D", Geometric Tools (the published region method).  The      */
/* reference's own variant (distance/distancetriangle.cpp:55-190,            */
/* cuda_addon/distance.cuh:130-251) departs from it in regions 4 and 6.      */
/* ------------------------------------------------------------------------ */
static inline double dot3d(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

double orc_point_triangle_dist2(const double q[3], const double B[3], const double V1[3],
                                const double V2[3], double cp[3]) {
  double E0[3], E1[3], D[3];
  for (int a = 0; a < 3; ++a) {
    E0[a] = V1[a] - B[a];
    E1[a] = V2[a] - B[a];
    D[a] = B[a] - q[a];
  }
  const double a = dot3d(E0, E0), b = dot3d(E0, E1), c = dot3d(E1, E1);
  const double d = dot3d(E0, D), e = dot3d(E1, D);
  const double det = a * c - b * b;
  double s = b * e - c * d, t = b * d - a * e;
  if (s + t <= det) {
    if (s < 0.0) {
      if (t < 0.0) { /* region 4 */
        if (d < 0.0) {
          t = 0.0;
          s = (-d >= a) ? 1.0 : -d / a;
        } else {
          s = 0.0;
          t = (e >= 0.0) ? 0.0 : ((-e >= c) ? 1.0 : -e / c);
        }
      } else { /* region 3 */
        s = 0.0;
        t = (e >= 0.0) ? 0.0 : ((-e >= c) ? 1.0 : -e / c);
      }
    } else if (t < 0.0) { /* region 5 */
      t = 0.0;
      s = (d >= 0.0) ? 0.0 : ((-d >= a) ? 1.0 : -d / a);
    } else { /* region 0 */
      const double inv = 1.0 / det;
      s *= inv;
      t *= inv;
    }
  } else {
    if (s < 0.0) { /* region 2 */
      const double tmp0 = b + d, tmp1 = c + e;
      if (tmp1 > tmp0) {
        const double numer = tmp1 - tmp0, denom = a - 2.0 * b + c;
        s = (numer >= denom) ? 1.0 : numer / denom;
        t = 1.0 - s;
      } else {
        s = 0.0;
        t = (tmp1 <= 0.0) ? 1.0 : ((e >= 0.0) ? 0.0 : -e / c);
      }
    } else if (t < 0.0) { /* region 6 */
      const double tmp0 = b + e, tmp1 = a + d;
      if (tmp1 > tmp0) {
        const double numer = tmp1 - tmp0, denom = a - 2.0 * b + c;
        t = (numer >= denom) ? 1.0 : numer / denom;
        s = 1.0 - t;
      } else {
        t = 0.0;
        s = (tmp1 <= 0.0) ? 1.0 : ((d >= 0.0) ? 0.0 : -d / a);
      }
    } else { /* region 1 */
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
  for (int k = 0; k < 3; ++k) {
    cp[k] = (B[k] + s * E0[k]) + t * E1[k];
    const double diff = cp[k] - q[k];
    r2 += diff * diff;
  }
  return r2;
}

/* Solid angle of triangle (a, b, c) seen from q (Van Oosterom & Strackee
 * 1983), used for the generalized winding number of a closed mesh. */
static double solid_angle(const double q[3], const double A[3], const double B[3],
                          const double C[3]) {
  double a[3], b[3], c[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = A[k] - q[k];
    b[k] = B[k] - q[k];
    c[k] = C[k] - q[k];
  }
  const double la = sqrt(dot3d(a, a)), lb = sqrt(dot3d(b, b)), lc = sqrt(dot3d(c, c));
  const double det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                     a[2] * (b[0] * c[1] - b[1] * c[0]);
  const double den = la * lb * lc + dot3d(a, b) * lc + dot3d(b, c) * la + dot3d(c, a) * lb;
  return 2.0 * atan2(det, den);
}

/* Signed distance of points to a closed, consistently oriented triangle
 * mesh: minimum Eberly distance over all triangles (first minimum in
 * triangle order), negative where the winding number exceeds 1/2; closest
 * point; normal = q - cp (distancegridcgal.hpp:44-58). */
void orc_mesh_sdf(const double* verts, const uint32_t* tris, uint32_t nt, const double* pts,
                  uint32_t np, double* sdf, double* cp, double* winding) {
  long i;
#pragma omp parallel for schedule(dynamic, 64)
  for (i = 0; i < (long)np; ++i) {
    const double* q = pts + 3 * i;
    double best = DBL_MAX, bcp[3] = {0, 0, 0}, w = 0.0;
    for (uint32_t t = 0; t < nt; ++t) {
      const double* A = verts + 3 * tris[3 * t];
      const double* B = verts + 3 * tris[3 * t + 1];
      const double* C = verts + 3 * tris[3 * t + 2];
      double c3[3];
      const double d2 = orc_point_triangle_dist2(q, A, B, C, c3);
      if (d2 < best) {
        best = d2;
        bcp[0] = c3[0]; bcp[1] = c3[1]; bcp[2] = c3[2];
      }
      w += solid_angle(q, A, B, C);
    }
    w /= 4.0 * 3.14159265358979323846;
    const double d = sqrt(best);
    sdf[i] = (w > 0.5) ? -d : d;
    if (cp) { cp[3 * i] = bcp[0]; cp[3 * i + 1] = bcp[1]; cp[3 * i + 2] = bcp[2]; }
    if (winding) winding[i] = w;
  }
}

/* One update with the distance-map boundary in the integrate stage
 * (CollisionPipeline<dem_gpu_unified>, dempipeline.hpp:115-122, in the
 * ParticleSystem::update order). */
void orc_update_distmap(orc_state* s, float dt, const orc_params* P, const orc_distmap* m) {
  orc_integrate_distmap(s->pos4, s->vel4, s->n, dt, P, m);
  orc_calc_hash(s->hash, s->index, s->pos4, s->n, P);
  orc_sort_pairs(s->hash, s->index, s->n);
  orc_reorder(s->cellStart, s->cellEnd, s->sortedPos4, s->sortedVel4, s->hash, s->index,
              s->pos4, s->vel4, s->n, P->numCells);
  orc_collide(s->vel4, s->sortedPos4, s->sortedVel4, s->index, s->cellStart, s->cellEnd, s->n,
              P);
}
