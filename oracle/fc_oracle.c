/*
 * fc_oracle.c - CPU restatement of FullC0ntact's particle-DEM step.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline); see
 * fc_oracle.h.  Built by oracle/Makefile with -ffp-contract=off and without
 * fast-math.  OpenMP parallelises the per-particle loops; every loop body is
 * independent per particle, so results do not depend on the thread count.
 */
#include "fc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_params_size(void) { return (int)sizeof(orc_params); }

static int g_threads = 1;

void orc_set_threads(int threads) {
  g_threads = threads < 1 ? 1 : threads;
#ifdef _OPENMP
  omp_set_num_threads(g_threads);
#endif
}

int orc_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------------ */
/* binary32 vector helpers: each operation is one IEEE op, left to right.    */
/* ------------------------------------------------------------------------ */
typedef struct { float x, y, z; } f3;

static inline f3 f3_make(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 f3_add(f3 a, f3 b) { return f3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 f3_sub(f3 a, f3 b) { return f3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
/* float3 * float and float * float3 both compute a.i * s
 * (cutil_math.cuh:327-334). */
static inline f3 f3_scale(f3 a, float s) { return f3_make(a.x * s, a.y * s, a.z * s); }
/* dot (cutil_math.cuh:379-382). */
static inline float f3_dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 f3_load(const float* p) { return f3_make(p[0], p[1], p[2]); }

/* calcGridPos (particles_kernel.cu:78-85): (int)floor((p - origin)/cell).
 * Out-of-range quotients saturate and NaN maps to 0 (a defined choice for
 * inputs the reference leaves undefined). */
static inline int grid_coord(float p, float origin, float cell) {
  float q = floorf((p - origin) / cell);
  if (isnan(q)) return 0;
  if (q >= 2147483648.0f) return 2147483647;
  if (q < -2147483648.0f) return -2147483647 - 1;
  return (int)q;
}

/* calcGridHash (particles_kernel.cu:88-94): wrap with & (G-1), then
 * (z*Gy)*Gx + y*Gx + x. */
static inline uint32_t grid_hash(int gx, int gy, int gz, const orc_params* P) {
  uint32_t x = (uint32_t)gx & (P->gridSize.x - 1u);
  uint32_t y = (uint32_t)gy & (P->gridSize.y - 1u);
  uint32_t z = (uint32_t)gz & (P->gridSize.z - 1u);
  return (z * P->gridSize.y) * P->gridSize.x + y * P->gridSize.x + x;
}

uint32_t orc_hash_point(float x, float y, float z, const orc_params* P) {
  return grid_hash(grid_coord(x, P->worldOrigin.x, P->cellSize.x),
                   grid_coord(y, P->worldOrigin.y, P->cellSize.y),
                   grid_coord(z, P->worldOrigin.z, P->cellSize.z), P);
}

/* integrate_functor (particles_kernel.cu:46-74). */
/* integrate_functor (particles_kernel.cu:39-75); with poly the wall
 * distance is each sphere's radius pos.w. */
static void integrate_impl(float* pos4, float* vel4, uint32_t n, float dt, const orc_params* P,
                           int poly) {
  const float bd = P->boundaryDamping;
  long i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < (long)n; ++i) {
    float* p = pos4 + 4 * i;
    float* v = vel4 + 4 * i;
    const float r = poly ? p[3] : P->particleRadius;
    f3 pos = f3_load(p), vel = f3_load(v);
    vel = f3_add(vel, f3_scale(f3_make(P->gravity.x, P->gravity.y, P->gravity.z), dt));
    vel = f3_scale(vel, P->globalDamping);
    pos = f3_add(pos, f3_scale(vel, dt));
    if (pos.x > 1.0f - r) { pos.x = 1.0f - r; vel.x *= bd; }
    if (pos.x < -1.0f + r) { pos.x = -1.0f + r; vel.x *= bd; }
    if (pos.y > 1.0f - r) { pos.y = 1.0f - r; vel.y *= bd; }
    if (pos.z > 1.0f - r) { pos.z = 1.0f - r; vel.z *= bd; }
    if (pos.z < -1.0f + r) { pos.z = -1.0f + r; vel.z *= bd; }
    if (pos.y < -1.0f + r) { pos.y = -1.0f + r; vel.y *= bd; }
    p[0] = pos.x; p[1] = pos.y; p[2] = pos.z;
    v[0] = vel.x; v[1] = vel.y; v[2] = vel.z;
  }
}

void orc_integrate(float* pos4, float* vel4, uint32_t n, float dt, const orc_params* P) {
  integrate_impl(pos4, vel4, n, dt, P, 0);
}

void orc_integrate_poly(float* pos4, float* vel4, uint32_t n, float dt, const orc_params* P) {
  integrate_impl(pos4, vel4, n, dt, P, 1);
}

/* calcHashD (particles_kernel.cu:97-115). */
void orc_calc_hash(uint32_t* hash, uint32_t* index, const float* pos4, uint32_t n,
                   const orc_params* P) {
  long i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < (long)n; ++i) {
    hash[i] = orc_hash_point(pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2], P);
    index[i] = (uint32_t)i;
  }
}

/* Stable sort of (key, value) by key: least-significant-digit counting sort,
 * 8 bits per pass over the bits present in the largest key.  Each pass is
 * stable, so the composition orders by key and keeps input order among
 * equal keys - the order thrust::sort_by_key's radix path produces
 * (particleSystem.cu:267).  Passes are parallelised by giving each thread a
 * contiguous chunk, counting per chunk, and scattering chunk by chunk. */
void orc_sort_pairs(uint32_t* hash, uint32_t* index, uint32_t n) {
  if (n < 2) return;
  uint32_t maxKey = 0;
  for (uint32_t i = 0; i < n; ++i) if (hash[i] > maxKey) maxKey = hash[i];
  int passes = 0;
  while (passes < 4 && (maxKey >> (8 * passes)) != 0u) ++passes;
  if (passes == 0) return;
  uint32_t* k2 = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
  uint32_t* v2 = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
  int T = g_threads;
  if ((uint32_t)T > n) T = (int)n;
  size_t* counts = (size_t*)malloc((size_t)T * 256 * sizeof(size_t));
  uint32_t *ks = hash, *vs = index, *kd = k2, *vd = v2;
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    int t;
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for (t = 0; t < T; ++t) {
      size_t lo = (size_t)n * t / T, hi = (size_t)n * (t + 1) / T;
      size_t* c = counts + (size_t)t * 256;
      memset(c, 0, 256 * sizeof(size_t));
      for (size_t i = lo; i < hi; ++i) c[(ks[i] >> shift) & 255u]++;
    }
    size_t run = 0;
    for (int d = 0; d < 256; ++d)
      for (int u = 0; u < T; ++u) {
        size_t c = counts[(size_t)u * 256 + d];
        counts[(size_t)u * 256 + d] = run;
        run += c;
      }
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for (t = 0; t < T; ++t) {
      size_t lo = (size_t)n * t / T, hi = (size_t)n * (t + 1) / T;
      size_t* c = counts + (size_t)t * 256;
      for (size_t i = lo; i < hi; ++i) {
        size_t dst = c[(ks[i] >> shift) & 255u]++;
        kd[dst] = ks[i];
        vd[dst] = vs[i];
      }
    }
    uint32_t* tk = ks; ks = kd; kd = tk;
    uint32_t* tv = vs; vs = vd; vd = tv;
  }
  if (ks != hash) {
    memcpy(hash, ks, (size_t)n * sizeof(uint32_t));
    memcpy(index, vs, (size_t)n * sizeof(uint32_t));
  }
  free(counts);
  free(k2);
  free(v2);
}

/* reorderDataAndFindCellStartD (particles_kernel.cu:119-182) after both
 * tables are set to EMPTY (particledem.cu:229-230). */
void orc_reorder(uint32_t* cellStart, uint32_t* cellEnd, float* sortedPos4, float* sortedVel4,
                 const uint32_t* hash, const uint32_t* index, const float* oldPos4,
                 const float* oldVel4, uint32_t n, uint32_t numCells) {
  long c;
#pragma omp parallel for schedule(static)
  for (c = 0; c < (long)numCells; ++c) {
    cellStart[c] = ORC_EMPTY;
    cellEnd[c] = ORC_EMPTY;
  }
  long k;
#pragma omp parallel for schedule(static)
  for (k = 0; k < (long)n; ++k) {
    uint32_t h = hash[k];
    if (k == 0 || h != hash[k - 1]) {
      cellStart[h] = (uint32_t)k;
      if (k > 0) cellEnd[hash[k - 1]] = (uint32_t)k;
    }
    if (k == (long)n - 1) cellEnd[h] = (uint32_t)n;
    uint32_t src = index[k];
    memcpy(sortedPos4 + 4 * k, oldPos4 + 4 * (size_t)src, 4 * sizeof(float));
    memcpy(sortedVel4 + 4 * k, oldVel4 + 4 * (size_t)src, 4 * sizeof(float));
  }
}

/* collideSpheres (particles_kernel.cu:185-218), binary32. */
static inline f3 pair_force(f3 posA, f3 posB, f3 velA, f3 velB, float radiusA, float radiusB,
                            float attraction, const orc_params* P) {
  f3 relPos = f3_sub(posB, posA);
  float dist = sqrtf(f3_dot(relPos, relPos));
  float collideDist = radiusA + radiusB;
  f3 force = f3_make(0.0f, 0.0f, 0.0f);
  if (dist < collideDist) {
    /* relPos / dist = relPos * (1/dist) (cutil_math.cuh:345-349). */
    f3 norm = f3_scale(relPos, 1.0f / dist);
    f3 relVel = f3_sub(velB, velA);
    f3 tanVel = f3_sub(relVel, f3_scale(norm, f3_dot(relVel, norm)));
    force = f3_scale(norm, -P->spring * (collideDist - dist));
    force = f3_add(force, f3_scale(relVel, P->damping));
    force = f3_add(force, f3_scale(tanVel, P->shear));
    force = f3_add(force, f3_scale(relPos, attraction));
  }
  return force;
}

void orc_pair_force(const float* pA, const float* pB, const float* vA, const float* vB, float rA,
                    float rB, float attraction, const orc_params* P, float* out3) {
  f3 f = pair_force(f3_load(pA), f3_load(pB), f3_load(vA), f3_load(vB), rA, rB, attraction, P);
  out3[0] = f.x; out3[1] = f.y; out3[2] = f.z;
}

/* collideD + collideCell (particles_kernel.cu:223-293); with poly each
 * pair uses the spheres' own radii pos.w (the collider pair uses rA).
 * collider == 0 with force3 != NULL is evalForces_launcher
 * (particledem_um.cu:538-589): the same 27-cell sum without the cursor
 * sphere, and the force is also stored as float3 at the original index. */
static void collide_impl(float* newVel4, const float* sortedPos4, const float* sortedVel4,
                         const uint32_t* index, const uint32_t* cellStart, const uint32_t* cellEnd,
                         uint32_t n, const orc_params* P, int poly, int collider, float* force3) {
  long k;
#pragma omp parallel for schedule(dynamic, 1024)
  for (k = 0; k < (long)n; ++k) {
    f3 pos = f3_load(sortedPos4 + 4 * k);
    f3 vel = f3_load(sortedVel4 + 4 * k);
    const float rA = poly ? sortedPos4[4 * k + 3] : P->particleRadius;
    int gx = grid_coord(pos.x, P->worldOrigin.x, P->cellSize.x);
    int gy = grid_coord(pos.y, P->worldOrigin.y, P->cellSize.y);
    int gz = grid_coord(pos.z, P->worldOrigin.z, P->cellSize.z);
    f3 force = f3_make(0.0f, 0.0f, 0.0f);
    for (int z = -1; z <= 1; ++z)
      for (int y = -1; y <= 1; ++y)
        for (int x = -1; x <= 1; ++x) {
          uint32_t h = grid_hash(gx + x, gy + y, gz + z, P);
          uint32_t start = cellStart[h];
          f3 cellForce = f3_make(0.0f, 0.0f, 0.0f);
          if (start != ORC_EMPTY) {
            uint32_t end = cellEnd[h];
            for (uint32_t j = start; j < end; ++j) {
              if (j != (uint32_t)k) {
                const float* pj = sortedPos4 + 4 * (size_t)j;
                const float rB = poly ? pj[3] : P->particleRadius;
                cellForce = f3_add(cellForce,
                                   pair_force(pos, f3_load(pj), vel,
                                              f3_load(sortedVel4 + 4 * (size_t)j), rA, rB,
                                              P->attraction, P));
              }
            }
          }
          force = f3_add(force, cellForce);
        }
    if (collider)
      force = f3_add(force, pair_force(pos, f3_make(P->colliderPos.x, P->colliderPos.y,
                                                    P->colliderPos.z),
                                       vel, f3_make(0.0f, 0.0f, 0.0f), rA, P->colliderRadius,
                                       0.0f, P));
    uint32_t orig = index[k];
    if (force3) {
      float* fo = force3 + 3 * (size_t)orig;
      fo[0] = force.x;
      fo[1] = force.y;
      fo[2] = force.z;
    }
    float* out = newVel4 + 4 * (size_t)orig;
    out[0] = vel.x + force.x;
    out[1] = vel.y + force.y;
    out[2] = vel.z + force.z;
    out[3] = 0.0f;
  }
}

void orc_collide(float* newVel4, const float* sortedPos4, const float* sortedVel4,
                 const uint32_t* index, const uint32_t* cellStart, const uint32_t* cellEnd,
                 uint32_t n, const orc_params* P) {
  collide_impl(newVel4, sortedPos4, sortedVel4, index, cellStart, cellEnd, n, P, 0, 1, NULL);
}

void orc_eval_forces(float* newVel4, float* force3, const float* sortedPos4,
                     const float* sortedVel4, const uint32_t* index, const uint32_t* cellStart,
                     const uint32_t* cellEnd, uint32_t n, const orc_params* P) {
  collide_impl(newVel4, sortedPos4, sortedVel4, index, cellStart, cellEnd, n, P, 0, 0, force3);
}

void orc_collide_poly(float* newVel4, const float* sortedPos4, const float* sortedVel4,
                      const uint32_t* index, const uint32_t* cellStart, const uint32_t* cellEnd,
                      uint32_t n, const orc_params* P) {
  collide_impl(newVel4, sortedPos4, sortedVel4, index, cellStart, cellEnd, n, P, 1, 1, NULL);
}

/* Contacts per sorted sphere: neighbours j != k in the 27 cells with
 * sqrtf(d.d) < rA + rB, the collide's contact test. */
void orc_contact_counts(uint32_t* counts, const float* sortedPos4, const uint32_t* cellStart,
                        const uint32_t* cellEnd, uint32_t n, const orc_params* P, int poly) {
  long k;
#pragma omp parallel for schedule(dynamic, 1024)
  for (k = 0; k < (long)n; ++k) {
    f3 pos = f3_load(sortedPos4 + 4 * k);
    const float rA = poly ? sortedPos4[4 * k + 3] : P->particleRadius;
    int gx = grid_coord(pos.x, P->worldOrigin.x, P->cellSize.x);
    int gy = grid_coord(pos.y, P->worldOrigin.y, P->cellSize.y);
    int gz = grid_coord(pos.z, P->worldOrigin.z, P->cellSize.z);
    uint32_t c = 0;
    for (int z = -1; z <= 1; ++z)
      for (int y = -1; y <= 1; ++y)
        for (int x = -1; x <= 1; ++x) {
          uint32_t h = grid_hash(gx + x, gy + y, gz + z, P);
          uint32_t start = cellStart[h];
          if (start == ORC_EMPTY) continue;
          for (uint32_t j = start; j < cellEnd[h]; ++j) {
            if (j == (uint32_t)k) continue;
            const float* pj = sortedPos4 + 4 * (size_t)j;
            const float rB = poly ? pj[3] : P->particleRadius;
            f3 rel = f3_sub(f3_load(pj), pos);
            if (sqrtf(f3_dot(rel, rel)) < rA + rB) ++c;
          }
        }
    counts[k] = c;
  }
}

/* ------------------------------------------------------------------------ */
/* binary64 twin of the collide, for error statements.                      */
/* ------------------------------------------------------------------------ */
typedef struct { double x, y, z; } d3;
static inline d3 d3_make(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static inline d3 d3_add(d3 a, d3 b) { return d3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 d3_sub(d3 a, d3 b) { return d3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 d3_scale(d3 a, double s) { return d3_make(a.x * s, a.y * s, a.z * s); }
static inline double d3_dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 d3_load(const float* p) { return d3_make(p[0], p[1], p[2]); }

static inline d3 pair_force_f64(d3 posA, d3 posB, d3 velA, d3 velB, double radiusA,
                                double radiusB, double attraction, const orc_params* P) {
  d3 relPos = d3_sub(posB, posA);
  double dist = sqrt(d3_dot(relPos, relPos));
  double collideDist = radiusA + radiusB;
  d3 force = d3_make(0.0, 0.0, 0.0);
  if (dist < collideDist) {
    d3 norm = d3_scale(relPos, 1.0 / dist);
    d3 relVel = d3_sub(velB, velA);
    d3 tanVel = d3_sub(relVel, d3_scale(norm, d3_dot(relVel, norm)));
    force = d3_scale(norm, -(double)P->spring * (collideDist - dist));
    force = d3_add(force, d3_scale(relVel, (double)P->damping));
    force = d3_add(force, d3_scale(tanVel, (double)P->shear));
    force = d3_add(force, d3_scale(relPos, attraction));
  }
  return force;
}

void orc_collide_f64(double* newVel3, const float* sortedPos4, const float* sortedVel4,
                     const uint32_t* index, const uint32_t* cellStart, const uint32_t* cellEnd,
                     uint32_t n, const orc_params* P) {
  const double r = P->particleRadius;
  long k;
#pragma omp parallel for schedule(dynamic, 1024)
  for (k = 0; k < (long)n; ++k) {
    const float* pk = sortedPos4 + 4 * k;
    d3 pos = d3_load(pk);
    d3 vel = d3_load(sortedVel4 + 4 * k);
    /* The cell of a particle is a binary32 computation in the reference;
     * the neighbour search is kept identical so both sums run over the same
     * pairs. */
    int gx = grid_coord(pk[0], P->worldOrigin.x, P->cellSize.x);
    int gy = grid_coord(pk[1], P->worldOrigin.y, P->cellSize.y);
    int gz = grid_coord(pk[2], P->worldOrigin.z, P->cellSize.z);
    d3 force = d3_make(0.0, 0.0, 0.0);
    for (int z = -1; z <= 1; ++z)
      for (int y = -1; y <= 1; ++y)
        for (int x = -1; x <= 1; ++x) {
          uint32_t h = grid_hash(gx + x, gy + y, gz + z, P);
          uint32_t start = cellStart[h];
          if (start == ORC_EMPTY) continue;
          uint32_t end = cellEnd[h];
          for (uint32_t j = start; j < end; ++j)
            if (j != (uint32_t)k)
              force = d3_add(force, pair_force_f64(pos, d3_load(sortedPos4 + 4 * (size_t)j), vel,
                                                   d3_load(sortedVel4 + 4 * (size_t)j), r, r,
                                                   (double)P->attraction, P));
        }
    force = d3_add(force, pair_force_f64(pos, d3_make(P->colliderPos.x, P->colliderPos.y,
                                                      P->colliderPos.z),
                                         vel, d3_make(0, 0, 0), r, (double)P->colliderRadius,
                                         0.0, P));
    uint32_t orig = index[k];
    newVel3[3 * (size_t)orig + 0] = vel.x + force.x;
    newVel3[3 * (size_t)orig + 1] = vel.y + force.y;
    newVel3[3 * (size_t)orig + 2] = vel.z + force.z;
  }
}

/* ParticleSystem::update (particleSystem.cpp:222-284). */
void orc_update(orc_state* s, float dt, const orc_params* P) {
  orc_integrate(s->pos4, s->vel4, s->n, dt, P);
  orc_calc_hash(s->hash, s->index, s->pos4, s->n, P);
  orc_sort_pairs(s->hash, s->index, s->n);
  orc_reorder(s->cellStart, s->cellEnd, s->sortedPos4, s->sortedVel4, s->hash, s->index,
              s->pos4, s->vel4, s->n, P->numCells);
  orc_collide(s->vel4, s->sortedPos4, s->sortedVel4, s->index, s->cellStart, s->cellEnd, s->n,
              P);
}

void orc_run(orc_state* s, float dt, uint32_t steps, const orc_params* P) {
  for (uint32_t i = 0; i < steps; ++i) orc_update(s, dt, P);
}

/* One update with the radius mode and boundary of a particle system: poly
 * takes radii from pos.w; a non-NULL map selects the distance-map boundary
 * (d_integrateSystem, particledem_um.cu:604-638). */
void orc_update_ex(orc_state* s, float dt, const orc_params* P, const orc_distmap* m, int poly) {
  if (m)
    orc_integrate_distmap_ex(s->pos4, s->vel4, s->n, dt, P, m, poly);
  else
    integrate_impl(s->pos4, s->vel4, s->n, dt, P, poly);
  orc_calc_hash(s->hash, s->index, s->pos4, s->n, P);
  orc_sort_pairs(s->hash, s->index, s->n);
  orc_reorder(s->cellStart, s->cellEnd, s->sortedPos4, s->sortedVel4, s->hash, s->index,
              s->pos4, s->vel4, s->n, P->numCells);
  collide_impl(s->vel4, s->sortedPos4, s->sortedVel4, s->index, s->cellStart, s->cellEnd, s->n, P,
               poly, 1, NULL);
}

/* SplitMix64 (Steele, Lea, Flood 2014). */
static inline uint64_t splitmix64(uint64_t* st) {
  uint64_t z = (*st += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline float unif01(uint64_t* st) {
  return (float)(splitmix64(st) >> 40) * (1.0f / 16777216.0f);
}

/* reset(CONFIG_GRID / CONFIG_RANDOM) (particleSystem.cpp:380-444). */
void orc_reset(float* pos4, float* vel4, uint32_t n, int config, uint64_t seed, float radius,
               float jitterFrac, float velAmp) {
  uint64_t st = seed;
  if (config == 1) {
    uint32_t s = 0;
    while ((uint64_t)s * s * s < n) ++s;
    float spacing = radius * 2.0f;
    float jitter = radius * jitterFrac;
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t c[3] = {i % s, (i / s) % s, i / (s * s)};
      for (int a = 0; a < 3; ++a) {
        float u = unif01(&st) * 2.0f - 1.0f;
        pos4[4 * (size_t)i + a] = ((spacing * (float)c[a] + radius) - 1.0f) + u * jitter;
      }
      pos4[4 * (size_t)i + 3] = 1.0f;
      for (int a = 0; a < 3; ++a) vel4[4 * (size_t)i + a] = (unif01(&st) * 2.0f - 1.0f) * velAmp;
      vel4[4 * (size_t)i + 3] = 0.0f;
    }
  } else {
    for (uint32_t i = 0; i < n; ++i) {
      for (int a = 0; a < 3; ++a) pos4[4 * (size_t)i + a] = 2.0f * (unif01(&st) - 0.5f);
      pos4[4 * (size_t)i + 3] = 1.0f;
      for (int a = 0; a < 3; ++a) vel4[4 * (size_t)i + a] = (unif01(&st) * 2.0f - 1.0f) * velAmp;
      vel4[4 * (size_t)i + 3] = 0.0f;
    }
  }
}

/* Neumaier-compensated binary64 sums of kinetic energy and momentum. */
void orc_energy_momentum(const float* vel4, uint32_t n, double* out4) {
  double sum[4] = {0, 0, 0, 0}, comp[4] = {0, 0, 0, 0};
  for (uint32_t i = 0; i < n; ++i) {
    double vx = vel4[4 * (size_t)i], vy = vel4[4 * (size_t)i + 1], vz = vel4[4 * (size_t)i + 2];
    double terms[4] = {0.5 * (vx * vx + vy * vy + vz * vz), vx, vy, vz};
    for (int a = 0; a < 4; ++a) {
      double t = sum[a] + terms[a];
      if (fabs(sum[a]) >= fabs(terms[a])) comp[a] += (sum[a] - t) + terms[a];
      else comp[a] += (terms[a] - t) + sum[a];
      sum[a] = t;
    }
  }
  for (int a = 0; a < 4; ++a) out4[a] = sum[a] + comp[a];
}
