/*
 * fc_oracle.h - CPU restatement of FullC0ntact's particle-DEM step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in fullc0ntact_amd/ and the CPU baseline timed beside it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path never calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the FullC0ntact tree).  Arithmetic is IEEE binary32 in the reference's
 * operation order; the library is compiled without FMA contraction and
 * without fast-math, so results are reproducible bit for bit.
 */
#ifndef FC_ORACLE_H
#define FC_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y, z; } orc_float3;
typedef struct { uint32_t x, y, z; } orc_uint3;

/* SimParams (cuda_addon/particles_kernel.cuh:27-48), 104 bytes. */
typedef struct orc_params {
  orc_float3 colliderPos;
  float colliderRadius;
  orc_float3 gravity;
  float globalDamping;
  float particleRadius;
  orc_uint3 gridSize;
  uint32_t numCells;
  orc_float3 worldOrigin;
  orc_float3 cellSize;
  uint32_t numBodies;
  uint32_t maxParticlesPerCell;
  float spring;
  float damping;
  float shear;
  float attraction;
  float boundaryDamping;
} orc_params;

#define ORC_EMPTY 0xffffffffu

int orc_params_size(void);
void orc_set_threads(int threads);
int orc_get_threads(void);

/* integrate_functor (particles_kernel.cu:39-75). */
void orc_integrate(float* pos4, float* vel4, uint32_t n, float dt, const orc_params* P);
/* calcGridPos/calcGridHash/calcHashD (particles_kernel.cu:78-115). */
uint32_t orc_hash_point(float x, float y, float z, const orc_params* P);
void orc_calc_hash(uint32_t* hash, uint32_t* index, const float* pos4, uint32_t n,
                   const orc_params* P);
/* thrust::sort_by_key (particleSystem.cu:265-270): stable by key. */
void orc_sort_pairs(uint32_t* hash, uint32_t* index, uint32_t n);
/* reorderDataAndFindCellStart (particleSystem.cu:183-222 with both tables
 * cleared as in particledem.cu:229-230; kernel particles_kernel.cu:119-182). */
void orc_reorder(uint32_t* cellStart, uint32_t* cellEnd, float* sortedPos4, float* sortedVel4,
                 const uint32_t* hash, const uint32_t* index, const float* oldPos4,
                 const float* oldVel4, uint32_t n, uint32_t numCells);
/* collideD (particles_kernel.cu:256-293) in binary32, reference order. */
void orc_collide(float* newVel4, const float* sortedPos4, const float* sortedVel4,
                 const uint32_t* index, const uint32_t* cellStart, const uint32_t* cellEnd,
                 uint32_t n, const orc_params* P);
/* Same algorithm in binary64 from the same binary32 inputs: newVel3 receives
 * (vel + force) per original index as doubles (3 per particle). */
/* evalForces_launcher (particledem_um.cu:538-589) in binary32: collideD's
 * 27-cell sum without the cursor sphere; newVel[orig] = (vel + force, 0)
 * and force3[orig] = force (float3, 12-byte stride). */
void orc_eval_forces(float* newVel4, float* force3, const float* sortedPos4,
                     const float* sortedVel4, const uint32_t* index, const uint32_t* cellStart,
                     const uint32_t* cellEnd, uint32_t n, const orc_params* P);
void orc_collide_f64(double* newVel3, const float* sortedPos4, const float* sortedVel4,
                     const uint32_t* index, const uint32_t* cellStart, const uint32_t* cellEnd,
                     uint32_t n, const orc_params* P);
/* collideSpheres (particles_kernel.cu:185-218) for one pair, binary32. */
void orc_pair_force(const float* pA, const float* pB, const float* vA, const float* vB, float rA,
                    float rB, float attraction, const orc_params* P, float* out3);

/* One ParticleSystem::update (particleSystem.cpp:222-284) on host arrays.
 * Work arrays: sortedPos4/sortedVel4 (4n floats), hash/index (n),
 * cellStart/cellEnd (numCells). */
typedef struct orc_state {
  float* pos4;
  float* vel4;
  float* sortedPos4;
  float* sortedVel4;
  uint32_t* hash;
  uint32_t* index;
  uint32_t* cellStart;
  uint32_t* cellEnd;
  uint32_t n;
} orc_state;
void orc_update(orc_state* s, float dt, const orc_params* P);
void orc_run(orc_state* s, float dt, uint32_t steps, const orc_params* P);

/* Polydisperse variants (radius of each sphere in pos.w; the reference has
 * none: with equal radii they reduce to orc_integrate / orc_collide). */
void orc_integrate_poly(float* pos4, float* vel4, uint32_t n, float dt, const orc_params* P);
void orc_collide_poly(float* newVel4, const float* sortedPos4, const float* sortedVel4,
                      const uint32_t* index, const uint32_t* cellStart, const uint32_t* cellEnd,
                      uint32_t n, const orc_params* P);
/* Contacts per sorted sphere (the collide's test sqrtf(d.d) < rA + rB). */
void orc_contact_counts(uint32_t* counts, const float* sortedPos4, const uint32_t* cellStart,
                        const uint32_t* cellEnd, uint32_t n, const orc_params* P, int poly);

/* reset(CONFIG_GRID / CONFIG_RANDOM) with SplitMix64 (see fc_dem.h). */
void orc_reset(float* pos4, float* vel4, uint32_t n, int config, uint64_t seed, float radius,
               float jitterFrac, float velAmp);

/* Observables in binary64 with compensated summation. out[0] = 1/2 sum |v|^2,
 * out[1..3] = sum v (momentum per unit mass). */
void orc_energy_momentum(const float* vel4, uint32_t n, double* out4);

/* ---- distance map (fc_oracle_distmap.c) ---- */

/* A DistanceMap<float> (grid/distancemap.h:36-143) on a tensor-product
 * lattice: vertex (ix,iy,iz) has coordinates (xs[ix], ys[iy], zs[iz]) and
 * index (iz*(cy+1) + iy)*(cx+1) + ix; cp and nrm hold 3 floats per vertex. */
typedef struct orc_distmap {
  int cells[3];
  float cellSize;
  float bmin[3];
  float bmax[3];
  const float* xs;
  const float* ys;
  const float* zs;
  const float* dist;
  const float* cp;
  const float* nrm;
} orc_distmap;

/* TransInfo (inshape3dcore/rigidbody.h:54-66), row-major 3x3 matrices. */
typedef struct orc_transinfo {
  float origin0[3];
  float origin1[3];
  float m2w0[9];
  float w2m0[9];
  float m2w1[9];
} orc_transinfo;

int orc_distmap_query(const orc_distmap* m, const float q[3], float* dist, float cp[3],
                      float nrm[3]);
void orc_distmap_query_batch(const orc_distmap* m, const float* q, uint32_t n, float* dist,
                             float* cp, float* nrm);
void orc_integrate_distmap(float* pos4, float* vel4, uint32_t n, float dt, const orc_params* P,
                           const orc_distmap* m);
void orc_distmap_eval(const orc_distmap* m, const float* v, uint32_t n, const orc_transinfo* t,
                      float* cps, float* normals, float* distance);
double orc_point_triangle_dist2(const double q[3], const double B[3], const double V1[3],
                                const double V2[3], double cp[3]);
void orc_mesh_sdf(const double* verts, const uint32_t* tris, uint32_t nt, const double* pts,
                  uint32_t np, double* sdf, double* cp, double* winding);
void orc_update_distmap(orc_state* s, float dt, const orc_params* P, const orc_distmap* m);
void orc_integrate_distmap_ex(float* pos4, float* vel4, uint32_t n, float dt, const orc_params* P,
                              const orc_distmap* m, int poly);
/* One update with a radius mode (poly) and an optional distance-map
 * boundary (m may be NULL: box walls). */
void orc_update_ex(orc_state* s, float dt, const orc_params* P, const orc_distmap* m, int poly);

/* ---- signed-distance surrogate MLP (fc_oracle_sdf.c) ---- */
int orc_sdf_num_params(void);
void orc_sdf_forward(const float* params, const float* freq, const float* frame,
                     const float* pts, uint32_t n, double* out);
void orc_sdf_loss_grad(const float* params, const float* freq, const float* frame,
                       const float* pts, const float* labels, uint32_t n, double* grad,
                       double* loss);
void orc_sdf_forward_bf16(const float* params, const float* freq, const float* frame,
                          const float* pts, uint32_t n, double* out);

#ifdef __cplusplus
}
#endif

#endif
