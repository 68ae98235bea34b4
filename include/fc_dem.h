/*
 * fc_dem.h - C ABI of the MI355X (gfx950) particle-DEM step.
 *
 * This is the drop-in boundary for the particle step of FullC0ntact's GPU
 * particle system.  Every entry point names the reference interface it
 * replaces (paths relative to the FullC0ntact tree).  All functions are
 * plain C: raw pointers and sizes, no C++ or torch types.  Device pointers
 * are HIP device allocations (fc_allocate_array or any hipMalloc'd buffer);
 * host pointers are ordinary host memory.
 *
 * Differences from the reference ABI, all deliberate:
 *   - Every call returns an fc_status instead of calling exit() on error
 *     (reference: cutilSafeCall, particleSystem.cu:67,180).
 *   - Parameters live in an explicit fc_context (one parameter set and one
 *     HIP stream per context) instead of the process-global __constant__
 *     SimParams (particles_kernel.cu:36).  A binding that wants the
 *     reference's global style keeps one context in a static.
 *   - Launches are asynchronous on the context's stream; copies to/from the
 *     host synchronise that stream.
 */
#ifndef FC_DEM_H
#define FC_DEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FC_DEM_ABI_VERSION 1

/* 12-byte vectors, unpadded, matching CUDA's float3/uint3 used inside
 * SimParams (particles_kernel.cuh:27-48). */
typedef struct fc_float3 { float x, y, z; } fc_float3;
typedef struct fc_uint3 { uint32_t x, y, z; } fc_uint3;

/* Same fields, order and 104-byte layout as SimParams
 * (cuda_addon/particles_kernel.cuh:27-48); host mirror of
 * SimulationParameters<T> (inshape3dcore/grid/uniformgrid.h:470-493). */
typedef struct fc_sim_params {
  fc_float3 colliderPos;
  float colliderRadius;
  fc_float3 gravity;
  float globalDamping;
  float particleRadius;
  fc_uint3 gridSize;
  uint32_t numCells;
  fc_float3 worldOrigin;
  fc_float3 cellSize;
  uint32_t numBodies;
  uint32_t maxParticlesPerCell;
  float spring;
  float damping;
  float shear;
  float attraction;
  float boundaryDamping;
} fc_sim_params;

typedef enum fc_status {
  FC_SUCCESS = 0,
  FC_ERROR_INVALID_VALUE = 1,    /* null pointer, zero-size grid, bad enum   */
  FC_ERROR_GRID_NOT_POW2 = 2,    /* hash wraps with & (G-1): G must be 2^k   */
  FC_ERROR_HIP = 3,              /* a HIP runtime call failed                */
  FC_ERROR_OUT_OF_MEMORY = 4,
  FC_ERROR_NO_DEVICE = 5,
  FC_ERROR_SORT_STALLED = 6,     /* radix-sort look-back hit its spin bound  */
  FC_ERROR_CAPACITY = 7,         /* size beyond a documented maximum         */
  FC_ERROR_NOT_SUPPORTED = 8
} fc_status;

/* Particle arrays (ParticleSystem::ParticleArray, particleSystem.h:35-39). */
typedef enum fc_particle_array { FC_POSITION = 0, FC_VELOCITY = 1 } fc_particle_array;

/* Initial configurations (ParticleSystem::ParticleConfig,
 * particleSystem.h:28-33). */
typedef enum fc_particle_config { FC_CONFIG_RANDOM = 0, FC_CONFIG_GRID = 1 } fc_particle_config;

#define FC_CELL_EMPTY 0xffffffffu

typedef struct fc_context fc_context;
typedef struct fc_system fc_system;
typedef struct fc_distmap fc_distmap;

/* TransInfo (inshape3dcore/rigidbody.h:54-66): rigid transforms of two
 * bodies, 3x3 matrices row major. */
typedef struct fc_trans_info {
  fc_float3 origin0;
  fc_float3 origin1;
  float m2w0[9]; /* model -> world of body 0 */
  float w2m0[9]; /* world -> model of body 0 */
  float m2w1[9]; /* model -> world of body 1 */
} fc_trans_info;

/* Boundary applied by the integrate stage of a particle system. */
typedef enum fc_boundary {
  /* Box walls at +-1 (integrate_functor, particles_kernel.cu:39-75). */
  FC_BOUNDARY_BOX = 0,
  /* Distance-map push-out, cylinder r = 1 around z, planes z = -1 and 10
   * (d_integrateSystem, particledem_um.cu:604-638; boundaries.cuh:38-89). */
  FC_BOUNDARY_DISTMAP = 1
} fc_boundary;

/* ------------------------------------------------------------------------- */
/* Runtime, errors, memory.                                                  */
/* ------------------------------------------------------------------------- */

/* Replaces cudaInit(argc, argv) (particleSystem.cuh:14): selects a device and
 * creates a context with its own HIP stream and the ParticleSystem default
 * parameters (particleSystem.cpp:48-68).  device < 0 keeps the current one. */
fc_status fc_context_create(int device, fc_context** out);
/* Objects created on a context (systems, maps, batches, surrogates) must not
 * be used after it is destroyed; destroying them afterwards is safe. */
fc_status fc_context_destroy(fc_context* ctx);
/* Text of the last error raised on this thread (never NULL). */
const char* fc_last_error(void);
int fc_abi_version(void);
/* Number of HIP devices visible (0 when none), without creating a context. */
int fc_device_count(void);

/* allocateArray / freeArray (particleSystem.cuh:16-17; the reference declares
 * the size as int but defines it as size_t, particleSystem.cu:65). */
fc_status fc_allocate_array(void** devPtr, size_t size);
fc_status fc_free_array(void* devPtr);
/* threadSync (particleSystem.cuh:19): waits for the context's stream and
 * reports any deferred device-side error (e.g. FC_ERROR_SORT_STALLED). */
fc_status fc_thread_sync(fc_context* ctx);
/* copyArrayToDevice(device, host, offset, size) (particleSystem.cuh:22). */
fc_status fc_copy_array_to_device(fc_context* ctx, void* device, const void* host, size_t offset, size_t size);
/* copyArrayFromDevice (particleSystem.cuh:21) without the GL resource: copies
 * size bytes starting at byte offset of the device buffer. */
fc_status fc_copy_array_from_device(fc_context* ctx, void* host, const void* device, size_t offset, size_t size);

/* ------------------------------------------------------------------------- */
/* Per-stage kernels on caller-owned device arrays (particleSystem.cuh:29-61) */
/* ------------------------------------------------------------------------- */

/* setParameters(SimParams*) (particleSystem.cuh:29).  Validates that every
 * grid dimension is a power of two and numCells == x*y*z. */
fc_status fc_set_parameters(fc_context* ctx, const fc_sim_params* params);
fc_status fc_get_parameters(const fc_context* ctx, fc_sim_params* params);

/* integrateSystem(pos, vel, dt, n) (particleSystem.cuh:31-34; kernel
 * integrate_functor, particles_kernel.cu:39-75): box walls at +-1. */
fc_status fc_integrate_system(fc_context* ctx, float* pos, float* vel, float deltaTime, uint32_t numParticles);

/* calcHash(hash, index, pos, n) (particleSystem.cuh:36-39; calcHashD,
 * particles_kernel.cu:97-115). */
fc_status fc_calc_hash(fc_context* ctx, uint32_t* gridParticleHash, uint32_t* gridParticleIndex,
                       const float* pos, uint32_t numParticles);

/* sortParticles(hash, index, n) (particleSystem.cuh:61; thrust::sort_by_key
 * at particleSystem.cu:265-270): stable ascending sort of the (hash, index)
 * pairs in place, by an LSD onesweep radix sort over the significant bits of
 * numCells-1 taken from the context parameters. */
fc_status fc_sort_particles(fc_context* ctx, uint32_t* gridParticleHash, uint32_t* gridParticleIndex,
                            uint32_t numParticles);

/* No counterpart in the reference: the sort of fc_sort_particles for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It calls no
 * allocating routine: its scratch is one module array per device (about
 * 16.6 MiB, held until the process ends), and it sorts at most
 * fc_sort_particles_noalloc_max() = 1048576 pairs (FC_ERROR_CAPACITY
 * beyond).  Every call on a device shares that device's array: calls that
 * overlap on one device (other streams or threads) race; calls on one
 * stream are ordered.  The results equal fc_sort_particles'. */
fc_status fc_sort_particles_noalloc(fc_context* ctx, uint32_t* gridParticleHash,
                                    uint32_t* gridParticleIndex, uint32_t numParticles);
size_t fc_sort_particles_noalloc_max(void);

/* reorderDataAndFindCellStart (particleSystem.cuh:41-50; kernel
 * particles_kernel.cu:119-182).  Clears cellStart AND cellEnd to
 * FC_CELL_EMPTY first (as particledem.cu:229-230 does) so every table entry
 * is defined. */
fc_status fc_reorder_data_and_find_cell_start(fc_context* ctx, uint32_t* cellStart, uint32_t* cellEnd,
                                              float* sortedPos, float* sortedVel,
                                              const uint32_t* gridParticleHash,
                                              const uint32_t* gridParticleIndex,
                                              const float* oldPos, const float* oldVel,
                                              uint32_t numParticles, uint32_t numCells);

/* collide (particleSystem.cuh:52-59; collideD, particles_kernel.cu:256-293):
 * 27-cell spring-dashpot-shear force plus the collider sphere, written as
 * newVel[index[k]] = (vel + force, 0).  Above 16384 particles the context
 * first builds, in buffers of its own, each sorted particle's velocity with
 * its cell hash and its 16-bit coordinates (the arrays are assumed to come
 * from fc_reorder_data_and_find_cell_start of the hashes fc_calc_hash gives
 * for these positions, as in the reference's update); the buffers grow on
 * the first call with more particles than they hold, which allocates, so
 * such a call cannot be captured in a graph. */
fc_status fc_collide(fc_context* ctx, float* newVel, const float* sortedPos, const float* sortedVel,
                     const uint32_t* gridParticleIndex, const uint32_t* cellStart,
                     const uint32_t* cellEnd, uint32_t numParticles, uint32_t numCells);

/* evalForces (particledem_um.cuh:19; evalForces_launcher,
 * particledem_um.cu:538-589), the force stage of the FullC0ntact plug-in
 * pipeline: the same 27-cell sum as fc_collide but without the cursor
 * sphere, written as newVel[index[k]] = (vel + force, 0) and
 * force[3 * index[k] + 0..2] = force (float3 array, 12-byte stride, the
 * ParticleWorld::force_ layout).  The reference reads the grid and the
 * particle arrays from its HashGrid/ParticleWorld structs; here they are
 * passed as device pointers, with the parameters of the context.  Large
 * systems use the context's buffers as fc_collide does. */
fc_status fc_eval_forces(fc_context* ctx, float* newVel, float* force, const float* sortedPos,
                         const float* sortedVel, const uint32_t* gridParticleIndex,
                         const uint32_t* cellStart, const uint32_t* cellEnd,
                         uint32_t numParticles, uint32_t numCells);

/* ------------------------------------------------------------------------- */
/* Particle system (ParticleSystem class, particleSystem.h:22-136).          */
/* The system owns every device buffer; update() runs the reference order     */
/* integrate -> calcHash -> sort -> reorder -> collide                        */
/* (particleSystem.cpp:222-284).                                              */
/* ------------------------------------------------------------------------- */

/* ParticleSystem(numParticles, gridSize, false) (particleSystem.cpp:30-72):
 * the system starts with the constructor's default parameters (radius 1/16,
 * cell 2r, origin (-1,-1,-1), spring .5, damping .02, shear .1, ...) and all
 * particles at the origin with zero velocity. */
fc_status fc_system_create(fc_context* ctx, uint32_t numParticles, fc_uint3 gridSize, fc_system** out);
fc_status fc_system_destroy(fc_system* sys);
uint32_t fc_system_num_particles(const fc_system* sys);

/* m_params writes/reads (the set* setters of particleSystem.h:64-72 are
 * field writes on a copy; see fc_particle_system.hpp). */
fc_status fc_system_set_params(fc_system* sys, const fc_sim_params* params);
fc_status fc_system_get_params(const fc_system* sys, fc_sim_params* params);

/* setArray(array, data, start, count) (particleSystem.cpp:350-373) for both
 * POSITION and VELOCITY (the reference uploads nothing for POSITION in
 * non-GL mode; here both upload). data holds count float4 records. */
fc_status fc_system_set_array(fc_system* sys, fc_particle_array array, const float* data, int start, int count);
/* getArray(array) (particleSystem.cpp:323-348): copies all numParticles
 * float4 records into out (4*numParticles floats). */
fc_status fc_system_get_array(fc_system* sys, fc_particle_array array, float* out);

/* reset(config) (particleSystem.cpp:404-444) with a portable SplitMix64
 * stream instead of srand/rand: CONFIG_GRID puts particles on the
 * s x s x s lattice (s = smallest integer with s^3 >= n) of spacing 2r from
 * -1+r, jittered by U(-jitterFrac*r, jitterFrac*r) per axis, velocities
 * U(-velAmp, velAmp); CONFIG_RANDOM draws positions U(-1,1)^3 and
 * velocities U(-velAmp, velAmp).  pos.w = 1, vel.w = 0. */
fc_status fc_system_reset(fc_system* sys, fc_particle_config config, uint64_t seed, float jitterFrac, float velAmp);

/* update(deltaTime) (particleSystem.cpp:222-284). */
fc_status fc_system_update(fc_system* sys, float deltaTime);
/* steps consecutive update(deltaTime) calls.  Results are bit-identical to
 * calling fc_system_update steps times. */
fc_status fc_system_run(fc_system* sys, float deltaTime, uint32_t steps);

/* Device pointers of the system's buffers (for zero-copy consumers).  Any
 * pointer argument may be NULL.  float arrays are float4[numParticles]. */
fc_status fc_system_device_arrays(fc_system* sys, float** pos, float** vel, float** sortedPos,
                                  float** sortedVel, uint32_t** hash, uint32_t** index,
                                  uint32_t** cellStart, uint32_t** cellEnd);

/* dumpGrid (particleSystem.cpp:286-301) without the printf: the largest
 * cellEnd-cellStart over non-empty cells, plus the number of occupied cells. */
fc_status fc_system_dump_grid(fc_system* sys, uint32_t* maxParticlesPerCell, uint32_t* occupiedCells);

/* Boundary of the integrate stage (default FC_BOUNDARY_BOX).  The map must
 * outlive its use by the system.  Setting a map synchronises the context's
 * stream and, for a map on the uniform lattice bmin + i cellSize, allocates
 * the bounds of its 8^3-cell blocks (8 bytes per block), with which the
 * integrate skips the query of spheres that cannot touch the boundary (the
 * results are those of querying every sphere). */
fc_status fc_system_set_boundary(fc_system* sys, fc_boundary mode, const fc_distmap* map);

/* Polydisperse spheres.  No counterpart in the reference, which writes
 * pos.w = 1 and ignores it (particleSystem.cpp:392,422).  When enabled, pos.w
 * holds each sphere's radius: the collide uses rA + rB per pair
 * (collideSpheres, particles_kernel.cu:185-218, with per-sphere radii) and
 * the walls and the distance-map boundary use the sphere's own radius.
 * the parameters' particleRadius must be the largest radius (it sizes the
 * neighbour search and the contact pre-filter; the cell size must be at
 * least twice it); a larger pos.w is reported as FC_ERROR_INVALID_VALUE at
 * the next synchronising call.  With every pos.w equal to particleRadius
 * the results are bit-identical to the monodisperse step. */
fc_status fc_system_set_polydisperse(fc_system* sys, int enable);

/* ------------------------------------------------------------------------- */
/* Distance maps (DistanceMap<float,gpu>, cuda_addon/distancemap.cuh:6-303;  */
/* difi.cu).                                                                  */
/* ------------------------------------------------------------------------- */

/* DistanceMap<float,gpu>::transferData (distancemap.cuh:50-95) /
 * allocate_dmap (difi.cu:774-835): uploads a map given in the reference's
 * host layout: (cells+1)^3 vertices indexed (iz*(cy+1)+iy)*(cx+1)+ix, 3
 * floats each of vertexCoords, normals and contactPoints, one distance.  The
 * vertex coordinates must form a tensor-product lattice.  bmin/bmax are the
 * bounding box vertices_[0] / vertices_[1]. */
fc_status fc_distmap_create(fc_context* ctx, const int32_t cells[3], float cellSize,
                            const float bmin[3], const float bmax[3], const float* vertexCoords,
                            const float* normals, const float* contactPoints,
                            const float* distance, fc_distmap** out);
/* Builds a map of a closed, consistently oriented triangle mesh (host
 * arrays: 3 doubles per vertex, 3 indices per triangle) on the lattice
 * bmin + (ix,iy,iz)*cellSize, computed in binary64 and stored in binary32:
 * signed distance (negative inside), closest point, normal = vertex - closest
 * point (DistanceMapMesh::ComputeDistance, distancegridcgal.hpp:36-71, with
 * the exact distance of D. Eberly's point-triangle method and the sign of the
 * generalized winding number instead of CGAL). */
fc_status fc_distmap_build(fc_context* ctx, const double* vertices, uint32_t numVertices,
                           const uint32_t* triangles, uint32_t numTriangles,
                           const int32_t cells[3], float cellSize, const float bmin[3],
                           fc_distmap** out);
fc_status fc_distmap_destroy(fc_distmap* map);
/* Map files of RigidBody::buildDistanceMapFromFile / storeDistanceMapToFile
 * (rigidbody.cpp:1214-1376): text lines "dim0 dim1", "cx cy cz",
 * "cellSize", then per lattice vertex (x fastest) the coordinates, closest
 * points and normals (3 values each), the distances and the FBM states.
 *
 * fc_distmap_file_read parses one into host binary64 arrays (the reference
 * reads into DistanceMap<Real>): with all array pointers NULL it returns
 * only the header, so callers can size the arrays ((cx+1)(cy+1)(cz+1)
 * vertices). */
fc_status fc_distmap_file_read(const char* path, int32_t cells[3], double* cellSize,
                               int32_t dim[2], double* vertexCoords, double* contactPoints,
                               double* normals, double* distance, int32_t* stateFBM);
/* Loads a map file onto the device, converting binary64 to binary32 as
 * DistanceMap<float,cpu>(DistanceMap<Real,cpu>*) does (distancemap.cpp:
 * 129-178).  The file holds no bounding box (the reference takes it from
 * the body's AABB): pass bmin/bmax, or NULL for both to use the first and
 * the last lattice vertex.  The FBM states are kept with the map. */
fc_status fc_distmap_load(fc_context* ctx, const char* path, const float bmin[3],
                          const float bmax[3], fc_distmap** out);
/* Writes a map in the same format, with nine significant digits so that
 * every binary32 value reads back exactly (the reference writes six). */
fc_status fc_distmap_save(fc_context* ctx, const fc_distmap* map, const char* path);
/* FBM states of the lattice vertices (host, one int per vertex; 0 when
 * none were loaded or set; NULL in set clears them). */
fc_status fc_distmap_get_fbm(const fc_distmap* map, int32_t* stateFBM);
fc_status fc_distmap_set_fbm(fc_distmap* map, const int32_t* stateFBM);
/* queryFBM (distancemap.cuh:195-225) for n points (device arrays, 3 floats
 * per point): 1 where the interpolated distance of a point inside the box
 * is <= 0, else 0. */
fc_status fc_distmap_query_fbm(fc_context* ctx, const fc_distmap* map, const float* points,
                               uint32_t n, int32_t* out);
fc_status fc_distmap_info(const fc_distmap* map, int32_t cells[3], float* cellSize, float bmin[3],
                          float bmax[3]);
/* Copies a map back in the reference's host layout (any pointer may be
 * NULL). */
fc_status fc_distmap_download(fc_context* ctx, const fc_distmap* map, float* vertexCoords,
                              float* normals, float* contactPoints, float* distance);
/* queryMap (distancemap.cuh:154-191) for n points (device arrays, 3 floats
 * per point): distance (5.0 outside the box), closest point and normal
 * (cp, normal may be NULL). */
fc_status fc_distmap_query(fc_context* ctx, const fc_distmap* map, const float* points, uint32_t n,
                           float* dist, float* cp, float* normal);
/* eval_distmap (difi.cu:427-500) on device arrays: vertices of body 1 into
 * body 0's model frame, query, world-space normal and midpoint contact;
 * distance 1000 (and cps/normals untouched) outside the box. */
fc_status fc_distmap_eval(fc_context* ctx, const fc_distmap* map, const float* v, uint32_t n,
                          const fc_trans_info* info, float* cps, float* normals, float* distance);
/* ColliderMeshMesh<gpu>::collide (collidermeshmesh.cuh:89-172): eval_distmap
 * of body 1's model-frame vertices (device, 3 floats each) in body 0's map,
 * keeping the vertices with distance < threshold (the reference uses
 * 0.0015) as contacts, in vertex order: world contact point, unit normal
 * and vertex index into cps/normals/vertexIds (device, capacity entries).
 * *numContacts (host) receives the number of contacts, which may exceed
 * capacity (only the first capacity are written).  Synchronises. */
fc_status fc_mesh_mesh_contacts(fc_context* ctx, const fc_distmap* map, const float* v, uint32_t n,
                                const fc_trans_info* info, float threshold, float* cps,
                                float* normals, uint32_t* vertexIds, uint32_t capacity,
                                uint32_t* numContacts);
/* integrateSystem with the distance-map boundary on caller arrays (device
 * float4 pos/vel), using the context parameters. */
fc_status fc_integrate_system_distmap(fc_context* ctx, const fc_distmap* map, float* pos,
                                      float* vel, float deltaTime, uint32_t numParticles);
/* Exact signed distance (and closest point, may be NULL) of numPoints
 * points to a closed triangle mesh, all arrays on the device (binary64,
 * 3 per vertex/point, 3 indices per triangle; indices must be < numVertices). */
fc_status fc_mesh_signed_distance(fc_context* ctx, const double* vertices, uint32_t numVertices,
                                  const uint32_t* triangles, uint32_t numTriangles,
                                  const double* points, uint32_t numPoints, double* sdf,
                                  double* closest);

/* ------------------------------------------------------------------------- */
/* Global observables.  No counterpart in the reference.  FP64 is used only   */
/* here, in the reductions; the particle state stays FP32.                   */
/* ------------------------------------------------------------------------- */

/* Kinetic energy 1/2 sum |v|^2 (unit mass) and momentum sum v of numParticles
 * float4 velocities (device array), accumulated in binary64 by a fixed-order
 * two-level reduction (deterministic for a given n).  out: host double[4] =
 * {energy, px, py, pz}; synchronises the context stream. */
fc_status fc_stats_fp64(fc_context* ctx, const float* vel, uint32_t numParticles, double out[4]);
/* The same for a system's current velocities. */
fc_status fc_system_stats_fp64(fc_system* sys, double out[4]);
/* Contact-count histogram of a sorted state (device arrays as produced by
 * fc_reorder_data_and_find_cell_start): for every sphere, the number of
 * neighbours in its 27 cells with sqrtf(d.d) < rA + rB (the collide's test);
 * hist[c] (device, nbins <= 4096 counters) receives the number of spheres
 * with c contacts, the last bin collecting c >= nbins-1.  Radii are the
 * context's particleRadius, or pos.w when polydisperse != 0. */
fc_status fc_contact_histogram(fc_context* ctx, const float* sortedPos, const uint32_t* cellStart,
                               const uint32_t* cellEnd, uint32_t numParticles, int polydisperse,
                               uint32_t* hist, uint32_t nbins);
/* The same for a system's state at its last collide; hist is a host array. */
fc_status fc_system_contact_histogram(fc_system* sys, uint32_t* hist, uint32_t nbins);

/* ------------------------------------------------------------------------- */
/* Batched independent scenes.  No counterpart in the reference: many        */
/* ParticleSystem::update steps of independent scenes of uneven sizes run    */
/* as one launch per stage over the whole batch.                              */
/* ------------------------------------------------------------------------- */

/* Scene s owns particles [offsets[s], offsets[s+1]) of the concatenated
 * float4 arrays and its own fc_sim_params (every field may differ except
 * gridSize/numCells, which are the batch's).  The hash is scene-segmented
 * (key = s * cells + cell) so each scene's results are bit-identical to the
 * same scene run alone as an fc_system.  Requires gridSize.x >= 4,
 * numScenes * cells < 2^32 and fewer than 2^27 particles in total. */
typedef struct fc_batch fc_batch;
/* Launch shape of the tuned batch kernels. */
typedef struct fc_batch_tuning {
  int integratePerThread; /* particles per thread of integrate+hash: 1, 2, 4, 8 */
  int collideBlock;       /* threads per block of the collide: 64, 128, 256     */
  float stepMs;           /* measured step time of this choice (autotune)       */
} fc_batch_tuning;

fc_status fc_batch_create(fc_context* ctx, uint32_t numScenes, const uint32_t* sceneCounts,
                          fc_uint3 gridSize, fc_batch** out);
fc_status fc_batch_destroy(fc_batch* batch);
uint32_t fc_batch_num_particles(const fc_batch* batch);
/* numScenes + 1 offsets into the concatenated arrays. */
fc_status fc_batch_scene_offsets(const fc_batch* batch, uint32_t* offsets);
/* numScenes parameter blocks (defaults: the ParticleSystem constructor's). */
fc_status fc_batch_set_params(fc_batch* batch, const fc_sim_params* params);
fc_status fc_batch_get_params(const fc_batch* batch, fc_sim_params* params);
/* All particles of all scenes, float4 each (host arrays). */
fc_status fc_batch_set_array(fc_batch* batch, fc_particle_array array, const float* data);
fc_status fc_batch_get_array(fc_batch* batch, fc_particle_array array, float* out);
/* fc_system_reset per scene with seed + s and the scene's radius. */
fc_status fc_batch_reset(fc_batch* batch, fc_particle_config config, uint64_t seed,
                         float jitterFrac, float velAmp);
/* One update of every scene. */
fc_status fc_batch_update(fc_batch* batch, float deltaTime);
/* steps updates, identical to repeated fc_batch_update: between its steps
 * the state stays in sorted order, each collide fused with the next
 * integrate, the later steps' sorts row resorts, eight steps per graph
 * launch; the state is back in particle order on return.  The first call
 * allocates the row tables (a call that cannot be captured in a graph). */
fc_status fc_batch_run(fc_batch* batch, float deltaTime, uint32_t steps);
/* Per scene {kinetic energy, px, py, pz} in binary64 (host double[4 *
 * numScenes]). */
fc_status fc_batch_stats_fp64(fc_batch* batch, double* out);
/* Launch shapes; results do not depend on them. */
fc_status fc_batch_set_tuning(fc_batch* batch, const fc_batch_tuning* tuning);
fc_status fc_batch_get_tuning(const fc_batch* batch, fc_batch_tuning* tuning);
/* Times every launch shape over stepsPerCandidate updates of a copy of the
 * current state (restored afterwards), keeps the fastest and returns it;
 * candidateMs (may be NULL) receives the 12 per-step times in the order
 * integratePerThread {1,2,4,8} x collideBlock {64,128,256}. */
fc_status fc_batch_autotune(fc_batch* batch, float deltaTime, uint32_t stepsPerCandidate,
                            fc_batch_tuning* best, float* candidateMs, uint32_t* numCandidates);

/* ------------------------------------------------------------------------- */
/* Learned signed-distance surrogate.  No counterpart in the reference: one  */
/* small MLP per rigid body approximating its signed distance, trained on    */
/* the exact distances of fc_mesh_signed_distance / a distance map, and     */
/* evaluated with BF16 MFMA (FP32 accumulation) to cull far-field boundary   */
/* queries before the exact distance-map query.                              */
/* ------------------------------------------------------------------------- */

/* Network per body (model-frame query q):
 *   x = (q - center) / halfExtent
 *   f = [sin(2 pi B x), cos(2 pi B x)], B = 16 x 3 fixed frequencies
 *   h1 = SiLU(f W1 + b1), h2 = SiLU(h1 W2 + b2), h3 = SiLU(h2 W3 + b3)
 *   distance = (h3 w4 + b4) * halfExtent
 * with widths 32 -> 64 -> 64 -> 64 -> 1.  Parameters per body are
 * fc_sdf_num_params() floats in the order W1[32][64] b1 W2[64][64] b2
 * W3[64][64] b3 w4[64] b4 (W[input][output]), stored with a per-body stride
 * of fc_sdf_param_stride() floats. */
typedef struct fc_sdf fc_sdf;
#define FC_SDF_FP32 0 /* FP32 forward (the training arithmetic) */
#define FC_SDF_BF16 1 /* BF16 MFMA forward, FP32 accumulation   */

uint32_t fc_sdf_num_params(void);
uint32_t fc_sdf_param_stride(void);
/* numBodies networks; centers (3 per body) and halfExtents normalise the
 * inputs; frequencies ~ N(0, freqSigma^2) and Glorot-uniform weights from a
 * SplitMix64 stream seeded with seed; zero biases. */
fc_status fc_sdf_create(fc_context* ctx, uint32_t numBodies, const float* centers,
                        const float* halfExtents, float freqSigma, uint64_t seed, fc_sdf** out);
fc_status fc_sdf_destroy(fc_sdf* sdf);
/* Host copies (any pointer may be NULL): params numBodies * stride floats,
 * freqs numBodies * 48, frames numBodies * 4 (center xyz, halfExtent). */
fc_status fc_sdf_get_params(fc_sdf* sdf, float* params, float* freqs, float* frames);
fc_status fc_sdf_set_params(fc_sdf* sdf, const float* params);
/* Gradient of L_b = 1/n sum_i (y_i - label_i / halfExtent_b)^2 for every
 * body b over its samples (device arrays: points numBodies * samplesPerBody
 * * 3 in the model frame, labels numBodies * samplesPerBody signed
 * distances); grad (host, numBodies * stride) and loss (host, numBodies).
 * FP32 forward/backward on the GPU, deterministic. */
fc_status fc_sdf_loss_grad(fc_sdf* sdf, const float* points, const float* labels,
                           uint32_t samplesPerBody, float* grad, float* loss);
/* Adam (beta 0.9/0.999, eps 1e-8) over minibatches of batchPerBody
 * consecutive samples of every body, all bodies in one launch per step; the
 * learning rate is lr in the first epoch and multiplied by lrDecay after
 * each; lossHistory (host, epochs * numBodies, may be NULL) receives each
 * epoch's mean minibatch loss. */
fc_status fc_sdf_train(fc_sdf* sdf, const float* points, const float* labels,
                       uint32_t samplesPerBody, uint32_t batchPerBody, uint32_t epochs, float lr,
                       float lrDecay, float* lossHistory);
/* Calibration statistics of network `net` against a distance map of its
 * shape (model frame), over the map's lattice vertices v with
 * d(v) < radius + band, the network evaluated by the BF16 MFMA inference of
 * the cull: out[0] = max (y(v) - radius), out[1] = max secant slope
 * |y(v') - y(v)| / |v' - v| between lattice neighbours along x, y and z,
 * out[2] = the number of such vertices, out[3] = max |y(v) - d(v)|.  No
 * reference counterpart (the surrogate is an addition). */
fc_status fc_sdf_map_margin(fc_sdf* sdf, uint32_t net, const fc_distmap* map, float radius,
                            float band, float out[4]);
/* Signed distance of n model-frame points (device, 3 floats each) of the
 * given bodies (device u32 per point; NULL = body 0) into out (device);
 * precision FC_SDF_FP32 or FC_SDF_BF16.  A body id >= numBodies gives NaN. */
fc_status fc_sdf_eval(fc_sdf* sdf, const uint32_t* bodies, const float* points, uint32_t n,
                      float* out, int precision);

/* Rigid bodies of a boundary.  Body b has a distance map in its model frame
 * and a pose: model2world (row-major rotation) and the world position of
 * its model origin; bodyNet[b] (NULL: b) is the surrogate network of its
 * shape (bodies of one shape may share a map and a network).  A coarse
 * lattice (lattice[3] cells of size latticeCell from latticeOrigin,
 * cellBody[(z*ly + y)*lx + x] = body or -1) assigns at most one body to
 * each region; a sphere is tested against the body of its lattice cell.
 * The maps must outlive the set. */
typedef struct fc_body_set fc_body_set;
fc_status fc_body_set_create(fc_context* ctx, uint32_t numBodies, const fc_distmap* const* maps,
                             const float* model2world, const float* origins,
                             const int32_t* bodyNet, const int32_t lattice[3],
                             const float latticeOrigin[3], float latticeCell,
                             const int32_t* cellBody, fc_body_set** out);
fc_status fc_body_set_destroy(fc_body_set* set);
/* Contacts of n spheres (device float4 positions, common radius) with the
 * bodies: contact[i] = 1 when sphere i lies in its body's map box and the
 * map distance (queryMap of the model-frame centre) is < radius, the test of
 * checkDistMap (boundaries.cuh:38-89) per body.  With surrogate == NULL
 * every sphere in a box gets the exact query.  With a surrogate (network
 * bodyNet[b] for body b) the spheres are grouped by network, the BF16 MFMA
 * network predicts every in-box sphere's distance, candidate[i] = 1 marks
 * those within radius + margins[body] (host array, NULL = 0), and only the
 * candidates get the exact query.
 * counts (host, may be NULL; synchronises): spheres in a box, exact queries,
 * contacts. */
fc_status fc_boundary_contacts(fc_context* ctx, fc_body_set* set, fc_sdf* surrogate,
                               const float* margins, const float* pos, uint32_t n, float radius,
                               uint8_t* candidate, uint8_t* contact, uint32_t counts[3]);

/* ------------------------------------------------------------------------- */
/* Rigid bodies moved by the particle step (CollisionPipelineGPU).           */
/* ------------------------------------------------------------------------- */

typedef struct fc_coupling fc_coupling;
/* initParticleSystem (fortrancppinterface/cppinterface.cpp:1262-1307): a
 * particle system with one sphere per rigid body at its centre of mass
 * (host, 3 doubles per body), at rest, with the coupling's parameters
 * (globalDamping 1, gravity (0,0,-9.81), spring 2.75, damping 0.02, shear
 * 0.1, attraction 0; the rest the ParticleSystem defaults).  The bodies'
 * binary64 positions and velocities are kept on the host. */
fc_status fc_coupling_create(fc_context* ctx, uint32_t numBodies, const double* com,
                             fc_uint3 gridSize, fc_coupling** out);
fc_status fc_coupling_destroy(fc_coupling* coupling);
/* The coupling's particle system (for parameters, arrays, boundaries). */
fc_system* fc_coupling_system(fc_coupling* coupling);
/* Time step of fc_coupling_step callers that take it from the coupling
 * (startcollisionpipeline_). */
fc_status fc_coupling_set_time_step(fc_coupling* coupling, double dt);
/* CollisionPipelineGPU::startPipeline (collision/collisionpipelinegpu.cpp:
 * 34-61): subSteps updates of (float)dt / subSteps (the reference uses 10),
 * then every sphere's position and velocity into its body. */
fc_status fc_coupling_step(fc_coupling* coupling, double dt, uint32_t subSteps);
/* A body's position and velocity (host, 3 doubles each; either may be
 * NULL); set also moves or accelerates its sphere. */
fc_status fc_coupling_get_body(const fc_coupling* coupling, uint32_t id, double pos[3],
                               double vel[3]);
fc_status fc_coupling_set_body(fc_coupling* coupling, uint32_t id, const double pos[3],
                               const double vel[3]);
/* The coupling the FEATFLOW entry points of fc_featflow.h act on (NULL
 * unbinds; destroying the bound coupling unbinds it). */
void fc_featflow_bind(fc_coupling* coupling);

/* ------------------------------------------------------------------------- */
/* Output files.                                                              */
/* ------------------------------------------------------------------------- */

/* CVtkWriter::WriteGPUParticleFile (vtkwriter.cpp:3399-3446): legacy ASCII
 * VTK unstructured grid of n sphere centres (host float4 array) with the
 * point scalars "radius" and "myid", every value printed with %f.  The
 * reference writes the radius 1/64 for every sphere: pass radius = 1/64 and
 * radiusFromW = 0 for identical files, or radiusFromW != 0 to write pos.w. */
fc_status fc_write_particle_vtk(const char* path, const float* pos4, uint32_t n, float radius,
                                int radiusFromW);

/* ------------------------------------------------------------------------- */
/* Per-stage timing with HIP events on the context stream.                   */
/* ------------------------------------------------------------------------- */

typedef struct fc_stage_timing {
  char name[32];
  uint64_t launches;
  double total_ms;
} fc_stage_timing;

/* When enabled, every stage launched by fc_system_update/run records a HIP
 * event pair around it on the context stream. */
fc_status fc_context_enable_timing(fc_context* ctx, int enable);
fc_status fc_context_reset_timing(fc_context* ctx);
/* Synchronises, then fills up to max entries; *count receives the number of
 * stages with data. */
fc_status fc_context_get_timing(fc_context* ctx, fc_stage_timing* out, int max, int* count);

/* ------------------------------------------------------------------------- */
/* Diagnostics.                                                              */
/* ------------------------------------------------------------------------- */

/* No counterpart in the reference: copies `bytes` (a multiple of 16) from
 * src to dst `iterations` times with a plain copy kernel (one 16-byte load
 * and store per lane) on the context stream, and returns the average kernel
 * time measured with HIP events.  Calibrates the HBM byte counters of a
 * profiler and measures the achievable copy bandwidth of the device. */
fc_status fc_copy_probe(fc_context* ctx, void* dst, const void* src, size_t bytes, int iterations,
                        float* avgMs);
/* No counterpart in the reference: the cost of separating the stages of a
 * step by a grid-wide barrier inside one persistent (cooperative) launch of
 * `blocks` blocks of 256 threads, against a kernel boundary inside a
 * captured graph of `iterations` such kernels; out[0] = microseconds per
 * grid barrier, out[1] = microseconds per graph kernel.  Measures whether a
 * single persistent launch per batched step could beat the graph-launched
 * stages. */
fc_status fc_sync_probe(fc_context* ctx, uint32_t blocks, int iterations, float out[2]);
/* Lane-layout check of the 16x16x32 BF16 MFMA as the surrogate kernels use
 * it: d = a b for a 16x32 and b 32x16 row-major bf16 matrices, d 16x16
 * row-major f32 (device arrays).  No counterpart in the reference. */
fc_status fc_mfma_probe(fc_context* ctx, const uint16_t* a, const uint16_t* b, float* d);
/* No counterpart in the reference: checks, on every float of their ranges,
 * the correctly rounded square root (arguments in [2^-96, 2^96]) and
 * reciprocal (arguments in [2^-48, 2^48]) that the contact loop uses in
 * place of the general sequences, against sqrtf and 1.0f / x on the
 * device.  out[0] = square-root mismatches, out[1] = reciprocal mismatches
 * (arguments with an all-ones significand excluded; the contact loop takes
 * the general division for those), out[2] = reciprocal mismatches at
 * all-ones significands, out[3] / out[4] = bits of the smallest mismatching
 * square-root / reciprocal argument (0xffffffff if none).  Synchronises. */
fc_status fc_selftest_pair_math(fc_context* ctx, uint64_t out[5]);

#ifdef __cplusplus
}
#endif

#endif /* FC_DEM_H */
