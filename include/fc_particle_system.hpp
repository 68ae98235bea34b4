// fc_particle_system.hpp - C++ host API over the C ABI in fc_dem.h.
//
// fc::ParticleSystem keeps the names and meaning of the reference's host
// class (cuda_addon/particleSystem.h:22-136, particleSystem.cpp:30-509):
// constructor defaults, update(dt), reset(config), setParticles/setVelocity,
// getArray/setArray, the set*/get* parameter accessors, dumpGrid,
// dumpParticles, transferArrays, addSphere and the public host mirrors
// m_hPos/m_hVel.  Parameters are edited on the m_params copy and uploaded
// by update(), as in the reference (particleSystem.cpp:236).
//
// fc::DemPipeline is the phase driver of CollisionPipeline<dem_gpu_unified>
// (inshape3dcore/collision/dempipeline.hpp:74-121): broad phase = calcHash +
// sortParticles, middle phase = reorderDataAndFindCellStart, contact
// problem = collide, integrate = integrateSystem with the system's boundary.
//
// Errors raise fc::Error carrying fc_last_error() (the reference aborts:
// cutilSafeCall, particleSystem.cu:67,180).  Header only; link libfc_dem.so.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "fc_dem.h"

namespace fc {

class Error : public std::runtime_error {
 public:
  Error(fc_status s, const std::string& what) : std::runtime_error(what), status(s) {}
  fc_status status;
};

inline void check(fc_status s) {
  if (s != FC_SUCCESS) throw Error(s, std::string("fc status ") + std::to_string((int)s) + ": " +
                                          fc_last_error());
}

// One device, one HIP stream, one parameter set (replaces cudaInit and the
// process-global SimParams).
class Context {
 public:
  explicit Context(int device = 0) { check(fc_context_create(device, &ctx_)); }
  ~Context() {
    if (ctx_) fc_context_destroy(ctx_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  fc_context* get() const { return ctx_; }
  void sync() const { check(fc_thread_sync(ctx_)); }

 private:
  fc_context* ctx_ = nullptr;
};

// A signed distance map resident on the device (DistanceMap<float,gpu>).
class DistanceMap {
 public:
  // Exact map of a closed triangle mesh on the lattice bmin + i*cellSize.
  static DistanceMap fromMesh(Context& ctx, const std::vector<double>& vertices,
                              const std::vector<uint32_t>& triangles, const int32_t cells[3],
                              float cellSize, const float bmin[3]) {
    fc_distmap* m = nullptr;
    check(fc_distmap_build(ctx.get(), vertices.data(), (uint32_t)(vertices.size() / 3),
                           triangles.data(), (uint32_t)(triangles.size() / 3), cells, cellSize,
                           bmin, &m));
    return DistanceMap(m);
  }
  // Upload of a map in the reference's host layout (transferData).
  static DistanceMap fromArrays(Context& ctx, const int32_t cells[3], float cellSize,
                                const float bmin[3], const float bmax[3], const float* coords,
                                const float* normals, const float* contactPoints,
                                const float* distance) {
    fc_distmap* m = nullptr;
    check(fc_distmap_create(ctx.get(), cells, cellSize, bmin, bmax, coords, normals,
                            contactPoints, distance, &m));
    return DistanceMap(m);
  }
  DistanceMap(DistanceMap&& o) noexcept : m_(std::exchange(o.m_, nullptr)) {}
  DistanceMap& operator=(DistanceMap&& o) noexcept {
    std::swap(m_, o.m_);
    return *this;
  }
  ~DistanceMap() {
    if (m_) fc_distmap_destroy(m_);
  }
  const fc_distmap* get() const { return m_; }

 private:
  explicit DistanceMap(fc_distmap* m) : m_(m) {}
  fc_distmap* m_ = nullptr;
};

class ParticleSystem {
 public:
  enum ParticleConfig { CONFIG_RANDOM = FC_CONFIG_RANDOM, CONFIG_GRID = FC_CONFIG_GRID, _NUM_CONFIGS };
  enum ParticleArray { POSITION = FC_POSITION, VELOCITY = FC_VELOCITY };

  // ParticleSystem(numParticles, gridSize, bUseOpenGL): OpenGL interop is
  // not part of this library; the system uses its own context on device 0
  // unless one is given.
  ParticleSystem(uint32_t numParticles, fc_uint3 gridSize, bool bUseOpenGL = false,
                 Context* ctx = nullptr)
      : m_numParticles(numParticles) {
    if (bUseOpenGL) throw Error(FC_ERROR_NOT_SUPPORTED, "OpenGL interop is not supported");
    if (!ctx) {
      own_ = new Context(0);
      ctx = own_;
    }
    ctx_ = ctx;
    check(fc_system_create(ctx_->get(), numParticles, gridSize, &sys_));
    check(fc_system_get_params(sys_, &m_params));
    m_hPos.assign((size_t)numParticles * 4, 0.0f);
    m_hVel.assign((size_t)numParticles * 4, 0.0f);
  }
  ~ParticleSystem() {
    if (sys_) fc_system_destroy(sys_);
    delete own_;
  }
  ParticleSystem(const ParticleSystem&) = delete;
  ParticleSystem& operator=(const ParticleSystem&) = delete;

  // update(deltaTime): uploads m_params, then integrate -> calcHash -> sort
  // -> reorder -> collide (particleSystem.cpp:222-284).
  void update(float deltaTime) {
    check(fc_system_set_params(sys_, &m_params));
    check(fc_system_update(sys_, deltaTime));
  }
  // steps consecutive update(deltaTime) calls, graph-launched.
  void run(float deltaTime, uint32_t steps) {
    check(fc_system_set_params(sys_, &m_params));
    check(fc_system_run(sys_, deltaTime, steps));
  }

  // reset(config) (particleSystem.cpp:404-444) with the reference's jitter
  // (0.01 r) and zero velocities; the portable generator replaces rand().
  void reset(ParticleConfig config) { reset(config, 1973u, 0.01f, 0.0f); }
  void reset(ParticleConfig config, uint64_t seed, float jitterFrac, float velAmp) {
    check(fc_system_set_params(sys_, &m_params));
    check(fc_system_reset(sys_, (fc_particle_config)config, seed, jitterFrac, velAmp));
  }

  // setParticles (particleSystem.cpp:446-463): w of positions forced to 1,
  // of velocities to 0.
  void setParticles(const float* positions, const float* velocities) {
    for (uint32_t i = 0; i < m_numParticles; ++i) {
      for (int a = 0; a < 3; ++a) {
        m_hPos[4 * i + a] = positions[4 * i + a];
        m_hVel[4 * i + a] = velocities[4 * i + a];
      }
      m_hPos[4 * i + 3] = 1.0f;
      m_hVel[4 * i + 3] = 0.0f;
    }
    setArray(POSITION, m_hPos.data(), 0, (int)m_numParticles);
    setArray(VELOCITY, m_hVel.data(), 0, (int)m_numParticles);
  }
  // setVelocity (particleSystem.cpp:466-477).
  void setVelocity(const float* velocities) {
    for (uint32_t i = 0; i < m_numParticles; ++i) {
      for (int a = 0; a < 3; ++a) m_hVel[4 * i + a] = velocities[4 * i + a];
      m_hVel[4 * i + 3] = 0.0f;
    }
    setArray(VELOCITY, m_hVel.data(), 0, (int)m_numParticles);
  }

  // getArray: copies the device array into m_hPos/m_hVel and returns it.
  float* getArray(ParticleArray array) {
    float* h = array == POSITION ? m_hPos.data() : m_hVel.data();
    check(fc_system_get_array(sys_, (fc_particle_array)array, h));
    return h;
  }
  // setArray: data holds count float4 records for particles start..start+count-1.
  void setArray(ParticleArray array, const float* data, int start, int count) {
    check(fc_system_set_array(sys_, (fc_particle_array)array, data, start, count));
  }

  int getNumParticles() const { return (int)m_numParticles; }

  // dumpGrid (particleSystem.cpp:286-301): prints and returns the largest
  // particle count of a cell.
  uint32_t dumpGrid() {
    uint32_t maxPerCell = 0, occupied = 0;
    check(fc_system_dump_grid(sys_, &maxPerCell, &occupied));
    std::printf("maximum particles per cell = %u\n", maxPerCell);
    return maxPerCell;
  }
  // dumpParticles (particleSystem.cpp:303-314).
  void dumpParticles(uint32_t start, uint32_t count) {
    transferArrays(start, count);
    for (uint32_t i = start; i < start + count && i < m_numParticles; ++i) {
      std::printf("pos: (%.4f, %.4f, %.4f, %.4f)\n", m_hPos[4 * i], m_hPos[4 * i + 1],
                  m_hPos[4 * i + 2], m_hPos[4 * i + 3]);
      std::printf("vel: (%.4f, %.4f, %.4f, %.4f)\n", m_hVel[4 * i], m_hVel[4 * i + 1],
                  m_hVel[4 * i + 2], m_hVel[4 * i + 3]);
    }
  }
  // The particle dump of particles_gpu (CVtkWriter::WriteGPUParticleFile,
  // vtkwriter.cpp:3399-3446) of the current positions; the radius column is
  // the system's radius, or pos.w for polydisperse systems.
  void writeParticleVtk(const char* path, bool radiusFromW = false) {
    getArray(POSITION);
    check(fc_write_particle_vtk(path, m_hPos.data(), m_numParticles, m_params.particleRadius,
                                radiusFromW ? 1 : 0));
  }
  // transferArrays (particleSystem.cpp:316-321): refreshes the host mirrors.
  void transferArrays(uint32_t /*start*/, uint32_t /*count*/) {
    getArray(POSITION);
    getArray(VELOCITY);
  }

  // addSphere (particleSystem.cpp:479-509): a ball of particles of radius r
  // lattice steps at pos, starting at index `start`, jittered by 0.01 r with
  // the portable generator (rand() in the reference).
  void addSphere(int start, const float* pos, const float* vel, int r, float spacing,
                 uint64_t seed = 1973u) {
    uint32_t index = (uint32_t)start;
    uint64_t state = seed;
    auto frand = [&state]() {
      state += 0x9E3779B97F4A7C15ull;
      uint64_t z = state;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      return (float)(z >> 40) * (1.0f / 16777216.0f);
    };
    const float jitter = m_params.particleRadius * 0.01f;
    for (int z = -r; z <= r; ++z)
      for (int y = -r; y <= r; ++y)
        for (int x = -r; x <= r; ++x) {
          const float dx = x * spacing, dy = y * spacing, dz = z * spacing;
          const float l = std::sqrt(dx * dx + dy * dy + dz * dz);
          if (l <= m_params.particleRadius * 2.0f * r && index < m_numParticles) {
            m_hPos[4 * index] = pos[0] + dx + (frand() * 2.0f - 1.0f) * jitter;
            m_hPos[4 * index + 1] = pos[1] + dy + (frand() * 2.0f - 1.0f) * jitter;
            m_hPos[4 * index + 2] = pos[2] + dz + (frand() * 2.0f - 1.0f) * jitter;
            m_hPos[4 * index + 3] = pos[3];
            for (int a = 0; a < 4; ++a) m_hVel[4 * index + a] = vel[a];
            ++index;
          }
        }
    const int count = (int)index - start;
    if (count > 0) {
      setArray(POSITION, m_hPos.data() + 4 * (size_t)start, start, count);
      setArray(VELOCITY, m_hVel.data() + 4 * (size_t)start, start, count);
    }
  }

  // Boundary of the integrate stage: the distance-map boundary of
  // d_integrateSystem (particledem_um.cu:604-638), or the box walls.
  void setBoundary(const DistanceMap* map) {
    check(fc_system_set_boundary(sys_, map ? FC_BOUNDARY_DISTMAP : FC_BOUNDARY_BOX,
                                 map ? map->get() : nullptr));
  }

  // Radii from pos.w (m_params.particleRadius is then the largest radius).
  void setPolydisperse(bool enable) { check(fc_system_set_polydisperse(sys_, enable ? 1 : 0)); }
  // {kinetic energy, px, py, pz} of the current velocities, FP64 on the GPU.
  void statsFp64(double out[4]) { check(fc_system_stats_fp64(sys_, out)); }
  // Spheres per contact count at the last collide (the last bin collects >=).
  std::vector<uint32_t> contactHistogram(uint32_t nbins) {
    std::vector<uint32_t> h(nbins, 0u);
    check(fc_system_contact_histogram(sys_, h.data(), nbins));
    return h;
  }

  void setDamping(float x) { m_params.globalDamping = x; }
  void setGravity(float x) { m_params.gravity = fc_float3{0.0f, 0.0f, x}; }
  void setCollideSpring(float x) { m_params.spring = x; }
  void setCollideDamping(float x) { m_params.damping = x; }
  void setCollideShear(float x) { m_params.shear = x; }
  void setCollideAttraction(float x) { m_params.attraction = x; }
  void setColliderPos(fc_float3 x) { m_params.colliderPos = x; }

  float getParticleRadius() const { return m_params.particleRadius; }
  fc_float3 getColliderPos() const { return m_params.colliderPos; }
  float getColliderRadius() const { return m_params.colliderRadius; }
  fc_uint3 getGridSize() const { return m_params.gridSize; }
  fc_float3 getWorldOrigin() const { return m_params.worldOrigin; }
  fc_float3 getCellSize() const { return m_params.cellSize; }

  fc_system* handle() const { return sys_; }
  Context& context() const { return *ctx_; }

  // Host mirrors (particleSystem.h:84-85), float4 per particle.
  std::vector<float> m_hPos;
  std::vector<float> m_hVel;
  // Parameters uploaded by update()/run()/reset() (particleSystem.h:125).
  fc_sim_params m_params{};

 private:
  uint32_t m_numParticles = 0;
  Context* ctx_ = nullptr;
  Context* own_ = nullptr;
  fc_system* sys_ = nullptr;
};

// Phase driver over a system's device arrays, in the FullC0ntact pipeline
// order: hash + sort, reorder, collide, integrate.
class DemPipeline {
 public:
  DemPipeline(ParticleSystem& ps, float timeStep, const DistanceMap* boundary = nullptr)
      : ps_(ps), dt_(timeStep), map_(boundary) {
    check(fc_system_device_arrays(ps.handle(), &pos_, &vel_, &spos_, &svel_, &hash_, &index_,
                                  &cs_, &ce_));
  }

  void startBroadPhase() {
    upload_params();
    const uint32_t n = (uint32_t)ps_.getNumParticles();
    check(fc_calc_hash(ctx(), hash_, index_, pos_, n));
    check(fc_sort_particles(ctx(), hash_, index_, n));
  }
  void startMiddlePhase() {
    check(fc_reorder_data_and_find_cell_start(ctx(), cs_, ce_, spos_, svel_, hash_, index_, pos_,
                                              vel_, (uint32_t)ps_.getNumParticles(),
                                              ps_.m_params.numCells));
  }
  void startNarrowPhase() {}
  void solveContactProblem() {
    check(fc_collide(ctx(), vel_, spos_, svel_, index_, cs_, ce_,
                     (uint32_t)ps_.getNumParticles(), ps_.m_params.numCells));
  }
  void integrateDynamics() {
    const uint32_t n = (uint32_t)ps_.getNumParticles();
    if (map_)
      check(fc_integrate_system_distmap(ctx(), map_->get(), pos_, vel_, dt_, n));
    else
      check(fc_integrate_system(ctx(), pos_, vel_, dt_, n));
    ps_.context().sync();
  }
  void startPipeline() {
    startBroadPhase();
    startMiddlePhase();
    startNarrowPhase();
    solveContactProblem();
    integrateDynamics();
  }

 private:
  fc_context* ctx() const { return ps_.context().get(); }
  void upload_params() { check(fc_set_parameters(ctx(), &ps_.m_params)); }

  ParticleSystem& ps_;
  float dt_;
  const DistanceMap* map_;
  float* pos_ = nullptr;
  float* vel_ = nullptr;
  float* spos_ = nullptr;
  float* svel_ = nullptr;
  uint32_t* hash_ = nullptr;
  uint32_t* index_ = nullptr;
  uint32_t* cs_ = nullptr;
  uint32_t* ce_ = nullptr;
};

}  // namespace fc
