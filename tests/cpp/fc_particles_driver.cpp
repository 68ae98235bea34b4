// Drives the C++ host API (fc_particle_system.hpp) the way the reference's
// applications drive ParticleSystem and CollisionPipeline<dem_gpu_unified>,
// and writes the final state for the parity tests.
//
//   fc_particles_driver MODE N GRID RADIUS_DEN STEPS DT SEED JITTER VAMP OUT
//
// MODE: update   - STEPS x ParticleSystem::update(DT)
//       run      - ParticleSystem::run(DT, STEPS)
//       pipeline - STEPS x DemPipeline::startPipeline()
//       sphere   - addSphere of radius 3 lattice steps at the origin, then
//                  STEPS updates
// OUT receives positions then velocities, float32, 4 per particle.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fc_particle_system.hpp"

int main(int argc, char** argv) {
  if (argc != 11) {
    std::fprintf(stderr, "usage: %s MODE N GRID RADIUS_DEN STEPS DT SEED JITTER VAMP OUT\n",
                 argv[0]);
    return 2;
  }
  const std::string mode = argv[1];
  const uint32_t n = (uint32_t)std::strtoul(argv[2], nullptr, 10);
  const uint32_t grid = (uint32_t)std::strtoul(argv[3], nullptr, 10);
  const float radius = 1.0f / (float)std::atof(argv[4]);
  const uint32_t steps = (uint32_t)std::strtoul(argv[5], nullptr, 10);
  const float dt = (float)std::atof(argv[6]);
  const uint64_t seed = std::strtoull(argv[7], nullptr, 10);
  const float jitter = (float)std::atof(argv[8]);
  const float vamp = (float)std::atof(argv[9]);
  try {
    fc::ParticleSystem ps(n, fc_uint3{grid, grid, grid});
    ps.m_params.particleRadius = radius;
    const float cell = radius * 2.0f;
    ps.m_params.cellSize = fc_float3{cell, cell, cell};
    std::printf("driver %s n=%u grid=%u steps=%u\n", mode.c_str(), n, grid, steps);
    std::fflush(stdout);
    if (mode == "sphere") {
      ps.reset(fc::ParticleSystem::CONFIG_GRID, seed, jitter, vamp);
      const float p[4] = {0.0f, 0.0f, 0.0f, 1.0f}, v[4] = {0.0f, 0.0f, -0.01f, 0.0f};
      ps.addSphere(0, p, v, 3, radius * 2.0f, seed);
      for (uint32_t s = 0; s < steps; ++s) ps.update(dt);
    } else {
      ps.reset(fc::ParticleSystem::CONFIG_GRID, seed, jitter, vamp);
      if (mode == "update") {
        for (uint32_t s = 0; s < steps; ++s) ps.update(dt);
      } else if (mode == "run") {
        ps.run(dt, steps);
      } else if (mode == "pipeline") {
        fc::DemPipeline pipe(ps, dt);
        for (uint32_t s = 0; s < steps; ++s) pipe.startPipeline();
      } else {
        std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
        return 2;
      }
    }
    const float* pos = ps.getArray(fc::ParticleSystem::POSITION);
    std::vector<float> out(pos, pos + 4 * (size_t)n);
    const float* vel = ps.getArray(fc::ParticleSystem::VELOCITY);
    out.insert(out.end(), vel, vel + 4 * (size_t)n);
    FILE* f = std::fopen(argv[10], "wb");
    if (!f || std::fwrite(out.data(), sizeof(float), out.size(), f) != out.size()) {
      std::fprintf(stderr, "cannot write %s\n", argv[10]);
      return 1;
    }
    std::fclose(f);
    ps.dumpGrid();
  } catch (const fc::Error& e) {
    std::fprintf(stderr, "fc error: %s\n", e.what());
    return 1;
  }
  return 0;
}
