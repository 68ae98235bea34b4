// Rigid bodies moved by the particle step: the GPU coupling of FullC0ntact's
// rigid-body world (one sphere per body), and the Fortran entry points of
// the FEATFLOW interface that read and write the bodies' state.
//
// The reference keeps the world and the particle system in process-global
// objects (myWorld, myWorld.psystem); the Fortran entry points here act on
// the coupling bound with fc_featflow_bind.  They return nothing, like the
// reference's; on a bad argument they leave their outputs unchanged and
// record the reason for fc_last_error.
#include "fc_internal.hpp"

#include <vector>

struct fc_coupling {
  fc_context* ctx = nullptr;
  fc_system* sys = nullptr;
  uint32_t n = 0;
  double timeStep = 0.01;  // WorldParameters' default time step
  std::vector<double> com, vel;  // per body, 3 each (binary64, as RigidBody)
  std::vector<float> hPos, hVel;
};

namespace {

fc_coupling* g_bound = nullptr;

bool valid_id(const fc_coupling* c, const int* iid) {
  if (!c) {
    fc::set_error(FC_ERROR_INVALID_VALUE, "no coupling bound (fc_featflow_bind)");
    return false;
  }
  if (!iid || *iid < 0 || (uint32_t)*iid >= c->n) {
    fc::set_error(FC_ERROR_INVALID_VALUE, "body id out of range");
    return false;
  }
  return true;
}

}  // namespace

extern "C" {

fc_status fc_coupling_create(fc_context* ctx, uint32_t numBodies, const double* com,
                             fc_uint3 gridSize, fc_coupling** out) {
  if (!ctx || !out || (numBodies > 0 && !com))
    return fc::set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  *out = nullptr;
  fc_system* sys = nullptr;
  FC_TRY(fc_system_create(ctx, numBodies, gridSize, &sys));
  fc_coupling* c = new fc_coupling();
  c->ctx = ctx;
  c->sys = sys;
  c->n = numBodies;
  c->com.assign(com, com + 3 * (size_t)numBodies);
  c->vel.assign(3 * (size_t)numBodies, 0.0);
  c->hPos.assign(4 * (size_t)numBodies, 0.0f);
  c->hVel.assign(4 * (size_t)numBodies, 0.0f);
  // initParticleSystem (cppinterface.cpp:1262-1307): spheres at the bodies'
  // centres of mass, at rest, with the coupling's parameters.
  for (uint32_t i = 0; i < numBodies; ++i) {
    for (int a = 0; a < 3; ++a) c->hPos[4 * (size_t)i + a] = (float)com[3 * (size_t)i + a];
    c->hPos[4 * (size_t)i + 3] = 1.0f;
  }
  fc_sim_params P;
  fc_status s = fc_system_get_params(sys, &P);
  if (s == FC_SUCCESS) {
    P.globalDamping = 1.0f;
    P.gravity = fc_float3{0.0f, 0.0f, -9.81f};
    P.spring = 2.75f;
    P.damping = 0.02f;
    P.shear = 0.1f;
    P.attraction = 0.0f;
    s = fc_system_set_params(sys, &P);
  }
  if (s == FC_SUCCESS && numBodies > 0)
    s = fc_system_set_array(sys, FC_POSITION, c->hPos.data(), 0, (int)numBodies);
  if (s == FC_SUCCESS && numBodies > 0)
    s = fc_system_set_array(sys, FC_VELOCITY, c->hVel.data(), 0, (int)numBodies);
  if (s != FC_SUCCESS) {
    fc_system_destroy(sys);
    delete c;
    return s;
  }
  *out = c;
  return FC_SUCCESS;
}

fc_status fc_coupling_destroy(fc_coupling* c) {
  if (!c) return FC_SUCCESS;
  if (g_bound == c) g_bound = nullptr;
  fc_system_destroy(c->sys);
  delete c;
  return FC_SUCCESS;
}

fc_system* fc_coupling_system(fc_coupling* c) { return c ? c->sys : nullptr; }

fc_status fc_coupling_set_time_step(fc_coupling* c, double dt) {
  if (!c) return fc::set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  c->timeStep = dt;
  return FC_SUCCESS;
}

fc_status fc_coupling_step(fc_coupling* c, double dt, uint32_t subSteps) {
  if (!c) return fc::set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (subSteps == 0) return fc::set_error(FC_ERROR_INVALID_VALUE, "subSteps must be positive");
  // CollisionPipelineGPU::startPipeline (collisionpipelinegpu.cpp:34-61):
  // float dt / float(steps), then that many updates, then the state back
  // into the bodies.
  float sub = (float)dt;
  sub /= (float)subSteps;
  FC_TRY(fc_system_run(c->sys, sub, subSteps));
  if (c->n == 0) return FC_SUCCESS;
  FC_TRY(fc_system_get_array(c->sys, FC_POSITION, c->hPos.data()));
  FC_TRY(fc_system_get_array(c->sys, FC_VELOCITY, c->hVel.data()));
  for (uint32_t i = 0; i < c->n; ++i)
    for (int a = 0; a < 3; ++a) {
      c->com[3 * (size_t)i + a] = c->hPos[4 * (size_t)i + a];
      c->vel[3 * (size_t)i + a] = c->hVel[4 * (size_t)i + a];
    }
  return FC_SUCCESS;
}

fc_status fc_coupling_get_body(const fc_coupling* c, uint32_t id, double pos[3], double vel[3]) {
  if (!c) return fc::set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (id >= c->n) return fc::set_error(FC_ERROR_INVALID_VALUE, "body id out of range");
  for (int a = 0; a < 3; ++a) {
    if (pos) pos[a] = c->com[3 * (size_t)id + a];
    if (vel) vel[a] = c->vel[3 * (size_t)id + a];
  }
  return FC_SUCCESS;
}

fc_status fc_coupling_set_body(fc_coupling* c, uint32_t id, const double pos[3],
                               const double vel[3]) {
  if (!c) return fc::set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (id >= c->n) return fc::set_error(FC_ERROR_INVALID_VALUE, "body id out of range");
  if (pos) {
    float rec[4];
    for (int a = 0; a < 3; ++a) {
      c->com[3 * (size_t)id + a] = pos[a];
      rec[a] = (float)pos[a];
    }
    rec[3] = c->hPos[4 * (size_t)id + 3];  // pos.w is not changed by the step
    FC_TRY(fc_system_set_array(c->sys, FC_POSITION, rec, (int)id, 1));
  }
  if (vel) {
    float rec[4];
    for (int a = 0; a < 3; ++a) {
      c->vel[3 * (size_t)id + a] = vel[a];
      rec[a] = (float)vel[a];
    }
    rec[3] = 0.0f;
    FC_TRY(fc_system_set_array(c->sys, FC_VELOCITY, rec, (int)id, 1));
  }
  return FC_SUCCESS;
}

void fc_featflow_bind(fc_coupling* c) { g_bound = c; }

// update_particle_state_ (cppwrapper_linux.h:28-42, interface_default.hpp:
// 103-115): position, velocity, Euler angles and angular velocity of body
// *iid.  The particle step moves spheres without rotation, so the angles
// and angular velocities are 0.
void update_particle_state_(double* px, double* py, double* pz, double* vx, double* vy,
                            double* vz, double* ax, double* ay, double* az, double* avx,
                            double* avy, double* avz, int* iid) {
  if (!valid_id(g_bound, iid)) return;
  const double* p = &g_bound->com[3 * (size_t)*iid];
  const double* v = &g_bound->vel[3 * (size_t)*iid];
  *px = p[0];
  *py = p[1];
  *pz = p[2];
  *vx = v[0];
  *vy = v[1];
  *vz = v[2];
  *ax = *ay = *az = 0.0;
  *avx = *avy = *avz = 0.0;
}

// getpos_ / getvel_ (cppwrapper_linux.h:717-725, set_get_func.hpp:159-177).
void getpos_(double* dx, double* dy, double* dz, int* iID) {
  if (!valid_id(g_bound, iID)) return;
  const double* p = &g_bound->com[3 * (size_t)*iID];
  *dx = p[0];
  *dy = p[1];
  *dz = p[2];
}

void getvel_(double* dx, double* dy, double* dz, int* iID) {
  if (!valid_id(g_bound, iID)) return;
  const double* v = &g_bound->vel[3 * (size_t)*iID];
  *dx = v[0];
  *dy = v[1];
  *dz = v[2];
}

// setpositionid_ / setvelocityid_ (cppwrapper_linux.h:702-716,
// set_get_func.hpp:72-107): the body's state, and its sphere's.
void setpositionid_(double* dx, double* dy, double* dz, int* iID) {
  if (!valid_id(g_bound, iID)) return;
  const double p[3] = {*dx, *dy, *dz};
  fc_coupling_set_body(g_bound, (uint32_t)*iID, p, nullptr);
}

void setvelocityid_(double* dx, double* dy, double* dz, int* iID) {
  if (!valid_id(g_bound, iID)) return;
  const double v[3] = {*dx, *dy, *dz};
  fc_coupling_set_body(g_bound, (uint32_t)*iID, nullptr, v);
}

// settimestep_ (cppwrapper_linux.h:368-372, set_get_func.hpp:251-259).
void settimestep_(double* dTime) {
  if (!g_bound || !dTime) {
    fc::set_error(FC_ERROR_INVALID_VALUE, "no coupling bound (fc_featflow_bind)");
    return;
  }
  g_bound->timeStep = *dTime;
}

// getnumparticles_ (cppwrapper_linux.h:413-417): the number of bodies.
void getnumparticles_(int* nParts) {
  if (!g_bound || !nParts) {
    fc::set_error(FC_ERROR_INVALID_VALUE, "no coupling bound (fc_featflow_bind)");
    return;
  }
  *nParts = (int)g_bound->n;
}

// getradius_ (cppwrapper_linux.h:418-422): the sphere's radius (pos.w for
// polydisperse systems is not tracked here: the system's radius).
void getradius_(double* drad, int* iid) {
  if (!valid_id(g_bound, iid)) return;
  fc_sim_params P;
  if (fc_system_get_params(g_bound->sys, &P) == FC_SUCCESS) *drad = P.particleRadius;
}

// startcollisionpipeline_ (cppwrapper_linux.h:453-457): one time step of
// the bound coupling in 10 particle updates (collisionpipelinegpu.cpp:40).
void startcollisionpipeline_() {
  if (!g_bound) {
    fc::set_error(FC_ERROR_INVALID_VALUE, "no coupling bound (fc_featflow_bind)");
    return;
  }
  fc_coupling_step(g_bound, g_bound->timeStep, 10);
}

}  // extern "C"
