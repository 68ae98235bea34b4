// Host side of the C ABI declared in include/fc_dem.h: contexts, memory
// helpers, per-stage entry points and the particle system object.
#include "fc_internal.hpp"

#include <cmath>
#include <cstring>
#include <algorithm>

namespace fc {

static thread_local std::string g_lastError;

fc_status set_error(fc_status s, const std::string& msg) {
  g_lastError = msg;
  return s;
}

fc_status hip_error(hipError_t e, const char* what) {
  std::string msg = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return set_error(FC_ERROR_OUT_OF_MEMORY, msg);
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
    return set_error(FC_ERROR_NO_DEVICE, msg);
  return set_error(FC_ERROR_HIP, msg);
}

static bool is_pow2(uint32_t v) { return v != 0u && (v & (v - 1u)) == 0u; }

fc_status validate_params(const fc_sim_params& P) {
  if (!is_pow2(P.gridSize.x) || !is_pow2(P.gridSize.y) || !is_pow2(P.gridSize.z))
    return set_error(FC_ERROR_GRID_NOT_POW2,
                     "gridSize components must be powers of two (the hash wraps with & (G-1))");
  uint64_t cells = (uint64_t)P.gridSize.x * P.gridSize.y * P.gridSize.z;
  if (cells > 0xffffffffull)
    return set_error(FC_ERROR_CAPACITY, "gridSize.x*y*z must be below 2^32");
  if (cells != P.numCells)
    return set_error(FC_ERROR_INVALID_VALUE, "numCells must equal gridSize.x*gridSize.y*gridSize.z");
  const float cs[3] = {P.cellSize.x, P.cellSize.y, P.cellSize.z};
  for (float c : cs)
    if (!(c > 0.0f) || !std::isfinite(c))
      return set_error(FC_ERROR_INVALID_VALUE, "cellSize components must be finite and positive");
  return FC_SUCCESS;
}

// ParticleSystem constructor defaults (particleSystem.cpp:30-68).
fc_sim_params default_params(uint32_t numParticles, fc_uint3 gridSize) {
  fc_sim_params P;
  std::memset(&P, 0, sizeof(P));
  P.gridSize = gridSize;
  P.numCells = gridSize.x * gridSize.y * gridSize.z;
  P.numBodies = numParticles;
  P.particleRadius = 1.0f / 16.0f;
  P.colliderPos = fc_float3{-1.2f, -0.8f, 0.8f};
  P.colliderRadius = 0.2f;
  P.worldOrigin = fc_float3{-1.0f, -1.0f, -1.0f};
  const float cell = P.particleRadius * 2.0f;
  P.cellSize = fc_float3{cell, cell, cell};
  P.spring = 0.5f;
  P.damping = 0.02f;
  P.shear = 0.1f;
  P.attraction = 0.0f;
  P.boundaryDamping = -0.5f;
  P.gravity = fc_float3{0.0f, 0.0f, -0.0003f};
  P.globalDamping = 1.0f;
  P.maxParticlesPerCell = 0u;
  return P;
}

// ---------------------------------------------------------------------------
// StageTimer
// ---------------------------------------------------------------------------
int StageTimer::stage_id(const char* name) {
  for (size_t i = 0; i < stages.size(); ++i)
    if (stages[i].name == name) return (int)i;
  stages.push_back(Stage{name, 0, 0.0});
  return (int)stages.size() - 1;
}

hipEvent_t StageTimer::get_event() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void StageTimer::begin(hipStream_t stream, const char* name) {
  if (!enabled) return;
  if (pending.size() > 1024) flush();
  open_stage = stage_id(name);
  open_event = get_event();
  if (open_event) hipEventRecord(open_event, stream);
}

void StageTimer::end(hipStream_t stream) {
  if (!enabled || open_stage < 0 || open_event == nullptr) return;
  hipEvent_t e = get_event();
  if (e == nullptr) return;
  hipEventRecord(e, stream);
  pending.push_back(Pending{open_stage, open_event, e});
  open_stage = -1;
  open_event = nullptr;
}

hipError_t StageTimer::flush() {
  hipError_t err = hipSuccess;
  for (auto& p : pending) {
    hipError_t e = hipEventSynchronize(p.end);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, p.begin, p.end);
    if (e == hipSuccess) {
      stages[p.stage].launches += 1;
      stages[p.stage].total_ms += (double)ms;
    } else {
      err = e;
    }
    pool.push_back(p.begin);
    pool.push_back(p.end);
  }
  pending.clear();
  return err;
}

void StageTimer::reset() {
  flush();
  stages.clear();
}

void StageTimer::destroy() {
  flush();
  for (auto e : pool) hipEventDestroy(e);
  pool.clear();
  stages.clear();
}

// ---------------------------------------------------------------------------
// Sort scratch management
// ---------------------------------------------------------------------------
fc_status ensure_sort_capacity(fc_context* ctx, size_t n) {
  SortScratch& s = ctx->sort;
  size_t tiles = (n + RS_TILE - 1) / RS_TILE;
  if (tiles == 0) tiles = 1;
  if (s.capacity >= n && s.statusCapacity >= tiles * RS_RADIX &&
      s.groupStatusCapacity >= lookback_groups(tiles) * RS_RADIX)
    return FC_SUCCESS;
  FC_HIP_TRY(hipStreamSynchronize(ctx->stream));
  hipFree(s.keysA);
  hipFree(s.valsA);
  hipFree(s.keysB);
  hipFree(s.valsB);
  hipFree(s.status);
  hipFree(s.groupStatus);
  s.keysA = s.valsA = s.keysB = s.valsB = nullptr;
  s.status = s.groupStatus = nullptr;
  s.capacity = s.statusCapacity = s.groupStatusCapacity = 0;
  size_t cap = std::max<size_t>(n, RS_TILE);
  FC_HIP_TRY(hipMalloc(&s.keysA, cap * sizeof(uint32_t)));
  FC_HIP_TRY(hipMalloc(&s.valsA, cap * sizeof(uint32_t)));
  FC_HIP_TRY(hipMalloc(&s.keysB, cap * sizeof(uint32_t)));
  FC_HIP_TRY(hipMalloc(&s.valsB, cap * sizeof(uint32_t)));
  size_t words = ((cap + RS_TILE - 1) / RS_TILE) * RS_RADIX;
  FC_HIP_TRY(hipMalloc(&s.status, words * sizeof(uint64_t)));
  FC_HIP_TRY(hipMemset(s.status, 0, words * sizeof(uint64_t)));
  const size_t gwords = lookback_groups((cap + RS_TILE - 1) / RS_TILE) * RS_RADIX;
  FC_HIP_TRY(hipMalloc(&s.groupStatus, gwords * sizeof(uint64_t)));
  FC_HIP_TRY(hipMemset(s.groupStatus, 0, gwords * sizeof(uint64_t)));
  s.capacity = cap;
  s.statusCapacity = words;
  s.groupStatusCapacity = gwords;
  return FC_SUCCESS;
}

// The context's stage timer when timing is enabled (fc_context_enable_timing),
// so that the per-stage entry points can be timed like a system's stages.
static StageTimer* stage_timer(fc_context* ctx) {
  return ctx->timer.enabled ? &ctx->timer : nullptr;
}

static fc_status sort_pairs(fc_context* ctx, const uint32_t* keysIn, const uint32_t* valsIn,
                            uint32_t* keysOut, uint32_t* valsOut, uint32_t n, uint32_t numCells,
                            StageTimer* timer) {
  if (n == 0) return FC_SUCCESS;
  FC_TRY(ensure_sort_capacity(ctx, n));
  int bits = key_bits_for_cells(numCells);
  hipError_t e = radix_sort_pairs(ctx->stream, ctx->sort, ctx->devFlags, keysIn, valsIn, keysOut,
                                  valsOut, n, bits, timer);
  if (e != hipSuccess) return hip_error(e, "radix_sort_pairs");
  return FC_SUCCESS;
}

fc_status check_device_flags(fc_context* ctx) {
  uint32_t flags = 0;
  FC_HIP_TRY(hipMemcpyAsync(&flags, ctx->devFlags, sizeof(uint32_t), hipMemcpyDeviceToHost,
                            ctx->stream));
  FC_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (flags == 0) return FC_SUCCESS;
  FC_HIP_TRY(hipMemsetAsync(ctx->devFlags, 0, sizeof(uint32_t), ctx->stream));
  FC_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (flags & DEVERR_SORT_STALLED)
    return set_error(FC_ERROR_SORT_STALLED, "radix sort look-back exceeded its spin bound");
  if (flags & DEVERR_HIST_STALLED)
    return set_error(FC_ERROR_SORT_STALLED, "radix sort histogram never completed");
  if (flags & DEVERR_ROW_COUNT)
    return set_error(FC_ERROR_INVALID_VALUE,
                     "row resort: a grid row's spheres and its counts disagree");
  if (flags & DEVERR_RADIUS)
    return set_error(FC_ERROR_INVALID_VALUE,
                     "a sphere radius (pos.w) is not <= particleRadius in polydisperse mode");
  return set_error(FC_ERROR_INVALID_VALUE,
                   "a sort key had bits above log2(numCells): hashes must be < numCells");
}

}  // namespace fc

using namespace fc;

// ===========================================================================
// C ABI
// ===========================================================================
namespace {

// The reference-style collide (collideD; with `force`, evalForces_launcher)
// on the caller's sorted arrays.  Large systems take the row-range kernel
// on the state launch_rows_state builds into the context's buffers (the
// caller's arrays carry neither the sorted hash nor the fixed-point
// coordinates), small ones the row-split kernel on the caller's arrays,
// grids that neither walk takes the cell-table walk.  The buffers grow on
// the first call with more spheres than they hold (an allocation, so not
// inside a graph capture).
fc_status collide_reference_arrays(fc_context* ctx, float* newVel, float* force,
                                   const float* sortedPos, const float* sortedVel,
                                   const uint32_t* index, const uint32_t* cellStart,
                                   const uint32_t* cellEnd, uint32_t n) {
  const fc_sim_params& P = ctx->params;
  const float4* vel = reinterpret_cast<const float4*>(sortedVel);
  GatheredState gs;
  const GatheredState* gathered = nullptr;
  if (collide_rows_state_pays(P, n)) {
    if (ctx->rowsCapacity < n) {
      FC_HIP_TRY(hipStreamSynchronize(ctx->stream));  // earlier calls may still read them
      hipFree(ctx->rowsVel);
      hipFree(ctx->rowsQ);
      ctx->rowsVel = nullptr;
      ctx->rowsQ = nullptr;
      ctx->rowsCapacity = 0;
      FC_HIP_TRY(hipMalloc(&ctx->rowsVel, (size_t)n * sizeof(float4)));
      FC_HIP_TRY(hipMalloc(&ctx->rowsQ, ((size_t)n + kSortedPosPad) * sizeof(uint2)));
      ctx->rowsCapacity = n;
    }
    gs.sortedQ = ctx->rowsQ;
    gs.frame = quant_frame_for(false);  // integrate's box walls
    FC_HIP_TRY(launch_rows_state(ctx->stream, P, gs.frame, (const float4*)sortedPos, vel, n,
                                 ctx->rowsVel, ctx->rowsQ));
    vel = ctx->rowsVel;
    gathered = &gs;
  }
  const hipError_t e =
      launch_collide(ctx->stream, P, false, ctx->devFlags, (float4*)newVel,
                     (const float4*)sortedPos, vel, index, cellStart, cellEnd, n, nullptr,
                     gathered, force);
  if (e != hipSuccess) return hip_error(e, force ? "evalForces" : "collide");
  return FC_SUCCESS;
}

}  // namespace

extern "C" {

const char* fc_last_error(void) { return g_lastError.c_str(); }
int fc_abi_version(void) { return FC_DEM_ABI_VERSION; }

int fc_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

fc_status fc_context_create(int device, fc_context** out) {
  if (!out) return set_error(FC_ERROR_INVALID_VALUE, "out is NULL");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return set_error(FC_ERROR_NO_DEVICE, "no HIP device is visible");
  if (device >= count) return set_error(FC_ERROR_NO_DEVICE, "device index out of range");
  if (device >= 0) FC_HIP_TRY(hipSetDevice(device));
  fc_context* ctx = new fc_context();
  FC_HIP_TRY(hipGetDevice(&ctx->device));
  {
    const fc_status st = init_noalloc_scratch(ctx->device);
    if (st != FC_SUCCESS) {
      delete ctx;
      return st;
    }
  }
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return hip_error(e, "hipStreamCreateWithFlags");
  }
  ctx->params = default_params(0, fc_uint3{64, 64, 64});
  ctx->paramsSet = true;
  SortScratch& s = ctx->sort;
  bool ok = hipMalloc(&ctx->devFlags, sizeof(uint32_t)) == hipSuccess &&
            hipMalloc(&s.counters, RSC_COUNT * sizeof(uint32_t)) == hipSuccess &&
            hipMalloc(&s.histogram, (size_t)RS_HIST_COPIES * RS_MAX_PASSES * RS_RADIX *
                                        sizeof(uint32_t)) == hipSuccess &&
            hipMemset(s.histogram, 0, (size_t)RS_HIST_COPIES * RS_MAX_PASSES * RS_RADIX *
                                          sizeof(uint32_t)) == hipSuccess &&
            hipMemset(ctx->devFlags, 0, sizeof(uint32_t)) == hipSuccess &&
            hipMemset(s.counters, 0, RSC_COUNT * sizeof(uint32_t)) == hipSuccess;
  if (!ok) {
    fc_context_destroy(ctx);
    return set_error(FC_ERROR_OUT_OF_MEMORY, "context allocation failed");
  }
  *out = ctx;
  return FC_SUCCESS;
}

fc_status fc_context_destroy(fc_context* ctx) {
  if (!ctx) return FC_SUCCESS;
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  ctx->timer.destroy();
  SortScratch& s = ctx->sort;
  hipFree(s.keysA);
  hipFree(s.valsA);
  hipFree(s.keysB);
  hipFree(s.valsB);
  hipFree(s.status);
  hipFree(s.groupStatus);
  hipFree(s.histogram);
  hipFree(s.counters);
  hipFree(ctx->devFlags);
  hipFree(ctx->statsScratch);
  hipFree(ctx->rowsVel);
  hipFree(ctx->rowsQ);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return FC_SUCCESS;
}

fc_status fc_allocate_array(void** devPtr, size_t size) {
  if (!devPtr) return set_error(FC_ERROR_INVALID_VALUE, "devPtr is NULL");
  *devPtr = nullptr;
  if (size == 0) return FC_SUCCESS;
  FC_HIP_TRY(hipMalloc(devPtr, size));
  return FC_SUCCESS;
}

fc_status fc_free_array(void* devPtr) {
  if (!devPtr) return FC_SUCCESS;
  FC_HIP_TRY(hipFree(devPtr));
  return FC_SUCCESS;
}

fc_status fc_thread_sync(fc_context* ctx) {
  if (!ctx) return set_error(FC_ERROR_INVALID_VALUE, "ctx is NULL");
  FC_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return check_device_flags(ctx);
}

fc_status fc_copy_array_to_device(fc_context* ctx, void* device, const void* host, size_t offset,
                                  size_t size) {
  if (!ctx) return set_error(FC_ERROR_INVALID_VALUE, "ctx is NULL");
  if (size == 0) return FC_SUCCESS;
  if (!device || !host) return set_error(FC_ERROR_INVALID_VALUE, "NULL pointer");
  FC_HIP_TRY(hipMemcpyAsync((char*)device + offset, host, size, hipMemcpyHostToDevice,
                            ctx->stream));
  FC_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return FC_SUCCESS;
}

fc_status fc_copy_array_from_device(fc_context* ctx, void* host, const void* device, size_t offset,
                                    size_t size) {
  if (!ctx) return set_error(FC_ERROR_INVALID_VALUE, "ctx is NULL");
  if (size == 0) return FC_SUCCESS;
  if (!device || !host) return set_error(FC_ERROR_INVALID_VALUE, "NULL pointer");
  FC_HIP_TRY(hipMemcpyAsync(host, (const char*)device + offset, size, hipMemcpyDeviceToHost,
                            ctx->stream));
  FC_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return FC_SUCCESS;
}

fc_status fc_set_parameters(fc_context* ctx, const fc_sim_params* params) {
  if (!ctx || !params) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  FC_TRY(validate_params(*params));
  ctx->params = *params;
  ctx->paramsSet = true;
  return FC_SUCCESS;
}

fc_status fc_get_parameters(const fc_context* ctx, fc_sim_params* params) {
  if (!ctx || !params) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  *params = ctx->params;
  return FC_SUCCESS;
}

fc_status fc_integrate_system(fc_context* ctx, float* pos, float* vel, float deltaTime,
                              uint32_t numParticles) {
  if (!ctx) return set_error(FC_ERROR_INVALID_VALUE, "ctx is NULL");
  if (numParticles == 0) return FC_SUCCESS;
  if (!pos || !vel) return set_error(FC_ERROR_INVALID_VALUE, "NULL array");
  StageTimer* t = stage_timer(ctx);
  if (t) t->begin(ctx->stream, "drop_in_integrate");
  hipError_t e = launch_integrate(ctx->stream, ctx->params, (float4*)pos, (float4*)vel, deltaTime,
                                  numParticles);
  if (t) t->end(ctx->stream);
  if (e != hipSuccess) return hip_error(e, "integrate");
  return FC_SUCCESS;
}

fc_status fc_calc_hash(fc_context* ctx, uint32_t* hash, uint32_t* index, const float* pos,
                       uint32_t numParticles) {
  if (!ctx) return set_error(FC_ERROR_INVALID_VALUE, "ctx is NULL");
  if (numParticles == 0) return FC_SUCCESS;
  if (!hash || !index || !pos) return set_error(FC_ERROR_INVALID_VALUE, "NULL array");
  StageTimer* t = stage_timer(ctx);
  if (t) t->begin(ctx->stream, "drop_in_calc_hash");
  hipError_t e = launch_calc_hash(ctx->stream, ctx->params, hash, index, (const float4*)pos,
                                  numParticles);
  if (t) t->end(ctx->stream);
  if (e != hipSuccess) return hip_error(e, "calcHash");
  return FC_SUCCESS;
}

fc_status fc_sort_particles(fc_context* ctx, uint32_t* hash, uint32_t* index,
                            uint32_t numParticles) {
  if (!ctx) return set_error(FC_ERROR_INVALID_VALUE, "ctx is NULL");
  if (numParticles == 0) return FC_SUCCESS;
  if (!hash || !index) return set_error(FC_ERROR_INVALID_VALUE, "NULL array");
  return sort_pairs(ctx, hash, index, hash, index, numParticles, ctx->params.numCells,
                    stage_timer(ctx));
}

fc_status fc_reorder_data_and_find_cell_start(fc_context* ctx, uint32_t* cellStart,
                                              uint32_t* cellEnd, float* sortedPos,
                                              float* sortedVel, const uint32_t* hash,
                                              const uint32_t* index, const float* oldPos,
                                              const float* oldVel, uint32_t numParticles,
                                              uint32_t numCells) {
  if (!ctx) return set_error(FC_ERROR_INVALID_VALUE, "ctx is NULL");
  if (numCells > 0 && (!cellStart || !cellEnd))
    return set_error(FC_ERROR_INVALID_VALUE, "NULL cell table");
  StageTimer* t = stage_timer(ctx);
  if (t) t->begin(ctx->stream, "drop_in_clear_tables");
  hipError_t e = launch_fill_u32(ctx->stream, cellStart, CELL_EMPTY, numCells, cellEnd);
  if (t) t->end(ctx->stream);
  if (e != hipSuccess) return hip_error(e, "clear cell tables");
  if (numParticles == 0) return FC_SUCCESS;
  if (!sortedPos || !sortedVel || !hash || !index || !oldPos || !oldVel)
    return set_error(FC_ERROR_INVALID_VALUE, "NULL array");
  if (t) t->begin(ctx->stream, "drop_in_reorder");
  e = launch_reorder(ctx->stream, cellStart, cellEnd, (float4*)sortedPos, (float4*)sortedVel, hash,
                     index, (const float4*)oldPos, (const float4*)oldVel, numParticles, numCells);
  if (t) t->end(ctx->stream);
  if (e != hipSuccess) return hip_error(e, "reorderDataAndFindCellStart");
  return FC_SUCCESS;
}

fc_status fc_collide(fc_context* ctx, float* newVel, const float* sortedPos, const float* sortedVel,
                     const uint32_t* index, const uint32_t* cellStart, const uint32_t* cellEnd,
                     uint32_t numParticles, uint32_t numCells) {
  if (!ctx) return set_error(FC_ERROR_INVALID_VALUE, "ctx is NULL");
  if (numParticles == 0) return FC_SUCCESS;
  if (!newVel || !sortedPos || !sortedVel || !index || !cellStart || !cellEnd)
    return set_error(FC_ERROR_INVALID_VALUE, "NULL array");
  if (numCells != ctx->params.numCells)
    return set_error(FC_ERROR_INVALID_VALUE, "numCells differs from the context parameters");
  StageTimer* t = stage_timer(ctx);
  if (t) t->begin(ctx->stream, "drop_in_collide");
  const fc_status st = collide_reference_arrays(ctx, newVel, nullptr, sortedPos, sortedVel, index,
                                                cellStart, cellEnd, numParticles);
  if (t) t->end(ctx->stream);
  return st;
}

fc_status fc_eval_forces(fc_context* ctx, float* newVel, float* force, const float* sortedPos,
                         const float* sortedVel, const uint32_t* index, const uint32_t* cellStart,
                         const uint32_t* cellEnd, uint32_t numParticles, uint32_t numCells) {
  if (!ctx) return set_error(FC_ERROR_INVALID_VALUE, "ctx is NULL");
  if (numParticles == 0) return FC_SUCCESS;
  if (!newVel || !force || !sortedPos || !sortedVel || !index || !cellStart || !cellEnd)
    return set_error(FC_ERROR_INVALID_VALUE, "NULL array");
  if (numCells != ctx->params.numCells)
    return set_error(FC_ERROR_INVALID_VALUE, "numCells differs from the context parameters");
  return collide_reference_arrays(ctx, newVel, force, sortedPos, sortedVel, index, cellStart,
                                  cellEnd, numParticles);
}

fc_status fc_context_enable_timing(fc_context* ctx, int enable) {
  if (!ctx) return set_error(FC_ERROR_INVALID_VALUE, "ctx is NULL");
  ctx->timer.enabled = enable != 0;
  return FC_SUCCESS;
}

fc_status fc_context_reset_timing(fc_context* ctx) {
  if (!ctx) return set_error(FC_ERROR_INVALID_VALUE, "ctx is NULL");
  ctx->timer.reset();
  return FC_SUCCESS;
}

fc_status fc_context_get_timing(fc_context* ctx, fc_stage_timing* out, int max, int* count) {
  if (!ctx || !count) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  hipError_t e = ctx->timer.flush();
  if (e != hipSuccess) return hip_error(e, "timing flush");
  int n = (int)ctx->timer.stages.size();
  *count = n;
  for (int i = 0; i < n && i < max && out; ++i) {
    std::memset(out[i].name, 0, sizeof(out[i].name));
    std::strncpy(out[i].name, ctx->timer.stages[i].name.c_str(), sizeof(out[i].name) - 1);
    out[i].launches = ctx->timer.stages[i].launches;
    out[i].total_ms = ctx->timer.stages[i].total_ms;
  }
  return FC_SUCCESS;
}

fc_status fc_copy_probe(fc_context* ctx, void* dst, const void* src, size_t bytes, int iterations,
                        float* avgMs) {
  if (!ctx || !dst || !src || !avgMs || iterations < 1 || bytes % 16 != 0)
    return set_error(FC_ERROR_INVALID_VALUE, "bad copy-probe arguments");
  hipEvent_t a = nullptr, b = nullptr;
  FC_HIP_TRY(hipEventCreate(&a));
  FC_HIP_TRY(hipEventCreate(&b));
  hipError_t e = launch_copy16(ctx->stream, dst, src, bytes);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(a, ctx->stream);
  for (int i = 0; i < iterations && e == hipSuccess; ++i)
    e = launch_copy16(ctx->stream, dst, src, bytes);
  if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (e != hipSuccess) return hip_error(e, "copy probe");
  *avgMs = ms / (float)iterations;
  return FC_SUCCESS;
}

}  // extern "C"
