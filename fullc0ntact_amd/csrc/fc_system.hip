// The particle-system object behind fc_system_* (ParticleSystem,
// particleSystem.h:22-136 / particleSystem.cpp:30-509): buffer ownership,
// initial configurations, update(dt) and graph-launched multi-step runs.
#include "fc_internal.hpp"
#include "row_resort.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace fc {

namespace {

// Steps per launch of the multi-step graph of a run (the kernel boundaries
// inside a graph cost less than the boundary between two graph launches).
#ifndef FC_GRAPH_STEPS
#define FC_GRAPH_STEPS 8
#endif
constexpr uint32_t kGraphSteps = FC_GRAPH_STEPS;

// SplitMix64 (Steele, Lea, Flood 2014).  Portable across compilers and
// hosts, unlike srand/rand (particleSystem.cpp:375-383).
struct SplitMix64 {
  uint64_t state;
  explicit SplitMix64(uint64_t seed) : state(seed) {}
  uint64_t next() {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // 24 random bits scaled to [0, 1): exactly representable floats.
  float uniform01() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
  // U(-1, 1) as 2u - 1 (the reference's frand()*2-1).
  float symmetric() { return uniform01() * 2.0f - 1.0f; }
};

uint32_t lattice_side(uint32_t n) {
  uint32_t s = 0;
  while ((uint64_t)s * s * s < n) ++s;
  return s;
}

void free_system_buffers(fc_system* sys) {
  hipFree(sys->pos);
  hipFree(sys->vel);
  hipFree(sys->sortedPos);
  hipFree(sys->sortedVel);
  hipFree(sys->sortedQ);
  hipFree(sys->hash);
  hipFree(sys->index);
  hipFree(sys->keys);
  hipFree(sys->slotOf);
  hipFree(sys->idOfSlot);
  hipFree(sys->cellStart);
  hipFree(sys->cellEnd);
  sys->pos = sys->vel = sys->sortedPos = sys->sortedVel = nullptr;
  sys->sortedQ = nullptr;
  sys->hash = sys->index = sys->keys = sys->slotOf = sys->idOfSlot = nullptr;
  sys->cellStart = sys->cellEnd = nullptr;
  free_row_tables(sys->rowTables);
}

void drop_graph(fc_system* sys) {
  for (int i = 0; i < 2; ++i) {
    if (sys->graphExec[i]) hipGraphExecDestroy(sys->graphExec[i]);
    if (sys->graph[i]) hipGraphDestroy(sys->graph[i]);
    sys->graphExec[i] = nullptr;
    sys->graph[i] = nullptr;
  }
}

fc_status alloc_tables(fc_system* sys, uint32_t numCells) {
  hipFree(sys->cellStart);
  hipFree(sys->cellEnd);
  sys->cellStart = sys->cellEnd = nullptr;
  FC_HIP_TRY(hipMalloc(&sys->cellStart, (size_t)numCells * sizeof(uint32_t)));
  FC_HIP_TRY(hipMalloc(&sys->cellEnd, (size_t)numCells * sizeof(uint32_t)));
  sys->tablesTrackHash = false;
  return FC_SUCCESS;
}

// One update(dt) (particleSystem.cpp:222-284): integrate -> calcHash ->
// sort -> reorder/findCellStart -> collide, enqueued on the context stream.
// Nothing here allocates or synchronises, so it can be stream-captured.
fc_status enqueue_update(fc_system* sys, float dt, StageTimer* timer) {
  fc_context* ctx = sys->ctx;
  hipStream_t st = ctx->stream;
  const fc_sim_params& P = sys->params;
  const uint32_t n = sys->n;
  hipError_t e = hipSuccess;
  const uint32_t* prev = nullptr;
  if (sys->tablesTrackHash) {
    prev = sys->hash;
  } else {
    if (timer) timer->begin(st, "clear_tables");
    e = launch_fill_u32(st, sys->cellStart, CELL_EMPTY, P.numCells, sys->cellEnd);
    if (timer) timer->end(st);
    if (e != hipSuccess) return hip_error(e, "clear cell tables");
  }
  if (n == 0) {
    sys->tablesTrackHash = false;
    return FC_SUCCESS;
  }
  const int bits = key_bits_for_cells(P.numCells);
  if (timer) timer->begin(st, "integrate_hash");
  e = launch_integrate_hash(st, P, sys->boundary_arg(), sys->pos, sys->vel, dt, sys->keys, prev,
                            sys->cellStart, sys->cellEnd, n, ctx->sort, radix_passes(bits, false),
                            ctx->devFlags);
  if (timer) timer->end(st);
  if (e != hipSuccess) return hip_error(e, "integrate_hash");
  // The reorder's gather rides on the sort's last pass; its cell tables
  // follow in their own launch.
  SortGather ga;
  ga.pos = sys->pos;
  ga.vel = sys->vel;
  ga.sortedPos = sys->sortedPos;
  ga.sortedVel = sys->sortedVel;
  ga.sortedQ = sys->sortedQ;
  ga.frame = sys->quant_frame();
  if (n <= (uint32_t)RS_TILE) {
    if (timer) timer->begin(st, "sort_small");
    e = launch_sort_small(st, ctx->sort, sys->keys, nullptr, sys->hash, sys->index, n,
                          radix_passes(bits, false), nullptr, sys->cellStart, sys->cellEnd,
                          P.numCells, ga);
    if (timer) timer->end(st);
    if (e != hipSuccess) return hip_error(e, "sort");
  } else {
    e = radix_sort_pairs(st, ctx->sort, ctx->devFlags, sys->keys, nullptr, sys->hash, sys->index,
                         n, bits, timer, /*histogramReady=*/true, 1, nullptr, nullptr, nullptr,
                         &ga);
    if (e != hipSuccess) return hip_error(e, "sort");
    if (timer) timer->begin(st, "cell_bounds");
    e = launch_cell_bounds(st, sys->cellStart, sys->cellEnd, sys->hash, n, P.numCells);
    if (timer) timer->end(st);
    if (e != hipSuccess) return hip_error(e, "cell bounds");
  }
  if (timer) timer->begin(st, "collide");
  const GatheredState gs{sys->sortedQ, ga.frame};
  e = launch_collide(st, P, sys->polydisperse != 0, ctx->devFlags, sys->vel, sys->sortedPos,
                     sys->sortedVel, sys->index, sys->cellStart, sys->cellEnd, n, nullptr, &gs);
  if (timer) timer->end(st);
  if (e != hipSuccess) return hip_error(e, "collide");
  sys->tablesTrackHash = true;
  return FC_SUCCESS;
}

// Multi-step runs chain the updates so that each collide also performs the
// next update's integrate + hash + histogram (launch_collide_fused) and the
// next sort's first pass clears the previous update's cells.  One run of
// `steps` updates is
//   head:   [integrate+hash+hist] sort reorder [collide+next integrate]
//   middle: sort(+clear) reorder [collide+next integrate]    x (steps - 2)
//   tail:   sort(+clear) reorder [collide]
// which computes exactly what `steps` separate updates compute.
// Between the fused steps the particle state stays in the previous sort's
// order: the fused collide writes pos/vel at each particle's sorted slot k
// (coalesced) and idOfSlot[k] = id.  For large systems (run_row_mode) it
// stores the next hash at the slot too, with the per-row counts, and the
// next sort is a row resort (row_resort.hip), whose result is the (cell,
// id) order of the separate updates; the run's first sort (keys by id,
// from the head) is a radix sort whose cell-table pass also records the
// row starts.  Otherwise it stores the hash and slotOf[id] = k at the
// particle's id and the radix sort carries the slots as its values, in id
// order, so its input and tie order are those of the separate updates.
// Either way the next order's state is gathered from the sorted slots
// (nearly in order) with the ids through idOfSlot.  The tail's collide
// scatters the positions (with the velocities) back to id order.
int run_passes(const fc_sim_params& P) {
  const int p = radix_passes(key_bits_for_cells(P.numCells), false);
  return p < 2 ? 2 : p;
}

// Whether a run's steps after its first sort resort the spheres row by row
// (row_resort.hip) instead of radix-sorting them: then the fused collide
// stores the next keys at the sorted slots.
bool run_row_mode(const fc_system* sys) { return collide_slot_keys(sys->params, sys->n); }

fc_status enqueue_run_head(fc_system* sys, float dt, StageTimer* timer) {
  fc_context* ctx = sys->ctx;
  hipStream_t st = ctx->stream;
  const fc_sim_params& P = sys->params;
  const uint32_t n = sys->n;
  const int passes = run_passes(P);
  hipError_t e = hipSuccess;
  const uint32_t* prev = nullptr;
  if (sys->tablesTrackHash) {
    prev = sys->hash;
  } else {
    if (timer) timer->begin(st, "clear_tables");
    e = launch_fill_u32(st, sys->cellStart, CELL_EMPTY, P.numCells, sys->cellEnd);
    if (timer) timer->end(st);
    if (e != hipSuccess) return hip_error(e, "clear cell tables");
  }
  if (run_row_mode(sys)) {
    // the row counts start from zero (a run that stopped early may have
    // left counts behind)
    e = launch_row_clear(st, sys->rowTables);
    if (e != hipSuccess) return hip_error(e, "clear row counts");
  }
  if (timer) timer->begin(st, "integrate_hash");
  e = launch_integrate_hash(st, P, sys->boundary_arg(), sys->pos, sys->vel, dt, sys->keys, prev,
                            sys->cellStart, sys->cellEnd, n, ctx->sort, passes, ctx->devFlags);
  if (timer) timer->end(st);
  if (e != hipSuccess) return hip_error(e, "integrate_hash");
  sys->tablesTrackHash = true;
  return FC_SUCCESS;
}

// sort (clearing the previous cells in its first pass unless the head just
// did) + reorder + collide, fused with the next integrate unless `last`.
// fromSlots: pos/vel hold the sorted-slot layout of the previous fused
// collide (every body but the first after the head).
fc_status enqueue_run_body(fc_system* sys, float dt, bool clearInSort, bool last, bool fromSlots,
                           StageTimer* timer) {
  fc_context* ctx = sys->ctx;
  hipStream_t st = ctx->stream;
  const fc_sim_params& P = sys->params;
  const uint32_t n = sys->n;
  const int passes = run_passes(P);
  // The sort's last pass gathers the state into sorted order (at the
  // sorted slots, writing the ids through idOfSlot, when fromSlots).
  SortGather ga;
  ga.pos = sys->pos;
  ga.vel = sys->vel;
  ga.sortedPos = sys->sortedPos;
  ga.sortedVel = sys->sortedVel;
  ga.sortedQ = sys->sortedQ;
  ga.frame = sys->quant_frame();
  if (fromSlots) {
    ga.idOfSlot = sys->idOfSlot;
    ga.indexOut = sys->index;
  }
  // rowMode: the fused collide stores the next keys at the sorted slots
  // with per-row counts; keysAtSlots: this sort's keys were stored that way,
  // so it is a row resort.
  const bool rowMode = run_row_mode(sys);
  const bool keysAtSlots = fromSlots && rowMode;
  const RowHists rh = rowMode ? row_hists(sys->rowTables) : RowHists{};
  const uint32_t* valsIn = (fromSlots && !keysAtSlots) ? sys->slotOf : nullptr;
  uint32_t* valsOut = fromSlots ? sys->slotOf : sys->index;
  const uint32_t* clearHash = clearInSort ? sys->hash : nullptr;
  hipError_t e;
  if (rowMode && keysAtSlots) {
    if (timer) timer->begin(st, "row_scan");
    e = launch_row_scan(st, sys->rowTables, ctx->devFlags);
    if (timer) timer->end(st);
    if (e != hipSuccess) return hip_error(e, "row scan");
    if (timer) timer->begin(st, "row_changers");
    e = launch_row_changers(st, sys->rowTables, sys->keys, sys->idOfSlot, sys->hash,
                            sys->cellStart, sys->cellEnd, n, P.numCells, ctx->devFlags);
    if (timer) timer->end(st);
    if (e != hipSuccess) return hip_error(e, "row changers");
    RowOutputs O;
    O.hash = sys->hash;
    O.index = sys->index;
    O.pos = sys->pos;
    O.vel = sys->vel;
    O.sortedPos = sys->sortedPos;
    O.sortedVel = sys->sortedVel;
    O.sortedQ = sys->sortedQ;
    O.cellStart = sys->cellStart;
    O.cellEnd = sys->cellEnd;
    O.frame = ga.frame;
    if (timer) timer->begin(st, "row_sort");
    e = launch_row_sort(st, sys->rowTables, sys->keys, sys->idOfSlot, O, ctx->devFlags);
    if (timer) timer->end(st);
    if (e != hipSuccess) return hip_error(e, "row sort");
  } else if (n <= (uint32_t)RS_TILE) {
    // One block sorts, gathers and writes the cell tables.
    if (timer) timer->begin(st, "sort_small");
    e = launch_sort_small(st, ctx->sort, sys->keys, valsIn, sys->hash, valsOut, n, passes,
                          clearHash, sys->cellStart, sys->cellEnd, P.numCells, ga);
    if (timer) timer->end(st);
    if (e != hipSuccess) return hip_error(e, "sort");
  } else {
    e = radix_sort_pairs(st, ctx->sort, ctx->devFlags, sys->keys, valsIn, sys->hash, valsOut, n,
                         key_bits_for_cells(P.numCells), timer, true, passes, clearHash,
                         sys->cellStart, sys->cellEnd, &ga);
    if (e != hipSuccess) return hip_error(e, "sort");
    if (timer) timer->begin(st, "cell_bounds");
    e = launch_cell_bounds(st, sys->cellStart, sys->cellEnd, sys->hash, n, P.numCells);
    // the next step's row resort starts from this order's rows
    if (e == hipSuccess && rowMode) e = launch_row_starts(st, sys->rowTables, sys->hash, n);
    if (timer) timer->end(st);
    if (e != hipSuccess) return hip_error(e, "cell bounds");
  }
  if (timer) timer->begin(st, last ? "collide" : "collide_integrate");
  const GatheredState gs{sys->sortedQ, ga.frame};
  if (last)
    e = launch_collide(st, P, sys->polydisperse != 0, ctx->devFlags, sys->vel, sys->sortedPos,
                       sys->sortedVel, sys->index, sys->cellStart, sys->cellEnd, n,
                       fromSlots ? sys->pos : nullptr, &gs);
  else
    e = launch_collide_fused(st, P, sys->boundary_arg(), sys->pos, sys->vel, sys->sortedPos,
                             sys->sortedVel, sys->index, sys->cellStart, sys->cellEnd, n, dt,
                             sys->keys, ctx->sort, passes, ctx->devFlags, sys->slotOf,
                             sys->idOfSlot, &gs, rowMode, rowMode ? &rh : nullptr);
  if (timer) timer->end(st);
  if (e != hipSuccess) return hip_error(e, "collide");
  return FC_SUCCESS;
}

// Captures `steps` steady-state run bodies back to back into graph[slot].
fc_status capture_steps(fc_system* sys, float dt, int slot, uint32_t steps) {
  hipStream_t st = sys->ctx->stream;
  FC_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  fc_status s = FC_SUCCESS;
  for (uint32_t i = 0; i < steps && s == FC_SUCCESS; ++i)
    s = enqueue_run_body(sys, dt, /*clearInSort=*/true, /*last=*/false, /*fromSlots=*/true,
                         nullptr);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(st, &g);
  if (s != FC_SUCCESS) {
    if (g) hipGraphDestroy(g);
    return s;
  }
  if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
  sys->graph[slot] = g;
  e = hipGraphInstantiate(&sys->graphExec[slot], g, nullptr, nullptr, 0);
  if (e != hipSuccess) return hip_error(e, "hipGraphInstantiate");
  return FC_SUCCESS;
}

fc_status capture_graph(fc_system* sys, float dt) {
  drop_graph(sys);
  fc_status s = capture_steps(sys, dt, 0, 1);
  if (s == FC_SUCCESS) s = capture_steps(sys, dt, 1, kGraphSteps);
  if (s != FC_SUCCESS) {
    drop_graph(sys);
    return s;
  }
  sys->graphDt = dt;
  sys->graphParams = sys->params;
  return FC_SUCCESS;
}

}  // namespace

// reset(CONFIG_GRID / CONFIG_RANDOM) (particleSystem.cpp:380-444) into host
// float4 arrays with a SplitMix64 stream: CONFIG_GRID is the s^3 lattice of
// spacing 2r from -1+r (s = smallest integer with s^3 >= n) jittered by
// U(-jitterFrac r, jitterFrac r) per axis; CONFIG_RANDOM is U(-1,1)^3;
// velocities U(-velAmp, velAmp); pos.w = 1, vel.w = 0.
void generate_initial_state(float* pos, float* vel, uint32_t n, fc_particle_config config,
                            uint64_t seed, float r, float jitterFrac, float velAmp) {
  SplitMix64 rng(seed);
  if (config == FC_CONFIG_GRID) {
    const uint32_t s = lattice_side(n);
    const float spacing = r * 2.0f;
    const float jitter = r * jitterFrac;
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t x = i % s, y = (i / s) % s, z = i / (s * s);
      const uint32_t c[3] = {x, y, z};
      for (int a = 0; a < 3; ++a)
        pos[(size_t)i * 4 + a] = ((spacing * (float)c[a] + r) - 1.0f) + rng.symmetric() * jitter;
      pos[(size_t)i * 4 + 3] = 1.0f;
      for (int a = 0; a < 3; ++a) vel[(size_t)i * 4 + a] = rng.symmetric() * velAmp;
      vel[(size_t)i * 4 + 3] = 0.0f;
    }
  } else {
    for (uint32_t i = 0; i < n; ++i) {
      for (int a = 0; a < 3; ++a) pos[(size_t)i * 4 + a] = 2.0f * (rng.uniform01() - 0.5f);
      pos[(size_t)i * 4 + 3] = 1.0f;
      for (int a = 0; a < 3; ++a) vel[(size_t)i * 4 + a] = rng.symmetric() * velAmp;
      vel[(size_t)i * 4 + 3] = 0.0f;
    }
  }
}

}  // namespace fc

using namespace fc;

// The region integrate keeps the spheres in: the box [-1, 1]^3 of the box
// walls; with a distance map, x and y within the unit cylinder and z in
// [-1, 10].  One scale for all axes keeps fixed-point distances isotropic.
fc::QuantFrame fc_system::quant_frame() const {
  return fc::quant_frame_for(boundary == FC_BOUNDARY_DISTMAP && boundaryMap != nullptr);
}

fc::BoundaryArg fc_system::boundary_arg() const {
  fc::BoundaryArg b{};
  b.hasMap = (boundary == FC_BOUNDARY_DISTMAP && boundaryMap != nullptr) ? 1 : 0;
  b.poly = polydisperse ? 1 : 0;
  if (b.hasMap) {
    b.map = distmap_view(boundaryMap);
    b.map.blockBounds = boundaryBounds;
    b.map.bx = boundaryBx;
    b.map.by = boundaryBy;
  }
  return b;
}

extern "C" {

fc_status fc_system_set_boundary(fc_system* sys, fc_boundary mode, const fc_distmap* map) {
  if (!sys) return set_error(FC_ERROR_INVALID_VALUE, "sys is NULL");
  if (mode != FC_BOUNDARY_BOX && mode != FC_BOUNDARY_DISTMAP)
    return set_error(FC_ERROR_INVALID_VALUE, "unknown boundary mode");
  if (mode == FC_BOUNDARY_DISTMAP && !map)
    return set_error(FC_ERROR_INVALID_VALUE, "the distance-map boundary needs a map");
  FC_HIP_TRY(hipStreamSynchronize(sys->ctx->stream));
  float2* bounds = nullptr;
  int bx = 0, by = 0;
  if (mode == FC_BOUNDARY_DISTMAP)
    FC_TRY(distmap_block_bounds(map, sys->ctx->stream, &bounds, &bx, &by));
  (void)hipFree(sys->boundaryBounds);
  sys->boundaryBounds = bounds;
  sys->boundaryBx = bx;
  sys->boundaryBy = by;
  sys->boundary = mode;
  sys->boundaryMap = mode == FC_BOUNDARY_DISTMAP ? map : nullptr;
  drop_graph(sys);  // the captured step carries the boundary in its arguments
  return FC_SUCCESS;
}

fc_status fc_system_set_polydisperse(fc_system* sys, int enable) {
  if (!sys) return set_error(FC_ERROR_INVALID_VALUE, "sys is NULL");
  FC_HIP_TRY(hipStreamSynchronize(sys->ctx->stream));
  sys->polydisperse = enable ? 1 : 0;
  drop_graph(sys);  // the captured step carries the radius mode in its kernels
  return FC_SUCCESS;
}

fc_status fc_system_create(fc_context* ctx, uint32_t numParticles, fc_uint3 gridSize,
                           fc_system** out) {
  if (!ctx || !out) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  *out = nullptr;
  fc_sim_params P = default_params(numParticles, gridSize);
  uint64_t cells = (uint64_t)gridSize.x * gridSize.y * gridSize.z;
  if (cells > 0xffffffffull) return set_error(FC_ERROR_CAPACITY, "grid has 2^32 or more cells");
  FC_TRY(validate_params(P));
  fc_system* sys = new fc_system();
  sys->ctx = ctx;
  sys->n = numParticles;
  sys->params = P;
  const size_t n = numParticles ? numParticles : 1;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMalloc(&sys->pos, n * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc(&sys->vel, n * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc(&sys->sortedPos, (n + kSortedPosPad) * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc(&sys->sortedVel, n * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc(&sys->sortedQ, (n + kSortedPosPad) * sizeof(uint2));
  if (e == hipSuccess) e = hipMalloc(&sys->hash, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&sys->index, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&sys->keys, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&sys->slotOf, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&sys->idOfSlot, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemset(sys->pos, 0, n * sizeof(float4));
  if (e == hipSuccess) e = hipMemset(sys->vel, 0, n * sizeof(float4));
  if (e != hipSuccess) {
    free_system_buffers(sys);
    delete sys;
    return hip_error(e, "fc_system_create allocation");
  }
  fc_status s = alloc_tables(sys, P.numCells);
  // The sort's scratch is sized here so that updates never allocate.
  if (s == FC_SUCCESS) s = ensure_sort_capacity(ctx, n);
  if (s != FC_SUCCESS) {
    free_system_buffers(sys);
    delete sys;
    return s;
  }
  *out = sys;
  return FC_SUCCESS;
}

fc_status fc_system_destroy(fc_system* sys) {
  if (!sys) return FC_SUCCESS;
  // No use of sys->ctx here: the context may already be gone.  hipFree
  // waits for the device before releasing memory.
  drop_graph(sys);
  free_system_buffers(sys);
  (void)hipFree(sys->boundaryBounds);
  delete sys;
  return FC_SUCCESS;
}

uint32_t fc_system_num_particles(const fc_system* sys) { return sys ? sys->n : 0u; }

fc_status fc_system_set_params(fc_system* sys, const fc_sim_params* params) {
  if (!sys || !params) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  FC_TRY(validate_params(*params));
  if (params->numCells != sys->params.numCells) {
    FC_HIP_TRY(hipStreamSynchronize(sys->ctx->stream));
    drop_graph(sys);
    FC_TRY(alloc_tables(sys, params->numCells));
  }
  sys->params = *params;
  return FC_SUCCESS;
}

fc_status fc_system_get_params(const fc_system* sys, fc_sim_params* params) {
  if (!sys || !params) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  *params = sys->params;
  return FC_SUCCESS;
}

fc_status fc_system_set_array(fc_system* sys, fc_particle_array array, const float* data, int start,
                              int count) {
  if (!sys) return set_error(FC_ERROR_INVALID_VALUE, "sys is NULL");
  if (array != FC_POSITION && array != FC_VELOCITY)
    return set_error(FC_ERROR_INVALID_VALUE, "unknown particle array");
  if (start < 0 || count < 0 || (uint64_t)start + (uint64_t)count > sys->n)
    return set_error(FC_ERROR_INVALID_VALUE, "start/count outside [0, numParticles)");
  if (count == 0) return FC_SUCCESS;
  if (!data) return set_error(FC_ERROR_INVALID_VALUE, "data is NULL");
  float4* dst = array == FC_POSITION ? sys->pos : sys->vel;
  FC_HIP_TRY(hipMemcpyAsync(dst + start, data, (size_t)count * sizeof(float4),
                            hipMemcpyHostToDevice, sys->ctx->stream));
  FC_HIP_TRY(hipStreamSynchronize(sys->ctx->stream));
  return FC_SUCCESS;
}

fc_status fc_system_get_array(fc_system* sys, fc_particle_array array, float* out) {
  if (!sys || !out) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (array != FC_POSITION && array != FC_VELOCITY)
    return set_error(FC_ERROR_INVALID_VALUE, "unknown particle array");
  if (sys->n == 0) return FC_SUCCESS;
  const float4* src = array == FC_POSITION ? sys->pos : sys->vel;
  FC_HIP_TRY(hipMemcpyAsync(out, src, (size_t)sys->n * sizeof(float4), hipMemcpyDeviceToHost,
                            sys->ctx->stream));
  FC_HIP_TRY(hipStreamSynchronize(sys->ctx->stream));
  return check_device_flags(sys->ctx);
}

fc_status fc_system_reset(fc_system* sys, fc_particle_config config, uint64_t seed,
                          float jitterFrac, float velAmp) {
  if (!sys) return set_error(FC_ERROR_INVALID_VALUE, "sys is NULL");
  if (config != FC_CONFIG_GRID && config != FC_CONFIG_RANDOM)
    return set_error(FC_ERROR_INVALID_VALUE, "unknown particle config");
  const uint32_t n = sys->n;
  if (n == 0) return FC_SUCCESS;
  std::vector<float> pos((size_t)n * 4), vel((size_t)n * 4);
  generate_initial_state(pos.data(), vel.data(), n, config, seed, sys->params.particleRadius,
                         jitterFrac, velAmp);
  FC_TRY(fc_system_set_array(sys, FC_POSITION, pos.data(), 0, (int)n));
  FC_TRY(fc_system_set_array(sys, FC_VELOCITY, vel.data(), 0, (int)n));
  return FC_SUCCESS;
}

fc_status fc_system_update(fc_system* sys, float deltaTime) {
  if (!sys) return set_error(FC_ERROR_INVALID_VALUE, "sys is NULL");
  StageTimer* t = sys->ctx->timer.enabled ? &sys->ctx->timer : nullptr;
  return enqueue_update(sys, deltaTime, t);
}

fc_status fc_system_run(fc_system* sys, float deltaTime, uint32_t steps) {
  if (!sys) return set_error(FC_ERROR_INVALID_VALUE, "sys is NULL");
  if (steps == 0) return FC_SUCCESS;
  fc_context* ctx = sys->ctx;
  if (steps == 1 || !collide_can_fuse(sys->params, sys->n)) {
    for (uint32_t i = 0; i < steps; ++i) FC_TRY(fc_system_update(sys, deltaTime));
    return FC_SUCCESS;
  }
  StageTimer* timer = ctx->timer.enabled ? &ctx->timer : nullptr;
  if (run_row_mode(sys)) FC_TRY(ensure_row_tables(sys->rowTables, sys->params, sys->n));
  FC_TRY(enqueue_run_head(sys, deltaTime, timer));
  FC_TRY(enqueue_run_body(sys, deltaTime, /*clearInSort=*/false, /*last=*/false,
                          /*fromSlots=*/false, timer));
  const uint32_t middle = steps - 2;
  if (middle > 0) {
    if (timer) {
      for (uint32_t i = 0; i < middle; ++i)
        FC_TRY(enqueue_run_body(sys, deltaTime, true, false, true, timer));
    } else {
      if (!sys->graphExec[0] || sys->graphDt != deltaTime ||
          std::memcmp(&sys->graphParams, &sys->params, sizeof(fc_sim_params)) != 0) {
        FC_TRY(capture_graph(sys, deltaTime));
      }
      uint32_t i = 0;
      for (; i + kGraphSteps <= middle; i += kGraphSteps)
        FC_HIP_TRY(hipGraphLaunch(sys->graphExec[1], ctx->stream));
      for (; i < middle; ++i) FC_HIP_TRY(hipGraphLaunch(sys->graphExec[0], ctx->stream));
    }
  }
  FC_TRY(enqueue_run_body(sys, deltaTime, /*clearInSort=*/true, /*last=*/true,
                          /*fromSlots=*/true, timer));
  return FC_SUCCESS;
}

fc_status fc_system_device_arrays(fc_system* sys, float** pos, float** vel, float** sortedPos,
                                  float** sortedVel, uint32_t** hash, uint32_t** index,
                                  uint32_t** cellStart, uint32_t** cellEnd) {
  if (!sys) return set_error(FC_ERROR_INVALID_VALUE, "sys is NULL");
  if (pos) *pos = (float*)sys->pos;
  if (vel) *vel = (float*)sys->vel;
  if (sortedPos) *sortedPos = (float*)sys->sortedPos;
  if (sortedVel) *sortedVel = (float*)sys->sortedVel;
  if (hash) *hash = sys->hash;
  if (index) *index = sys->index;
  if (cellStart) *cellStart = sys->cellStart;
  if (cellEnd) *cellEnd = sys->cellEnd;
  return FC_SUCCESS;
}

fc_status fc_system_dump_grid(fc_system* sys, uint32_t* maxParticlesPerCell,
                              uint32_t* occupiedCells) {
  if (!sys) return set_error(FC_ERROR_INVALID_VALUE, "sys is NULL");
  const uint32_t C = sys->params.numCells;
  std::vector<uint32_t> cs(C), ce(C);
  FC_HIP_TRY(hipMemcpyAsync(cs.data(), sys->cellStart, (size_t)C * 4, hipMemcpyDeviceToHost,
                            sys->ctx->stream));
  FC_HIP_TRY(hipMemcpyAsync(ce.data(), sys->cellEnd, (size_t)C * 4, hipMemcpyDeviceToHost,
                            sys->ctx->stream));
  FC_HIP_TRY(hipStreamSynchronize(sys->ctx->stream));
  uint32_t mx = 0, occ = 0;
  for (uint32_t i = 0; i < C; ++i) {
    if (cs[i] != CELL_EMPTY) {
      ++occ;
      uint32_t c = ce[i] - cs[i];
      if (c > mx) mx = c;
    }
  }
  if (maxParticlesPerCell) *maxParticlesPerCell = mx;
  if (occupiedCells) *occupiedCells = occ;
  return FC_SUCCESS;
}

}  // extern "C"
