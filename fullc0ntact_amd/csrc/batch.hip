// Batched independent scenes: many particle systems of uneven sizes stepped
// together, one launch per stage for the whole batch.
//
// Scene s owns the particles [offsets[s], offsets[s+1]) of the concatenated
// arrays and its own parameter block (spring, damping, radius, ...; every
// scene uses the batch's grid dimensions).  The hash is scene-segmented:
// key = s * C + cell with C = cells per scene (a power of two), so one
// stable sort of (key, global id) orders the batch scene by scene and, inside
// a scene, by (cell, id) exactly as the scene's own sort would.  The cell
// tables are S * C entries, scene-major, cleared incrementally.  Each stage
// therefore computes, per scene, exactly what ParticleSystem::update
// (particleSystem.cpp:222-284) computes for that scene alone.
#include "fc_internal.hpp"
#include "fc_device.hpp"
#include "radix_common.hpp"
#include "collide.hpp"
#include "row_resort.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace fc {
// Per-scene constants of the row-range collide (the single system computes
// them per launch): exact reciprocal cell sizes (or zeros), the contact
// pre-test bound on d^2 and the fixed-point candidate filter's bound.
struct SceneCollide {
  v3f invCell;
  float contactD2;
  uint32_t qThreshold;
};
}  // namespace fc

struct fc_batch {
  fc_context* ctx = nullptr;
  uint32_t S = 0;
  uint32_t n = 0;
  fc_uint3 grid{0, 0, 0};
  uint32_t C = 0;
  int cellBits = 0;
  std::vector<uint32_t> offsets;  // S + 1
  std::vector<fc_sim_params> params;
  uint32_t* dOffsets = nullptr;
  fc_sim_params* dParams = nullptr;
  float* dContactD2 = nullptr;
  fc::SceneCollide* dSceneCollide = nullptr;
  uint2* sortedQ = nullptr;  // 16-bit coordinates of sortedPos (the box frame)
  float4* pos = nullptr;
  float4* vel = nullptr;
  float4* sortedPos = nullptr;
  float4* sortedVel = nullptr;
  uint32_t* hash = nullptr;
  uint32_t* index = nullptr;
  uint32_t* keys = nullptr;
  uint32_t* slotOf = nullptr;    // multi-step runs: slot of each particle id (sort values)
  uint32_t* idOfSlot = nullptr;  // multi-step runs: particle id held by each slot
  uint32_t* cellStart = nullptr;
  uint32_t* cellEnd = nullptr;
  double* stats = nullptr;  // S x 4
  bool tablesTrackHash = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graphExec = nullptr;
  float graphDt = 0.0f;
  // Row resort of a multi-step run's steady steps (rows: S x gridSize.y x
  // gridSize.z, each gridSize.x consecutive keys)
  fc::RowTables rowTables;
  // kBatchRunSteps steady-state steps of a multi-step run
  hipGraph_t runGraph = nullptr;
  hipGraphExec_t runGraphExec = nullptr;
  float runGraphDt = 0.0f;
  fc_batch_tuning tuning{4, 256, 0.0f};
};

namespace fc {

namespace {

constexpr int kBB = 256;  // block of the non-tuned batch kernels

inline unsigned blocks_for(size_t n, int per) {
  size_t g = (n + (size_t)per - 1) / (size_t)per;
  return (unsigned)(g == 0 ? 1 : g);
}

// Scene of particle i: the largest s with offsets[s] <= i (binary search).
__device__ __forceinline__ uint32_t scene_of(const uint32_t* __restrict__ offsets, uint32_t S,
                                             uint32_t i) {
  uint32_t lo = 0, hi = S;  // offsets[lo] <= i < offsets[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (offsets[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// integrate (box walls) + scene-segmented hash + digit-0 histogram, PER
// particles per thread; clears the previous update's occupied cells.
template <int PER>
__global__ __launch_bounds__(kBB) void k_batch_integrate_hash(
    const fc_sim_params* __restrict__ prm, const uint32_t* __restrict__ offsets, uint32_t S,
    int cellBits, float4* __restrict__ pos, float4* __restrict__ vel, float dt,
    uint32_t* __restrict__ keysOut, const uint32_t* __restrict__ prevSortedHash,
    uint32_t* __restrict__ cellStart, uint32_t* __restrict__ cellEnd, uint32_t n, int passes,
    uint32_t* gHist, uint32_t* counters) {
  __shared__ uint32_t s_hist[RS_RADIX];
  for (int i = threadIdx.x; i < RS_RADIX; i += kBB) s_hist[i] = 0u;
  histogram_producer_prologue(gHist, passes, counters);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const uint32_t i = blockIdx.x * (kBB * PER) + u * kBB + threadIdx.x;
    uint32_t key = 0;
    if (i < n) {
      if (prevSortedHash != nullptr) {
        const uint32_t h = prevSortedHash[i];
        if (i == 0 || prevSortedHash[i - 1] != h) {
          cellStart[h] = CELL_EMPTY;
          cellEnd[h] = CELL_EMPTY;
        }
      }
      const uint32_t s = scene_of(offsets, S, i);
      const fc_sim_params P = prm[s];
      const float4 p4 = pos[i];
      const float4 v4 = vel[i];
      v3f p = xyz(p4), v = xyz(v4);
      integrate_box(P, P.particleRadius, dt, p, v);
      pos[i] = make_float4(p.x, p.y, p.z, p4.w);
      vel[i] = make_float4(v.x, v.y, v.z, v4.w);
      key = (s << cellBits) | hash_of(P, p.x, p.y, p.z);
      keysOut[i] = key;
    }
    histogram_add_digit(s_hist, key & (RS_RADIX - 1), i < n);
  }
  __syncthreads();
  histogram_flush(s_hist, gHist, 0);
}

// collideD per scene: the scene comes from the sorted key, the cell tables
// are the scene's slice of the batch tables.
template <int BS>
__global__ __launch_bounds__(BS) void k_batch_collide(
    const fc_sim_params* __restrict__ prm, const float* __restrict__ contactD2, int cellBits,
    uint32_t C, float4* __restrict__ newVel, const float4* __restrict__ sortedPos,
    const float4* __restrict__ sortedVel, const uint32_t* __restrict__ hash,
    const uint32_t* __restrict__ index, const uint32_t* __restrict__ cellStart,
    const uint32_t* __restrict__ cellEnd, uint32_t n) {
  __shared__ uint32_t s_list[kListCap][BS];
  const uint32_t k =
      xcd_grouped_block(blockIdx.x, gridDim.x, kCollideXcdGroup * (256 / BS)) * BS + threadIdx.x;
  if (k >= n) return;
  const uint32_t s = hash[k] >> cellBits;
  const fc_sim_params P = prm[s];
  const size_t base = (size_t)s * C;
  const float4 p4 = sortedPos[k];
  const v3f pos = xyz(p4), vel = xyz(sortedVel[k]);
  const v3f force = add_collider(
      P, pos, vel, P.particleRadius,
      collide_one_v3<false, BS>(P, contactD2[s], k, pos, vel, P.particleRadius, sortedPos,
                                sortedVel, cellStart + base, cellEnd + base, n, s_list));
  newVel[index[k]] = make_float4(vel.x + force.x, vel.y + force.y, vel.z + force.z, 0.0f);
}

// The scene of the first lane of `pending`, wave-uniform (so that what is
// indexed by it is read with scalar loads into scalar registers).
__device__ __forceinline__ uint32_t lead_scene(uint64_t pending, uint32_t s) {
  const int first = __ffsll((unsigned long long)pending) - 1;
  return (uint32_t)__builtin_amdgcn_readfirstlane(__shfl((int)s, first));
}

// The per-scene blocks as read-only (constant address space) memory, which
// the kernels never write: a load at a wave-uniform index then goes to
// scalar registers.
template <typename T>
__device__ __forceinline__ T scene_block(const T* p, uint32_t s) {
  static_assert(sizeof(T) % 4 == 0, "whole words");
  typedef const __attribute__((address_space(4))) uint32_t* ConstWords;
  const ConstWords w = (ConstWords)(const void*)(p + s);
  uint32_t words[sizeof(T) / 4];
#pragma unroll
  for (size_t i = 0; i < sizeof(T) / 4; ++i) words[i] = w[i];
  T out;
  std::memcpy(&out, words, sizeof(T));
  return out;
}

// collideD per scene on the arrays gathered by the sort (the row-range walk
// of the single system, collide_one_rows): the scene comes from the sorted
// key in sortedVel.w, the cell tables are the scene's slice of the batch
// tables, the parameters and row-walk constants the scene's own.  The lanes
// of a wave take their scenes one at a time (nearly every wave holds one
// scene), each with its parameters in scalar registers, which keeps them
// out of the vector registers the walk needs.
template <int BS>
__global__ __launch_bounds__(BS) void k_batch_collide_rows(
    const fc_sim_params* __restrict__ prm, const SceneCollide* __restrict__ sc, int cellBits,
    uint32_t C, float4* __restrict__ newVel, const float4* __restrict__ sortedPos,
    const float4* __restrict__ sortedVel, const uint2* __restrict__ sortedQ,
    const uint32_t* __restrict__ index, const uint32_t* __restrict__ cellStart,
    const uint32_t* __restrict__ cellEnd, uint32_t n) {
  __shared__ uint32_t s_list[kListCap + 1][BS];
  const uint32_t k =
      xcd_grouped_block(blockIdx.x, gridDim.x, kCollideXcdGroup * (256 / BS)) * BS + threadIdx.x;
  if (k >= n) return;
  const float4 v4 = sortedVel[k];
  const uint32_t s = __float_as_uint(v4.w) >> cellBits;
  bool pending = true;
  for (uint64_t m = __ballot(pending); m != 0ull; m = __ballot(pending)) {
    const uint32_t lead = lead_scene(m, s);
    const fc_sim_params P = scene_block(prm, lead);  // (before the divergent branch)
    const SceneCollide c = scene_block(sc, lead);
    if (!pending || s != lead) continue;
    pending = false;
    const size_t base = (size_t)lead * C;
    const float4 p4 = sortedPos[k];
    const v3f pos = xyz(p4), vel = xyz(v4);
    const v3f force = add_collider(
        P, pos, vel, P.particleRadius,
        collide_one_rows<false, BS>(P, c.invCell, c.contactD2, k, pos, vel, P.particleRadius,
                                    sortedPos, sortedVel, sortedQ, c.qThreshold,
                                    cellStart + base, cellEnd + base, n, s_list));
    newVel[index[k]] = make_float4(vel.x + force.x, vel.y + force.y, vel.z + force.z, 0.0f);
  }
}

// k_batch_collide_rows inside a multi-step run (fc_batch_run), where the
// state between steps lives at the sorted slots: fused with the scene's next
// integrate and hash, which the update order runs next on the same values.
// The integrated state is written at the slot k, the next key at the
// particle's id (keysOut[id], so the stable sort still orders ties by id),
// slotOf[id] = k (the value the sort carries) and idOfSlot[k] = id, and the
// keys' digit-0 histogram for the next sort.  kLast (the run's last step):
// newVel and the positions go back to the ids instead.
template <int BS, bool kLast>
__global__ __launch_bounds__(BS) void k_batch_collide_run(
    const fc_sim_params* __restrict__ prm, const SceneCollide* __restrict__ sc, int cellBits,
    uint32_t C, float4* __restrict__ pos, float4* __restrict__ vel,
    const float4* __restrict__ sortedPos, const float4* __restrict__ sortedVel,
    const uint2* __restrict__ sortedQ, const uint32_t* __restrict__ index,
    const uint32_t* __restrict__ cellStart, const uint32_t* __restrict__ cellEnd, uint32_t n,
    float dt, uint32_t* __restrict__ keysOut, uint32_t* __restrict__ slotOf,
    uint32_t* __restrict__ idOfSlot, int passes, uint32_t* gHist, uint32_t* counters,
    RowHists rows) {
  __shared__ uint32_t s_list[kListCap + 1][BS];
  __shared__ uint32_t s_hist[kLast ? 1 : RS_RADIX];
  const uint32_t k =
      xcd_grouped_block(blockIdx.x, gridDim.x, kCollideXcdGroup * (256 / BS)) * BS + threadIdx.x;
  const bool valid = k < n;
  // rows.rowHist: the next sort is a row resort (next keys at the slots,
  // per-row counts) instead of a radix sort (keys at the ids, digit counts)
  const bool slotKeys = rows.rowHist != nullptr;
  if (!kLast && !slotKeys) {
    for (int i = threadIdx.x; i < RS_RADIX; i += BS) s_hist[i] = 0u;
    histogram_producer_prologue(gHist, passes, counters);
  }
  uint32_t key = 0;
  const float4 v4 = valid ? sortedVel[k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const uint32_t s = __float_as_uint(v4.w) >> cellBits;
  bool pending = valid;
  for (uint64_t m = __ballot(pending); m != 0ull; m = __ballot(pending)) {
    const uint32_t lead = lead_scene(m, s);
    const fc_sim_params P = scene_block(prm, lead);  // (before the divergent branch)
    const SceneCollide c = scene_block(sc, lead);
    if (!pending || s != lead) continue;
    pending = false;
    const size_t base = (size_t)lead * C;
    const float4 p4 = sortedPos[k];
    const v3f p0 = xyz(p4), v0 = xyz(v4);
    const v3f force = add_collider(
        P, p0, v0, P.particleRadius,
        collide_one_rows<false, BS>(P, c.invCell, c.contactD2, k, p0, v0, P.particleRadius,
                                    sortedPos, sortedVel, sortedQ, c.qThreshold,
                                    cellStart + base, cellEnd + base, n, s_list));
    const uint32_t orig = index[k];
    const v3f nv = make_v3(v0.x + force.x, v0.y + force.y, v0.z + force.z);
    if (kLast) {
      vel[orig] = make_float4(nv.x, nv.y, nv.z, 0.0f);
      pos[orig] = p4;
    } else {
      v3f p = p0, v = nv;
      integrate_box(P, P.particleRadius, dt, p, v);
      pos[k] = make_float4(p.x, p.y, p.z, p4.w);
      vel[k] = make_float4(v.x, v.y, v.z, 0.0f);
      key = (lead << cellBits) | hash_of(P, p.x, p.y, p.z);
      if (slotKeys) {
        // at the slot, with whether the sphere changed cell (v4.w holds its
        // current key)
        keysOut[k] = key != __float_as_uint(v4.w) ? (key | kMovedBit) : key;
      } else {
        keysOut[orig] = key;
        slotOf[orig] = k;
      }
      idOfSlot[k] = orig;
    }
  }
  if (!kLast) {
    if (slotKeys) {
      row_histograms(rows, k, key, __float_as_uint(v4.w), valid);
    } else {
      __syncthreads();
      histogram_add_digit(s_hist, key & (RS_RADIX - 1), valid);
      __syncthreads();
      histogram_flush(s_hist, gHist, 0);
    }
  }
}

// Kinetic energy and momentum per scene in FP64, one block per scene, fixed
// reduction order.
__global__ __launch_bounds__(kBB) void k_batch_energy(const float4* __restrict__ vel,
                                                      const uint32_t* __restrict__ offsets,
                                                      double* __restrict__ out) {
  __shared__ double sh[4][kBB];
  const uint32_t s = blockIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint32_t i = offsets[s] + threadIdx.x; i < offsets[s + 1]; i += kBB) {
    const float4 v = vel[i];
    const double x = v.x, y = v.y, z = v.z;
    acc[0] += 0.5 * (x * x + y * y + z * z);
    acc[1] += x;
    acc[2] += y;
    acc[3] += z;
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) sh[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int w = kBB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int c = 0; c < 4; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) out[4 * s + threadIdx.x] = sh[threadIdx.x][0];
}

hipError_t launch_batch_integrate_hash(fc_batch* b, float dt, const uint32_t* prev, int passes) {
  fc_context* ctx = b->ctx;
  const int per = b->tuning.integratePerThread;
#define FC_BIH(PER)                                                                            \
  hipLaunchKernelGGL(k_batch_integrate_hash<PER>, dim3(blocks_for(b->n, kBB * PER)), dim3(kBB), \
                     0, ctx->stream, b->dParams, b->dOffsets, b->S, b->cellBits, b->pos, b->vel, \
                     dt, b->keys, prev, b->cellStart, b->cellEnd, b->n, passes,                 \
                     ctx->sort.histogram, ctx->sort.counters)
  if (per == 1) FC_BIH(1);
  else if (per == 2) FC_BIH(2);
  else if (per == 8) FC_BIH(8);
  else FC_BIH(4);
#undef FC_BIH
  return hipGetLastError();
}

hipError_t launch_batch_collide(fc_batch* b) {
  hipStream_t st = b->ctx->stream;
  const int bs = b->tuning.collideBlock;
#define FC_BC(BS)                                                                              \
  hipLaunchKernelGGL(k_batch_collide<BS>, dim3(blocks_for(b->n, BS)), dim3(BS), 0, st,        \
                     b->dParams, b->dContactD2, b->cellBits, b->C, b->vel, b->sortedPos,      \
                     b->sortedVel, b->hash, b->index, b->cellStart, b->cellEnd, b->n)
  if (bs == 64) FC_BC(64);
  else if (bs == 128) FC_BC(128);
  else FC_BC(256);
#undef FC_BC
  return hipGetLastError();
}

hipError_t launch_batch_collide_rows(fc_batch* b) {
  hipStream_t st = b->ctx->stream;
  const int bs = b->tuning.collideBlock;
#define FC_BCR(BS)                                                                             \
  hipLaunchKernelGGL(k_batch_collide_rows<BS>, dim3(blocks_for(b->n, BS)), dim3(BS), 0, st,   \
                     b->dParams, b->dSceneCollide, b->cellBits, b->C, b->vel, b->sortedPos,   \
                     b->sortedVel, b->sortedQ, b->index, b->cellStart, b->cellEnd, b->n)
  if (bs == 64) FC_BCR(64);
  else if (bs == 128) FC_BCR(128);
  else FC_BCR(256);
#undef FC_BCR
  return hipGetLastError();
}

hipError_t launch_batch_collide_run(fc_batch* b, float dt, bool last, int passes,
                                   const RowHists& rows) {
  fc_context* ctx = b->ctx;
  const int bs = b->tuning.collideBlock;
#define FC_BCRUN(BS, LAST)                                                                     \
  hipLaunchKernelGGL((k_batch_collide_run<BS, LAST>), dim3(blocks_for(b->n, BS)), dim3(BS), 0, \
                     ctx->stream, b->dParams, b->dSceneCollide, b->cellBits, b->C, b->pos,     \
                     b->vel, b->sortedPos, b->sortedVel, b->sortedQ, b->index, b->cellStart,   \
                     b->cellEnd, b->n, dt, b->keys, b->slotOf, b->idOfSlot, passes,            \
                     ctx->sort.histogram, ctx->sort.counters, rows)
  if (last) {
    if (bs == 64) FC_BCRUN(64, true);
    else if (bs == 128) FC_BCRUN(128, true);
    else FC_BCRUN(256, true);
  } else {
    if (bs == 64) FC_BCRUN(64, false);
    else if (bs == 128) FC_BCRUN(128, false);
    else FC_BCRUN(256, false);
  }
#undef FC_BCRUN
  return hipGetLastError();
}

// S * C < 2^32 is checked at creation.
int batch_key_bits(const fc_batch* b) { return key_bits_for_cells(b->S * b->C); }

// The row-range walk needs every grid dimension >= 4 (27 distinct cells,
// rows of three) and keys below 2^30 (its "no cell" marker).
bool batch_rows_ok(const fc_batch* b) {
  return b->grid.x >= 4 && b->grid.y >= 4 && b->grid.z >= 4 &&
         (uint64_t)b->S * b->C <= (1ull << 30);
}

// Frame of the fixed-point coordinates: the box walls at +-1 of every scene.
QuantFrame batch_frame() {
  QuantFrame f;
  f.lo[0] = f.lo[1] = f.lo[2] = -1.0f;
  f.scale = 65535.0f / 2.0f;
  return f;
}

void drop_batch_graph(fc_batch* b) {
  if (b->graphExec) hipGraphExecDestroy(b->graphExec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->graphExec = nullptr;
  b->graph = nullptr;
  if (b->runGraphExec) hipGraphExecDestroy(b->runGraphExec);
  if (b->runGraph) hipGraphDestroy(b->runGraph);
  b->runGraphExec = nullptr;
  b->runGraph = nullptr;
}

// One update of every scene: integrate+hash -> sort -> reorder -> collide.
fc_status enqueue_batch_update(fc_batch* b, float dt) {
  fc_context* ctx = b->ctx;
  hipStream_t st = ctx->stream;
  const uint64_t cells = (uint64_t)b->S * b->C;
  const uint32_t* prev = nullptr;
  if (b->tablesTrackHash) {
    prev = b->hash;
  } else {
    FC_HIP_TRY(launch_fill_u32(st, b->cellStart, CELL_EMPTY, cells, b->cellEnd));
  }
  if (b->n == 0) return FC_SUCCESS;
  const int bits = batch_key_bits(b);
  FC_HIP_TRY(launch_batch_integrate_hash(b, dt, prev, radix_passes(bits, false)));
  if (batch_rows_ok(b)) {
    // The sort's last pass gathers the state into sorted order (the key
    // bits in sortedVel.w, the fixed-point coordinates), a small kernel
    // writes the cell tables, and the row-range collide runs per scene.
    SortGather ga;
    ga.pos = b->pos;
    ga.vel = b->vel;
    ga.sortedPos = b->sortedPos;
    ga.sortedVel = b->sortedVel;
    ga.sortedQ = b->sortedQ;
    ga.frame = batch_frame();
    FC_HIP_TRY(radix_sort_pairs(st, ctx->sort, ctx->devFlags, b->keys, nullptr, b->hash,
                                b->index, b->n, bits, nullptr, /*histogramReady=*/true, 1,
                                nullptr, nullptr, nullptr, &ga));
    FC_HIP_TRY(launch_cell_bounds(st, b->cellStart, b->cellEnd, b->hash, b->n, (uint32_t)cells));
    FC_HIP_TRY(launch_batch_collide_rows(b));
  } else {
    FC_HIP_TRY(radix_sort_pairs(st, ctx->sort, ctx->devFlags, b->keys, nullptr, b->hash,
                                b->index, b->n, bits, nullptr, /*histogramReady=*/true));
    FC_HIP_TRY(launch_reorder(st, b->cellStart, b->cellEnd, b->sortedPos, b->sortedVel, b->hash,
                              b->index, b->pos, b->vel, b->n, (uint32_t)cells));
    FC_HIP_TRY(launch_batch_collide(b));
  }
  b->tablesTrackHash = true;
  return FC_SUCCESS;
}

fc_status capture_batch_graph(fc_batch* b, float dt) {
  drop_batch_graph(b);
  hipStream_t st = b->ctx->stream;
  FC_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  fc_status s = enqueue_batch_update(b, dt);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(st, &g);
  if (s != FC_SUCCESS) {
    if (g) hipGraphDestroy(g);
    return s;
  }
  if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
  b->graph = g;
  e = hipGraphInstantiate(&b->graphExec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    drop_batch_graph(b);
    return hip_error(e, "hipGraphInstantiate");
  }
  b->graphDt = dt;
  return FC_SUCCESS;
}

// The batch's rows as the row resort sees them: kBatchRowGridRows
// consecutive grid rows of a scene form one row (a contiguous range of
// keys), so row = key >> log2(gridSize.x * kBatchRowGridRows).  The
// resort's cost is a chain of dependent loads per row, and a scene's grid
// rows hold few spheres (about 20 on the freshly dropped 64^3 ensemble):
// per ensemble step, 1 grid row per row measured 1.876 ms, 4 1.391 ms,
// 8 1.323 ms (radix passes: 1.663 ms); 16 and 32 (4.6 and 9.1 ms) put
// rows past the row sort's LDS capacity (768 members) into its counting
// path.  A settled bed (about 2.2 spheres per occupied cell, some 140 per
// 64-cell grid row) would do that to 8 as well; 4 keeps such rows near
// 570 members.
constexpr uint32_t kBatchRowGridRows = 4;

fc_sim_params batch_row_grid(const fc_batch* b) {
  const uint32_t f = b->grid.y < kBatchRowGridRows ? b->grid.y : kBatchRowGridRows;
  fc_sim_params P = b->params[0];
  P.gridSize = fc_uint3{b->grid.x * f, b->grid.y / f, b->grid.z * b->S};
  return P;
}

// Whether a run's steady steps are row resorts (the row tables fit).
bool batch_row_mode(const fc_batch* b) {
  return batch_rows_ok(b) && (uint64_t)b->grid.z * b->S <= 0xffffffffull &&
         row_resort_grid_ok(batch_row_grid(b));
}

// A multi-step run of the row-range path, as fc_system_run runs a system:
// the head integrates and hashes from the ids; each step sorts (gathering
// the state), writes the cell tables and collides, fused with the next
// integrate and hash except in the last step, which returns the state to
// the ids.  The first step's sort is a radix sort from the ids; in row mode
// each later one is a row resort of the next keys the fused collide stored
// at the slots (row_resort.hip), else a radix sort gathering from the
// previous step's slots.  Every value is the one the update path computes.
constexpr uint32_t kBatchRunSteps = 8;  // steady-state steps per graph launch

// keysAtSlots: the keys of this step's sort were stored at the slots (a
// steady step in row mode).  fromSlots: the state is at the previous step's
// slots (every step but the first).
fc_status enqueue_batch_run_step(fc_batch* b, float dt, bool fromSlots, bool last) {
  fc_context* ctx = b->ctx;
  hipStream_t st = ctx->stream;
  const uint64_t cells = (uint64_t)b->S * b->C;
  const int bits = batch_key_bits(b);
  const int passes = radix_passes(bits, false);
  const bool rowMode = batch_row_mode(b);
  RowTables& T = b->rowTables;
  if (rowMode && fromSlots) {
    FC_HIP_TRY(launch_row_scan(st, T, ctx->devFlags));
    FC_HIP_TRY(launch_row_changers(st, T, b->keys, b->idOfSlot, b->hash, b->cellStart,
                                   b->cellEnd, b->n, (uint32_t)cells, ctx->devFlags));
    RowOutputs O;
    O.hash = b->hash;
    O.index = b->index;
    O.pos = b->pos;
    O.vel = b->vel;
    O.sortedPos = b->sortedPos;
    O.sortedVel = b->sortedVel;
    O.sortedQ = b->sortedQ;
    O.cellStart = b->cellStart;
    O.cellEnd = b->cellEnd;
    O.frame = batch_frame();
    FC_HIP_TRY(launch_row_sort(st, T, b->keys, b->idOfSlot, O, ctx->devFlags));
  } else {
    SortGather ga;
    ga.pos = b->pos;
    ga.vel = b->vel;
    ga.sortedPos = b->sortedPos;
    ga.sortedVel = b->sortedVel;
    ga.sortedQ = b->sortedQ;
    ga.frame = batch_frame();
    if (fromSlots) {
      // the values are the slots; the ids come through idOfSlot; the first
      // pass clears the previous step's cells
      ga.idOfSlot = b->idOfSlot;
      ga.indexOut = b->index;
      FC_HIP_TRY(radix_sort_pairs(st, ctx->sort, ctx->devFlags, b->keys, b->slotOf, b->hash,
                                  b->slotOf, b->n, bits, nullptr, /*histogramReady=*/true, 2,
                                  b->hash, b->cellStart, b->cellEnd, &ga));
    } else {
      FC_HIP_TRY(radix_sort_pairs(st, ctx->sort, ctx->devFlags, b->keys, nullptr, b->hash,
                                  b->index, b->n, bits, nullptr, /*histogramReady=*/true, 1,
                                  nullptr, nullptr, nullptr, &ga));
    }
    FC_HIP_TRY(launch_cell_bounds(st, b->cellStart, b->cellEnd, b->hash, b->n, (uint32_t)cells));
    // the next step's row resort starts from this order's rows
    if (rowMode) FC_HIP_TRY(launch_row_starts(st, T, b->hash, b->n));
  }
  FC_HIP_TRY(launch_batch_collide_run(b, dt, last, passes,
                                      rowMode && !last ? row_hists(T) : RowHists{}));
  return FC_SUCCESS;
}

fc_status capture_batch_run_graph(fc_batch* b, float dt) {
  if (b->runGraphExec) hipGraphExecDestroy(b->runGraphExec);
  if (b->runGraph) hipGraphDestroy(b->runGraph);
  b->runGraphExec = nullptr;
  b->runGraph = nullptr;
  hipStream_t st = b->ctx->stream;
  FC_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  fc_status s = FC_SUCCESS;
  for (uint32_t i = 0; i < kBatchRunSteps && s == FC_SUCCESS; ++i)
    s = enqueue_batch_run_step(b, dt, /*fromSlots=*/true, /*last=*/false);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(st, &g);
  if (s != FC_SUCCESS) {
    if (g) hipGraphDestroy(g);
    return s;
  }
  if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
  b->runGraph = g;
  e = hipGraphInstantiate(&b->runGraphExec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(g);
    b->runGraph = nullptr;
    b->runGraphExec = nullptr;
    return hip_error(e, "hipGraphInstantiate");
  }
  b->runGraphDt = dt;
  return FC_SUCCESS;
}

fc_status enqueue_batch_run(fc_batch* b, float dt, uint32_t steps) {
  fc_context* ctx = b->ctx;
  hipStream_t st = ctx->stream;
  const uint64_t cells = (uint64_t)b->S * b->C;
  const bool rowMode = batch_row_mode(b);
  if (rowMode) FC_TRY(ensure_row_tables(b->rowTables, batch_row_grid(b), b->n));
  // head: integrate + hash from the ids, clearing the previous cells
  const uint32_t* prev = nullptr;
  if (b->tablesTrackHash) {
    prev = b->hash;
  } else {
    FC_HIP_TRY(launch_fill_u32(st, b->cellStart, CELL_EMPTY, cells, b->cellEnd));
  }
  // the row counts start from zero (a run that stopped early may have left
  // counts behind)
  if (rowMode) FC_HIP_TRY(launch_row_clear(st, b->rowTables));
  FC_HIP_TRY(launch_batch_integrate_hash(b, dt, prev, radix_passes(batch_key_bits(b), false)));
  b->tablesTrackHash = true;
  FC_TRY(enqueue_batch_run_step(b, dt, /*fromSlots=*/false, /*last=*/steps == 1));
  if (steps == 1) return FC_SUCCESS;
  uint32_t middle = steps - 2;
  if (middle >= kBatchRunSteps && (!b->runGraphExec || b->runGraphDt != dt))
    FC_TRY(capture_batch_run_graph(b, dt));
  for (; middle >= kBatchRunSteps; middle -= kBatchRunSteps)
    FC_HIP_TRY(hipGraphLaunch(b->runGraphExec, st));
  for (; middle > 0; --middle)
    FC_TRY(enqueue_batch_run_step(b, dt, /*fromSlots=*/true, /*last=*/false));
  return enqueue_batch_run_step(b, dt, /*fromSlots=*/true, /*last=*/true);
}

fc_status upload_batch_params(fc_batch* b) {
  std::vector<float> d2(b->S);
  std::vector<SceneCollide> sc(b->S);
  for (uint32_t s = 0; s < b->S; ++s) {
    const fc_sim_params& P = b->params[s];
    d2[s] = contact_d2_threshold(P.particleRadius + P.particleRadius);
    sc[s].invCell = inv_cell(P);
    sc[s].contactD2 = d2[s];
    sc[s].qThreshold = quant_threshold(P.particleRadius + P.particleRadius, batch_frame().scale);
  }
  FC_HIP_TRY(hipMemcpyAsync(b->dSceneCollide, sc.data(), b->S * sizeof(SceneCollide),
                            hipMemcpyHostToDevice, b->ctx->stream));
  FC_HIP_TRY(hipMemcpyAsync(b->dParams, b->params.data(), b->S * sizeof(fc_sim_params),
                            hipMemcpyHostToDevice, b->ctx->stream));
  FC_HIP_TRY(hipMemcpyAsync(b->dContactD2, d2.data(), b->S * sizeof(float),
                            hipMemcpyHostToDevice, b->ctx->stream));
  FC_HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  return FC_SUCCESS;
}

void free_batch(fc_batch* b) {
  drop_batch_graph(b);
  hipFree(b->dOffsets);
  hipFree(b->dParams);
  hipFree(b->dContactD2);
  hipFree(b->dSceneCollide);
  hipFree(b->sortedQ);
  hipFree(b->pos);
  hipFree(b->vel);
  hipFree(b->sortedPos);
  hipFree(b->sortedVel);
  hipFree(b->hash);
  hipFree(b->index);
  hipFree(b->keys);
  hipFree(b->slotOf);
  hipFree(b->idOfSlot);
  hipFree(b->cellStart);
  hipFree(b->cellEnd);
  hipFree(b->stats);
  free_row_tables(b->rowTables);
}

bool valid_tuning(const fc_batch_tuning& t) {
  const int p = t.integratePerThread, c = t.collideBlock;
  return (p == 1 || p == 2 || p == 4 || p == 8) && (c == 64 || c == 128 || c == 256);
}

}  // namespace

}  // namespace fc

using namespace fc;

extern "C" {

fc_status fc_batch_create(fc_context* ctx, uint32_t numScenes, const uint32_t* sceneCounts,
                          fc_uint3 gridSize, fc_batch** out) {
  if (!ctx || !sceneCounts || !out) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  *out = nullptr;
  if (numScenes == 0) return set_error(FC_ERROR_INVALID_VALUE, "a batch needs a scene");
  fc_sim_params P0 = default_params(0, gridSize);
  FC_TRY(validate_params(P0));
  if (gridSize.x < 4)
    return set_error(FC_ERROR_INVALID_VALUE, "batched scenes need gridSize.x >= 4");
  const uint64_t C = (uint64_t)gridSize.x * gridSize.y * gridSize.z;
  if ((uint64_t)numScenes * C > 0xffffffffull)
    return set_error(FC_ERROR_CAPACITY, "numScenes * cells per scene must be < 2^32");
  uint64_t total = 0;
  for (uint32_t s = 0; s < numScenes; ++s) total += sceneCounts[s];
  if (total >= (1ull << 27))
    return set_error(FC_ERROR_CAPACITY, "a batch holds fewer than 2^27 particles");
  fc_batch* b = new fc_batch();
  b->ctx = ctx;
  b->S = numScenes;
  b->n = (uint32_t)total;
  b->grid = gridSize;
  b->C = (uint32_t)C;
  b->cellBits = key_bits_for_cells(b->C);
  b->offsets.resize(numScenes + 1);
  b->offsets[0] = 0;
  for (uint32_t s = 0; s < numScenes; ++s) b->offsets[s + 1] = b->offsets[s] + sceneCounts[s];
  b->params.resize(numScenes);
  for (uint32_t s = 0; s < numScenes; ++s) b->params[s] = default_params(sceneCounts[s], gridSize);
  const size_t n = std::max<size_t>(b->n, 1);
  const size_t cells = (size_t)numScenes * C;
  hipError_t e = hipMalloc(&b->dOffsets, (numScenes + 1) * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&b->dParams, numScenes * sizeof(fc_sim_params));
  if (e == hipSuccess) e = hipMalloc(&b->dContactD2, numScenes * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&b->dSceneCollide, numScenes * sizeof(SceneCollide));
  if (e == hipSuccess) e = hipMalloc(&b->pos, n * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc(&b->vel, n * sizeof(float4));
  // sortedPos and sortedQ carry the row walk's padding past the last sphere
  if (e == hipSuccess) e = hipMalloc(&b->sortedPos, (n + kSortedPosPad) * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc(&b->sortedQ, (n + kSortedPosPad) * sizeof(uint2));
  if (e == hipSuccess) e = hipMalloc(&b->sortedVel, n * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc(&b->hash, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&b->index, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&b->keys, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&b->slotOf, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&b->idOfSlot, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&b->cellStart, cells * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&b->cellEnd, cells * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&b->stats, numScenes * 4 * sizeof(double));
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->dOffsets, b->offsets.data(), (numScenes + 1) * sizeof(uint32_t),
                       hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(b->pos, 0, n * sizeof(float4), ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(b->vel, 0, n * sizeof(float4), ctx->stream);
  if (e != hipSuccess) {
    free_batch(b);
    delete b;
    return hip_error(e, "batch allocation");
  }
  fc_status s = ensure_sort_capacity(ctx, b->n);
  if (s == FC_SUCCESS) s = upload_batch_params(b);
  if (s != FC_SUCCESS) {
    free_batch(b);
    delete b;
    return s;
  }
  *out = b;
  return FC_SUCCESS;
}

fc_status fc_batch_destroy(fc_batch* b) {
  if (!b) return FC_SUCCESS;
  // No use of b->ctx: the context may already be gone; hipFree waits for
  // the device.
  free_batch(b);
  delete b;
  return FC_SUCCESS;
}

uint32_t fc_batch_num_particles(const fc_batch* b) { return b ? b->n : 0u; }

fc_status fc_batch_scene_offsets(const fc_batch* b, uint32_t* offsets) {
  if (!b || !offsets) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  std::memcpy(offsets, b->offsets.data(), (b->S + 1) * sizeof(uint32_t));
  return FC_SUCCESS;
}

fc_status fc_batch_set_params(fc_batch* b, const fc_sim_params* params) {
  if (!b || !params) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  for (uint32_t s = 0; s < b->S; ++s) {
    FC_TRY(validate_params(params[s]));
    if (params[s].gridSize.x != b->grid.x || params[s].gridSize.y != b->grid.y ||
        params[s].gridSize.z != b->grid.z)
      return set_error(FC_ERROR_INVALID_VALUE, "every scene must use the batch's grid size");
  }
  FC_HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  std::memcpy(b->params.data(), params, b->S * sizeof(fc_sim_params));
  drop_batch_graph(b);
  return upload_batch_params(b);
}

fc_status fc_batch_get_params(const fc_batch* b, fc_sim_params* params) {
  if (!b || !params) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  std::memcpy(params, b->params.data(), b->S * sizeof(fc_sim_params));
  return FC_SUCCESS;
}

fc_status fc_batch_set_array(fc_batch* b, fc_particle_array array, const float* data) {
  if (!b || !data) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (array != FC_POSITION && array != FC_VELOCITY)
    return set_error(FC_ERROR_INVALID_VALUE, "unknown particle array");
  if (b->n == 0) return FC_SUCCESS;
  FC_HIP_TRY(hipMemcpyAsync(array == FC_POSITION ? b->pos : b->vel, data,
                            (size_t)b->n * sizeof(float4), hipMemcpyHostToDevice,
                            b->ctx->stream));
  FC_HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  return FC_SUCCESS;
}

fc_status fc_batch_get_array(fc_batch* b, fc_particle_array array, float* out) {
  if (!b || !out) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (array != FC_POSITION && array != FC_VELOCITY)
    return set_error(FC_ERROR_INVALID_VALUE, "unknown particle array");
  if (b->n == 0) return FC_SUCCESS;
  FC_HIP_TRY(hipMemcpyAsync(out, array == FC_POSITION ? b->pos : b->vel,
                            (size_t)b->n * sizeof(float4), hipMemcpyDeviceToHost,
                            b->ctx->stream));
  FC_HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  return check_device_flags(b->ctx);
}

fc_status fc_batch_reset(fc_batch* b, fc_particle_config config, uint64_t seed, float jitterFrac,
                         float velAmp) {
  if (!b) return set_error(FC_ERROR_INVALID_VALUE, "batch is NULL");
  if (config != FC_CONFIG_GRID && config != FC_CONFIG_RANDOM)
    return set_error(FC_ERROR_INVALID_VALUE, "unknown particle config");
  std::vector<float> pos((size_t)b->n * 4), vel((size_t)b->n * 4);
  for (uint32_t s = 0; s < b->S; ++s) {
    const uint32_t o = b->offsets[s], cnt = b->offsets[s + 1] - o;
    generate_initial_state(pos.data() + (size_t)o * 4, vel.data() + (size_t)o * 4, cnt, config,
                           seed + s, b->params[s].particleRadius, jitterFrac, velAmp);
  }
  FC_TRY(fc_batch_set_array(b, FC_POSITION, pos.data()));
  return fc_batch_set_array(b, FC_VELOCITY, vel.data());
}

fc_status fc_batch_update(fc_batch* b, float deltaTime) {
  if (!b) return set_error(FC_ERROR_INVALID_VALUE, "batch is NULL");
  return enqueue_batch_update(b, deltaTime);
}

fc_status fc_batch_run(fc_batch* b, float deltaTime, uint32_t steps) {
  if (!b) return set_error(FC_ERROR_INVALID_VALUE, "batch is NULL");
  if (steps == 0 || b->n == 0) return FC_SUCCESS;
  if (batch_rows_ok(b)) return enqueue_batch_run(b, deltaTime, steps);
  uint32_t i = 0;
  if (steps > 0 && !b->tablesTrackHash) {
    FC_TRY(enqueue_batch_update(b, deltaTime));
    ++i;
  }
  if (i < steps && (!b->graphExec || b->graphDt != deltaTime))
    FC_TRY(capture_batch_graph(b, deltaTime));
  for (; i < steps; ++i) FC_HIP_TRY(hipGraphLaunch(b->graphExec, b->ctx->stream));
  return FC_SUCCESS;
}

fc_status fc_batch_stats_fp64(fc_batch* b, double* out) {
  if (!b || !out) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  hipLaunchKernelGGL(k_batch_energy, dim3(b->S), dim3(kBB), 0, b->ctx->stream, b->vel,
                     b->dOffsets, b->stats);
  FC_HIP_TRY(hipGetLastError());
  FC_HIP_TRY(hipMemcpyAsync(out, b->stats, (size_t)b->S * 4 * sizeof(double),
                            hipMemcpyDeviceToHost, b->ctx->stream));
  FC_HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  return FC_SUCCESS;
}

fc_status fc_batch_set_tuning(fc_batch* b, const fc_batch_tuning* t) {
  if (!b || !t) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (!valid_tuning(*t))
    return set_error(FC_ERROR_INVALID_VALUE,
                     "integratePerThread must be 1, 2, 4 or 8 and collideBlock 64, 128 or 256");
  FC_HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  b->tuning.integratePerThread = t->integratePerThread;
  b->tuning.collideBlock = t->collideBlock;
  drop_batch_graph(b);
  return FC_SUCCESS;
}

fc_status fc_batch_get_tuning(const fc_batch* b, fc_batch_tuning* t) {
  if (!b || !t) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  *t = b->tuning;
  return FC_SUCCESS;
}

fc_status fc_batch_autotune(fc_batch* b, float deltaTime, uint32_t stepsPerCandidate,
                            fc_batch_tuning* best, float* candidateMs, uint32_t* numCandidates) {
  if (!b) return set_error(FC_ERROR_INVALID_VALUE, "batch is NULL");
  if (stepsPerCandidate == 0) stepsPerCandidate = 3;
  hipStream_t st = b->ctx->stream;
  const size_t bytes = (size_t)std::max<uint32_t>(b->n, 1) * sizeof(float4);
  // The trials step a copy of the state; the batch is restored afterwards.
  float4 *savePos = nullptr, *saveVel = nullptr;
  FC_HIP_TRY(hipMalloc(&savePos, bytes));
  hipError_t e = hipMalloc(&saveVel, bytes);
  if (e == hipSuccess) e = hipMemcpyAsync(savePos, b->pos, bytes, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(saveVel, b->vel, bytes, hipMemcpyDeviceToDevice, st);
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&t0);
  if (e == hipSuccess) e = hipEventCreate(&t1);
  const int pers[4] = {1, 2, 4, 8};
  const int blocks[3] = {64, 128, 256};
  fc_batch_tuning keep = b->tuning, top = b->tuning;
  float topMs = 0.0f;
  uint32_t c = 0;
  fc_status s = FC_SUCCESS;
  for (int pi = 0; pi < 4 && e == hipSuccess && s == FC_SUCCESS; ++pi)
    for (int bi = 0; bi < 3 && e == hipSuccess && s == FC_SUCCESS; ++bi, ++c) {
      b->tuning.integratePerThread = pers[pi];
      b->tuning.collideBlock = blocks[bi];
      e = hipMemcpyAsync(b->pos, savePos, bytes, hipMemcpyDeviceToDevice, st);
      if (e == hipSuccess) e = hipMemcpyAsync(b->vel, saveVel, bytes, hipMemcpyDeviceToDevice, st);
      b->tablesTrackHash = false;
      if (e != hipSuccess) break;
      s = enqueue_batch_update(b, deltaTime);  // warm-up (clears the tables)
      if (s != FC_SUCCESS) break;
      e = hipEventRecord(t0, st);
      for (uint32_t k = 0; k < stepsPerCandidate && s == FC_SUCCESS; ++k)
        s = enqueue_batch_update(b, deltaTime);
      if (e == hipSuccess) e = hipEventRecord(t1, st);
      if (e == hipSuccess) e = hipEventSynchronize(t1);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
      ms /= (float)stepsPerCandidate;
      if (candidateMs) candidateMs[c] = ms;
      if (c == 0 || ms < topMs) {
        topMs = ms;
        top = b->tuning;
        top.stepMs = ms;
      }
    }
  if (e == hipSuccess) e = hipMemcpyAsync(b->pos, savePos, bytes, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(b->vel, saveVel, bytes, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (t0) hipEventDestroy(t0);
  if (t1) hipEventDestroy(t1);
  hipFree(savePos);
  hipFree(saveVel);
  b->tablesTrackHash = false;
  drop_batch_graph(b);
  if (s != FC_SUCCESS) {
    b->tuning = keep;
    return s;
  }
  if (e != hipSuccess) {
    b->tuning = keep;
    return hip_error(e, "batch autotune");
  }
  b->tuning = top;
  if (best) *best = top;
  if (numCandidates) *numCandidates = c;
  return FC_SUCCESS;
}

}  // extern "C"
