// Particle-step kernels for gfx950: integrate, calcHash, reorder /
// find-cell-start and the 27-cell collide.
//
// Layout in HBM: positions and velocities are float4[N] arrays (x, y, z, w),
// read and written as one 16-byte vector per particle per lane, so a wave64
// touches 1 KiB of consecutive memory per access.  Hash, index and cell
// tables are uint32 arrays.  Parameters arrive by value as kernel arguments.
#include "fc_internal.hpp"
#include "fc_device.hpp"

namespace fc {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(size_t n, int block) {
  size_t g = (n + (size_t)block - 1) / (size_t)block;
  return (unsigned)(g == 0 ? 1 : g);
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* __restrict__ p, uint32_t value,
                                                     size_t count) {
  size_t tid = (size_t)blockIdx.x * kBlock + threadIdx.x;
  size_t stride = (size_t)gridDim.x * kBlock;
  // 16-byte stores over the aligned body, scalar stores for the ragged ends.
  size_t head = ((16 - ((uintptr_t)p & 15)) & 15) / 4;
  if (head > count) head = count;
  if (tid < head) p[tid] = value;
  uint32_t* body = p + head;
  size_t nvec = (count - head) / 4;
  uint4 v = make_uint4(value, value, value, value);
  for (size_t i = tid; i < nvec; i += stride) reinterpret_cast<uint4*>(body)[i] = v;
  size_t tail = head + nvec * 4;
  if (tid < count - tail) p[tail + tid] = value;
}

// integrate_functor (particles_kernel.cu:39-75) over zip(pos, vel).
__global__ __launch_bounds__(kBlock) void k_integrate(fc_sim_params P, float4* __restrict__ pos,
                                                      float4* __restrict__ vel, float dt,
                                                      uint32_t n) {
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float4 p4 = pos[i];
  float4 v4 = vel[i];
  v3f p = xyz(p4), v = xyz(v4);
  integrate_box(P, dt, p, v);
  pos[i] = make_float4(p.x, p.y, p.z, p4.w);
  vel[i] = make_float4(v.x, v.y, v.z, v4.w);
}

// calcHashD (particles_kernel.cu:97-115).
__global__ __launch_bounds__(kBlock) void k_calc_hash(fc_sim_params P, uint32_t* __restrict__ hash,
                                                      uint32_t* __restrict__ index,
                                                      const float4* __restrict__ pos, uint32_t n) {
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float4 p = pos[i];
  hash[i] = hash_of(P, p.x, p.y, p.z);
  index[i] = i;
}

// integrate + calcHash in one pass (the update order runs them back to back
// on the same arrays, so fusing them changes no value).  When
// prevSortedHash is given, the cells occupied after the previous update are
// reset to EMPTY: cell h was written iff h starts a run in the sorted hash
// array, so exactly those entries are cleared (instead of the reference's
// memset of the whole table, particleSystem.cu:198).
__global__ __launch_bounds__(kBlock) void k_integrate_hash(
    fc_sim_params P, float4* __restrict__ pos, float4* __restrict__ vel, float dt,
    uint32_t* __restrict__ keysOut, const uint32_t* __restrict__ prevSortedHash,
    uint32_t* __restrict__ cellStart, uint32_t* __restrict__ cellEnd, uint32_t n) {
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  if (prevSortedHash != nullptr) {
    uint32_t h = prevSortedHash[i];
    if (i == 0 || prevSortedHash[i - 1] != h) {
      cellStart[h] = CELL_EMPTY;
      cellEnd[h] = CELL_EMPTY;
    }
  }
  float4 p4 = pos[i];
  float4 v4 = vel[i];
  v3f p = xyz(p4), v = xyz(v4);
  integrate_box(P, dt, p, v);
  pos[i] = make_float4(p.x, p.y, p.z, p4.w);
  vel[i] = make_float4(v.x, v.y, v.z, v4.w);
  keysOut[i] = hash_of(P, p.x, p.y, p.z);
}

// reorderDataAndFindCellStartD (particles_kernel.cu:119-182).  The
// neighbour's hash comes from the previous lane through a cross-lane
// shuffle; lane 0 of each wave reads it from memory.
__global__ __launch_bounds__(kBlock) void k_reorder(
    uint32_t* __restrict__ cellStart, uint32_t* __restrict__ cellEnd,
    float4* __restrict__ sortedPos, float4* __restrict__ sortedVel,
    const uint32_t* __restrict__ hash, const uint32_t* __restrict__ index,
    const float4* __restrict__ oldPos, const float4* __restrict__ oldVel, uint32_t n,
    uint32_t numCells) {
  uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  uint32_t h = (k < n) ? hash[k] : 0u;
  uint32_t hprev = __shfl_up(h, 1);
  if (k >= n) return;
  if ((threadIdx.x & 63u) == 0u && k > 0) hprev = hash[k - 1];
  // Hashes produced by calcHash are < numCells; the guards only keep
  // caller-supplied out-of-range hashes from writing outside the tables.
  if (k == 0 || h != hprev) {
    if (h < numCells) cellStart[h] = k;
    if (k > 0 && hprev < numCells) cellEnd[hprev] = k;
  }
  if (k == n - 1 && h < numCells) cellEnd[h] = k + 1;
  uint32_t src = index[k];
  sortedPos[k] = oldPos[src];
  sortedVel[k] = oldVel[src];
}

// collideD / collideCell (particles_kernel.cu:223-293): one lane per sorted
// particle; for z, y, x in {-1,0,1} (that nesting is the summation order),
// each neighbour cell's forces are summed into a per-cell partial that is
// then added to the total, then the collider sphere is added, and
// vel + force is scattered to the particle's original slot.
__global__ __launch_bounds__(kBlock) void k_collide(
    fc_sim_params P, float4* __restrict__ newVel, const float4* __restrict__ sortedPos,
    const float4* __restrict__ sortedVel, const uint32_t* __restrict__ index,
    const uint32_t* __restrict__ cellStart, const uint32_t* __restrict__ cellEnd, uint32_t n) {
  uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  const v3f pos = xyz(sortedPos[k]);
  const v3f vel = xyz(sortedVel[k]);
  const int gx = cell_coord(pos.x, P.worldOrigin.x, P.cellSize.x);
  const int gy = cell_coord(pos.y, P.worldOrigin.y, P.cellSize.y);
  const int gz = cell_coord(pos.z, P.worldOrigin.z, P.cellSize.z);
  const float r = P.particleRadius;
  v3f force = make_v3(0.0f, 0.0f, 0.0f);
  for (int dz = -1; dz <= 1; ++dz) {
    for (int dy = -1; dy <= 1; ++dy) {
      for (int dx = -1; dx <= 1; ++dx) {
        uint32_t h = grid_hash(gx + dx, gy + dy, gz + dz, P.gridSize);
        uint32_t start = cellStart[h];
        v3f cf = make_v3(0.0f, 0.0f, 0.0f);
        if (start != CELL_EMPTY) {
          uint32_t end = cellEnd[h];
          if (end > n) end = n;  // consistent tables never exceed n
          for (uint32_t j = start; j < end; ++j) {
            if (j != k) {
              v3f f = collide_spheres(pos, xyz(sortedPos[j]), vel, xyz(sortedVel[j]), r, r,
                                      P.spring, P.damping, P.shear, P.attraction);
              cf = make_v3(cf.x + f.x, cf.y + f.y, cf.z + f.z);
            }
          }
        }
        force = make_v3(force.x + cf.x, force.y + cf.y, force.z + cf.z);
      }
    }
  }
  v3f cpos = make_v3(P.colliderPos.x, P.colliderPos.y, P.colliderPos.z);
  v3f f = collide_spheres(pos, cpos, vel, make_v3(0.0f, 0.0f, 0.0f), r, P.colliderRadius,
                          P.spring, P.damping, P.shear, 0.0f);
  force = make_v3(force.x + f.x, force.y + f.y, force.z + f.z);
  uint32_t orig = index[k];
  newVel[orig] = make_float4(vel.x + force.x, vel.y + force.y, vel.z + force.z, 0.0f);
}

}  // namespace

hipError_t launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t value, size_t count) {
  if (count == 0) return hipSuccess;
  size_t vec = count / 4 + 2;
  unsigned grid = grid_for(vec, kBlock);
  if (grid > 2048u) grid = 2048u;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(kBlock), 0, s, p, value, count);
  return hipGetLastError();
}

hipError_t launch_integrate(hipStream_t s, const fc_sim_params& P, float4* pos, float4* vel,
                            float dt, uint32_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_integrate, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, P, pos, vel, dt,
                     n);
  return hipGetLastError();
}

hipError_t launch_calc_hash(hipStream_t s, const fc_sim_params& P, uint32_t* hash,
                            uint32_t* index, const float4* pos, uint32_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_calc_hash, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, P, hash, index,
                     pos, n);
  return hipGetLastError();
}

hipError_t launch_integrate_hash(hipStream_t s, const fc_sim_params& P, float4* pos, float4* vel,
                                 float dt, uint32_t* keysOut, const uint32_t* prevSortedHash,
                                 uint32_t* cellStart, uint32_t* cellEnd, uint32_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_integrate_hash, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, P, pos, vel,
                     dt, keysOut, prevSortedHash, cellStart, cellEnd, n);
  return hipGetLastError();
}

hipError_t launch_reorder(hipStream_t s, uint32_t* cellStart, uint32_t* cellEnd, float4* sortedPos,
                          float4* sortedVel, const uint32_t* hash, const uint32_t* index,
                          const float4* oldPos, const float4* oldVel, uint32_t n,
                          uint32_t numCells) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_reorder, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, cellStart, cellEnd,
                     sortedPos, sortedVel, hash, index, oldPos, oldVel, n, numCells);
  return hipGetLastError();
}

hipError_t launch_collide(hipStream_t s, const fc_sim_params& P, float4* newVel,
                          const float4* sortedPos, const float4* sortedVel,
                          const uint32_t* index, const uint32_t* cellStart,
                          const uint32_t* cellEnd, uint32_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_collide, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, P, newVel,
                     sortedPos, sortedVel, index, cellStart, cellEnd, n);
  return hipGetLastError();
}

}  // namespace fc
