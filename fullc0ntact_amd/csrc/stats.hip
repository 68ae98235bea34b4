// Global observables of a particle state in FP64: kinetic energy and
// momentum (deterministic two-level reduction), and the contact-count
// histogram of the sorted state (exact integer counts).
#include "fc_internal.hpp"

namespace fc {

namespace {

constexpr int kStatBlock = 256;
constexpr uint32_t kStatMaxBlocks = 1024;

// Block-wide sum of four doubles in a fixed tree order; the result is
// valid in thread 0.
__device__ __forceinline__ void block_sum4(double v[4], double (*sh)[kStatBlock]) {
#pragma unroll
  for (int c = 0; c < 4; ++c) sh[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int s = kStatBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int c = 0; c < 4; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + s];
    }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) v[c] = sh[c][0];
}

// Partial sums per block over a grid-stride range: 1/2 |v|^2 and v, each
// float component widened to double before any arithmetic.
__global__ __launch_bounds__(kStatBlock) void k_energy_momentum_partial(
    const float4* __restrict__ vel, uint32_t n, double* __restrict__ partial) {
  __shared__ double sh[4][kStatBlock];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint32_t i = blockIdx.x * kStatBlock + threadIdx.x; i < n; i += gridDim.x * kStatBlock) {
    const float4 v = vel[i];
    const double x = v.x, y = v.y, z = v.z;
    acc[0] += 0.5 * (x * x + y * y + z * z);
    acc[1] += x;
    acc[2] += y;
    acc[3] += z;
  }
  block_sum4(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < 4; ++c) partial[4 * blockIdx.x + c] = acc[c];
  }
}

__global__ __launch_bounds__(kStatBlock) void k_energy_momentum_final(
    const double* __restrict__ partial, uint32_t blocks, double* __restrict__ out) {
  __shared__ double sh[4][kStatBlock];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint32_t b = threadIdx.x; b < blocks; b += kStatBlock) {
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] += partial[4 * b + c];
  }
  block_sum4(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < 4; ++c) out[c] = acc[c];
  }
}

// Contacts of sorted sphere k: neighbours j != k in the 27 cells (the
// collide's walk, wrapped cells included) with sqrtf(d.d) < rA + rB, the
// collide's contact test (particles_kernel.cu:185-218).  Radii are the
// parameter radius, or pos.w when kPoly.
template <bool kPoly>
__global__ __launch_bounds__(kStatBlock) void k_contact_histogram(
    fc_sim_params P, const float4* __restrict__ sortedPos, const uint32_t* __restrict__ cellStart,
    const uint32_t* __restrict__ cellEnd, uint32_t n, uint32_t* __restrict__ hist,
    uint32_t nbins) {
  extern __shared__ uint32_t s_hist[];
  for (uint32_t b = threadIdx.x; b < nbins; b += kStatBlock) s_hist[b] = 0u;
  __syncthreads();
  const uint32_t k = blockIdx.x * kStatBlock + threadIdx.x;
  if (k < n) {
    const float4 p4 = sortedPos[k];
    const v3f pos = xyz(p4);
    const float rA = kPoly ? p4.w : P.particleRadius;
    const int gx = cell_coord(pos.x, P.worldOrigin.x, P.cellSize.x);
    const int gy = cell_coord(pos.y, P.worldOrigin.y, P.cellSize.y);
    const int gz = cell_coord(pos.z, P.worldOrigin.z, P.cellSize.z);
    uint32_t count = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const uint32_t h = grid_hash(gx + dx, gy + dy, gz + dz, P.gridSize);
          const uint32_t st = cellStart[h];
          if (st == FC_CELL_EMPTY) continue;
          uint32_t e = cellEnd[h];
          if (e > n) e = n;
          for (uint32_t j = st; j < e; ++j) {
            if (j == k) continue;
            const float4 q4 = sortedPos[j];
            const v3f rel = make_v3(q4.x - pos.x, q4.y - pos.y, q4.z - pos.z);
            const float rB = kPoly ? q4.w : P.particleRadius;
            if (sqrtf(dot3(rel, rel)) < rA + rB) ++count;
          }
        }
    atomicAdd(&s_hist[count < nbins ? count : nbins - 1], 1u);
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < nbins; b += kStatBlock)
    if (s_hist[b]) atomicAdd(&hist[b], s_hist[b]);
}

}  // namespace

uint32_t energy_blocks(uint32_t n) {
  uint32_t b = (n + kStatBlock * 4 - 1) / (kStatBlock * 4);
  if (b == 0) b = 1;
  return b > kStatMaxBlocks ? kStatMaxBlocks : b;
}

// Context scratch of the statistics: per-block partials (kStatMaxBlocks x 4
// doubles), the 4 results, then kStatMaxBins u32 histogram bins.
constexpr uint32_t kStatMaxBins = 4096;

fc_status ensure_stats_scratch(fc_context* ctx) {
  if (ctx->statsScratch) return FC_SUCCESS;
  FC_HIP_TRY(hipMalloc(&ctx->statsScratch,
                       (kStatMaxBlocks * 4 + 4) * sizeof(double) + kStatMaxBins * sizeof(uint32_t)));
  return FC_SUCCESS;
}

static uint32_t* stats_hist(fc_context* ctx) {
  return reinterpret_cast<uint32_t*>(ctx->statsScratch + kStatMaxBlocks * 4 + 4);
}

hipError_t launch_energy_momentum(hipStream_t s, const float4* vel, uint32_t n, double* partial,
                                  double* out) {
  const uint32_t blocks = energy_blocks(n);
  hipLaunchKernelGGL(k_energy_momentum_partial, dim3(blocks), dim3(kStatBlock), 0, s, vel, n,
                     partial);
  hipLaunchKernelGGL(k_energy_momentum_final, dim3(1), dim3(kStatBlock), 0, s, partial, blocks,
                     out);
  return hipGetLastError();
}

hipError_t launch_contact_histogram(hipStream_t s, const fc_sim_params& P, bool poly,
                                    const float4* sortedPos, const uint32_t* cellStart,
                                    const uint32_t* cellEnd, uint32_t n, uint32_t* hist,
                                    uint32_t nbins) {
  hipError_t e = launch_fill_u32(s, hist, 0u, nbins);
  if (e != hipSuccess || n == 0) return e;
  const dim3 g((n + kStatBlock - 1) / kStatBlock), b(kStatBlock);
  const size_t lds = nbins * sizeof(uint32_t);
  if (poly)
    hipLaunchKernelGGL(k_contact_histogram<true>, g, b, lds, s, P, sortedPos, cellStart, cellEnd,
                       n, hist, nbins);
  else
    hipLaunchKernelGGL(k_contact_histogram<false>, g, b, lds, s, P, sortedPos, cellStart, cellEnd,
                       n, hist, nbins);
  return hipGetLastError();
}

}  // namespace fc

using namespace fc;

extern "C" {

fc_status fc_stats_fp64(fc_context* ctx, const float* vel, uint32_t numParticles, double out[4]) {
  if (!ctx || !out) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (numParticles && !vel) return set_error(FC_ERROR_INVALID_VALUE, "vel is NULL");
  FC_TRY(ensure_stats_scratch(ctx));
  double* partial = ctx->statsScratch;
  double* dOut = ctx->statsScratch + kStatMaxBlocks * 4;
  FC_HIP_TRY(launch_energy_momentum(ctx->stream, (const float4*)vel, numParticles, partial, dOut));
  FC_HIP_TRY(hipMemcpyAsync(out, dOut, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  FC_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return FC_SUCCESS;
}

fc_status fc_contact_histogram(fc_context* ctx, const float* sortedPos, const uint32_t* cellStart,
                               const uint32_t* cellEnd, uint32_t numParticles, int polydisperse,
                               uint32_t* hist, uint32_t nbins) {
  if (!ctx || !hist) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (nbins == 0 || nbins > kStatMaxBins)
    return set_error(FC_ERROR_INVALID_VALUE, "nbins must be in [1, 4096]");
  if (numParticles && (!sortedPos || !cellStart || !cellEnd))
    return set_error(FC_ERROR_INVALID_VALUE, "NULL array");
  FC_HIP_TRY(launch_contact_histogram(ctx->stream, ctx->params, polydisperse != 0,
                                      (const float4*)sortedPos, cellStart, cellEnd, numParticles,
                                      hist, nbins));
  return FC_SUCCESS;
}

fc_status fc_system_stats_fp64(fc_system* sys, double out[4]) {
  if (!sys) return set_error(FC_ERROR_INVALID_VALUE, "system is NULL");
  return fc_stats_fp64(sys->ctx, (const float*)sys->vel, sys->n, out);
}

fc_status fc_system_contact_histogram(fc_system* sys, uint32_t* hist, uint32_t nbins) {
  if (!sys || !hist) return set_error(FC_ERROR_INVALID_VALUE, "NULL argument");
  if (nbins == 0 || nbins > kStatMaxBins)
    return set_error(FC_ERROR_INVALID_VALUE, "nbins must be in [1, 4096]");
  fc_context* ctx = sys->ctx;
  FC_TRY(ensure_stats_scratch(ctx));
  uint32_t* dHist = stats_hist(ctx);
  FC_HIP_TRY(launch_contact_histogram(ctx->stream, sys->params, sys->polydisperse != 0,
                                      sys->sortedPos, sys->cellStart, sys->cellEnd, sys->n, dHist,
                                      nbins));
  FC_HIP_TRY(hipMemcpyAsync(hist, dHist, nbins * sizeof(uint32_t), hipMemcpyDeviceToHost,
                            ctx->stream));
  FC_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return FC_SUCCESS;
}

}  // extern "C"
