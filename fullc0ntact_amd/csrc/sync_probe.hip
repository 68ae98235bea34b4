// Cost of the two ways to separate the stages of a step: a grid-wide
// barrier inside one persistent (cooperative) launch, against a kernel
// boundary inside a captured graph.  A batched step has a stage boundary
// after integrate/hash, after every sort pass and after the cell tables;
// a persistent launch only pays off if its barriers cost less than the
// graph's kernel boundaries (the step's state, hundreds of MB, cannot stay
// on chip between stages either way).
#include "fc_internal.hpp"

#include <hip/hip_cooperative_groups.h>

namespace cg = cooperative_groups;

namespace fc {
namespace {

constexpr int kProbeBlock = 256;

// `barriers` grid-wide barriers; each block's thread 0 bumps its own
// counter between them, so the barrier orders real memory traffic.
__global__ __launch_bounds__(kProbeBlock) void k_grid_barriers(uint32_t* __restrict__ work,
                                                               int barriers) {
  cg::grid_group grid = cg::this_grid();
  for (int i = 0; i < barriers; ++i) {
    if (threadIdx.x == 0) work[blockIdx.x] += 1u;
    grid.sync();
  }
}

// One stage of the graph: the same per-block store, then the kernel ends.
__global__ __launch_bounds__(kProbeBlock) void k_stage(uint32_t* __restrict__ work) {
  if (threadIdx.x == 0) work[blockIdx.x] += 1u;
}

}  // namespace
}  // namespace fc

using namespace fc;

extern "C" {

fc_status fc_sync_probe(fc_context* ctx, uint32_t blocks, int iterations, float out[2]) {
  if (!ctx || !out || iterations < 1 || blocks == 0)
    return set_error(FC_ERROR_INVALID_VALUE, "bad sync-probe arguments");
  int coop = 0;
  FC_HIP_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, ctx->device));
  if (!coop) return set_error(FC_ERROR_NOT_SUPPORTED, "no cooperative launch on this device");
  int perCu = 0, cus = 0;
  FC_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, k_grid_barriers, kProbeBlock, 0));
  FC_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  if ((uint64_t)blocks > (uint64_t)perCu * (uint64_t)cus)
    return set_error(FC_ERROR_CAPACITY, "more blocks than can be resident at once");
  hipStream_t st = ctx->stream;
  uint32_t* work = nullptr;
  FC_HIP_TRY(hipMalloc(&work, blocks * sizeof(uint32_t)));
  hipEvent_t a = nullptr, b = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipError_t e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  if (e == hipSuccess) e = hipMemsetAsync(work, 0, blocks * sizeof(uint32_t), st);
  // persistent: one cooperative launch with `iterations` grid barriers
  int barriers = iterations;
  void* args[2] = {&work, &barriers};
  float msCoop = 0.0f, msGraph = 0.0f;
  for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {  // rep 0 warms up
    e = hipEventRecord(a, st);
    if (e == hipSuccess)
      e = hipLaunchCooperativeKernel((const void*)k_grid_barriers, dim3(blocks),
                                     dim3(kProbeBlock), args, 0, st);
    if (e == hipSuccess) e = hipEventRecord(b, st);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(&msCoop, a, b);
  }
  // graph: `iterations` single-stage kernels captured back to back
  if (e == hipSuccess) e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    for (int i = 0; i < iterations; ++i)
      hipLaunchKernelGGL(k_stage, dim3(blocks), dim3(kProbeBlock), 0, st, work);
    hipError_t le = hipGetLastError();
    e = hipStreamEndCapture(st, &graph);
    if (e == hipSuccess) e = le;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
    e = hipEventRecord(a, st);
    if (e == hipSuccess) e = hipGraphLaunch(exec, st);
    if (e == hipSuccess) e = hipEventRecord(b, st);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(&msGraph, a, b);
  }
  if (exec) hipGraphExecDestroy(exec);
  if (graph) hipGraphDestroy(graph);
  if (a) hipEventDestroy(a);
  if (b) hipEventDestroy(b);
  hipStreamSynchronize(st);
  hipFree(work);
  if (e != hipSuccess) return hip_error(e, "sync probe");
  out[0] = msCoop * 1000.0f / (float)iterations;   // us per grid barrier
  out[1] = msGraph * 1000.0f / (float)iterations;  // us per graph kernel boundary
  return FC_SUCCESS;
}

}  // extern "C"
