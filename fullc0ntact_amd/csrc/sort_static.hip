// The (hash, index) sort for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc): fc_sort_particles_noalloc runs the same onesweep radix sort as
// fc_sort_particles, with its scratch (ping-pong key/value buffers, look-back
// status words, digit histograms, counters) in one file-scope device array
// instead of the context's growable buffers.
//
// The array is module memory: every device the library runs on holds one
// copy of kNoAllocScratchBytes until the process ends.  Host code never uses
// the array's name as an address (in host code it is only a registration
// stand-in); each device's address is looked up once, with that device
// current, when the first context is created on it, and kept per device.
#include "fc_internal.hpp"

#include <mutex>

namespace fc {

namespace {

// Largest sort of the no-allocation entry point: 2^20 pairs (the 1M-sphere
// workload).  Up to 256 tiles of RS_TILE keys, so every tile of a pass is
// resident at once (the unticketed look-back).
constexpr size_t kNoAllocMaxKeys = (size_t)1 << 20;
constexpr size_t kNoAllocTiles = (kNoAllocMaxKeys + RS_TILE - 1) / RS_TILE;
constexpr size_t kStatusWords = kNoAllocTiles * RS_RADIX;
constexpr size_t kGroupWords = ((kNoAllocTiles + RS_LOOKBACK_GROUP - 1) / RS_LOOKBACK_GROUP) *
                               RS_RADIX;
constexpr size_t kHistWords = (size_t)RS_MAX_PASSES * RS_HIST_COPIES * RS_RADIX;

// Layout (byte offsets, each 256-byte aligned): the words that must start
// at zero come first (status, group status, histograms, counters), so one
// fill clears them; the key/value buffers follow.
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kOffStatus = 0;
constexpr size_t kOffGroup = align256(kOffStatus + kStatusWords * 8);
constexpr size_t kOffHist = align256(kOffGroup + kGroupWords * 8);
constexpr size_t kOffCounters = align256(kOffHist + kHistWords * 4);
constexpr size_t kZeroBytes = align256(kOffCounters + (size_t)RSC_COUNT * 4);
constexpr size_t kOffKeysA = kZeroBytes;
constexpr size_t kOffValsA = kOffKeysA + kNoAllocMaxKeys * 4;
constexpr size_t kOffKeysB = kOffValsA + kNoAllocMaxKeys * 4;
constexpr size_t kOffValsB = kOffKeysB + kNoAllocMaxKeys * 4;
constexpr size_t kNoAllocScratchBytes = kOffValsB + kNoAllocMaxKeys * 4;
static_assert(kZeroBytes % 16 == 0, "the clear runs in 16-byte stores");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
unsigned char* g_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

// Clears the words of the scratch that must start at zero, with vector
// stores from a kernel (no memset node lands in a captured graph).
__global__ __launch_bounds__(256) void k_clear_words(uint4* __restrict__ p, size_t count) {
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < count; i += stride)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
}

}  // namespace

// Resolves the device address of the scratch array on `device` (the
// current device) once; called by fc_context_create.
fc_status init_noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices)
    return set_error(FC_ERROR_NO_DEVICE, "device index beyond the no-allocation table");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_ptr[device] != nullptr) return FC_SUCCESS;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr) return hip_error(e, "hipGetSymbolAddress(scratch)");
  g_ptr[device] = static_cast<unsigned char*>(ptr);
  return FC_SUCCESS;
}

}  // namespace fc

using namespace fc;

extern "C" {

size_t fc_sort_particles_noalloc_max(void) { return kNoAllocMaxKeys; }

fc_status fc_sort_particles_noalloc(fc_context* ctx, uint32_t* hash, uint32_t* index,
                                    uint32_t numParticles) {
  if (!ctx) return set_error(FC_ERROR_INVALID_VALUE, "ctx is NULL");
  if (numParticles == 0) return FC_SUCCESS;
  if (!hash || !index) return set_error(FC_ERROR_INVALID_VALUE, "NULL array");
  if ((size_t)numParticles > kNoAllocMaxKeys)
    return set_error(FC_ERROR_CAPACITY,
                     "fc_sort_particles_noalloc sorts at most 1048576 pairs; use "
                     "fc_sort_particles for larger arrays");
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return set_error(FC_ERROR_NO_DEVICE, "context device beyond the no-allocation table");
  unsigned char* base;
  {
    std::lock_guard<std::mutex> lock(g_mutex);
    base = g_ptr[ctx->device];
  }
  if (base == nullptr)
    return set_error(FC_ERROR_INVALID_VALUE, "the no-allocation scratch of this device was "
                                             "not resolved at context creation");
  SortScratch s;
  s.status = reinterpret_cast<uint64_t*>(base + kOffStatus);
  s.statusCapacity = kStatusWords;
  s.groupStatus = reinterpret_cast<uint64_t*>(base + kOffGroup);
  s.groupStatusCapacity = kGroupWords;
  s.histogram = reinterpret_cast<uint32_t*>(base + kOffHist);
  s.counters = reinterpret_cast<uint32_t*>(base + kOffCounters);
  s.keysA = reinterpret_cast<uint32_t*>(base + kOffKeysA);
  s.valsA = reinterpret_cast<uint32_t*>(base + kOffValsA);
  s.keysB = reinterpret_cast<uint32_t*>(base + kOffKeysB);
  s.valsB = reinterpret_cast<uint32_t*>(base + kOffValsB);
  s.capacity = kNoAllocMaxKeys;
  // Status words, histograms and counters from zero: the look-back epoch
  // restarts, and no word of an earlier call reads as published.
  hipLaunchKernelGGL(k_clear_words, dim3(64), dim3(256), 0, ctx->stream,
                     reinterpret_cast<uint4*>(base), kZeroBytes / 16);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_error(e, "clear no-allocation scratch");
  e = radix_sort_pairs(ctx->stream, s, ctx->devFlags, hash, index, hash, index, numParticles,
                       key_bits_for_cells(ctx->params.numCells), nullptr);
  if (e != hipSuccess) return hip_error(e, "radix_sort_pairs (no allocation)");
  return FC_SUCCESS;
}

}  // extern "C"
