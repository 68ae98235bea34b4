// Internal declarations shared by the HIP translation units of libfc_dem.so.
// Not part of the public ABI (that is include/fc_dem.h).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>

#include "fc_dem.h"
#include "distmap.hpp"

struct fc_context;

namespace fc {

// ---------------------------------------------------------------------------
// Errors.  Every entry point returns an fc_status and leaves a thread-local
// message readable through fc_last_error().
// ---------------------------------------------------------------------------
fc_status set_error(fc_status s, const std::string& msg);
fc_status hip_error(hipError_t e, const char* what);

#define FC_HIP_TRY(call)                                   \
  do {                                                     \
    hipError_t fc_e_ = (call);                             \
    if (fc_e_ != hipSuccess) return ::fc::hip_error(fc_e_, #call); \
  } while (0)

#define FC_TRY(call)                                       \
  do {                                                     \
    fc_status fc_s_ = (call);                              \
    if (fc_s_ != FC_SUCCESS) return fc_s_;                 \
  } while (0)

// Bits of the device-side error word (ctx->devFlags).
enum : uint32_t {
  DEVERR_SORT_STALLED = 1u << 0,   // look-back exceeded its spin bound
  DEVERR_SORT_KEY_RANGE = 1u << 1, // a key had bits above the sorted width
  DEVERR_HIST_STALLED = 1u << 2,   // histogram totals never reached n
  DEVERR_RADIUS = 1u << 3,         // polydisperse radius above particleRadius
  DEVERR_ROW_COUNT = 1u << 4,      // row resort: a row's members and counts disagree
};

constexpr uint32_t CELL_EMPTY = 0xffffffffu;
// Entries allocated past the last sphere in the fixed-point coordinates
// sortedQ (read, and discarded, by the row-range collide's batched loads;
// a system's sortedPos is padded alike).
constexpr uint32_t kSortedPosPad = 8;

// ---------------------------------------------------------------------------
// Radix sort (radix_sort.hip).
// ---------------------------------------------------------------------------
constexpr int RS_THREADS = 1024;              // 16 wave64s
#ifndef FC_RS_ITEMS
#define FC_RS_ITEMS 4
#endif
constexpr int RS_ITEMS = FC_RS_ITEMS;         // keys per thread per tile
constexpr int RS_TILE = RS_THREADS * RS_ITEMS; // 4096 keys
constexpr int RS_RADIX_BITS = 8;
constexpr int RS_RADIX = 1 << RS_RADIX_BITS;
constexpr int RS_MAX_PASSES = 4;
constexpr int RS_MAX_HIST_BLOCKS = 512;
constexpr int RS_HIST_COPIES = 16;
// Tiles per group of the sort's two-level look-back.
#ifndef FC_RS_GROUP
#define FC_RS_GROUP 16
#endif
constexpr int RS_LOOKBACK_GROUP = FC_RS_GROUP;
inline size_t lookback_groups(size_t tiles) {
  return (tiles + RS_LOOKBACK_GROUP - 1) / RS_LOOKBACK_GROUP;
}  // privatised copies of the global histogram

// Device counters used by the sort; all self-resetting (see radix_sort.hip).
enum : int {
  RSC_TILE0 = 1,      // RSC_TILE0 + pass: tile-id ticket of each pass
  RSC_EPOCH = 8,      // bumped once per sort by the kernel producing the keys
  RSC_COUNT = 16
};

struct SortScratch {
  uint32_t* keysA = nullptr;
  uint32_t* valsA = nullptr;
  uint32_t* keysB = nullptr;
  uint32_t* valsB = nullptr;
  size_t capacity = 0;               // elements in keysA/valsA/keysB/valsB
  uint64_t* status = nullptr;        // tile x digit look-back words
  size_t statusCapacity = 0;         // in words
  uint64_t* groupStatus = nullptr;   // tile-group x digit look-back words
  size_t groupStatusCapacity = 0;    // in words
  uint32_t* histogram = nullptr;     // [RS_MAX_PASSES][RS_HIST_COPIES][RS_RADIX]
  uint32_t* counters = nullptr;      // [RSC_COUNT]
};

struct StageTimer;

// Number of significant bits needed for keys in [0, numCells).
int key_bits_for_cells(uint32_t numCells);

// Number of scatter passes for `bits`-bit keys.
int radix_passes(int bits, bool inPlace);

// Frame of the 16-bit fixed-point sphere coordinates that the row-range
// collide's candidate filter reads: u = rint((p - lo) * scale) clamped to
// [0, 65535] per axis.  Clamping only shortens distances, so a sphere
// outside the frame is still never dropped from a contact it has.
struct QuantFrame {
  float lo[3] = {0.0f, 0.0f, 0.0f};
  float scale = 0.0f;
};

// The frame of the region integrate keeps spheres in: the box [-1, 1]^3
// (box walls), or [-1, 1]^2 x [-1, 10] with a distance-map boundary, one
// scale for all three axes.
inline QuantFrame quant_frame_for(bool distanceMap) {
  QuantFrame f;
  f.lo[0] = f.lo[1] = f.lo[2] = -1.0f;
  f.scale = 65535.0f / (distanceMap ? 11.0f : 2.0f);
  return f;
}

// The packed coordinates of one sphere: (x | y << 16, z).  A NaN
// coordinate maps to 0.
__device__ __forceinline__ uint2 quantize_pos(const QuantFrame& f, float x, float y, float z) {
  auto q = [&](float p, float lo) {
    return (uint32_t)fminf(fmaxf(rintf((p - lo) * f.scale), 0.0f), 65535.0f);
  };
  return make_uint2(q(x, f.lo[0]) | (q(y, f.lo[1]) << 16), q(z, f.lo[2]));
}

// Particle state gathered by the sort's last pass into sorted order: for
// the output at position g with key h and value v, sortedPos[g] = pos[v]
// and sortedVel[g] = (vel[v].xyz, bits of h); with idOfSlot (values are
// slots) also indexOut[g] = idOfSlot[v]; with sortedQ also
// sortedQ[g] = quantize_pos(frame, pos[v]).
struct SortGather {
  const float4* pos = nullptr;
  const float4* vel = nullptr;
  const uint32_t* idOfSlot = nullptr;
  float4* sortedPos = nullptr;
  float4* sortedVel = nullptr;
  uint32_t* indexOut = nullptr;
  uint2* sortedQ = nullptr;
  QuantFrame frame;
};

// Stable ascending LSD sort of (key, value) pairs on the low `bits` bits.
// valsIn == nullptr means the values are the identity 0..n-1.  keysIn may
// equal keysOut and valsIn may equal valsOut (in-place); intermediate passes
// use the scratch buffers.  All launches go to `stream`, nothing is
// allocated, nothing synchronises: the call is graph-capturable.
// histogramReady: the digit histogram of keysIn was already produced (by
// the kernel that wrote the keys), so the histogram launch is skipped.
// clearHash: when given, the first pass also resets to EMPTY every cell of
// cellStart/cellEnd that starts a run in the sorted array clearHash (the
// previous update's cells); requires a pass count >= 2 (minPasses).
hipError_t radix_sort_pairs(hipStream_t stream, SortScratch& s, uint32_t* devFlags,
                            const uint32_t* keysIn, const uint32_t* valsIn,
                            uint32_t* keysOut, uint32_t* valsOut, uint32_t n, int bits,
                            StageTimer* timer, bool histogramReady = false, int minPasses = 1,
                            const uint32_t* clearHash = nullptr, uint32_t* cellStart = nullptr,
                            uint32_t* cellEnd = nullptr, const SortGather* gather = nullptr);
// Sort of n <= RS_TILE pairs by counting, in two launches that also do the
// cell-table work of a system step: every key's rank in the stable order
// (one wave per key over all keys staged in LDS) into keysA/valsA of the
// scratch, with the cells of clearHash reset; then the sorted keys and
// values, the gathered state (gather) and cellStart/cellEnd of the sorted
// keys.  Bit-identical to radix_sort_pairs (with the same clearHash and
// gather) followed by launch_cell_bounds, for keys below numCells (the
// cells of a system, at most 8 * passes bits wide); inputs and outputs may
// alias.  The digit-0 histogram of keysIn is not read; H[0] is zeroed for
// the next sort.
hipError_t launch_sort_small(hipStream_t stream, SortScratch& s, const uint32_t* keysIn,
                             const uint32_t* valsIn, uint32_t* keysOut, uint32_t* valsOut,
                             uint32_t n, int passes, const uint32_t* clearHash,
                             uint32_t* cellStart, uint32_t* cellEnd, uint32_t numCells,
                             const SortGather& gather);
// The cell-table half of reorderDataAndFindCellStart: cellStart/cellEnd of
// the sorted hashes (the tables already hold EMPTY for every other cell).
hipError_t launch_cell_bounds(hipStream_t s, uint32_t* cellStart, uint32_t* cellEnd,
                              const uint32_t* hash, uint32_t n, uint32_t numCells);

// ---------------------------------------------------------------------------
// Per-stage timing with HIP events.
// ---------------------------------------------------------------------------
struct StageTimer {
  struct Pending { int stage; hipEvent_t begin, end; };
  struct Stage { std::string name; uint64_t launches = 0; double total_ms = 0.0; };
  bool enabled = false;
  std::vector<Stage> stages;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;

  int stage_id(const char* name);
  hipEvent_t get_event();
  // begin/end bracket one launch on `stream`; no-ops when disabled.
  void begin(hipStream_t stream, const char* name);
  void end(hipStream_t stream);
  hipError_t flush();  // synchronises pending events and accumulates
  void reset();
  void destroy();
 private:
  int open_stage = -1;
  hipEvent_t open_event = nullptr;
};

// ---------------------------------------------------------------------------
// DEM stage launchers (dem_kernels.hip).  Parameters are passed by value as
// kernel arguments (no constant-memory upload per step).
// ---------------------------------------------------------------------------
// Fills count words of p, and of p2 when given, with value (one launch).
hipError_t launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t value, size_t count,
                           uint32_t* p2 = nullptr);
hipError_t launch_copy16(hipStream_t s, void* dst, const void* src, size_t bytes);
hipError_t launch_integrate(hipStream_t s, const fc_sim_params& P, float4* pos, float4* vel,
                            float dt, uint32_t n);
hipError_t launch_calc_hash(hipStream_t s, const fc_sim_params& P, uint32_t* hash,
                            uint32_t* index, const float4* pos, uint32_t n);
// integrate + calcHash fused, plus clearing of the cells listed in the
// previous step's sorted hashes (prevSortedHash may be null: no clearing),
// plus the radix sort's digit histogram of the new keys for `passes`
// passes (written into the sort scratch, as k_rs_histogram would).
hipError_t launch_integrate_hash(hipStream_t s, const fc_sim_params& P, const BoundaryArg& B,
                                 float4* pos, float4* vel, float dt, uint32_t* keysOut,
                                 const uint32_t* prevSortedHash, uint32_t* cellStart,
                                 uint32_t* cellEnd, uint32_t n, SortScratch& sort, int passes,
                                 uint32_t* devFlags);
struct RowHists;  // row_resort.hpp
hipError_t launch_integrate_distmap(hipStream_t s, const fc_sim_params& P, const DistMapView& m,
                                    float4* pos, float4* vel, float dt, uint32_t n);
DistMapView distmap_view(const fc_distmap* m);
// Per 8^3-cell block of m, its smallest and largest vertex distance (for
// distmap_far), in a new device array, or *out = nullptr when m's lattice
// coordinates are not the uniform bmin + i cellSize.
fc_status distmap_block_bounds(const fc_distmap* m, hipStream_t s, float2** out, int* bx,
                               int* by);
// index: the sorted values, particle ids; or, with idOfSlot, the slots of
// oldPos/oldVel holding each particle (whose ids idOfSlot gives, written to
// indexOut).
hipError_t launch_reorder(hipStream_t s, uint32_t* cellStart, uint32_t* cellEnd, float4* sortedPos,
                          float4* sortedVel, const uint32_t* hash, const uint32_t* index,
                          const float4* oldPos, const float4* oldVel, uint32_t n,
                          uint32_t numCells, const uint32_t* idOfSlot = nullptr,
                          uint32_t* indexOut = nullptr);
// poly: radii from pos.w (bounded by P.particleRadius, else DEVERR_RADIUS).
// posOut (optional): also scatter the sorted positions to id order.
// forceOut (optional): evalForces -- no cursor sphere, and the 27 cells'
// force stored as float3 at the original index.
// Gathered state (the row-range walk): sortedVel.w holds the sorted hash
// and sortedQ (kSortedPosPad entries of padding) the fixed-point
// coordinates in `frame`: a system's arrays written by its sort's
// SortGather, or the ones launch_rows_state builds for the reference-style
// collide.
struct GatheredState {
  const uint2* sortedQ = nullptr;
  QuantFrame frame;
};
hipError_t launch_collide(hipStream_t s, const fc_sim_params& P, bool poly, uint32_t* devFlags,
                          float4* newVel, const float4* sortedPos, const float4* sortedVel,
                          const uint32_t* index, const uint32_t* cellStart,
                          const uint32_t* cellEnd, uint32_t n, float4* posOut = nullptr,
                          const GatheredState* gathered = nullptr, float* forceOut = nullptr);

// For the reference-style fc_collide / fc_eval_forces: builds the state
// the row-range collide reads (velOut = sortedVel with the bits of each
// position's cell hash in w, qOut the 16-bit coordinates in frame, n +
// kSortedPosPad entries) from the reference's sorted arrays.
hipError_t launch_rows_state(hipStream_t s, const fc_sim_params& P, const QuantFrame& frame,
                             const float4* sortedPos, const float4* sortedVel, uint32_t n,
                             float4* velOut, uint2* qOut);
// Whether the reference-style collide takes the row-range kernel through
// launch_rows_state (large systems; small ones take the row-split kernel
// on the reference's arrays as they are).
bool collide_rows_state_pays(const fc_sim_params& P, uint32_t n);

float contact_d2_threshold(float R);
// Exact reciprocals of the cell sizes when all three are powers of two (the
// multiply then equals the division), else zeros (cell_quotients divides).
v3f inv_cell(const fc_sim_params& P);
// Bound on the squared fixed-point distance (units of 1/scale) of every pair
// closer than R, for the row-range collide's candidate filter.
uint32_t quant_threshold(float R, float scale);
// collide + the next update's integrate/hash/histogram in one launch (see
// dem_kernels.hip); usable when collide_can_fuse(P, n).
bool collide_can_fuse(const fc_sim_params& P, uint32_t n);
// Whether a run's fused collide stores the next key at the sorted slot
// (then the run's later sorts are row resorts).
bool collide_slot_keys(const fc_sim_params& P, uint32_t n);
hipError_t launch_collide_fused(hipStream_t s, const fc_sim_params& P, const BoundaryArg& B,
                                float4* pos, float4* vel,
                                const float4* sortedPos, const float4* sortedVel,
                                const uint32_t* index, const uint32_t* cellStart,
                                const uint32_t* cellEnd, uint32_t n, float dt, uint32_t* keysOut,
                                SortScratch& sort, int passes, uint32_t* devFlags,
                                uint32_t* slotOf, uint32_t* idOfSlot,
                                const GatheredState* gathered, bool slotKeys = false,
                                const RowHists* rows = nullptr);

fc_status validate_params(const fc_sim_params& P);
// Initial configurations of reset() into host float4 arrays (fc_system.hip).
void generate_initial_state(float* pos, float* vel, uint32_t n, fc_particle_config config,
                            uint64_t seed, float r, float jitterFrac, float velAmp);
// Grows the context's sort scratch to hold n keys (synchronises first).
fc_status ensure_sort_capacity(fc_context* ctx, size_t n);
// Synchronises the context stream and converts the device error word.
fc_status check_device_flags(fc_context* ctx);
fc_sim_params default_params(uint32_t numParticles, fc_uint3 gridSize);
// Resolves, once per device, the device address of the no-allocation sort's
// scratch array (sort_static.hip); called with `device` current.
fc_status init_noalloc_scratch(int device);

}  // namespace fc

// ---------------------------------------------------------------------------
// Opaque handles.
// ---------------------------------------------------------------------------
struct fc_context {
  int device = 0;
  hipStream_t stream = nullptr;
  fc_sim_params params{};
  bool paramsSet = false;
  fc::SortScratch sort;
  uint32_t* devFlags = nullptr;   // device error word
  fc::StageTimer timer;
  double* statsScratch = nullptr; // FP64 reduction partials + histogram bins
  // State built for the row-range walk by the reference-style collide
  // (launch_rows_state), grown to the largest system it has seen.
  float4* rowsVel = nullptr;
  uint2* rowsQ = nullptr;
  uint32_t rowsCapacity = 0;
};

namespace fc {
// Tables of the row resort of a multi-step run (row_resort.hip): one
// device block holding, per grid row (the gridSize.x cells sharing y, z),
// the start of the row in the current and in the next sorted order, the
// next keys' counts, and the bucket list of the spheres that change row.
struct RowTables {
  uint32_t rows = 0, shift = 0, groups = 0, capacity = 0;
  uint32_t* block = nullptr;
  uint32_t* count = nullptr;      // [rows] spheres per row in the current order
  uint32_t* startNext = nullptr;  // [rows + 1] row starts of the next order
  uint32_t* chgStart = nullptr;   // [rows x copies + 1] bucket starts of the row changers
  // [rows][2] one entry per row that holds spheres in the next order, in row
  // order: {row, next start, next count, current start}, {current count,
  // changers' first bucket start, changers moving in, 0}
  uint4* rowList = nullptr;
  uint64_t* groupWords = nullptr; // [groups][3] per 256-row group: spheres, changers, rows (epoch-tagged)
  uint32_t* listCount = nullptr;  // [1] entries of rowList
  uint32_t* epoch = nullptr;      // [1] steps resorted (tags the group words)
  uint32_t* zeroed = nullptr;     // the counts below, zero between steps
  size_t zeroedWords = 0;
  uint32_t* rowHist = nullptr;    // [rows x copies] next keys per row
  uint32_t* chgHist = nullptr;    // [rows x copies] of those, the spheres whose row changes
  uint32_t* chgCursor = nullptr;  // [rows x copies]
  // [capacity] the row changers, by new row: {current slot, next key, id, 0}
  uint4* chgList = nullptr;
};

// Where the row sort writes the sorted order: hash and id per position,
// the state gathered from the sorted slots, and the cell tables.
struct RowOutputs {
  uint32_t* hash = nullptr;
  uint32_t* index = nullptr;
  const float4* pos = nullptr;
  const float4* vel = nullptr;
  float4* sortedPos = nullptr;
  float4* sortedVel = nullptr;
  uint2* sortedQ = nullptr;
  uint32_t* cellStart = nullptr;
  uint32_t* cellEnd = nullptr;
  QuantFrame frame;
};

constexpr uint64_t kMaxRows = 1ull << 22;
bool row_resort_grid_ok(const fc_sim_params& P);
RowHists row_hists(const RowTables& T);
fc_status ensure_row_tables(RowTables& T, const fc_sim_params& P, uint32_t n);
void free_row_tables(RowTables& T);
hipError_t launch_row_clear(hipStream_t s, const RowTables& T);
hipError_t launch_row_scan(hipStream_t s, const RowTables& T, uint32_t* devFlags);
// Row starts of an order sorted by another path (the radix sort of a run's
// first step), read from its sorted hashes.
hipError_t launch_row_starts(hipStream_t s, const RowTables& T, const uint32_t* sortedHash,
                             uint32_t n);
hipError_t launch_row_changers(hipStream_t s, const RowTables& T, const uint32_t* keys,
                               const uint32_t* idOfSlot, const uint32_t* hash,
                               uint32_t* cellStart, uint32_t* cellEnd, uint32_t n,
                               uint32_t numCells, uint32_t* devFlags);
hipError_t launch_row_sort(hipStream_t s, const RowTables& T, const uint32_t* keys,
                           const uint32_t* idOfSlot, const RowOutputs& O, uint32_t* devFlags);
}  // namespace fc

struct fc_system {
  fc_context* ctx = nullptr;
  uint32_t n = 0;
  fc_sim_params params{};
  float4* pos = nullptr;
  float4* vel = nullptr;
  float4* sortedPos = nullptr;
  float4* sortedVel = nullptr;
  uint2* sortedQ = nullptr;       // 16-bit coordinates of sortedPos (quant_frame())
  uint32_t* hash = nullptr;       // sorted hashes of the last update
  uint32_t* index = nullptr;      // sorted original indices of the last update
  uint32_t* keys = nullptr;       // unsorted hashes written by integrate+hash
  uint32_t* slotOf = nullptr;     // multi-step runs: slot of each particle id (sort values)
  uint32_t* idOfSlot = nullptr;   // multi-step runs: particle id held by each slot
  uint32_t* cellStart = nullptr;
  uint32_t* cellEnd = nullptr;
  // True when the cell tables hold exactly the cells listed in `hash`, so the
  // next update may clear only those instead of the whole table.
  bool tablesTrackHash = false;
  // Captured steady-state steps, graph-launched by fc_system_run: [0] one
  // step, [1] kGraphSteps steps back to back (fewer launches per step).
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t graphExec[2] = {nullptr, nullptr};
  float graphDt = 0.0f;
  fc_sim_params graphParams{};
  // Boundary of the integrate stage.
  fc_boundary boundary = FC_BOUNDARY_BOX;
  const fc_distmap* boundaryMap = nullptr;
  // The map's block bounds (taken when it is set; maps do not change).
  float2* boundaryBounds = nullptr;
  int boundaryBx = 0, boundaryBy = 0;
  // Radii from pos.w (set by fc_system_set_polydisperse).
  int polydisperse = 0;
  // Row resort tables of multi-step runs (allocated by the first run that
  // uses them).
  fc::RowTables rowTables;
  fc::BoundaryArg boundary_arg() const;
  fc::QuantFrame quant_frame() const;
};
