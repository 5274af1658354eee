// exa_kernels.h -- device views and kernel launch wrappers shared by the
// sampler, renderer and C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <atomic>
#include "exa_device.h"
#include "exa_host.h"

namespace exa {

// Workgroups of `kernel` (at `block` threads) resident on the whole current
// device: occupancy per CU times the CU count, computed once per device and
// kept in `cache` (one slot per device ordinal; atomics, so concurrent first
// calls from several host threads at worst compute it twice).
constexpr int kMaxCachedDevices = 64;
template <typename K>
int residentWorkgroups(K kernel, int block, std::atomic<int> (&cache)[kMaxCachedDevices]) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = -1;
  if (dev >= 0 && dev < kMaxCachedDevices) {
    const int v = cache[dev].load(std::memory_order_relaxed);
    if (v > 0) return v;
  }
  int perCU = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kernel, block, 0) != hipSuccess || perCU < 1) perCU = 1;
  if (dev < 0 || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    cus = 256;
  const int v = perCU * cus;
  if (dev >= 0 && dev < kMaxCachedDevices) cache[dev].store(v, std::memory_order_relaxed);
  return v;
}

// Point-location data on the device (the ExaStitchSampler::LP of
// sampler/ExaStitchSampler.h:42-49 plus the geometry the OptiX programs
// read through their SBT records, programs/ExaStitchSampler.cu:25-35).
struct SamplerView {
  const Gridlet *gridlets;
  const float *gridletScalars;
  const float *gridletMaxOp;  // nullptr: no culling
  const BvhNode *gNodes;
  const int32_t *gRefs;
  int32_t numGridlets;
  const float4 *vertices;
  const int32_t *indices;
  const float *elemMaxOp;  // nullptr: no culling
  const BvhNode *eNodes;
  const int32_t *eRefs;
  int32_t numElements;
  // direct tables (LocateTables)
  const uint32_t *gStart;
  const int32_t *gList;
  int32_t numLevels;
  int32_t levelDims[LocateTables::kMaxLevels][3];
  int32_t levelOffset[LocateTables::kMaxLevels];
  const int32_t *gTop;    // nullptr: owner tables not built
  const int32_t *gOwner;
  const uint32_t *eStart;
  const uint3 *eCand;    // per list entry: id | vertex count << 28, quantised box lo, hi
  const float4 *eVerts;  // 8 per element
  float binLo[3];
  float binInv;
  int32_t binDims[3];
  float bandMargin;  // LocateTables::bandMargin / fastMargin
  float fastMargin;
};

enum Accel { kAccelTables = 0, kAccelBvh = 1 };

// GridTraversable (Grid.cuh:21-24) plus the majorant buffer.
struct GridView {
  int32_t dims[3];
  float bounds[6];
  const float *majorants;
};

constexpr int kLocateBlock = 256;
constexpr int kStackDepth = kBvhMaxDepth;

// Per-launch counters of the instrumented render/locate variants; the
// algorithmic-byte roofline is computed from these.
enum Counter {
  kCntGridletHit = 0,  // full point location ending in a gridlet
  kCntTetHit,
  kCntPyrHit,
  kCntWedgeHit,
  kCntHexHit,
  kCntMiss,
  kCntCachedHit,  // testSample hit on the remembered gridlet
  kCntDdaCells,
  kCntPixels,
  kCntTfFetch,
  kCntNeural,  // element-stage lookups answered by the neural field
  kNumCounters
};

// ---- sampler kernels (sampler.hip) ----------------------------------------
void launchGridletValueRanges(const Gridlet *g, const float *scalars, int64_t n, float2 *ranges,
                              hipStream_t s);
void launchGridBuild(const SamplerView &sv, const int32_t dims[3], const float bounds[6],
                     float2 *ranges, uint2 *scratch, hipStream_t s);
void launchMapMaxOpacity(const float2 *ranges, int64_t n, const float4 *cmap, int numColors,
                         float x0, float x1, float *out, hipStream_t s);
void launchElementMaxOpacity(const float4 *vertices, const int32_t *indices, int64_t n,
                             const float4 *cmap, int numColors, float x0, float x1, float *out,
                             hipStream_t s);
// queue: n + 1 uint32 of device scratch for the two-stage scheme, or nullptr
// for one pass per query
void launchLocate(const SamplerView &sv, int accel, const float *pos, int64_t n, Sample *out,
                  unsigned long long *counters, hipStream_t s, uint32_t *queue = nullptr);
void launchTestSample(const SamplerView &sv, const float *pos, const int32_t *primIDs, int64_t n,
                      Sample *out, hipStream_t s);

// ---- renderer (render.hip) ------------------------------------------------
constexpr int kLightsMax = 4;  // common.h:27 LIGHTS_MAX
struct LightParams {
  float pos[3];
  float intensity;
  int32_t on;
  int32_t type;  // 0 point, 1 directional
};

// LaunchParams (LaunchParams.h:56-119) restricted to the ExaStitch path.
struct RenderParams {
  int32_t W, H, spp, accumID;
  int32_t integrator;  // 0 path tracer, 1 direct lighting, 2 ray marcher
  int32_t shadeMode;   // 0 default, 1 gridlet checkerboard, 2 teaser
  int32_t maxBounces;
  int32_t debugCrosshair;
  float dt;                // ray-marcher step (voxel units)
  int32_t heatMapEnabled;  // per-pixel clock cycles shown over the image
  float heatMapScale;
  float camOrg[3], camDir00[3], camDu[3], camDv[3];
  float worldBounds[6];
  float voxelXfm[12];
  float voxelXfmInv[12];
  float lightXfm[12];
  float tfDomain[2];
  float opacityScale;
  int32_t tfN;
  const float4 *tf;
  int32_t clipEnabled;
  float clipN[3];
  float clipD;
  int32_t subImageActive;
  float subImage[4];  // lower.xy, upper.xy
  int32_t subImageSelecting;
  float subImageSel[4];
  LightParams lights[kLightsMax];
  float4 *accum;
  uint32_t *fb;
};

void launchRender(const RenderParams &p, const SamplerView &sv, int accel, const GridView &gv,
                  unsigned long long *counters, hipStream_t s);
// One request of a batched launch: its own camera, image, spp, integrator,
// transfer function (and the majorants and culling that follow from it) and
// accumulation / frame buffers; the model and point-location tables are
// shared by all requests.
struct BatchRequest {
  RenderParams P;
  const float *majorants;
  const float *gridletMaxOp;
  const float *elemMaxOp;
  int32_t numLights;
  int32_t tilesX;     // 8x8 pixel tiles per row
  uint32_t workBegin;  // first work id (64 per tile); ascending over requests
  uint32_t pad;
};
constexpr int kMaxBatchRequests = 1024;

// scheduling knobs of the persistent renderer (results do not depend on them)
struct PersistTuning {
  int32_t ctrlSteps = 4;   // tracking iterations per lane and round, >= 1
  int32_t elemBatch = 16;  // element pass when waiting lanes >= elemBatch/64 of live lanes
  int32_t rareBatch = 24;  // per-segment pass when waiting lanes >= rareBatch/64 of live lanes
  // where the element stage runs: kElemInWave (the waiting lanes' own wave,
  // gated by elemBatch), kElemWorkgroupQueue (any wave of the workgroup
  // answers up to 64 queries posted in LDS, once elemBatch are posted),
  // kElemQueueRounds (queue between launches: queueRounds rounds of
  // queueDepth entries, 0 = one per pixel)
  int32_t elementStage = 0;
  int32_t queueRounds = 1;
  int32_t queueDepth = 0;
  int32_t blockSize = 256; // threads per workgroup: 64, 128 or 256
};
enum ElementStage { kElemInWave = 0, kElemWorkgroupQueue = 1, kElemQueueRounds = 2 };
constexpr int kMaxQueueRounds = 32;

// Device buffers of the element-stage queue (two queues used in turn):
// entries of kElementQueueVec float4 per path, a sample per entry, and the
// counters (two queue counts, one hand-out counter).
constexpr int kElementQueueVec = 16;
struct ElementQueueBuffers {
  float4 *rec[2];
  Sample *samples[2];
  uint32_t *keys[2];   // work id of each entry's path
  uint32_t *order;     // launch order of a queue's entries (image order)
  uint32_t *hist;      // maxBuckets counters of the ordering
  uint32_t maxBuckets;
  uint32_t *count;  // 3 uint32
  uint32_t cap;     // entries per queue
};
// work: one device uint32 (reset by the launch itself)
struct NeuralView;
// nv / netFlags (optional, tables accelerator only): macrocells flagged 1
// answer element-stage lookups with the neural field
void launchRenderPersistent(const RenderParams &p, const SamplerView &sv, int accel, const GridView &gv,
                            uint32_t *work, const PersistTuning &t, unsigned long long *counters, hipStream_t s,
                            const NeuralView *nv = nullptr, const uint8_t *netFlags = nullptr,
                            const ElementQueueBuffers *qb = nullptr);
// Batched launch over reqs[0..numReqs) (device array); numWork = total work
// ids.  Afterwards every request's accumID is advanced by one on the device.
// With nv, the network answers the element stage of the macrocells flagged in
// netFlags (tables accelerator).
void launchRenderBatch(const SamplerView &sv, int accel, const GridView &gv, BatchRequest *reqs, int numReqs,
                       uint32_t numWork, uint32_t *work, const PersistTuning &t, unsigned long long *counters,
                       hipStream_t s, const NeuralView *nv = nullptr, const uint8_t *netFlags = nullptr,
                       const ElementQueueBuffers *qb = nullptr);

}  // namespace exa
