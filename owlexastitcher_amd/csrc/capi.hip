// capi.hip -- extern "C" boundary of libexastitch.so (declared in
// include/exa/exa.h) and the host objects behind it: ExaStitchModel,
// ExaStitchSampler (+ its majorant Grid) and OWLRenderer.
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <exception>
#include <fstream>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>
#include "../../include/exa/exa.h"
#include "exa_kernels.h"
#include "exa_io.h"
#include "neural_kernels.h"
#include "bricks.h"
#include "bricks_kernels.h"

namespace exa {

static thread_local std::string g_lastError;

struct Status : std::runtime_error {
  int code;
  Status(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(x)                                                                                     \
  do {                                                                                                   \
    hipError_t e_ = (x);                                                                                 \
    if (e_ != hipSuccess)                                                                                \
      throw Status(EXA_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));                         \
  } while (0)

template <typename F>
static int guarded(F &&f) {
  try {
    f();
    return EXA_OK;
  } catch (const Status &s) {
    g_lastError = s.what();
    return s.code;
  } catch (const std::bad_alloc &) {
    g_lastError = "host allocation failed";
    return EXA_ERR_ALLOC;
  } catch (const std::exception &e) {
    g_lastError = e.what();
    return EXA_ERR_INVALID;
  } catch (...) {
    g_lastError = "unknown error";
    return EXA_ERR_INVALID;
  }
}

static void require(bool c, const char *msg) {
  if (!c) throw Status(EXA_ERR_INVALID, msg);
}

template <typename T>
struct DeviceBuffer {
  T *ptr = nullptr;
  size_t count = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  ~DeviceBuffer() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    count = 0;
  }
  void alloc(size_t n) {
    if (n == count && ptr) return;
    release();
    if (n == 0) return;
    hipError_t e = hipMalloc((void **)&ptr, n * sizeof(T));
    if (e != hipSuccess) {
      ptr = nullptr;
      throw Status(EXA_ERR_ALLOC, std::string("hipMalloc of ") + std::to_string(n * sizeof(T)) +
                                      " bytes failed: " + hipGetErrorString(e));
    }
    count = n;
  }
  void upload(const T *h, size_t n, hipStream_t s) {
    alloc(n);
    if (n) HIP_CHECK(hipMemcpyAsync(ptr, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  size_t bytes() const { return count * sizeof(T); }
};

// ---------------------------------------------------------------------------
// affine helpers (owl affine3f: columns vx, vy, vz and translation p)
// ---------------------------------------------------------------------------
static void affineIdentity(float a[12]) {
  static const float I[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
  memcpy(a, I, sizeof(I));
}
// general inverse, evaluated in double and rounded
static void affineInverse(const float a[12], float out[12]) {
  const double m[3][3] = {{a[0], a[3], a[6]}, {a[1], a[4], a[7]}, {a[2], a[5], a[8]}};  // row-major
  const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                     m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                     m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  if (det == 0.0) throw Status(EXA_ERR_INVALID, "singular voxel-space transform");
  double inv[3][3];
  inv[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det;
  inv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
  inv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
  inv[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det;
  inv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
  inv[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
  inv[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det;
  inv[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
  inv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
  double t[3];
  for (int r = 0; r < 3; ++r) t[r] = -(inv[r][0] * a[9] + inv[r][1] * a[10] + inv[r][2] * a[11]);
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) out[3 * c + r] = (float)inv[r][c];
  for (int r = 0; r < 3; ++r) out[9 + r] = (float)t[r];
}
static void xfmPointHost(const float a[12], const float p[3], float out[3]) {
  for (int k = 0; k < 3; ++k) out[k] = a[k] * p[0] + (a[3 + k] * p[1] + (a[6 + k] * p[2] + a[9 + k]));
}

template <typename T>
static void download(const DeviceBuffer<T> &b, void *out) {
  if (!out || !b.count) return;
  HIP_CHECK(hipDeviceSynchronize());
  HIP_CHECK(hipMemcpy(out, b.ptr, b.bytes(), hipMemcpyDeviceToHost));
}

}  // namespace exa

using namespace exa;

// ---------------------------------------------------------------------------
// opaque objects
// ---------------------------------------------------------------------------
struct exa_amr {
  AmrScene scene;
};
struct exa_dual_mesh {
  DualMesh mesh;
};
struct exa_model {
  HostModel m;
};

struct exa_sampler {
  // model copy on the device
  DeviceBuffer<float4> vertices;
  DeviceBuffer<int32_t> indices;
  DeviceBuffer<Gridlet> gridlets;
  DeviceBuffer<float> gridletScalars;
  int64_t numElements = 0, numGridlets = 0;
  // point-location accelerators: direct tables (default) and BVHs
  int accel = kAccelTables;
  bool wantBvh = false, haveBvh = false, haveTables = false;
  DeviceBuffer<BvhNode> gNodes, eNodes;
  DeviceBuffer<int32_t> gRefs, eRefs;
  int gDepth = 0, eDepth = 0;
  LocateTables tabMeta;  // sizes/offsets only (vectors are released after upload)
  float bandMargin = 0.00390625f, fastMargin = 0.0078125f;  // gridletMargins of the model
  DeviceBuffer<uint32_t> tGStart, tEStart;
  DeviceBuffer<int32_t> tGList, tGTop, tGOwner;
  DeviceBuffer<uint3> tECand;
  DeviceBuffer<float4> tEVerts;
  // gridlet value ranges + per-primitive max opacities (culling)
  DeviceBuffer<float2> gridletRanges;
  DeviceBuffer<float> gridletMaxOp, elemMaxOp;
  bool haveMaxOp = false;
  // majorant grid (Grid.h:33-86)
  int32_t dims[3] = {0, 0, 0};
  float bounds[6] = {0, 0, 0, 0, 0, 0};
  DeviceBuffer<float2> valueRanges;
  DeviceBuffer<float> maxOpacities;
  DeviceBuffer<float4> colorMap;
  DeviceBuffer<unsigned long long> counters;
  bool built = false;

  SamplerView view() const {
    SamplerView v;
    v.gridlets = gridlets.ptr;
    v.gridletScalars = gridletScalars.ptr;
    v.gridletMaxOp = haveMaxOp ? gridletMaxOp.ptr : nullptr;
    v.gNodes = gNodes.ptr;
    v.gRefs = gRefs.ptr;
    v.numGridlets = (int32_t)numGridlets;
    v.vertices = vertices.ptr;
    v.indices = indices.ptr;
    v.elemMaxOp = haveMaxOp ? elemMaxOp.ptr : nullptr;
    v.eNodes = eNodes.ptr;
    v.eRefs = eRefs.ptr;
    v.numElements = (int32_t)numElements;
    v.gStart = tGStart.ptr;
    v.gList = tGList.ptr;
    v.numLevels = tabMeta.numLevels;
    for (int L = 0; L < LocateTables::kMaxLevels; ++L) {
      for (int k = 0; k < 3; ++k) v.levelDims[L][k] = tabMeta.levelDims[L][k];
      v.levelOffset[L] = (int32_t)tabMeta.levelOffset[L];
    }
    v.gTop = tGTop.count ? tGTop.ptr : nullptr;
    v.gOwner = tGOwner.ptr;
    v.eStart = tEStart.ptr;
    v.eCand = tECand.ptr;
    v.eVerts = tEVerts.ptr;
    for (int k = 0; k < 3; ++k) {
      v.binLo[k] = tabMeta.binLo[k];
      v.binDims[k] = tabMeta.binDims[k];
    }
    v.binInv = tabMeta.binInv;
    v.bandMargin = bandMargin;
    v.fastMargin = fastMargin;
    return v;
  }

  void buildBvhs(const HostModel &m, hipStream_t s) {
    std::vector<float> boxes;
    Bvh bvh;
    gridletBoxes(m.gridlets, boxes);
    buildBvh(boxes.data(), (int64_t)m.gridlets.size(), 4, bvh);
    gNodes.upload(bvh.nodes.data(), bvh.nodes.size(), s);
    gRefs.upload(bvh.primRefs.data(), bvh.primRefs.size(), s);
    gDepth = bvh.depth;
    HIP_CHECK(hipStreamSynchronize(s));
    elementBoxes(m.indices, m.vertices, boxes);
    buildBvh(boxes.data(), m.numElements(), 4, bvh);
    eNodes.upload(bvh.nodes.data(), bvh.nodes.size(), s);
    eRefs.upload(bvh.primRefs.data(), bvh.primRefs.size(), s);
    eDepth = bvh.depth;
    HIP_CHECK(hipStreamSynchronize(s));
    haveBvh = true;
  }

  void buildTables(const HostModel &m, hipStream_t s) {
    LocateTables t;
    buildLocateTables(m, t);
    tGStart.upload(t.gStart.data(), t.gStart.size(), s);
    tGList.upload(t.gList.data(), t.gList.size(), s);
    tGTop.upload(t.gTop.data(), t.gTop.size(), s);
    tGOwner.upload(t.gOwner.data(), t.gOwner.size(), s);
    tEStart.upload(t.eStart.data(), t.eStart.size(), s);
    tECand.upload((const uint3 *)t.eCand.data(), t.eCand.size() / 3, s);
    tEVerts.upload((const float4 *)t.eVerts.data(), t.eVerts.size() / 4, s);
    HIP_CHECK(hipStreamSynchronize(s));
    tabMeta = LocateTables();
    tabMeta.numLevels = t.numLevels;
    for (int L = 0; L < LocateTables::kMaxLevels; ++L) {
      for (int k = 0; k < 3; ++k) tabMeta.levelDims[L][k] = t.levelDims[L][k];
      tabMeta.levelOffset[L] = t.levelOffset[L];
    }
    for (int k = 0; k < 3; ++k) {
      tabMeta.binLo[k] = t.binLo[k];
      tabMeta.binDims[k] = t.binDims[k];
    }
    tabMeta.binInv = t.binInv;
    haveTables = true;
  }

  // selects the accelerator; builds it from m when missing
  void setAccel(int a, const HostModel *m, hipStream_t s) {
    if (a != kAccelTables && a != kAccelBvh) throw Status(EXA_ERR_INVALID, "unknown accelerator");
    if (built) {
      if (a == kAccelBvh && !haveBvh) {
        if (!m) throw Status(EXA_ERR_STATE, "BVH not built (select it before build)");
        buildBvhs(*m, s);
      }
      if (a == kAccelTables && !haveTables) {
        if (!m) throw Status(EXA_ERR_STATE, "tables not built (select them before build)");
        buildTables(*m, s);
      }
    }
    accel = a;
  }
  GridView gridView() const {
    GridView g;
    for (int k = 0; k < 3; ++k) g.dims[k] = dims[k];
    for (int k = 0; k < 6; ++k) g.bounds[k] = bounds[k];
    g.majorants = maxOpacities.ptr;
    return g;
  }

  // ExaStitchSampler::build (ExaStitchSampler.cpp:23-340) + Grid::build
  // (Grid.cu:457-594)
  void build(const HostModel &m, hipStream_t s) {
    built = false;
    haveMaxOp = false;
    if (m.gridlets.empty() && m.indices.empty())
      throw Status(EXA_ERR_INVALID, "model has neither gridlets nor elements");
    if (m.numMCs[0] <= 0 || m.numMCs[1] <= 0 || m.numMCs[2] <= 0)
      throw Status(EXA_ERR_STATE, "number of macrocells not set (Model::setNumGridCells)");
    if (m.numElements() >= (1ll << 31) || (int64_t)m.gridlets.size() >= (1ll << 31))
      throw Status(EXA_ERR_INVALID, "more than 2^31 primitives");
    numElements = m.numElements();
    numGridlets = (int64_t)m.gridlets.size();
    gridletMargins(m, bandMargin, fastMargin);
    vertices.upload((const float4 *)m.vertices.data(), m.vertices.size() / 4, s);
    indices.upload(m.indices.data(), m.indices.size(), s);
    gridlets.upload(m.gridlets.data(), m.gridlets.size(), s);
    gridletScalars.upload(m.gridletScalars.data(), m.gridletScalars.size(), s);
    // point-location accelerators (replace the OptiX BLASes over
    // GridletGeomBounds / StitchGeomBounds and the TLAS)
    haveBvh = haveTables = false;
    bool tablesOk = true;
    for (const Gridlet &g : m.gridlets)
      if (g.lower[0] < 0 || g.lower[1] < 0 || g.lower[2] < 0) tablesOk = false;
    if (accel == kAccelTables && !tablesOk) accel = kAccelBvh;
    if (accel == kAccelTables) buildTables(m, s);
    if (accel == kAccelBvh || wantBvh) buildBvhs(m, s);
    // computeGridletValueRanges (ExaStitchSampler.cu:229-277)
    gridletRanges.alloc(numGridlets);
    gridletMaxOp.alloc(numGridlets);
    elemMaxOp.alloc(numElements);
    launchGridletValueRanges(gridlets.ptr, gridletScalars.ptr, numGridlets, gridletRanges.ptr, s);
    // majorant grid over the model's cellBounds (ExaStitchSampler.cpp:329)
    for (int k = 0; k < 3; ++k) dims[k] = m.numMCs[k];
    for (int k = 0; k < 6; ++k) bounds[k] = m.cellBounds[k];
    const int64_t nmc = (int64_t)dims[0] * dims[1] * dims[2];
    valueRanges.alloc(nmc);
    maxOpacities.alloc(nmc);
    HIP_CHECK(hipMemsetAsync(maxOpacities.ptr, 0, nmc * sizeof(float), s));
    DeviceBuffer<uint2> scratch;
    scratch.alloc(nmc);
    launchGridBuild(view(), dims, bounds, valueRanges.ptr, scratch.ptr, s);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
    counters.alloc(kNumCounters);
    built = true;
  }

  // ExaStitchSampler::computeMaxOpacities (ExaStitchSampler.cu:279-376)
  void computeMaxOpacities(const float *rgba, int n, float lo, float hi, hipStream_t s) {
    if (!built) throw Status(EXA_ERR_STATE, "sampler not built");
    require(rgba && n > 0, "empty color map");
    colorMap.upload((const float4 *)rgba, n, s);
    const int64_t nmc = (int64_t)dims[0] * dims[1] * dims[2];
    launchMapMaxOpacity(valueRanges.ptr, nmc, colorMap.ptr, n, lo, hi, maxOpacities.ptr, s);
    launchMapMaxOpacity(gridletRanges.ptr, numGridlets, colorMap.ptr, n, lo, hi, gridletMaxOp.ptr, s);
    launchElementMaxOpacity(vertices.ptr, indices.ptr, numElements, colorMap.ptr, n, lo, hi, elemMaxOp.ptr, s);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
    haveMaxOp = true;
  }
};

// ExaBrickModel (model/ExaBrickModel.h:27-56) and ExaBrickSampler in its
// ABR mode (sampler/ExaBrickSampler.cpp:26-247): bricks, scalars, regions
// and their kd-tree on the device; per-region value ranges and max
// opacities (the majorants, and culling where 0).
struct exa_brick_model {
  BrickModel m;
};

struct exa_brick_sampler {
  DeviceBuffer<Brick> bricks;
  DeviceBuffer<float> scalars;
  DeviceBuffer<float4> abrLo, abrHi;
  DeviceBuffer<float> abrCellWidth;
  DeviceBuffer<int32_t> leafList;
  DeviceBuffer<KdNode> nodes;
  DeviceBuffer<float2> abrRanges;
  DeviceBuffer<float> maxOp;
  DeviceBuffer<float4> colorMap;
  DeviceBuffer<unsigned long long> counters;
  int32_t numAbrs = 0;
  bool haveMaxOp = false;

  void build(const BrickModel &m, hipStream_t s) {
    if (m.bricks.empty() || m.abrs.empty()) throw Status(EXA_ERR_INVALID, "brick model without bricks");
    if (m.abrs.size() >= (1u << 29) || m.nodes.size() >= (1u << 29))
      throw Status(EXA_ERR_INVALID, "too many regions for the kd-tree encoding");
    numAbrs = (int32_t)m.abrs.size();
    std::vector<float4> lo(numAbrs), hi(numAbrs);
    std::vector<float> cw(numAbrs);
    std::vector<float2> ranges(numAbrs);
    for (int32_t i = 0; i < numAbrs; ++i) {
      const Abr &a = m.abrs[i];
      int32_t b = a.leafListBegin, c = a.leafListSize;
      float fb, fc;
      memcpy(&fb, &b, 4);
      memcpy(&fc, &c, 4);
      lo[i] = make_float4(a.domain[0], a.domain[1], a.domain[2], fb);
      hi[i] = make_float4(a.domain[3], a.domain[4], a.domain[5], fc);
      cw[i] = a.finestLevelCellWidth;
      ranges[i] = make_float2(a.valueRange[0], a.valueRange[1]);
    }
    // every walk on the device ends: children come after their parent
    const int64_t numNodes = (int64_t)m.nodes.size();
    for (int64_t i = 0; i < numNodes; ++i) {
      const int32_t info = m.nodes[i].info, ref = info >> 2;
      const bool ok = (info & 3) == 3 ? (ref >= 0 && ref <= numAbrs) : (i + 1 < numNodes && ref > i + 1 && ref < numNodes);
      if (!ok) throw Status(EXA_ERR_INVALID, "invalid kd-tree node");
    }
    bricks.upload(m.bricks.data(), m.bricks.size(), s);
    scalars.upload(m.scalars.data(), m.scalars.size(), s);
    abrLo.upload(lo.data(), lo.size(), s);
    abrHi.upload(hi.data(), hi.size(), s);
    abrCellWidth.upload(cw.data(), cw.size(), s);
    leafList.upload(m.leafList.data(), m.leafList.size(), s);
    nodes.upload(m.nodes.data(), m.nodes.size(), s);
    abrRanges.upload(ranges.data(), ranges.size(), s);
    maxOp.alloc(numAbrs);
    counters.alloc(kNumCounters);
    HIP_CHECK(hipStreamSynchronize(s));
    haveMaxOp = false;
  }

  BrickView view() const {
    BrickView v;
    v.bricks = (const int4 *)bricks.ptr;
    v.scalars = scalars.ptr;
    v.abrLo = abrLo.ptr;
    v.abrHi = abrHi.ptr;
    v.abrCellWidth = abrCellWidth.ptr;
    v.leafList = leafList.ptr;
    v.nodes = (const int2 *)nodes.ptr;
    v.abrMaxOp = haveMaxOp ? maxOp.ptr : nullptr;
    v.numAbrs = numAbrs;
    return v;
  }

  // ExaBrickSampler::computeMaxOpacities in ABR mode (ExaBrickSampler.cu:65-98,
  // 110-118): computeMaxOpacitiesForABRs, the same mapping as the grid's
  void computeMaxOpacities(const float *rgba, int n, float lo, float hi, hipStream_t s) {
    require(rgba && n > 0, "empty color map");
    colorMap.upload((const float4 *)rgba, n, s);
    launchMapMaxOpacity(abrRanges.ptr, numAbrs, colorMap.ptr, n, lo, hi, maxOp.ptr, s);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
    haveMaxOp = true;
  }
};

// Neural field sampler state: parameters (float master, bf16 copies), Adam
// moments, the training pool, and the per-macrocell validation result.
struct exa_neural {
  int32_t levels = 16, log2T = 19;
  float nMin = 16.f, nMax = 512.f;
  int32_t res[kNeuralMaxLevels] = {};
  uint32_t dense[kNeuralMaxLevels] = {};
  int64_t entries = 0;  // levels << log2T
  int64_t numParams = 0;
  DeviceBuffer<float> master, grad, m, v;  // table (2 per entry) | W1 | W2 | w3
  DeviceBuffer<uint32_t> table;           // packed bf16 pairs
  DeviceBuffer<uint16_t> mlp;             // bf16 W1 | W2 | w3
  DeviceBuffer<float4> pool;
  int64_t poolSize = 0;
  DeviceBuffer<float> lossBuf;
  int step = 0;
  // model frame
  float lo[3] = {0, 0, 0}, invExtent[3] = {1, 1, 1};
  float vmin = 0.f, vrange = 1.f;
  // validation
  DeviceBuffer<uint8_t> flags;
  DeviceBuffer<float2> widened;
  float bound = 0.f;
  bool validated = false;
  uint64_t generation = 0;  // bumped whenever flags / widened ranges / parameters change

  static constexpr int64_t kW1 = kNeuralIn * kNeuralHidden, kW2 = kNeuralHidden * kNeuralHidden,
                           kW3 = kNeuralHidden;
  int64_t mlpOffset() const { return 2 * entries; }
  NeuralView view() const {
    NeuralView nv;
    nv.levels = levels;
    nv.log2T = log2T;
    for (int l = 0; l < kNeuralMaxLevels; ++l) {
      nv.res[l] = res[l];
      nv.dense[l] = dense[l];
    }
    nv.table = table.ptr;
    nv.W1 = mlp.ptr;
    nv.W2 = mlp.ptr + kW1;
    nv.w3 = mlp.ptr + kW1 + kW2;
    for (int k = 0; k < 3; ++k) {
      nv.lo[k] = lo[k];
      nv.invExtent[k] = invExtent[k];
    }
    nv.vmin = vmin;
    nv.vrange = vrange;
    nv.widened = nullptr;
    return nv;
  }
  void pack(hipStream_t s) {
    neuralPack(master.ptr, table.ptr, entries, master.ptr + mlpOffset(), mlp.ptr, kW1 + kW2 + kW3, s);
  }
};

// Buffers of the element-stage queue, sized on use: `depth` entries, or
// (depth 0) one per pixel of the frame, at most 2^21; a path that finds the
// queue full runs the renderer's element pass instead.
constexpr uint32_t kQueueCapMax = 1u << 21;
constexpr uint32_t kQueueBuckets = 1u << 16;
struct QueueStore {
  DeviceBuffer<float4> rec[2];
  DeviceBuffer<Sample> samples[2];
  DeviceBuffer<uint32_t> keys[2], order, hist;
  DeviceBuffer<uint32_t> count;
  ElementQueueBuffers view(int64_t pixels, int32_t depth) {
    const int64_t want = depth > 0 ? (int64_t)depth : std::min<int64_t>(pixels, kQueueCapMax);
    const uint32_t cap = (uint32_t)std::max<int64_t>(1, want);
    for (int k = 0; k < 2; ++k) {
      rec[k].alloc((size_t)cap * kElementQueueVec);
      samples[k].alloc(cap);
      keys[k].alloc(cap);
    }
    order.alloc(cap);
    hist.alloc(kQueueBuckets);
    count.alloc(3);
    ElementQueueBuffers b;
    for (int k = 0; k < 2; ++k) {
      b.rec[k] = rec[k].ptr;
      b.samples[k] = samples[k].ptr;
      b.keys[k] = keys[k].ptr;
    }
    b.order = order.ptr;
    b.hist = hist.ptr;
    b.maxBuckets = kQueueBuckets;
    b.count = count.ptr;
    b.cap = cap;
    return b;
  }
};

struct exa_renderer {
  HostModel model;
  exa_sampler sampler;
  RenderParams P;
  DeviceBuffer<float4> accum;
  DeviceBuffer<float4> tf;
  std::vector<float> colorMap;
  float absDomain[2] = {0.f, 1.f};
  float relDomain[2] = {0.f, 100.f};
  float modelBounds[6];
  float valueRange[2];
  int accumID = 0;
  int spp = 1;
  int scheme = EXA_RENDER_PERSISTENT;
  exa_neural *neural = nullptr;  // neural field for validated macrocells (not owned)
  PersistTuning tuning;
  DeviceBuffer<uint32_t> work;
  QueueStore queue;
  // set for a renderer over an ExaBricks model (then `sampler` stays unbuilt)
  std::unique_ptr<exa_brick_sampler> bricks;

  void updateMaxOpacities() {
    if (colorMap.empty()) return;
    // OWLRenderer.cpp:647-652
    const float lo = absDomain[0] + (relDomain[0] / 100.f) * (absDomain[1] - absDomain[0]);
    const float hi = absDomain[0] + (relDomain[1] / 100.f) * (absDomain[1] - absDomain[0]);
    if (bricks) {
      bricks->computeMaxOpacities(colorMap.data(), (int)colorMap.size() / 4, lo, hi, 0);
      return;
    }
    sampler.computeMaxOpacities(colorMap.data(), (int)colorMap.size() / 4, lo, hi, 0);
    if (neural && neural->validated) {
      // majorants of the macrocells the network answers: the TF maximum over
      // the exact value range widened by the validation bound
      const int64_t nmc = (int64_t)sampler.dims[0] * sampler.dims[1] * sampler.dims[2];
      launchMapMaxOpacity(neural->widened.ptr, nmc, sampler.colorMap.ptr, (int)colorMap.size() / 4, lo, hi,
                          sampler.maxOpacities.ptr, 0);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipStreamSynchronize(0));
    }
  }
};

// A batch of render requests sharing one renderer's model, sampler and
// tables; all of them run in one persistent launch per frame.
struct exa_batch {
  exa_renderer *r = nullptr;
  struct Tf {
    DeviceBuffer<float4> colorMap;
    DeviceBuffer<float> majorants, gridletMaxOp, elemMaxOp;
    float lo = 0.f, hi = 1.f, opacityScale = 1.f;
    int n = 0;
  };
  struct Req {
    int tf = 0;
    DeviceBuffer<float4> accum;
    DeviceBuffer<uint32_t> fb;
  };
  std::vector<std::unique_ptr<Tf>> tfs;
  std::vector<std::unique_ptr<Req>> reqs;
  std::vector<BatchRequest> host;  // mirror of the device array (accumIDs advanced on both)
  DeviceBuffer<BatchRequest> dev;
  // the neural field the TF majorants were widened for (nullptr: exact)
  const exa_neural *majNeural = nullptr;
  uint64_t majGeneration = 0;
  DeviceBuffer<uint32_t> work;
  uint32_t numWork = 0;
  int64_t numPixels = 0;
  QueueStore queue;
  bool dirty = true;
  PersistTuning tuning;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  ~exa_batch() { dropGraph(); }
  void dropGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  // uploads the request array when requests changed (synchronous)
  void sync() {
    if (!dirty) return;
    if (host.empty()) throw Status(EXA_ERR_STATE, "batch has no requests");
    uint32_t w = 0;
    numPixels = 0;
    for (BatchRequest &b : host) {
      numPixels += (int64_t)b.P.W * b.P.H;
      b.workBegin = w;
      const uint64_t tiles = (uint64_t)b.tilesX * (uint64_t)((b.P.H + 7) / 8);
      if ((uint64_t)w + tiles * 64 >= (1ull << 32)) throw Status(EXA_ERR_INVALID, "batch too large");
      w += (uint32_t)(tiles * 64);
    }
    numWork = w;
    dropGraph();
    dev.alloc(host.size());
    HIP_CHECK(hipDeviceSynchronize());
    HIP_CHECK(hipMemcpy(dev.ptr, host.data(), host.size() * sizeof(BatchRequest), hipMemcpyHostToDevice));
    work.alloc(1);
    dirty = false;
  }
  // With a validated neural field on the renderer, every TF's majorants come
  // from the macrocell value ranges widened by the validation bound where the
  // network answers (as the renderer's own); without one, from the exact
  // ranges.  Recomputed when that choice or the field's validation changes.
  const exa_neural *activeNeural() const {
    return r->neural && r->neural->validated && r->sampler.accel == kAccelTables ? r->neural : nullptr;
  }
  void syncMajorants() {
    const exa_neural *n = activeNeural();
    if (n == majNeural && (!n || n->generation == majGeneration)) return;
    const exa_sampler &S = r->sampler;
    const int64_t nmc = (int64_t)S.dims[0] * S.dims[1] * S.dims[2];
    for (auto &t : tfs)
      launchMapMaxOpacity(n ? n->widened.ptr : S.valueRanges.ptr, nmc, t->colorMap.ptr, t->n, t->lo, t->hi,
                          t->majorants.ptr, 0);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(0));
    majNeural = n;
    majGeneration = n ? n->generation : 0;
    dropGraph();  // a captured frame would keep the old network state
  }
  void launch(unsigned long long *counters, hipStream_t s) {
    const exa_neural *n = activeNeural();
    NeuralView nv{};
    if (n) {
      nv = n->view();
      nv.widened = n->widened.ptr;
    }
    ElementQueueBuffers qb{};
    const bool queued = tuning.elementStage == kElemQueueRounds && !n && r->sampler.accel == kAccelTables;
    if (queued) qb = queue.view(numPixels, tuning.queueDepth);
    launchRenderBatch(r->sampler.view(), r->sampler.accel, r->sampler.gridView(), dev.ptr, (int)host.size(),
                      numWork, work.ptr, tuning, counters, s, n ? &nv : nullptr, n ? n->flags.ptr : nullptr,
                      queued ? &qb : nullptr);
  }
  void advanced() {
    for (BatchRequest &b : host) b.P.accumID += 1;
  }
};

// ---------------------------------------------------------------------------
extern "C" {

const char *exa_get_last_error(void) { return g_lastError.c_str(); }
int exa_version(void) { return 1; }

// ---- AMR ------------------------------------------------------------------
int exa_amr_create_gaussians(int baseRes, int coarseLevel, int fineLo, int fineHi, int numGaussians,
                             uint64_t seed, exa_amr **out) {
  return guarded([&] {
    require(out, "out is NULL");
    auto a = std::make_unique<exa_amr>();
    amrGaussians(a->scene, baseRes, coarseLevel, fineLo, fineHi, numGaussians, seed);
    *out = a.release();
  });
}
int exa_amr_create_noise(int baseRes, int baseLevel, float t1, float t2, int octaves, float contrast,
                         uint64_t seed, exa_amr **out) {
  return guarded([&] {
    require(out, "out is NULL");
    auto a = std::make_unique<exa_amr>();
    amrNoise(a->scene, baseRes, baseLevel, t1, t2, octaves, contrast, seed);
    *out = a.release();
  });
}
int exa_amr_create_shock(int baseRes, int baseLevel, float radius, float thickness, float thetaCore,
                         float thetaWake, uint64_t seed, exa_amr **out) {
  return guarded([&] {
    require(out, "out is NULL");
    auto a = std::make_unique<exa_amr>();
    amrShock(a->scene, baseRes, baseLevel, radius, thickness, thetaCore, thetaWake, seed);
    *out = a.release();
  });
}
int exa_amr_count_shock(int baseRes, int baseLevel, float radius, float thickness, float thetaCore,
                        float thetaWake, uint64_t seed, int64_t counts[16]) {
  return guarded([&] {
    require(counts, "counts is NULL");
    amrShockCount(baseRes, baseLevel, radius, thickness, thetaCore, thetaWake, seed, counts);
  });
}
int exa_amr_create_from_cells(const exa_amr_cell *cells, const float *scalars, int64_t n, exa_amr **out) {
  return guarded([&] {
    require(out && cells && scalars && n > 0, "invalid cells");
    auto a = std::make_unique<exa_amr>();
    a->scene.cells.assign((const AmrCell *)cells, (const AmrCell *)cells + n);
    a->scene.scalars.assign(scalars, scalars + n);
    *out = a.release();
  });
}
int64_t exa_amr_num_cells(const exa_amr *a) { return a ? (int64_t)a->scene.cells.size() : -1; }
int exa_amr_get(const exa_amr *a, exa_amr_cell *cells, float *scalars) {
  return guarded([&] {
    require(a, "amr is NULL");
    if (cells) memcpy(cells, a->scene.cells.data(), a->scene.cells.size() * sizeof(AmrCell));
    if (scalars) memcpy(scalars, a->scene.scalars.data(), a->scene.scalars.size() * sizeof(float));
  });
}
int exa_amr_get_params(const exa_amr *a, float *out, int maxCount) {
  if (!a) return EXA_ERR_INVALID;
  const int n = (int)a->scene.params.size();
  if (out)
    for (int i = 0; i < n && i < maxCount; ++i) out[i] = a->scene.params[i];
  return n;
}
void exa_amr_destroy(exa_amr *a) { delete a; }

// ---- dual mesh -------------------------------------------------------------
int exa_dual_mesh_build(const exa_amr *a, exa_dual_mesh **out) {
  return guarded([&] {
    require(a && out, "invalid arguments");
    auto d = std::make_unique<exa_dual_mesh>();
    buildDualMesh(a->scene.cells.data(), a->scene.scalars.data(), (int64_t)a->scene.cells.size(), d->mesh);
    *out = d.release();
  });
}
int exa_dual_mesh_counts(const exa_dual_mesh *d, int64_t c[11]) {
  return guarded([&] {
    require(d && c, "invalid arguments");
    const DualMesh &m = d->mesh;
    for (int t = 0; t < 4; ++t) c[t] = m.counts[t];
    c[4] = (int64_t)m.indices.size() / 8;
    c[5] = (int64_t)m.vertices.size() / 4;
    c[6] = (int64_t)m.gridlets.size();
    c[7] = (int64_t)m.gridletScalars.size();
    c[8] = m.numCubes;
    c[9] = m.conflicts;
    c[10] = m.degenerate;
  });
}
int exa_dual_mesh_get(const exa_dual_mesh *d, int32_t *indices, float *vertices, exa_gridlet *gridlets,
                      float *gridletScalars) {
  return guarded([&] {
    require(d, "dual mesh is NULL");
    const DualMesh &m = d->mesh;
    if (indices) memcpy(indices, m.indices.data(), m.indices.size() * sizeof(int32_t));
    if (vertices) memcpy(vertices, m.vertices.data(), m.vertices.size() * sizeof(float));
    if (gridlets) memcpy(gridlets, m.gridlets.data(), m.gridlets.size() * sizeof(Gridlet));
    if (gridletScalars) memcpy(gridletScalars, m.gridletScalars.data(), m.gridletScalars.size() * sizeof(float));
  });
}
void exa_dual_mesh_destroy(exa_dual_mesh *d) { delete d; }

// ---- on-disk formats ----------------------------------------------------------
static std::string pathOf(const char *p) { return p ? std::string(p) : std::string(); }

int exa_amr_load(const char *cellsPath, const char *scalarsPath, exa_amr **out) {
  return guarded([&] {
    require(out && cellsPath && scalarsPath, "invalid arguments");
    auto a = std::make_unique<exa_amr>();
    readCells(cellsPath, a->scene.cells);
    readScalars(scalarsPath, a->scene.scalars);
    require(!a->scene.cells.empty(), "no cells in the cells file");
    require(a->scene.scalars.size() >= a->scene.cells.size(), "fewer scalars than cells");
    a->scene.scalars.resize(a->scene.cells.size());
    *out = a.release();
  });
}
int exa_amr_save(const exa_amr *a, const char *cellsPath, const char *scalarsPath) {
  return guarded([&] {
    require(a, "amr is NULL");
    if (cellsPath) writeCells(cellsPath, a->scene.cells);
    if (scalarsPath) writeScalars(scalarsPath, a->scene.scalars);
  });
}
int exa_amr_write_dual_mesh(const exa_amr *a, const char *elementMeshPath, const char *cubesPrefix,
                            const char *gridsPath, int64_t counts[11]) {
  return guarded([&] {
    require(a, "amr is NULL");
    exa_dual_mesh d;
    buildDualMesh(a->scene.cells.data(), a->scene.scalars.data(), (int64_t)a->scene.cells.size(), d.mesh, true);
    const DualMesh &m = d.mesh;
    if (elementMeshPath) {
      ElementMesh em;
      const int64_t nv = (int64_t)m.vertices.size() / 4;
      em.xyz.resize(3 * nv);
      em.values.resize(nv);
      for (int64_t i = 0; i < nv; ++i) {
        for (int k = 0; k < 3; ++k) em.xyz[3 * i + k] = m.vertices[4 * i + k];
        em.values[i] = m.vertices[4 * i + 3];
      }
      em.vertexTag = m.vertexTags;
      static const int nvs[4] = {4, 5, 6, 8};
      int64_t e = 0;
      for (int t = 0; t < 4; ++t)
        for (int64_t i = 0; i < m.counts[t]; ++i, ++e)
          for (int j = 0; j < nvs[t]; ++j) em.elems[t].push_back(m.indices[8 * e + j]);
      writeElementMesh(elementMeshPath, em);
    }
    if (cubesPrefix) writeCubes(cubesPrefix, m.cubes);
    if (gridsPath) writeGrids(gridsPath, m.gridlets, m.gridletScalarIDs, false);
    if (counts) {
      exa_dual_mesh_counts(&d, counts);
    }
  });
}
int exa_make_grids(const char *const *cubesPaths, int numPaths, const char *gridsPath, int64_t *conflicts) {
  return guarded([&] {
    require(gridsPath && numPaths >= 0 && (numPaths == 0 || cubesPaths), "invalid arguments");
    std::vector<std::string> paths;
    for (int i = 0; i < numPaths; ++i) {
      require(cubesPaths[i], "NULL cubes path");
      paths.push_back(cubesPaths[i]);
    }
    makeGridsFiles(paths, gridsPath, conflicts);
  });
}
int exa_model_load(const char *elementMeshPath, const char *gridsPath, const char *scalarsPath, exa_model **out) {
  return guarded([&] {
    require(out, "invalid arguments");
    auto m = std::make_unique<exa_model>();
    loadModelFiles(pathOf(elementMeshPath), pathOf(gridsPath), pathOf(scalarsPath), m->m);
    require(m->m.numElements() > 0 || !m->m.gridlets.empty(), "the model files hold no elements and no gridlets");
    *out = m.release();
  });
}
int exa_xf_write(const char *path, const float params[5], const float *rgba, int numColors) {
  return guarded([&] {
    require(path && params && numColors >= 0 && (numColors == 0 || rgba), "invalid arguments");
    XfFile xf;
    xf.opacityScale = params[0];
    xf.absDomain[0] = params[1];
    xf.absDomain[1] = params[2];
    xf.relDomain[0] = params[3];
    xf.relDomain[1] = params[4];
    xf.colorMap.assign(rgba, rgba + 4 * (size_t)numColors);
    writeXf(path, xf);
  });
}

// ---- model -----------------------------------------------------------------
int exa_model_create(const exa_stitch_model_desc *desc, exa_model **out) {
  return guarded([&] {
    require(desc && out, "invalid arguments");
    require(desc->numVertices >= 0 && desc->numElements >= 0 && desc->numGridlets >= 0 &&
                desc->numGridletScalars >= 0,
            "negative counts");
    require(!desc->numVertices || desc->vertices, "vertices missing");
    require(!desc->numElements || desc->indices, "indices missing");
    require(!desc->numGridlets || desc->gridlets, "gridlets missing");
    require(!desc->numGridletScalars || desc->gridletScalars, "gridlet scalars missing");
    auto m = std::make_unique<exa_model>();
    m->m.vertices.assign(desc->vertices, desc->vertices + 4 * desc->numVertices);
    m->m.indices.assign(desc->indices, desc->indices + 8 * desc->numElements);
    m->m.gridlets.assign((const Gridlet *)desc->gridlets, (const Gridlet *)desc->gridlets + desc->numGridlets);
    m->m.gridletScalars.assign(desc->gridletScalars, desc->gridletScalars + desc->numGridletScalars);
    finalizeModel(m->m);
    *out = m.release();
  });
}
int exa_model_create_from_dual_mesh(const exa_dual_mesh *d, exa_model **out) {
  return guarded([&] {
    require(d && out, "invalid arguments");
    auto m = std::make_unique<exa_model>();
    m->m.vertices = d->mesh.vertices;
    m->m.indices = d->mesh.indices;
    m->m.gridlets = d->mesh.gridlets;
    m->m.gridletScalars = d->mesh.gridletScalars;
    finalizeModel(m->m);
    *out = m.release();
  });
}
int exa_model_get_bounds(const exa_model *m, float cb[6], float vr[2]) {
  return guarded([&] {
    require(m, "model is NULL");
    if (cb) memcpy(cb, m->m.cellBounds, sizeof(float) * 6);
    if (vr) memcpy(vr, m->m.valueRange, sizeof(float) * 2);
  });
}
int exa_model_get_counts(const exa_model *m, int64_t c[4]) {
  return guarded([&] {
    require(m && c, "invalid arguments");
    c[0] = (int64_t)m->m.vertices.size() / 4;
    c[1] = m->m.numElements();
    c[2] = (int64_t)m->m.gridlets.size();
    c[3] = (int64_t)m->m.gridletScalars.size();
  });
}
int exa_model_get(const exa_model *m, float *vertices, int32_t *indices, exa_gridlet *gridlets,
                  float *gridletScalars) {
  return guarded([&] {
    require(m, "model is NULL");
    const HostModel &h = m->m;
    if (vertices) memcpy(vertices, h.vertices.data(), h.vertices.size() * sizeof(float));
    if (indices) memcpy(indices, h.indices.data(), h.indices.size() * sizeof(int32_t));
    if (gridlets) memcpy(gridlets, h.gridlets.data(), h.gridlets.size() * sizeof(Gridlet));
    if (gridletScalars) memcpy(gridletScalars, h.gridletScalars.data(), h.gridletScalars.size() * sizeof(float));
  });
}
int exa_model_set_num_grid_cells(exa_model *m, const int32_t numMCs[3]) {
  return guarded([&] {
    require(m && numMCs, "invalid arguments");
    require(numMCs[0] >= 0 && numMCs[1] >= 0 && numMCs[2] >= 0, "negative macrocell count");
    require((int64_t)numMCs[0] * numMCs[1] * numMCs[2] < (1ll << 31), "macrocell grid too large");
    for (int k = 0; k < 3; ++k) m->m.numMCs[k] = numMCs[k];
  });
}
// Model.cpp:35-48 voxelSpaceTransform = T(from.lo)S(from.span) * rcp(T(to.lo)S(to.span))
static void voxelTransform(const float from[6], const float to[6], float out[12]) {
  for (int k = 0; k < 9; ++k) out[k] = 0.f;
  for (int k = 0; k < 3; ++k) {
    const float sf = from[3 + k] - from[k], st = to[3 + k] - to[k];
    require(st != 0.f, "degenerate remap_to box");
    const float inv = 1.f / st;
    out[4 * k] = sf * inv;
    out[9 + k] = sf * (-(inv * to[k])) + from[k];
  }
}
int exa_model_set_voxel_space_transform(exa_model *m, const float from[6], const float to[6]) {
  return guarded([&] {
    require(m && from && to, "invalid arguments");
    voxelTransform(from, to, m->m.voxelXfm);
  });
}
void exa_model_destroy(exa_model *m) { delete m; }

// ---- sampler ---------------------------------------------------------------
int exa_sampler_create(exa_sampler **out) {
  return guarded([&] {
    require(out, "out is NULL");
    *out = new exa_sampler();
  });
}
int exa_sampler_build(exa_sampler *s, const exa_model *m, void *stream) {
  return guarded([&] {
    require(s && m, "invalid arguments");
    s->build(m->m, (hipStream_t)stream);
  });
}
int exa_sampler_compute_max_opacities(exa_sampler *s, const float *rgba, int n, float lo, float hi,
                                      void *stream) {
  return guarded([&] {
    require(s, "sampler is NULL");
    s->computeMaxOpacities(rgba, n, lo, hi, (hipStream_t)stream);
  });
}
// Point location of n queries: gridlet stage, then the element stage over a
// queue of the queries left (stream-ordered scratch of n + 1 uint32, so calls
// on different streams do not share it); one pass per query beyond 2^32 - 2.
static void locateTwoStage(const exa_sampler *s, const float *d_pos, int64_t n, Sample *d_out,
                           unsigned long long *counters, hipStream_t st) {
  if (n <= 0) return;
  uint32_t *queue = nullptr;
  if (n < (int64_t)0xFFFFFFFEll) HIP_CHECK(hipMallocAsync((void **)&queue, (size_t)(n + 1) * sizeof(uint32_t), st));
  launchLocate(s->view(), s->accel, d_pos, n, d_out, counters, st, queue);
  const hipError_t e = hipGetLastError();
  if (queue) HIP_CHECK(hipFreeAsync(queue, st));
  HIP_CHECK(e);
}
int exa_sampler_locate(const exa_sampler *s, const float *d_pos, int64_t n, exa_sample *d_out, void *stream) {
  return guarded([&] {
    require(s && (n == 0 || (d_pos && d_out)) && n >= 0, "invalid arguments");
    if (!s->built) throw Status(EXA_ERR_STATE, "sampler not built");
    locateTwoStage(s, d_pos, n, (Sample *)d_out, nullptr, (hipStream_t)stream);
  });
}
int exa_sampler_locate_counted(const exa_sampler *s, const float *d_pos, int64_t n, exa_sample *d_out,
                               uint64_t counters[6], void *stream) {
  return guarded([&] {
    require(s && counters && (n == 0 || (d_pos && d_out)) && n >= 0, "invalid arguments");
    if (!s->built) throw Status(EXA_ERR_STATE, "sampler not built");
    hipStream_t st = (hipStream_t)stream;
    HIP_CHECK(hipMemsetAsync(s->counters.ptr, 0, kNumCounters * sizeof(unsigned long long), st));
    locateTwoStage(s, d_pos, n, (Sample *)d_out, s->counters.ptr, st);
    unsigned long long h[kNumCounters];
    HIP_CHECK(hipMemcpyAsync(h, s->counters.ptr, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    for (int i = 0; i < 6; ++i) counters[i] = h[i];
  });
}

// Point location that allocates nothing: the element-stage queue lives in a
// file-scope device array instead of stream-ordered scratch.  Each device
// holds its own copy of the array; host code reaches it only through the
// address hipGetSymbolAddress gives for that device (the array's name in host
// code is a stand-in, not a device address).
static __device__ __align__(256) uint32_t scratch[EXA_LOCATE_NOALLOC_MAX + 1];

static uint32_t *noAllocQueue() {
  static std::mutex mu;
  static uint32_t *ptrs[kMaxCachedDevices] = {};
  int dev = -1;
  HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxCachedDevices) throw Status(EXA_ERR_STATE, "device ordinal beyond 63");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[dev]) {
    void *p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Status(EXA_ERR_HIP, "hipGetSymbolAddress returned no address for the locate queue");
    ptrs[dev] = (uint32_t *)p;
  }
  return ptrs[dev];
}

int exa_sampler_locate_noalloc(const exa_sampler *s, const float *d_pos, int64_t n, exa_sample *d_out,
                               void *stream) {
  return guarded([&] {
    require(s && (n == 0 || (d_pos && d_out)) && n >= 0, "invalid arguments");
    if (n > (int64_t)EXA_LOCATE_NOALLOC_MAX)
      throw Status(EXA_ERR_INVALID, "exa_sampler_locate_noalloc: more than EXA_LOCATE_NOALLOC_MAX (2^24) points; "
                                    "split the batch");
    if (!s->built) throw Status(EXA_ERR_STATE, "sampler not built");
    if (n == 0) return;
    uint32_t *queue = noAllocQueue();  // queue[0] is cleared by a kernel of the launch sequence
    launchLocate(s->view(), s->accel, d_pos, n, (Sample *)d_out, nullptr, (hipStream_t)stream, queue);
    HIP_CHECK(hipGetLastError());
  });
}
int exa_sampler_test_sample(const exa_sampler *s, const float *d_pos, const int32_t *d_prim, int64_t n,
                            exa_sample *d_out, void *stream) {
  return guarded([&] {
    require(s && (n == 0 || (d_pos && d_prim && d_out)) && n >= 0, "invalid arguments");
    if (!s->built) throw Status(EXA_ERR_STATE, "sampler not built");
    launchTestSample(s->view(), d_pos, d_prim, n, (Sample *)d_out, (hipStream_t)stream);
    HIP_CHECK(hipGetLastError());
  });
}
int exa_sampler_get_grid_info(const exa_sampler *s, int32_t dims[3], float bounds[6]) {
  return guarded([&] {
    require(s, "sampler is NULL");
    if (!s->built) throw Status(EXA_ERR_STATE, "sampler not built");
    if (dims) memcpy(dims, s->dims, sizeof(s->dims));
    if (bounds) memcpy(bounds, s->bounds, sizeof(s->bounds));
  });
}
int exa_sampler_get_value_ranges(const exa_sampler *s, float *out) {
  return guarded([&] {
    require(s && s->built, "sampler not built");
    download(s->valueRanges, out);
  });
}
int exa_sampler_get_max_opacities(const exa_sampler *s, float *out) {
  return guarded([&] {
    require(s && s->built, "sampler not built");
    download(s->maxOpacities, out);
  });
}
int exa_sampler_get_gridlet_value_ranges(const exa_sampler *s, float *out) {
  return guarded([&] {
    require(s && s->built, "sampler not built");
    download(s->gridletRanges, out);
  });
}
int exa_sampler_get_prim_max_opacities(const exa_sampler *s, float *g, float *e) {
  return guarded([&] {
    require(s && s->built, "sampler not built");
    if (!s->haveMaxOp) throw Status(EXA_ERR_STATE, "max opacities not computed");
    download(s->gridletMaxOp, g);
    download(s->elemMaxOp, e);
  });
}
int exa_sampler_get_info(const exa_sampler *s, int64_t info[8]) {
  return guarded([&] {
    require(s && info, "invalid arguments");
    info[0] = (int64_t)s->gNodes.count;
    info[1] = s->gDepth;
    info[2] = (int64_t)s->gRefs.count;
    info[3] = (int64_t)s->eNodes.count;
    info[4] = s->eDepth;
    info[5] = (int64_t)s->eRefs.count;
    info[6] = (int64_t)(s->vertices.bytes() + s->indices.bytes() + s->gridlets.bytes() + s->gridletScalars.bytes());
    info[7] = (int64_t)(s->gNodes.bytes() + s->eNodes.bytes() + s->gRefs.bytes() + s->eRefs.bytes() +
                        s->tGStart.bytes() + s->tGList.bytes() + s->tGTop.bytes() + s->tGOwner.bytes() +
                        s->tEStart.bytes() + s->tECand.bytes() + s->tEVerts.bytes());
  });
}
int exa_sampler_set_accel(exa_sampler *s, int accel) {
  return guarded([&] {
    require(s, "sampler is NULL");
    if (!s->built) {
      if (accel != kAccelTables && accel != kAccelBvh) throw Status(EXA_ERR_INVALID, "unknown accelerator");
      s->accel = accel;
      return;
    }
    s->setAccel(accel, nullptr, 0);
  });
}
int exa_sampler_get_accel(const exa_sampler *s) { return s ? s->accel : EXA_ERR_INVALID; }
void exa_sampler_destroy(exa_sampler *s) { delete s; }

// ---- renderer --------------------------------------------------------------
// entry points of the exaStitch sampler only
static void requireStitch(const exa_renderer *r) {
  if (r && r->bricks) throw Status(EXA_ERR_STATE, "not available for a renderer over an ExaBricks model");
}

// OWLRenderer constructors after the model load (OWLRenderer.cpp:166-181,
// 278-291) and the parameter defaults of initGPU / render (309-622).
static void initRenderer(exa_renderer *r, const float cellBounds[6], const float valueRange[2],
                         const float remapFrom[6], const float remapTo[6], float voxelXfm[12]) {
  float cellSize[3] = {cellBounds[3] - cellBounds[0], cellBounds[4] - cellBounds[1], cellBounds[5] - cellBounds[2]};
  float lightScale = 1.f;
  while (fmaxf(fmaxf(cellSize[0], cellSize[1]), cellSize[2]) > 1000.f) {
    for (int k = 0; k < 3; ++k) cellSize[k] /= 1000.f;
    lightScale /= 1000.f;
  }
  static const float unit[6] = {0, 0, 0, 1, 1, 1};
  voxelTransform(remapFrom ? remapFrom : unit, remapTo ? remapTo : unit, voxelXfm);
  RenderParams &P = r->P;
  memset(&P, 0, sizeof(P));
  memcpy(P.voxelXfm, voxelXfm, sizeof(P.voxelXfm));
  affineInverse(voxelXfm, P.voxelXfmInv);
  // Model::getBounds (Model.cpp:50-70)
  xfmPointHost(P.voxelXfmInv, cellBounds, r->modelBounds);
  xfmPointHost(P.voxelXfmInv, cellBounds + 3, r->modelBounds + 3);
  memcpy(r->valueRange, valueRange, sizeof(r->valueRange));
  affineIdentity(P.lightXfm);
  P.lightXfm[0] = P.lightXfm[4] = P.lightXfm[8] = lightScale;
  memcpy(P.worldBounds, r->modelBounds, sizeof(P.worldBounds));
  P.clipEnabled = 0;
  P.clipN[0] = 0.f;
  P.clipN[1] = 0.f;
  P.clipN[2] = 1.f;
  P.clipD = 0.5f * (r->modelBounds[2] + r->modelBounds[5]);
  P.subImage[0] = 0.f;
  P.subImage[1] = 0.f;
  P.subImage[2] = 0.f;
  P.subImage[3] = 1.f;
  P.subImageSel[2] = 1.f;
  P.subImageSel[3] = 1.f;
  P.integrator = 0;
  P.shadeMode = 0;
  P.maxBounces = 1024;
  P.debugCrosshair = 1;
  P.dt = 1.f;  // OWLRenderer.cpp:616 sets render.dt = 1 for every frame
  P.heatMapEnabled = 0;
  P.heatMapScale = 1.f;
  P.opacityScale = 1.f;
  P.spp = 1;
}

int exa_renderer_create(const exa_model *m, const int32_t numMCs[3], const float remapFrom[6],
                        const float remapTo[6], exa_renderer **out) {
  return guarded([&] {
    require(m && numMCs && out, "invalid arguments");
    auto r = std::make_unique<exa_renderer>();
    r->model = m->m;
    HostModel &M = r->model;
    initRenderer(r.get(), M.cellBounds, M.valueRange, remapFrom, remapTo, M.voxelXfm);
    for (int k = 0; k < 3; ++k) M.numMCs[k] = numMCs[k];
    // initGPU (OWLRenderer.cpp:309-582)
    r->sampler.build(M, 0);
    *out = r.release();
  });
}
// OWLRenderer(const ExaBrick *, const float *, size_t, ...) (OWLRenderer.cpp:
// 269-291) with the ExaBrickSampler in ABR mode; numMCs is kept for the
// signature (the ABR traversal uses no macrocell grid).
int exa_renderer_create_bricks(const exa_brick_model *m, const int32_t numMCs[3], const float remapFrom[6],
                               const float remapTo[6], exa_renderer **out) {
  return guarded([&] {
    require(m && out, "invalid arguments");
    auto r = std::make_unique<exa_renderer>();
    float voxelXfm[12];
    initRenderer(r.get(), m->m.cellBounds, m->m.valueRange, remapFrom, remapTo, voxelXfm);
    if (numMCs)
      for (int k = 0; k < 3; ++k) r->model.numMCs[k] = numMCs[k];
    memcpy(r->model.voxelXfm, voxelXfm, sizeof(voxelXfm));
    r->bricks = std::make_unique<exa_brick_sampler>();
    r->bricks->build(m->m, 0);
    *out = r.release();
  });
}
int exa_renderer_set_camera(exa_renderer *r, const float org[3], const float d00[3], const float du[3],
                            const float dv[3]) {
  return guarded([&] {
    require(r && org && d00 && du && dv, "invalid arguments");
    memcpy(r->P.camOrg, org, 12);
    memcpy(r->P.camDir00, d00, 12);
    memcpy(r->P.camDu, du, 12);
    memcpy(r->P.camDv, dv, 12);
  });
}
int exa_renderer_resize(exa_renderer *r, int w, int h) {
  return guarded([&] {
    require(r && w > 0 && h > 0 && (int64_t)w * h < (1ll << 31), "invalid size");
    if (w != r->P.W || h != r->P.H) {
      r->accum.alloc((size_t)w * h);
      r->P.W = w;
      r->P.H = h;
    }
  });
}
static void doRender(exa_renderer *r, uint32_t *fb, unsigned long long *counters, hipStream_t s) {
  if (r->P.W <= 0 || r->P.H <= 0) throw Status(EXA_ERR_STATE, "resize() not called");
  if (r->tf.count == 0) throw Status(EXA_ERR_STATE, "setColorMap() not called");
  RenderParams P = r->P;
  P.accum = r->accum.ptr;
  P.fb = fb;
  P.accumID = r->accumID;
  P.spp = r->spp < 1 ? 1 : r->spp;
  P.tf = r->tf.ptr;
  P.tfN = (int32_t)r->tf.count;
  if (r->bricks) {
    if (!r->bricks->haveMaxOp) throw Status(EXA_ERR_STATE, "region max opacities not computed");
    uint32_t *work = nullptr;
    if (r->scheme == EXA_RENDER_PERSISTENT) {
      r->work.alloc(1);
      work = r->work.ptr;
    }
    launchRenderBricks(P, r->bricks->view(), work, counters, s);
    HIP_CHECK(hipGetLastError());
    r->accumID++;
    return;
  }
  // the ray marcher and the heat map (a per-pixel clock) run in the
  // one-thread-per-pixel kernel
  const bool mega = r->scheme != EXA_RENDER_PERSISTENT || P.integrator == 2 || P.heatMapEnabled;
  if (!mega) {
    r->work.alloc(1);
    const bool useNet = r->neural && r->neural->validated;
    if (useNet && r->sampler.accel != kAccelTables)
      throw Status(EXA_ERR_STATE, "the neural field runs with the table accelerator");
    NeuralView nv{};
    if (useNet) {
      nv = r->neural->view();
      nv.widened = r->neural->widened.ptr;
    }
    ElementQueueBuffers qb{};
    const bool queued = r->tuning.elementStage == kElemQueueRounds && !useNet && r->sampler.accel == kAccelTables;
    if (queued) qb = r->queue.view((int64_t)P.W * P.H, r->tuning.queueDepth);
    launchRenderPersistent(P, r->sampler.view(), r->sampler.accel, r->sampler.gridView(), r->work.ptr, r->tuning,
                           counters, s, useNet ? &nv : nullptr, useNet ? r->neural->flags.ptr : nullptr,
                           queued ? &qb : nullptr);
  } else {
    if (r->neural && r->neural->validated && P.integrator != 2)
      throw Status(EXA_ERR_STATE, "the neural field needs the persistent scheme");
    launchRender(P, r->sampler.view(), r->sampler.accel, r->sampler.gridView(), counters, s);
  }
  HIP_CHECK(hipGetLastError());
  r->accumID++;
}
int exa_renderer_render(exa_renderer *r, uint32_t *d_fb, void *stream) {
  return guarded([&] {
    require(r, "renderer is NULL");
    doRender(r, d_fb, nullptr, (hipStream_t)stream);
  });
}
int exa_renderer_render_counted(exa_renderer *r, uint32_t *d_fb, uint64_t counters[11], void *stream) {
  return guarded([&] {
    require(r && counters, "invalid arguments");
    hipStream_t st = (hipStream_t)stream;
    unsigned long long *dc = r->bricks ? r->bricks->counters.ptr : r->sampler.counters.ptr;
    HIP_CHECK(hipMemsetAsync(dc, 0, kNumCounters * sizeof(unsigned long long), st));
    doRender(r, d_fb, dc, st);
    unsigned long long h[kNumCounters];
    HIP_CHECK(hipMemcpyAsync(h, dc, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    for (int i = 0; i < kNumCounters; ++i) counters[i] = h[i];
  });
}
int exa_renderer_get_accum(const exa_renderer *r, float *out) {
  return guarded([&] {
    require(r && out, "invalid arguments");
    download(r->accum, out);
  });
}
int exa_renderer_set_type(exa_renderer *r, int type) {
  return guarded([&] {
    require(r, "renderer is NULL");
    require(type >= 0 && type <= 2, "integrators: 0 path tracer, 1 direct lighting, 2 ray marcher");
    r->P.integrator = type;
    r->accumID = 0;
  });
}
int exa_renderer_set_color_map(exa_renderer *r, const float *rgba, int n) {
  return guarded([&] {
    require(r && rgba && n > 0, "invalid color map");
    r->colorMap.assign(rgba, rgba + 4 * (size_t)n);
    r->tf.upload((const float4 *)rgba, n, 0);
    r->updateMaxOpacities();
    r->accumID = 0;
  });
}
int exa_renderer_set_range(exa_renderer *r, float lo, float hi) {
  return guarded([&] {
    require(r, "renderer is NULL");
    r->absDomain[0] = lo;
    r->absDomain[1] = hi;
    // OWLRenderer.cpp:717-733
    r->P.tfDomain[0] = r->absDomain[0] + (r->relDomain[0] / 100.f) * (r->absDomain[1] - r->absDomain[0]);
    r->P.tfDomain[1] = r->absDomain[0] + (r->relDomain[1] / 100.f) * (r->absDomain[1] - r->absDomain[0]);
    r->updateMaxOpacities();
  });
}
int exa_renderer_set_rel_domain(exa_renderer *r, float lo, float hi) {
  if (!r) return EXA_ERR_INVALID;
  r->relDomain[0] = lo;
  r->relDomain[1] = hi;
  return exa_renderer_set_range(r, r->absDomain[0], r->absDomain[1]);
}
int exa_renderer_set_opacity_scale(exa_renderer *r, float scale) {
  return guarded([&] {
    require(r, "renderer is NULL");
    r->P.opacityScale = powf(1.1f, scale - 100);  // OWLRenderer.cpp:741-744
  });
}
int exa_renderer_set_clip_plane(exa_renderer *r, int id, int enabled, const float N[3], float d) {
  return guarded([&] {
    require(r && N, "invalid arguments");
    require(id == 0, "clip plane id must be 0 (CLIP_PLANES_MAX == 1)");
    r->P.clipEnabled = enabled ? 1 : 0;
    memcpy(r->P.clipN, N, 12);
    r->P.clipD = d;
    r->accumID = 0;
  });
}
int exa_renderer_set_shade_mode(exa_renderer *r, int mode) {
  return guarded([&] {
    require(r, "renderer is NULL");
    require(mode >= 0 && mode <= 2, "shade modes: 0 default, 1 gridlets, 2 teaser");
    r->P.shadeMode = mode;
    r->accumID = 0;
  });
}
int exa_renderer_set_dt(exa_renderer *r, float dt) {
  return guarded([&] {
    require(r, "renderer is NULL");
    require(dt > 0.f && dt < 1e30f, "dt must be positive and finite");
    r->P.dt = dt;
    r->accumID = 0;
  });
}
int exa_renderer_set_heat_map(exa_renderer *r, int enabled, float scale) {
  return guarded([&] {
    require(r, "renderer is NULL");
    r->P.heatMapEnabled = enabled ? 1 : 0;
    r->P.heatMapScale = scale;
    r->accumID = 0;
  });
}
int exa_renderer_set_light_source(exa_renderer *r, int id, const float pos[3], float intensity, int on) {
  return guarded([&] {
    require(r && pos && id >= 0 && id < kLightsMax, "invalid light");
    LightParams &L = r->P.lights[id];
    memcpy(L.pos, pos, 12);
    L.intensity = intensity;
    L.on = on ? 1 : 0;
    L.type = 0;
    r->accumID = 0;
  });
}
int exa_renderer_set_directional_light(exa_renderer *r, int id, const float dir[3], float intensity, int on) {
  return guarded([&] {
    require(r && dir && id >= 0 && id < kLightsMax, "invalid light");
    LightParams &L = r->P.lights[id];
    memcpy(L.pos, dir, 12);
    L.intensity = intensity;
    L.on = on ? 1 : 0;
    L.type = 1;
    r->accumID = 0;
  });
}
int exa_renderer_set_light_space_transform(exa_renderer *r, const float xfm[12]) {
  return guarded([&] {
    require(r && xfm, "invalid arguments");
    memcpy(r->P.lightXfm, xfm, sizeof(r->P.lightXfm));
    r->accumID = 0;
  });
}
int exa_renderer_set_sub_image(exa_renderer *r, const float box[4], int active) {
  return guarded([&] {
    require(r && box, "invalid arguments");
    memcpy(r->P.subImage, box, 16);
    r->P.subImageActive = active ? 1 : 0;
    r->accumID = 0;
  });
}
int exa_renderer_set_sub_image_selection(exa_renderer *r, const float box[4], int active) {
  return guarded([&] {
    require(r && box, "invalid arguments");
    memcpy(r->P.subImageSel, box, 16);
    r->P.subImageSelecting = active ? 1 : 0;
    r->accumID = 0;
  });
}
int exa_renderer_reset_accum(exa_renderer *r) {
  if (!r) return EXA_ERR_INVALID;
  r->accumID = 0;
  return EXA_OK;
}
int exa_renderer_set_spp(exa_renderer *r, int spp) {
  if (!r) return EXA_ERR_INVALID;
  r->spp = spp;
  return EXA_OK;
}
int exa_renderer_set_max_bounces(exa_renderer *r, int n) {
  if (!r || n < 0) return EXA_ERR_INVALID;
  r->P.maxBounces = n;
  r->accumID = 0;
  return EXA_OK;
}
int exa_renderer_set_debug_crosshair(exa_renderer *r, int on) {
  if (!r) return EXA_ERR_INVALID;
  r->P.debugCrosshair = on ? 1 : 0;
  return EXA_OK;
}
int exa_renderer_get_bounds(const exa_renderer *r, float b[6]) {
  if (!r || !b) return EXA_ERR_INVALID;
  memcpy(b, r->modelBounds, sizeof(r->modelBounds));
  return EXA_OK;
}
int exa_renderer_get_value_range(const exa_renderer *r, float v[2]) {
  if (!r || !v) return EXA_ERR_INVALID;
  memcpy(v, r->valueRange, sizeof(r->valueRange));
  return EXA_OK;
}
int exa_renderer_get_accum_id(const exa_renderer *r) { return r ? r->accumID : EXA_ERR_INVALID; }
int exa_renderer_set_accum_id(exa_renderer *r, int accumID) {
  if (!r || accumID < 0) return EXA_ERR_INVALID;
  r->accumID = accumID;
  return EXA_OK;
}
const exa_sampler *exa_renderer_get_sampler(const exa_renderer *r) { return r ? &r->sampler : nullptr; }
int exa_renderer_set_accel(exa_renderer *r, int accel) {
  return guarded([&] {
    require(r, "renderer is NULL");
    requireStitch(r);
    r->sampler.setAccel(accel, &r->model, 0);
  });
}
int exa_renderer_set_scheme(exa_renderer *r, int scheme) {
  return guarded([&] {
    require(r, "renderer is NULL");
    require(scheme == EXA_RENDER_MEGAKERNEL || scheme == EXA_RENDER_PERSISTENT, "unknown render scheme");
    r->scheme = scheme;
  });
}
int exa_renderer_get_scheme(const exa_renderer *r) { return r ? r->scheme : EXA_ERR_INVALID; }
int exa_renderer_set_persistent_tuning(exa_renderer *r, int ctrlSteps, int elemBatch, int rareBatch) {
  return guarded([&] {
    require(r, "renderer is NULL");
    require(ctrlSteps >= 1 && ctrlSteps <= 1024, "ctrlSteps must be in [1, 1024]");
    require(elemBatch >= 0 && elemBatch <= 64, "elemBatch must be in [0, 64]");
    require(rareBatch >= 0 && rareBatch <= 64, "rareBatch must be in [0, 64]");
    r->tuning.ctrlSteps = ctrlSteps;
    r->tuning.elemBatch = elemBatch;
    r->tuning.rareBatch = rareBatch;
  });
}
static void checkLaunchTuning(int elementStage, int queueRounds, int queueDepth, int blockSize) {
  require(elementStage >= EXA_ELEMENT_IN_WAVE && elementStage <= EXA_ELEMENT_QUEUE_ROUNDS,
          "elementStage must be EXA_ELEMENT_IN_WAVE, EXA_ELEMENT_WORKGROUP_QUEUE or EXA_ELEMENT_QUEUE_ROUNDS");
  require(queueRounds >= 1 && queueRounds <= kMaxQueueRounds, "queueRounds must be in [1, 32]");
  require(queueDepth >= 0 && queueDepth <= (1 << 24), "queueDepth must be in [0, 2^24]");
  require(blockSize == 64 || blockSize == 128 || blockSize == 256, "blockSize must be 64, 128 or 256");
}
static void setLaunchTuning(PersistTuning &t, int elementStage, int queueRounds, int queueDepth, int blockSize) {
  checkLaunchTuning(elementStage, queueRounds, queueDepth, blockSize);
  t.elementStage = elementStage;
  t.queueRounds = queueRounds;
  t.queueDepth = queueDepth;
  t.blockSize = blockSize;
}
int exa_renderer_set_launch_tuning(exa_renderer *r, int elementStage, int queueRounds, int queueDepth,
                                   int blockSize) {
  return guarded([&] {
    require(r, "renderer is NULL");
    setLaunchTuning(r->tuning, elementStage, queueRounds, queueDepth, blockSize);
  });
}
int exa_renderer_get_tuning(const exa_renderer *r, int tuning[7]) {
  if (!r || !tuning) return EXA_ERR_INVALID;
  tuning[0] = r->tuning.ctrlSteps;
  tuning[1] = r->tuning.elemBatch;
  tuning[2] = r->tuning.rareBatch;
  tuning[3] = r->tuning.elementStage;
  tuning[4] = r->tuning.queueRounds;
  tuning[5] = r->tuning.queueDepth;
  tuning[6] = r->tuning.blockSize;
  return EXA_OK;
}
// ---- batched requests -------------------------------------------------------
int exa_batch_create(exa_renderer *r, exa_batch **out) {
  return guarded([&] {
    require(r && out, "invalid arguments");
    requireStitch(r);
    auto b = std::make_unique<exa_batch>();
    b->r = r;
    *out = b.release();
  });
}
void exa_batch_destroy(exa_batch *b) { delete b; }
int exa_batch_add_transfer_function(exa_batch *b, const float *rgba, int numColors, float lo, float hi,
                                    float opacityScale, int *index) {
  return guarded([&] {
    require(b && rgba && numColors > 0 && index, "invalid arguments");
    require(b->tfs.size() < 256, "too many transfer functions");
    exa_renderer *r = b->r;
    auto t = std::make_unique<exa_batch::Tf>();
    t->colorMap.upload((const float4 *)rgba, numColors, 0);
    t->lo = lo;
    t->hi = hi;
    t->opacityScale = opacityScale;
    t->n = numColors;
    // majorants and culling of this TF, computed by the shared sampler, then
    // the renderer's own state is put back
    exa_sampler &S = r->sampler;
    const bool hadMaxOp = S.haveMaxOp;
    S.computeMaxOpacities(rgba, numColors, lo, hi, 0);
    t->majorants.alloc(S.maxOpacities.count);
    t->gridletMaxOp.alloc(S.gridletMaxOp.count);
    t->elemMaxOp.alloc(S.elemMaxOp.count);
    if (S.maxOpacities.count)
      HIP_CHECK(hipMemcpy(t->majorants.ptr, S.maxOpacities.ptr, S.maxOpacities.bytes(), hipMemcpyDeviceToDevice));
    if (S.gridletMaxOp.count)
      HIP_CHECK(hipMemcpy(t->gridletMaxOp.ptr, S.gridletMaxOp.ptr, S.gridletMaxOp.bytes(), hipMemcpyDeviceToDevice));
    if (S.elemMaxOp.count)
      HIP_CHECK(hipMemcpy(t->elemMaxOp.ptr, S.elemMaxOp.ptr, S.elemMaxOp.bytes(), hipMemcpyDeviceToDevice));
    if (!r->colorMap.empty()) r->updateMaxOpacities();
    else S.haveMaxOp = hadMaxOp;
    *index = (int)b->tfs.size();
    b->tfs.push_back(std::move(t));
    b->majNeural = nullptr;  // the new TF has exact majorants; widen all on the next frame if needed
    b->majGeneration = 0;
  });
}
int exa_batch_add_request(exa_batch *b, const exa_batch_request *q, int *index) {
  return guarded([&] {
    require(b && q && index, "invalid arguments");
    require(q->width > 0 && q->height > 0 && (int64_t)q->width * q->height < (1ll << 31), "invalid image size");
    require(q->spp >= 1, "spp must be >= 1");
    require(q->integrator == 0 || q->integrator == 1, "supported integrators: 0 path tracer, 1 direct lighting");
    require(q->tf >= 0 && q->tf < (int)b->tfs.size(), "unknown transfer function");
    require((int)b->host.size() < kMaxBatchRequests, "too many requests in one batch");
    const exa_batch::Tf &t = *b->tfs[q->tf];
    auto rq = std::make_unique<exa_batch::Req>();
    rq->tf = q->tf;
    rq->accum.alloc((size_t)q->width * q->height);
    rq->fb.alloc((size_t)q->width * q->height);
    BatchRequest br;
    memset(&br, 0, sizeof(br));
    br.P = b->r->P;  // model transforms, lights, clip plane, bounce cap, crosshair
    RenderParams &P = br.P;
    P.W = q->width;
    P.H = q->height;
    P.spp = q->spp;
    P.accumID = 0;
    P.integrator = q->integrator;
    memcpy(P.camOrg, q->camOrg, 12);
    memcpy(P.camDir00, q->camDir00, 12);
    memcpy(P.camDu, q->camDu, 12);
    memcpy(P.camDv, q->camDv, 12);
    P.tfDomain[0] = t.lo;
    P.tfDomain[1] = t.hi;
    P.opacityScale = t.opacityScale;
    P.tf = t.colorMap.ptr;
    P.tfN = t.n;
    P.subImageActive = 0;
    P.subImageSelecting = 0;
    P.accum = rq->accum.ptr;
    P.fb = rq->fb.ptr;
    br.majorants = t.majorants.ptr;
    br.gridletMaxOp = t.gridletMaxOp.ptr;
    br.elemMaxOp = t.elemMaxOp.ptr;
    int nl = 0;
    for (int i = 0; i < kLightsMax; ++i) nl += P.lights[i].on ? 1 : 0;
    br.numLights = nl;
    br.tilesX = (q->width + 7) / 8;
    *index = (int)b->host.size();
    b->host.push_back(br);
    b->reqs.push_back(std::move(rq));
    b->dirty = true;
  });
}
int exa_batch_num_requests(const exa_batch *b) { return b ? (int)b->host.size() : EXA_ERR_INVALID; }
int exa_batch_set_tuning(exa_batch *b, int ctrlSteps, int elemBatch, int rareBatch) {
  return guarded([&] {
    require(b, "batch is NULL");
    require(ctrlSteps >= 1 && ctrlSteps <= 1024 && elemBatch >= 0 && elemBatch <= 64 && rareBatch >= 0 &&
                rareBatch <= 64,
            "tuning out of range");
    b->tuning.ctrlSteps = ctrlSteps;
    b->tuning.elemBatch = elemBatch;
    b->tuning.rareBatch = rareBatch;
    b->dropGraph();
  });
}
int exa_batch_get_tuning(const exa_batch *b, int tuning[3]) {
  if (!b || !tuning) return EXA_ERR_INVALID;
  tuning[0] = b->tuning.ctrlSteps;
  tuning[1] = b->tuning.elemBatch;
  tuning[2] = b->tuning.rareBatch;
  return EXA_OK;
}
int exa_batch_set_launch_tuning(exa_batch *b, int elementStage, int queueRounds, int queueDepth, int blockSize) {
  return guarded([&] {
    require(b, "batch is NULL");
    setLaunchTuning(b->tuning, elementStage, queueRounds, queueDepth, blockSize);
    b->dropGraph();
  });
}
int exa_batch_get_launch_tuning(const exa_batch *b, int tuning[4]) {
  if (!b || !tuning) return EXA_ERR_INVALID;
  tuning[0] = b->tuning.elementStage;
  tuning[1] = b->tuning.queueRounds;
  tuning[2] = b->tuning.queueDepth;
  tuning[3] = b->tuning.blockSize;
  return EXA_OK;
}
int exa_batch_render(exa_batch *b, void *stream) {
  return guarded([&] {
    require(b, "batch is NULL");
    if (b->r->tf.count == 0 && b->tfs.empty()) throw Status(EXA_ERR_STATE, "no transfer function");
    b->sync();
    b->syncMajorants();
    b->launch(nullptr, (hipStream_t)stream);
    HIP_CHECK(hipGetLastError());
    b->advanced();
  });
}
int exa_batch_render_counted(exa_batch *b, uint64_t counters[11], void *stream) {
  return guarded([&] {
    require(b && counters, "invalid arguments");
    b->sync();
    b->syncMajorants();
    hipStream_t st = (hipStream_t)stream;
    unsigned long long *dc = b->r->sampler.counters.ptr;
    HIP_CHECK(hipMemsetAsync(dc, 0, kNumCounters * sizeof(unsigned long long), st));
    b->launch(dc, st);
    HIP_CHECK(hipGetLastError());
    b->advanced();
    unsigned long long h[kNumCounters];
    HIP_CHECK(hipMemcpyAsync(h, dc, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    for (int i = 0; i < kNumCounters; ++i) counters[i] = h[i];
  });
}
int exa_batch_capture_graph(exa_batch *b, void *stream) {
  return guarded([&] {
    require(b && stream, "a non-default stream is needed for capture");
    b->sync();
    b->syncMajorants();
    b->dropGraph();
    hipStream_t st = (hipStream_t)stream;
    if (b->tuning.elementStage == kElemQueueRounds) b->queue.view(b->numPixels, b->tuning.queueDepth);  // before the capture
    HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    b->launch(nullptr, st);
    hipGraph_t g = nullptr;
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipStreamEndCapture(st, &g);
    if (e1 != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      throw Status(EXA_ERR_HIP, std::string("graph capture failed: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2));
    }
    b->graph = g;
    HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  });
}
int exa_batch_render_graph(exa_batch *b, void *stream) {
  return guarded([&] {
    require(b, "batch is NULL");
    if (!b->exec || b->dirty) throw Status(EXA_ERR_STATE, "no captured graph (capture after the last change)");
    const exa_neural *n = b->activeNeural();
    if (n != b->majNeural || (n && n->generation != b->majGeneration))
      throw Status(EXA_ERR_STATE, "the neural field changed since the capture (capture again)");
    HIP_CHECK(hipGraphLaunch(b->exec, (hipStream_t)stream));
    b->advanced();
  });
}
int exa_batch_reset_accum(exa_batch *b) {
  return guarded([&] {
    require(b, "batch is NULL");
    for (BatchRequest &q : b->host) q.P.accumID = 0;
    b->dirty = true;
  });
}
int exa_batch_get_accum_id(const exa_batch *b, int req) {
  if (!b || req < 0 || req >= (int)b->host.size()) return EXA_ERR_INVALID;
  return b->host[req].P.accumID;
}
int exa_batch_get_accum(const exa_batch *b, int req, float *out) {
  return guarded([&] {
    require(b && out && req >= 0 && req < (int)b->reqs.size(), "invalid arguments");
    download(b->reqs[req]->accum, out);
  });
}
int exa_batch_get_frame(const exa_batch *b, int req, uint32_t *out) {
  return guarded([&] {
    require(b && out && req >= 0 && req < (int)b->reqs.size(), "invalid arguments");
    download(b->reqs[req]->fb, out);
  });
}
// Tries every combination of the candidate knob values on one frame each
// (timed with HIP events on `stream`), keeps the fastest and resets the
// accumulation (the trial frames accumulated).
// One timed frame per combination of the candidate values of every axis;
// keeps the fastest.  Axes: ctrlSteps, elemBatch, rareBatch, elementStage,
// queueRounds, queueDepth, blockSize.
constexpr int kTuneAxes = 7;
static PersistTuning autotuneBatch(exa_batch *b, const int *const axes[kTuneAxes], const int counts[kTuneAxes],
                                   hipStream_t st, float *bestMs) {
  for (int a = 0; a < kTuneAxes; ++a) require(axes[a] && counts[a] > 0, "every axis needs at least one candidate");
  b->sync();
  b->syncMajorants();
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  float bestT = 3.4e38f;
  PersistTuning bestK = b->tuning;
  b->launch(nullptr, st);  // warm-up
  b->advanced();
  int idx[kTuneAxes] = {};
  for (;;) {
    PersistTuning t;
    t.ctrlSteps = axes[0][idx[0]];
    t.elemBatch = axes[1][idx[1]];
    t.rareBatch = axes[2][idx[2]];
    require(t.ctrlSteps >= 1 && t.ctrlSteps <= 1024 && t.elemBatch >= 0 && t.elemBatch <= 64 &&
                t.rareBatch >= 0 && t.rareBatch <= 64,
            "tuning candidate out of range");
    setLaunchTuning(t, axes[3][idx[3]], axes[4][idx[4]], axes[5][idx[5]], axes[6][idx[6]]);
    b->tuning = t;
    if (t.elementStage == kElemQueueRounds) b->queue.view(b->numPixels, t.queueDepth);  // outside the timed frame
    HIP_CHECK(hipEventRecord(e0, st));
    b->launch(nullptr, st);
    HIP_CHECK(hipEventRecord(e1, st));
    HIP_CHECK(hipEventSynchronize(e1));
    b->advanced();
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    if (ms < bestT) {
      bestT = ms;
      bestK = t;
    }
    int a = 0;
    while (a < kTuneAxes && ++idx[a] == counts[a]) idx[a++] = 0;
    if (a == kTuneAxes) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->tuning = bestK;
  b->dropGraph();
  for (BatchRequest &q : b->host) q.P.accumID = 0;
  b->dirty = true;
  if (bestMs) *bestMs = bestT;
  return bestK;
}
int exa_batch_autotune_launch(exa_batch *b, const int *ctrlSteps, int nCtrl, const int *elemBatch, int nElem,
                              const int *rareBatch, int nRare, const int *elementStage, int nStage,
                              const int *queueRounds, int nRounds, const int *queueDepth, int nDepth,
                              const int *blockSize, int nBlock, void *stream, int best[7], float *bestMs) {
  return guarded([&] {
    require(b && best, "invalid arguments");
    const int *axes[kTuneAxes] = {ctrlSteps, elemBatch, rareBatch, elementStage, queueRounds, queueDepth, blockSize};
    const int counts[kTuneAxes] = {nCtrl, nElem, nRare, nStage, nRounds, nDepth, nBlock};
    const PersistTuning k = autotuneBatch(b, axes, counts, (hipStream_t)stream, bestMs);
    best[0] = k.ctrlSteps;
    best[1] = k.elemBatch;
    best[2] = k.rareBatch;
    best[3] = k.elementStage;
    best[4] = k.queueRounds;
    best[5] = k.queueDepth;
    best[6] = k.blockSize;
  });
}
int exa_batch_autotune(exa_batch *b, const int *ctrlSteps, int nCtrl, const int *elemBatch, int nElem,
                       const int *rareBatch, int nRare, void *stream, int best[3], float *bestMs) {
  return guarded([&] {
    require(b && ctrlSteps && elemBatch && rareBatch && nCtrl > 0 && nElem > 0 && nRare > 0 && best,
            "invalid arguments");
    // the launch shape stays as set
    const int e = b->tuning.elementStage, q = b->tuning.queueRounds, d = b->tuning.queueDepth,
              bs = b->tuning.blockSize;
    const int *axes[kTuneAxes] = {ctrlSteps, elemBatch, rareBatch, &e, &q, &d, &bs};
    const int counts[kTuneAxes] = {nCtrl, nElem, nRare, 1, 1, 1, 1};
    const PersistTuning k = autotuneBatch(b, axes, counts, (hipStream_t)stream, bestMs);
    best[0] = k.ctrlSteps;
    best[1] = k.elemBatch;
    best[2] = k.rareBatch;
  });
}
// ---- neural field sampler ---------------------------------------------------------
int exa_neural_create(exa_renderer *r, int levels, int log2T, float nMin, float nMax, uint32_t seed,
                      exa_neural **out) {
  return guarded([&] {
    require(r && out, "invalid arguments");
    requireStitch(r);
    require(levels >= 1 && levels <= kNeuralMaxLevels, "levels must be in [1, 16]");
    require(log2T >= 4 && log2T <= 24, "log2T must be in [4, 24]");
    require(nMin >= 1.f && nMax >= nMin && nMax <= 65536.f, "need 1 <= nMin <= nMax <= 65536");
    auto n = std::make_unique<exa_neural>();
    n->levels = levels;
    n->log2T = log2T;
    n->nMin = nMin;
    n->nMax = nMax;
    const double b = levels > 1 ? exp((log((double)nMax) - log((double)nMin)) / (levels - 1)) : 1.0;
    for (int l = 0; l < levels; ++l) {
      n->res[l] = (int32_t)floor((double)nMin * pow(b, (double)l) + 1e-9);
      const double v = (double)(n->res[l] + 1);
      n->dense[l] = v * v * v <= (double)(1ll << log2T) ? 1u : 0u;
    }
    n->entries = (int64_t)levels << log2T;
    n->numParams = 2 * n->entries + exa_neural::kW1 + exa_neural::kW2 + exa_neural::kW3;
    n->master.alloc(n->numParams);
    n->grad.alloc(n->numParams);
    n->m.alloc(n->numParams);
    n->v.alloc(n->numParams);
    n->table.alloc(n->entries);
    n->mlp.alloc(exa_neural::kW1 + exa_neural::kW2 + exa_neural::kW3);
    float *p = n->master.ptr;
    neuralUniform(p, 2 * n->entries, -1e-4f, 1e-4f, seed, 0);
    p += 2 * n->entries;
    const float a1 = sqrtf(6.f / (kNeuralIn + kNeuralHidden)), a2 = sqrtf(6.f / (2 * kNeuralHidden)),
                a3 = sqrtf(6.f / (kNeuralHidden + 1));
    neuralUniform(p, exa_neural::kW1, -a1, a1, seed + 1, 0);
    neuralUniform(p + exa_neural::kW1, exa_neural::kW2, -a2, a2, seed + 2, 0);
    neuralUniform(p + exa_neural::kW1 + exa_neural::kW2, exa_neural::kW3, -a3, a3, seed + 3, 0);
    neuralZero(n->m.ptr, n->numParams, 0);
    neuralZero(n->v.ptr, n->numParams, 0);
    n->pack(0);
    // model frame: positions normalised over the cell bounds, values over the value range
    const HostModel &M = r->model;
    for (int k = 0; k < 3; ++k) {
      n->lo[k] = M.cellBounds[k];
      const float ext = M.cellBounds[3 + k] - M.cellBounds[k];
      n->invExtent[k] = ext > 0.f ? 1.f / ext : 1.f;
    }
    n->vmin = M.valueRange[0];
    n->vrange = M.valueRange[1] > M.valueRange[0] ? M.valueRange[1] - M.valueRange[0] : 1.f;
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipDeviceSynchronize());
    n->validated = false;
    n->generation++;
    *out = n.release();
  });
}
void exa_neural_destroy(exa_neural *n) { delete n; }
int exa_neural_get_config(const exa_neural *n, int32_t info[4], int32_t res[16]) {
  return guarded([&] {
    require(n && info, "invalid arguments");
    info[0] = n->levels;
    info[1] = n->log2T;
    info[2] = n->step;
    info[3] = n->validated ? 1 : 0;
    if (res)
      for (int l = 0; l < kNeuralMaxLevels; ++l) res[l] = n->res[l];
  });
}
// parameters as float: table (levels x 2^log2T x 2), W1 (32 x 64), W2 (64 x 64), w3 (64)
int exa_neural_get_params(const exa_neural *n, float *table, float *W1, float *W2, float *w3) {
  return guarded([&] {
    require(n, "neural field is NULL");
    HIP_CHECK(hipDeviceSynchronize());
    const float *p = n->master.ptr;
    if (table) HIP_CHECK(hipMemcpy(table, p, 2 * n->entries * sizeof(float), hipMemcpyDeviceToHost));
    p += 2 * n->entries;
    if (W1) HIP_CHECK(hipMemcpy(W1, p, exa_neural::kW1 * sizeof(float), hipMemcpyDeviceToHost));
    if (W2) HIP_CHECK(hipMemcpy(W2, p + exa_neural::kW1, exa_neural::kW2 * sizeof(float), hipMemcpyDeviceToHost));
    if (w3)
      HIP_CHECK(hipMemcpy(w3, p + exa_neural::kW1 + exa_neural::kW2, exa_neural::kW3 * sizeof(float),
                          hipMemcpyDeviceToHost));
  });
}
int exa_neural_set_params(exa_neural *n, const float *table, const float *W1, const float *W2, const float *w3) {
  return guarded([&] {
    require(n && table && W1 && W2 && w3, "invalid arguments");
    float *p = n->master.ptr;
    HIP_CHECK(hipMemcpy(p, table, 2 * n->entries * sizeof(float), hipMemcpyHostToDevice));
    p += 2 * n->entries;
    HIP_CHECK(hipMemcpy(p, W1, exa_neural::kW1 * sizeof(float), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(p + exa_neural::kW1, W2, exa_neural::kW2 * sizeof(float), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(p + exa_neural::kW1 + exa_neural::kW2, w3, exa_neural::kW3 * sizeof(float),
                        hipMemcpyHostToDevice));
    n->pack(0);
    neuralZero(n->m.ptr, n->numParams, 0);
    neuralZero(n->v.ptr, n->numParams, 0);
    n->step = 0;
    n->validated = false;
    n->generation++;
    HIP_CHECK(hipDeviceSynchronize());
  });
}
// Training pool of n samples drawn with the exact sampler (culling off):
// elemFrac of the points inside random stitching elements, the rest inside
// random gridlets.  Returns the number of points the exact sampler found.
int exa_neural_make_pool(exa_neural *n, exa_renderer *r, int64_t count, float elemFrac, uint32_t seed,
                         int64_t *numValid) {
  return guarded([&] {
    require(n && r && count > 0 && count < (1ll << 31), "invalid arguments");
    requireStitch(r);
    exa_sampler &S = r->sampler;
    DeviceBuffer<float> pos;
    DeviceBuffer<Sample> smp;
    pos.alloc(3 * count);
    smp.alloc(count);
    neuralPoolPositions(S.indices.ptr, S.vertices.ptr, S.numElements, S.gridlets.ptr, S.numGridlets, count,
                        elemFrac, seed, pos.ptr, 0);
    SamplerView sv = S.view();
    sv.gridletMaxOp = nullptr;
    sv.elemMaxOp = nullptr;
    launchLocate(sv, S.accel, pos.ptr, count, smp.ptr, nullptr, 0);
    n->pool.alloc(count);
    neuralPoolFinish(pos.ptr, smp.ptr, count, n->view(), n->pool.ptr, 0);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipDeviceSynchronize());
    n->poolSize = count;
    if (numValid) {
      std::vector<float4> h(count);
      HIP_CHECK(hipMemcpy(h.data(), n->pool.ptr, count * sizeof(float4), hipMemcpyDeviceToHost));
      int64_t c = 0;
      for (const float4 &e : h) c += e.w == e.w ? 1 : 0;
      *numValid = c;
    }
  });
}
static void neuralStep(exa_neural *n, const float4 *pool, int64_t poolSize, int batch, bool identity,
                       uint32_t seed, float *lossSlot, hipStream_t s) {
  neuralZero(n->grad.ptr, n->numParams, s);
  TrainArgs a;
  a.nv = n->view();
  a.pool = pool;
  a.poolSize = poolSize;
  a.batch = batch;
  a.tilesPerWave = identity ? 1 : 8;
  a.identity = identity ? 1 : 0;
  a.stepSeed = seed;
  a.gTable = n->grad.ptr;
  a.gW1 = n->grad.ptr + n->mlpOffset();
  a.gW2 = a.gW1 + exa_neural::kW1;
  a.gw3 = a.gW2 + exa_neural::kW2;
  a.lossSum = lossSlot;
  neuralTrain(a, s);
}
// steps of Adam on batches drawn from the pool, the learning rate decaying
// exponentially from lr to lrEnd; lossOut (may be NULL) receives the mean
// squared error of each step's batch (normalised values)
int exa_neural_train(exa_neural *n, int steps, int batch, float lr, float lrEnd, float *lossOut) {
  return guarded([&] {
    require(n && steps >= 0 && batch >= 64 && batch % 64 == 0, "batch must be a positive multiple of 64");
    require(lr > 0.f && lrEnd > 0.f, "learning rates must be positive");
    require(n->poolSize > 0, "no training pool (exa_neural_make_pool)");
    n->lossBuf.alloc(std::max(steps, 1));
    neuralZero(n->lossBuf.ptr, std::max(steps, 1), 0);
    for (int i = 0; i < steps; ++i) {
      n->step++;
      // exponential decay from lr to lrEnd over the run
      const float li = steps > 1 ? lr * powf(lrEnd / lr, (float)i / (float)(steps - 1)) : lr;
      neuralStep(n, n->pool.ptr, n->poolSize, batch, false, 0x85ebca6bu * (uint32_t)n->step, n->lossBuf.ptr + i, 0);
      neuralAdam(n->master.ptr, n->grad.ptr, n->m.ptr, n->v.ptr, n->numParams, li, 0.9f, 0.99f, 1e-10f, n->step,
                 0);
      n->pack(0);
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipDeviceSynchronize());
    n->validated = false;
    n->generation++;
    if (lossOut && steps > 0) {
      HIP_CHECK(hipMemcpy(lossOut, n->lossBuf.ptr, steps * sizeof(float), hipMemcpyDeviceToHost));
      for (int i = 0; i < steps; ++i) lossOut[i] /= (float)batch;
    }
  });
}
// Loss and gradients of one batch given on the host (x in [0,1]^3, targets
// normalised; NaN target = no sample), rows used in order.
int exa_neural_gradients(exa_neural *n, const float *x01, const float *target, int count, float *loss,
                         float *dTable, float *dW1, float *dW2, float *dw3) {
  return guarded([&] {
    require(n && x01 && target && count > 0 && count % 64 == 0, "count must be a positive multiple of 64");
    std::vector<float4> h(count);
    for (int i = 0; i < count; ++i) h[i] = make_float4(x01[3 * i], x01[3 * i + 1], x01[3 * i + 2], target[i]);
    DeviceBuffer<float4> pool;
    pool.upload(h.data(), count, 0);
    DeviceBuffer<float> ls;
    ls.alloc(1);
    neuralZero(ls.ptr, 1, 0);
    neuralStep(n, pool.ptr, count, count, true, 0, ls.ptr, 0);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipDeviceSynchronize());
    float l = 0.f;
    HIP_CHECK(hipMemcpy(&l, ls.ptr, sizeof(float), hipMemcpyDeviceToHost));
    if (loss) *loss = l / (float)count;
    const float *g = n->grad.ptr;
    if (dTable) HIP_CHECK(hipMemcpy(dTable, g, 2 * n->entries * sizeof(float), hipMemcpyDeviceToHost));
    g += 2 * n->entries;
    if (dW1) HIP_CHECK(hipMemcpy(dW1, g, exa_neural::kW1 * sizeof(float), hipMemcpyDeviceToHost));
    if (dW2) HIP_CHECK(hipMemcpy(dW2, g + exa_neural::kW1, exa_neural::kW2 * sizeof(float), hipMemcpyDeviceToHost));
    if (dw3)
      HIP_CHECK(hipMemcpy(dw3, g + exa_neural::kW1 + exa_neural::kW2, exa_neural::kW3 * sizeof(float),
                          hipMemcpyDeviceToHost));
  });
}
// one Adam step with the gradients of the last exa_neural_gradients call
int exa_neural_apply_adam(exa_neural *n, float lr) {
  return guarded([&] {
    require(n, "neural field is NULL");
    n->step++;
    neuralAdam(n->master.ptr, n->grad.ptr, n->m.ptr, n->v.ptr, n->numParams, lr, 0.9f, 0.99f, 1e-10f, n->step, 0);
    n->pack(0);
    HIP_CHECK(hipDeviceSynchronize());
  });
}
int exa_neural_evaluate(const exa_neural *n, const float *d_pos, int64_t count, float *d_out, void *stream) {
  return guarded([&] {
    require(n && (count == 0 || (d_pos && d_out)) && count >= 0, "invalid arguments");
    neuralEval(n->view(), d_pos, count, d_out, (hipStream_t)stream);
    HIP_CHECK(hipGetLastError());
  });
}
// Per-macrocell validation: pointsPerMC uniform points in every macrocell; a
// macrocell is answered by the network when it had points that reach the
// element stage, none that the exact sampler misses, and a max |error| <=
// bound (scalar units).  out[3]: macrocells using the network, macrocells with
// element-stage points, max error over the accepted macrocells.
int exa_neural_validate(exa_neural *n, exa_renderer *r, int pointsPerMC, float bound, uint32_t seed,
                        float out[3]) {
  return guarded([&] {
    require(n && r && pointsPerMC > 0 && bound >= 0.f, "invalid arguments");
    requireStitch(r);
    exa_sampler &S = r->sampler;
    const GridView gv = S.gridView();
    const int64_t nmc = (int64_t)gv.dims[0] * gv.dims[1] * gv.dims[2];
    const int64_t np = nmc * pointsPerMC;
    require(np < (1ll << 31), "too many validation points");
    DeviceBuffer<float> pos, net;
    DeviceBuffer<Sample> smp;
    DeviceBuffer<uint32_t> cnt, miss, err;
    pos.alloc(3 * np);
    net.alloc(np);
    smp.alloc(np);
    cnt.alloc(nmc);
    miss.alloc(nmc);
    err.alloc(nmc);
    HIP_CHECK(hipMemset(cnt.ptr, 0, nmc * 4));
    HIP_CHECK(hipMemset(miss.ptr, 0, nmc * 4));
    HIP_CHECK(hipMemset(err.ptr, 0, nmc * 4));
    neuralValidate(gv, pointsPerMC, seed, pos.ptr, np, 0);
    SamplerView sv = S.view();
    sv.gridletMaxOp = nullptr;
    sv.elemMaxOp = nullptr;
    launchLocate(sv, S.accel, pos.ptr, np, smp.ptr, nullptr, 0);
    neuralEval(n->view(), pos.ptr, np, net.ptr, 0);
    neuralValReduce(smp.ptr, net.ptr, np, pointsPerMC, cnt.ptr, miss.ptr, err.ptr, 0);
    n->flags.alloc(nmc);
    neuralValFlags(cnt.ptr, miss.ptr, err.ptr, nmc, bound, n->flags.ptr, 0);
    n->widened.alloc(nmc);
    neuralWidenRanges(S.valueRanges.ptr, n->flags.ptr, nmc, bound, n->widened.ptr, 0);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipDeviceSynchronize());
    std::vector<uint32_t> hc(nmc), he(nmc);
    std::vector<uint8_t> hf(nmc);
    HIP_CHECK(hipMemcpy(hc.data(), cnt.ptr, nmc * 4, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(he.data(), err.ptr, nmc * 4, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(hf.data(), n->flags.ptr, nmc, hipMemcpyDeviceToHost));
    int64_t used = 0, withElems = 0;
    float maxErr = 0.f;
    for (int64_t i = 0; i < nmc; ++i) {
      withElems += hc[i] > 0 ? 1 : 0;
      if (hf[i]) {
        used++;
        float e;
        memcpy(&e, &he[i], 4);
        maxErr = std::max(maxErr, e);
      }
    }
    n->bound = bound;
    n->validated = true;
    n->generation++;
    if (out) {
      out[0] = (float)used;
      out[1] = (float)withElems;
      out[2] = maxErr;
    }
    if (r->neural == n) r->updateMaxOpacities();
  });
}
int exa_neural_get_flags(const exa_neural *n, uint8_t *out) {
  return guarded([&] {
    require(n && out && n->validated, "not validated");
    download(n->flags, out);
  });
}
// element-stage lookups of validated macrocells answered by the network
// (NULL: exact sampler everywhere); majorants are rebuilt accordingly
int exa_renderer_set_neural(exa_renderer *r, exa_neural *n) {
  return guarded([&] {
    require(r, "renderer is NULL");
    requireStitch(r);
    require(!n || n->validated, "the neural field must be validated first (exa_neural_validate)");
    r->neural = n;
    r->updateMaxOpacities();
    r->accumID = 0;
  });
}
int exa_renderer_get_state(const exa_renderer *r, exa_render_state *o) {
  return guarded([&] {
    require(r && o, "invalid arguments");
    const RenderParams &P = r->P;
    memset(o, 0, sizeof(*o));
    o->W = P.W;
    o->H = P.H;
    o->spp = r->spp < 1 ? 1 : r->spp;
    o->accumID = r->accumID;
    o->integrator = P.integrator;
    o->shadeMode = P.shadeMode;
    o->maxBounces = P.maxBounces;
    o->debugCrosshair = P.debugCrosshair;
    memcpy(o->camOrg, P.camOrg, 12);
    memcpy(o->camDir00, P.camDir00, 12);
    memcpy(o->camDu, P.camDu, 12);
    memcpy(o->camDv, P.camDv, 12);
    memcpy(o->worldBounds, P.worldBounds, 24);
    memcpy(o->voxelXfm, P.voxelXfm, 48);
    memcpy(o->voxelXfmInv, P.voxelXfmInv, 48);
    memcpy(o->lightXfm, P.lightXfm, 48);
    memcpy(o->gridDims, r->sampler.dims, 12);
    memcpy(o->gridBounds, r->sampler.bounds, 24);
    memcpy(o->tfDomain, P.tfDomain, 8);
    o->opacityScale = P.opacityScale;
    o->tfN = (int32_t)r->tf.count;
    o->clipEnabled = P.clipEnabled;
    memcpy(o->clipN, P.clipN, 12);
    o->clipD = P.clipD;
    for (int i = 0; i < 4; ++i) {
      memcpy(o->lightPos[i], P.lights[i].pos, 12);
      o->lightIntensity[i] = P.lights[i].intensity;
      o->lightOn[i] = P.lights[i].on;
      o->lightType[i] = P.lights[i].type;
    }
    o->dt = P.dt;
    o->heatMapEnabled = P.heatMapEnabled;
    o->heatMapScale = P.heatMapScale;
    o->subImageActive = P.subImageActive;
    memcpy(o->subImage, P.subImage, 16);
    o->subImageSelecting = P.subImageSelecting;
    memcpy(o->subImageSel, P.subImageSel, 16);
  });
}
void exa_renderer_destroy(exa_renderer *r) { delete r; }

// ---- helpers ---------------------------------------------------------------
int exa_camera_look_at(const float from[3], const float at[3], const float up[3], float fovyDeg, int W, int H,
                       float org[3], float dir00[3], float du[3], float dv[3]) {
  return guarded([&] {
    require(from && at && up && org && dir00 && du && dv && W > 0 && H > 0, "invalid arguments");
    auto sub = [](const float *a, const float *b, float *o) { for (int k = 0; k < 3; ++k) o[k] = a[k] - b[k]; };
    auto cross3 = [](const float *a, const float *b, float *o) {
      o[0] = a[1] * b[2] - a[2] * b[1];
      o[1] = a[2] * b[0] - a[0] * b[2];
      o[2] = a[0] * b[1] - a[1] * b[0];
    };
    auto norm = [](float *v) {
      const float l = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      require(l > 0.f, "degenerate camera frame");
      for (int k = 0; k < 3; ++k) v[k] /= l;
    };
    // camera frame: vz points from the point of interest to the eye
    float vz[3], vx[3], vy[3], d[3];
    sub(from, at, vz);
    const float focal = sqrtf(vz[0] * vz[0] + vz[1] * vz[1] + vz[2] * vz[2]);
    norm(vz);
    cross3(up, vz, vx);
    norm(vx);
    cross3(vz, vx, vy);
    norm(vy);
    (void)d;
    // viewer.cpp:79-107 SimpleCamera
    auto eps = [](float f) { return fabsf(f) * (float)(1. / (1 << 21)); };
    const float minFocal = fmaxf(fmaxf(fmaxf(eps(from[0]), eps(from[1])), eps(from[2])),
                                 fmaxf(fmaxf(eps(vx[0]), eps(vx[1])), eps(vx[2])));
    const float fd = fmaxf(minFocal, focal);
    const float aspect = W / (float)H;
    const float screen_height = 2.f * tanf(fovyDeg / 2.f * (float)M_PI / 180.f) * fd;
    float vert[3], horiz[3];
    for (int k = 0; k < 3; ++k) {
      vert[k] = screen_height * vy[k];
      horiz[k] = screen_height * aspect * vx[k];
    }
    for (int k = 0; k < 3; ++k) {
      org[k] = from[k];
      dir00[k] = -fd * vz[k] - 0.5f * vert[k] - 0.5f * horiz[k];
      // viewer.cpp:550-560 cameraChanged
      du[k] = horiz[k] / (float)W;
      dv[k] = vert[k] / (float)H;
    }
  });
}

int exa_xf_load(const char *path, float *opacityScale, float absDomain[2], float relDomain[2], int *numColors,
                float *rgba, int maxColors) {
  return guarded([&] {
    require(path, "path is NULL");
    XfFile xf;
    readXf(path, xf);
    const int n = (int)(xf.colorMap.size() / 4);
    if (opacityScale) *opacityScale = xf.opacityScale;
    if (absDomain) memcpy(absDomain, xf.absDomain, 8);
    if (relDomain) memcpy(relDomain, xf.relDomain, 8);
    if (numColors) *numColors = n;
    if (rgba) {
      require(maxColors >= n, "rgba buffer too small");
      memcpy(rgba, xf.colorMap.data(), xf.colorMap.size() * 4);
    }
  });
}

int exa_tf_rasterize(const float *pts, int np, int N, float *rgba) {
  return guarded([&] {
    require(pts && rgba && np >= 1 && N >= 2, "invalid arguments");
    for (int i = 1; i < np; ++i) require(pts[5 * i] >= pts[5 * (i - 1)], "control points must ascend in x");
    for (int i = 0; i < N; ++i) {
      const float x = (float)i / (float)(N - 1);
      int k = 0;
      while (k + 1 < np && pts[5 * (k + 1)] < x) ++k;
      const float *a = pts + 5 * k;
      const float *b = pts + 5 * (k + 1 < np ? k + 1 : k);
      float t = 0.f;
      if (x <= a[0]) t = 0.f;
      else if (b[0] > a[0]) t = fminf(1.f, (x - a[0]) / (b[0] - a[0]));
      for (int c = 0; c < 4; ++c) rgba[4 * i + c] = a[1 + c] + t * (b[1 + c] - a[1 + c]);
    }
  });
}

// ---- ExaBricks comparison sampler -------------------------------------------
int exa_brick_model_create(const exa_brick *bricks, const float *scalars, size_t numBricks, exa_brick_model **out) {
  return guarded([&] {
    require(bricks && scalars && numBricks > 0 && out, "invalid arguments");
    auto m = std::make_unique<exa_brick_model>();
    brickModelFromArrays((const Brick *)bricks, scalars, numBricks, m->m);
    *out = m.release();
  });
}
int exa_brick_model_load(const char *bricksPath, const char *scalarsPath, exa_brick_model **out) {
  return guarded([&] {
    require(bricksPath && scalarsPath && out, "invalid arguments");
    auto m = std::make_unique<exa_brick_model>();
    loadBrickModel(bricksPath, scalarsPath, m->m);
    *out = m.release();
  });
}
int exa_brick_model_from_amr(const exa_amr *a, int blockCells, exa_brick_model **out) {
  return guarded([&] {
    require(a && out, "invalid arguments");
    std::vector<Brick> bricks;
    std::vector<int32_t> ids;
    bricksFromCells(a->scene.cells.data(), (int64_t)a->scene.cells.size(), blockCells, bricks, ids);
    auto m = std::make_unique<exa_brick_model>();
    brickModelFromCellIDs(bricks, ids, a->scene.scalars.data(), (int64_t)a->scene.scalars.size(), m->m);
    *out = m.release();
  });
}
int exa_amr_write_bricks(const exa_amr *a, int blockCells, const char *bricksPath, const char *scalarsPath) {
  return guarded([&] {
    require(a && bricksPath, "invalid arguments");
    std::vector<Brick> bricks;
    std::vector<int32_t> ids;
    bricksFromCells(a->scene.cells.data(), (int64_t)a->scene.cells.size(), blockCells, bricks, ids);
    writeBrickFile(bricksPath, bricks, ids);
    if (scalarsPath) writeScalars(scalarsPath, a->scene.scalars);
  });
}
int exa_brick_model_get_counts(const exa_brick_model *m, int64_t c[6]) {
  return guarded([&] {
    require(m && c, "invalid arguments");
    c[0] = (int64_t)m->m.bricks.size();
    c[1] = (int64_t)m->m.scalars.size();
    c[2] = (int64_t)m->m.abrs.size();
    c[3] = (int64_t)m->m.leafList.size();
    c[4] = (int64_t)m->m.nodes.size();
    c[5] = m->m.kdDepth;
  });
}
int exa_brick_model_get(const exa_brick_model *m, exa_brick *bricks, float *scalars, exa_abr *abrs,
                        int32_t *leafList) {
  return guarded([&] {
    require(m, "model is NULL");
    const BrickModel &M = m->m;
    if (bricks) memcpy(bricks, M.bricks.data(), M.bricks.size() * sizeof(Brick));
    if (scalars) memcpy(scalars, M.scalars.data(), M.scalars.size() * sizeof(float));
    if (abrs) memcpy(abrs, M.abrs.data(), M.abrs.size() * sizeof(Abr));
    if (leafList) memcpy(leafList, M.leafList.data(), M.leafList.size() * sizeof(int32_t));
  });
}
int exa_brick_model_get_kd_nodes(const exa_brick_model *m, int32_t *nodes) {
  return guarded([&] {
    require(m && nodes, "invalid arguments");
    memcpy(nodes, m->m.nodes.data(), m->m.nodes.size() * sizeof(KdNode));
  });
}
int exa_brick_model_get_bounds(const exa_brick_model *m, float cb[6], float vr[2]) {
  return guarded([&] {
    require(m && cb && vr, "invalid arguments");
    memcpy(cb, m->m.cellBounds, 24);
    memcpy(vr, m->m.valueRange, 8);
  });
}
void exa_brick_model_destroy(exa_brick_model *m) { delete m; }

int exa_brick_sampler_create(const exa_brick_model *m, void *stream, exa_brick_sampler **out) {
  return guarded([&] {
    require(m && out, "invalid arguments");
    auto s = std::make_unique<exa_brick_sampler>();
    s->build(m->m, (hipStream_t)stream);
    *out = s.release();
  });
}
int exa_brick_sampler_compute_max_opacities(exa_brick_sampler *s, const float *rgba, int numColors, float lo,
                                            float hi, void *stream) {
  return guarded([&] {
    require(s, "sampler is NULL");
    s->computeMaxOpacities(rgba, numColors, lo, hi, (hipStream_t)stream);
  });
}
int exa_brick_sampler_get_max_opacities(const exa_brick_sampler *s, float *out) {
  return guarded([&] {
    require(s && out, "invalid arguments");
    require(s->haveMaxOp, "max opacities not computed");
    download(s->maxOp, out);
  });
}
int exa_brick_sampler_locate(const exa_brick_sampler *s, const float *d_pos, int64_t n, exa_sample *d_out,
                             void *stream) {
  return guarded([&] {
    require(s && n >= 0 && (n == 0 || (d_pos && d_out)), "invalid arguments");
    launchBrickLocate(s->view(), d_pos, n, (Sample *)d_out, (hipStream_t)stream);
    HIP_CHECK(hipGetLastError());
  });
}
void exa_brick_sampler_destroy(exa_brick_sampler *s) { delete s; }
const exa_brick_sampler *exa_renderer_get_brick_sampler(const exa_renderer *r) {
  return r && r->bricks ? r->bricks.get() : nullptr;
}

}  // extern "C"
