// model.cpp -- in-memory ExaStitchModel (model/ExaStitchModel.cpp:24-257)
// and the point-location BVH builder (replaces OptiX owlGroupBuildAccel over
// the user-geometry bounds programs in
// sampler/programs/ExaStitchSampler.cu:29-129).
#include <math.h>
#include <string.h>
#include <algorithm>
#include <atomic>
#include <thread>
#include <numeric>
#include <stdexcept>
#include <string>
#include "exa_host.h"

namespace exa {

void finalizeModel(HostModel &m) {
  if (m.vertices.size() % 4) throw std::runtime_error("model: vertex array is not float4");
  if (m.indices.size() % 8) throw std::runtime_error("model: index array is not int[8] rows");
  const int64_t numVerts = (int64_t)m.vertices.size() / 4;
  const int64_t numElems = m.numElements();
  for (int64_t e = 0; e < numElems; ++e) {
    const int32_t *r = &m.indices[8 * e];
    int k = 0;
    while (k < 8 && r[k] >= 0) ++k;
    for (int j = k; j < 8; ++j)
      if (r[j] != -1)
        throw std::runtime_error("model: element " + std::to_string(e) +
                                 " has -1 padding before a valid index");
    if (!(k == 4 || k == 5 || k == 6 || k == 8))
      throw std::runtime_error("model: element " + std::to_string(e) + " has " + std::to_string(k) +
                               " vertices (expected 4, 5, 6 or 8)");
    for (int j = 0; j < k; ++j)
      if (r[j] >= numVerts)
        throw std::runtime_error("model: element " + std::to_string(e) + " index out of range");
  }
  const int64_t numScalars = (int64_t)m.gridletScalars.size();
  for (size_t g = 0; g < m.gridlets.size(); ++g) {
    const Gridlet &G = m.gridlets[g];
    if (G.level < 0 || G.level > 20 || G.dims[0] < 1 || G.dims[1] < 1 || G.dims[2] < 1 ||
        G.dims[0] > 1024 || G.dims[1] > 1024 || G.dims[2] > 1024)
      throw std::runtime_error("model: gridlet " + std::to_string(g) + " has invalid level/dims");
    const int64_t ns = (int64_t)(G.dims[0] + 1) * (G.dims[1] + 1) * (G.dims[2] + 1);
    if ((int64_t)G.begin + ns > numScalars)
      throw std::runtime_error("model: gridlet " + std::to_string(g) + " scalars out of range");
  }
  // cellBounds / valueRange as ExaStitchModel::load builds them
  // (ExaStitchModel.cpp:78-99 elements, 224-245 gridlets); std::min(acc, v)
  // and std::max(acc, v) leave the accumulator unchanged for a NaN v.
  float cb[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  float vr[2] = {1e30f, -1e30f};
  for (int64_t e = 0; e < numElems; ++e)
    for (int j = 0; j < 8; ++j) {
      const int32_t idx = m.indices[8 * e + j];
      if (idx < 0) continue;
      const float *v = &m.vertices[4 * (int64_t)idx];
      for (int k = 0; k < 3; ++k) {
        cb[k] = fminf(cb[k], v[k]);
        cb[3 + k] = fmaxf(cb[3 + k], v[k]);
      }
      vr[0] = (v[3] < vr[0]) ? v[3] : vr[0];
      vr[1] = (vr[1] < v[3]) ? v[3] : vr[1];
    }
  for (const Gridlet &G : m.gridlets) {
    const float cw = (float)(1 << G.level);
    for (int k = 0; k < 3; ++k) {
      const float lo = ((float)G.lower[k] + 0.5f) * cw;
      const float hi = ((float)G.lower[k] + (float)G.dims[k] + 0.5f) * cw;
      cb[k] = fminf(fminf(cb[k], lo), hi);
      cb[3 + k] = fmaxf(fmaxf(cb[3 + k], lo), hi);
    }
    const int64_t ns = (int64_t)(G.dims[0] + 1) * (G.dims[1] + 1) * (G.dims[2] + 1);
    for (int64_t i = 0; i < ns; ++i) {
      const float v = m.gridletScalars[G.begin + i];
      vr[0] = (v < vr[0]) ? v : vr[0];
      vr[1] = (vr[1] < v) ? v : vr[1];
    }
  }
  memcpy(m.cellBounds, cb, sizeof(cb));
  memcpy(m.valueRange, vr, sizeof(vr));
  static const float ident[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
  memcpy(m.voxelXfm, ident, sizeof(ident));
}

void gridletBoxes(const std::vector<Gridlet> &gl, std::vector<float> &boxes) {
  boxes.resize(6 * gl.size());
  for (size_t g = 0; g < gl.size(); ++g) {
    const Gridlet &G = gl[g];
    const float cw = (float)(1 << G.level);
    for (int k = 0; k < 3; ++k) {
      boxes[6 * g + k] = ((float)G.lower[k] + 0.5f) * cw;
      boxes[6 * g + 3 + k] = ((float)G.lower[k] + (float)G.dims[k] + 0.5f) * cw;
    }
  }
}

// StitchGeomBounds_Impl (programs/ExaStitchSampler.cu:73-94)
void elementBoxes(const std::vector<int32_t> &indices, const std::vector<float> &vertices,
                  std::vector<float> &boxes) {
  const int64_t n = (int64_t)indices.size() / 8;
  boxes.resize(6 * n);
  for (int64_t e = 0; e < n; ++e) {
    float *b = &boxes[6 * e];
    b[0] = b[1] = b[2] = INFINITY;
    b[3] = b[4] = b[5] = -INFINITY;
    for (int j = 0; j < 8; ++j) {
      const int32_t idx = indices[8 * e + j];
      if (idx < 0) break;
      for (int k = 0; k < 3; ++k) {
        b[k] = fminf(b[k], vertices[4 * (int64_t)idx + k]);
        b[3 + k] = fmaxf(b[3 + k], vertices[4 * (int64_t)idx + k]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// binned-SAH BVH over primitive boxes; depth capped at kBvhMaxDepth so the
// device traversal stack (one entry per level) can never overflow.
// ---------------------------------------------------------------------------
namespace {

struct Builder {
  const float *boxes;
  int leafSize;
  std::vector<int32_t> ids;
  std::vector<float> cent;  // 3 per prim
  std::atomic<int> maxDepth{0};

  void boxOf(int64_t b, int64_t e, float lo[3], float hi[3]) const {
    for (int k = 0; k < 3; ++k) { lo[k] = INFINITY; hi[k] = -INFINITY; }
    for (int64_t i = b; i < e; ++i) {
      const float *bx = boxes + 6 * (int64_t)ids[i];
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], bx[k]);
        hi[k] = std::max(hi[k], bx[3 + k]);
      }
    }
  }
  static float area(const float lo[3], const float hi[3]) {
    const float d0 = std::max(0.f, hi[0] - lo[0]), d1 = std::max(0.f, hi[1] - lo[1]),
                d2 = std::max(0.f, hi[2] - lo[2]);
    return 2.f * (d0 * d1 + d1 * d2 + d2 * d0);
  }
  static int levelsNeeded(int64_t count, int leafSize) {
    int l = 0;
    int64_t c = (count + leafSize - 1) / leafSize;
    while (c > 1) { c = (c + 1) / 2; ++l; }
    return l;
  }
  void noteDepth(int d) {
    int cur = maxDepth.load();
    while (d > cur && !maxDepth.compare_exchange_weak(cur, d)) {
    }
  }

  // Splits ids[b,e) (binned SAH over 16 bins on each axis; object median when
  // SAH finds no split or the depth budget requires halving); returns mid.
  int64_t split(int64_t b, int64_t e, int depth) {
    const int64_t n = e - b;
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = b; i < e; ++i)
      for (int k = 0; k < 3; ++k) {
        clo[k] = std::min(clo[k], cent[3 * (int64_t)ids[i] + k]);
        chi[k] = std::max(chi[k], cent[3 * (int64_t)ids[i] + k]);
      }
    int64_t mid = -1;
    const bool forceMedian = depth + levelsNeeded(n, leafSize) >= kBvhMaxDepth - 1;
    if (!forceMedian) {
      constexpr int NB = 16;
      float bestCost = INFINITY;
      int bestAxis = -1, bestSplit = -1;
      for (int axis = 0; axis < 3; ++axis) {
        const float ext = chi[axis] - clo[axis];
        if (!(ext > 0.f)) continue;
        int64_t bc[NB] = {0};
        float blo[NB][3], bhi[NB][3];
        for (int i = 0; i < NB; ++i)
          for (int k = 0; k < 3; ++k) { blo[i][k] = INFINITY; bhi[i][k] = -INFINITY; }
        for (int64_t i = b; i < e; ++i) {
          const int64_t p = ids[i];
          int bi = (int)((cent[3 * p + axis] - clo[axis]) / ext * NB);
          bi = std::min(NB - 1, std::max(0, bi));
          bc[bi]++;
          for (int k = 0; k < 3; ++k) {
            blo[bi][k] = std::min(blo[bi][k], boxes[6 * p + k]);
            bhi[bi][k] = std::max(bhi[bi][k], boxes[6 * p + 3 + k]);
          }
        }
        float rlo[NB][3], rhi[NB][3];
        int64_t rc[NB];
        float alo[3] = {INFINITY, INFINITY, INFINITY}, ahi[3] = {-INFINITY, -INFINITY, -INFINITY};
        int64_t ac = 0;
        for (int i = NB - 1; i >= 0; --i) {
          for (int k = 0; k < 3; ++k) {
            alo[k] = std::min(alo[k], blo[i][k]);
            ahi[k] = std::max(ahi[k], bhi[i][k]);
            rlo[i][k] = alo[k];
            rhi[i][k] = ahi[k];
          }
          ac += bc[i];
          rc[i] = ac;
        }
        for (int k = 0; k < 3; ++k) { alo[k] = INFINITY; ahi[k] = -INFINITY; }
        ac = 0;
        for (int s = 0; s < NB - 1; ++s) {
          for (int k = 0; k < 3; ++k) {
            alo[k] = std::min(alo[k], blo[s][k]);
            ahi[k] = std::max(ahi[k], bhi[s][k]);
          }
          ac += bc[s];
          if (ac == 0 || rc[s + 1] == 0) continue;
          const float cost = area(alo, ahi) * (float)ac + area(rlo[s + 1], rhi[s + 1]) * (float)rc[s + 1];
          if (cost < bestCost) {
            bestCost = cost;
            bestAxis = axis;
            bestSplit = s;
          }
        }
      }
      if (bestAxis >= 0) {
        const float ext = chi[bestAxis] - clo[bestAxis];
        auto it = std::partition(ids.begin() + b, ids.begin() + e, [&](int32_t p) {
          int bi = (int)((cent[3 * (int64_t)p + bestAxis] - clo[bestAxis]) / ext * NB);
          bi = std::min(NB - 1, std::max(0, bi));
          return bi <= bestSplit;
        });
        mid = it - ids.begin();
        if (mid == b || mid == e) mid = -1;
      }
    }
    if (mid < 0) {
      int axis = 0;
      for (int k = 1; k < 3; ++k)
        if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
      mid = b + n / 2;
      std::nth_element(ids.begin() + b, ids.begin() + mid, ids.begin() + e, [&](int32_t p, int32_t q) {
        const float cp = cent[3 * (int64_t)p + axis], cq = cent[3 * (int64_t)q + axis];
        return cp < cq || (cp == cq && p < q);
      });
    }
    return mid;
  }

  // Builds the subtree over ids[b,e) into `nodes` (indices local to that
  // vector); returns the child reference (inner node index, or first prim of
  // a leaf) and the count (0 for an inner node).  Large subtrees near the top
  // are built on their own threads and appended with their indices shifted.
  void build(int64_t b, int64_t e, int depth, std::vector<BvhNode> &nodes, int32_t &ref, int32_t &cnt) {
    const int64_t n = e - b;
    noteDepth(depth);
    if (n <= leafSize) {
      std::sort(ids.begin() + b, ids.begin() + e);
      ref = (int32_t)b;
      cnt = (int32_t)n;
      return;
    }
    const int64_t mid = split(b, e, depth);
    const int32_t nodeIdx = (int32_t)nodes.size();
    nodes.emplace_back();
    float lo0[3], hi0[3], lo1[3], hi1[3];
    boxOf(b, mid, lo0, hi0);
    boxOf(mid, e, lo1, hi1);
    int32_t r0, c0, r1, c1;
    if (n > (1 << 16) && depth < 6) {
      std::vector<BvhNode> left, right;
      std::thread t([&] { build(b, mid, depth + 1, left, r0, c0); });
      build(mid, e, depth + 1, right, r1, c1);
      t.join();
      auto append = [&](std::vector<BvhNode> &sub, int32_t &r, int32_t c) {
        const int32_t off = (int32_t)nodes.size();
        for (BvhNode &nd : sub) {
          if (nd.cnt0 == 0) nd.ref0 += off;
          if (nd.cnt1 == 0) nd.ref1 += off;
        }
        nodes.insert(nodes.end(), sub.begin(), sub.end());
        if (c == 0) r += off;
      };
      append(left, r0, c0);
      append(right, r1, c1);
    } else {
      build(b, mid, depth + 1, nodes, r0, c0);
      build(mid, e, depth + 1, nodes, r1, c1);
    }
    BvhNode &nd = nodes[nodeIdx];
    for (int k = 0; k < 3; ++k) {
      nd.lo0[k] = lo0[k]; nd.hi0[k] = hi0[k];
      nd.lo1[k] = lo1[k]; nd.hi1[k] = hi1[k];
    }
    nd.ref0 = r0; nd.cnt0 = c0;
    nd.ref1 = r1; nd.cnt1 = c1;
    ref = nodeIdx;
    cnt = 0;
  }
};

}  // namespace

void buildBvh(const float *boxes, int64_t n, int leafSize, Bvh &out) {
  out.nodes.clear();
  out.primRefs.clear();
  out.depth = 0;
  if (n <= 0) return;
  if (n >= ((int64_t)1 << 31)) throw std::runtime_error("buildBvh: too many primitives");
  Builder B;
  B.boxes = boxes;
  B.leafSize = std::max(1, std::min(leafSize, 16));
  B.ids.resize(n);
  std::iota(B.ids.begin(), B.ids.end(), 0);
  B.cent.resize(3 * n);
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) B.cent[3 * i + k] = 0.5f * (boxes[6 * i + k] + boxes[6 * i + 3 + k]);
  // the first inner node built is the root (index 0); a single-leaf tree
  // gets a root whose child 0 is that leaf and whose child 1 is empty
  int32_t r, c;
  B.build(0, n, 1, out.nodes, r, c);
  out.depth = B.maxDepth.load();
  if (c > 0) {
    float lo[3], hi[3];
    B.boxOf(0, n, lo, hi);
    out.nodes.emplace_back();
    BvhNode &root = out.nodes[0];
    for (int k = 0; k < 3; ++k) {
      root.lo0[k] = lo[k]; root.hi0[k] = hi[k];
      root.lo1[k] = INFINITY; root.hi1[k] = -INFINITY;
    }
    root.ref0 = r; root.cnt0 = c;
    root.ref1 = -1; root.cnt1 = -1;
  }
  if (out.depth >= kBvhMaxDepth) throw std::runtime_error("buildBvh: depth bound violated");
  out.primRefs = B.ids;
}

}  // namespace exa
