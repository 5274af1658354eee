// locate_tables.cpp -- host build of the direct point-location tables (the
// default replacement for the OptiX sampling BVH of
// sampler/ExaStitchSampler.cpp:71-323).  See LocateTables in exa_host.h.
#include <math.h>
#include <string.h>
#include <algorithm>
#include <atomic>
#include <stdexcept>
#include <thread>
#include "exa_host.h"

namespace exa {

// budgets of the element bins (coarsened past any of them): 2^29 bins (2 GB
// of list offsets), at most 32 bins per element (at least 2^20), so that a
// few small elements in a large domain do not allocate mostly empty bins,
// and 2^30 list entries (12 GB)
static constexpr int64_t kMaxBins = 1ll << 29;
static constexpr int64_t kBinsPerElement = 32;
static constexpr int64_t kMaxCandidates = 1ll << 30;

// bin index of coordinate x: the same float expression runs on the device
static inline int32_t binOf(float x, float lo, float inv, int32_t dims) {
  const float f = floorf((x - lo) * inv);
  int32_t i = f != f ? 0 : (f < 0.f ? 0 : (f > (float)(dims - 1) ? dims - 1 : (int32_t)f));
  return i;
}

// Sub-cell 0..511 of coordinate x inside bin b of its axis (512 sub-cells
// per bin, clamped).  The same float expression runs on the device.  It is
// monotone in x, so lo <= x <= hi gives subCellOf(lo) <= subCellOf(x) <=
// subCellOf(hi) in every bin: a box quantised this way never rejects a point
// the float box holds, and a point whose sub-cell lies strictly between the
// quantised bounds lies strictly inside the float box.
static inline uint32_t subCellOf(float x, float lo, float inv512, int32_t b) {
  const float g = floorf((x - lo) * inv512) - (float)b * 512.f;
  return g != g ? 0u : (g < 0.f ? 0u : (g > 511.f ? 511u : (uint32_t)g));
}

// Dual-cell owner tables.  Gridlet IDs ascend with level, so per level the
// owner of a dual cell is the first gridlet (in ID order) that has a valid
// cell there; blocks are allocated for macrocells that hold any valid cell.
static void buildOwnerTables(const HostModel &m, LocateTables &t) {
  if (t.fastMargin > 0.0625f) return;
  const int64_t numMC = t.levelOffset[t.numLevels];
  const float *S = m.gridletScalars.data();
  auto valid = [&](const Gridlet &g, int x, int y, int z) {
    const int sx = g.dims[0] + 1, sy = g.dims[1] + 1;
    const float *s = S + g.begin;
    for (int dz = 0; dz < 2; ++dz)
      for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
          const float f = s[((int64_t)(z + dz) * sy + (y + dy)) * sx + (x + dx)];
          if (f != f) return false;
        }
    return true;
  };
  auto mcOf = [&](const Gridlet &g, int cx, int cy, int cz) {
    const int32_t *D = t.levelDims[g.level];
    return t.levelOffset[g.level] + ((int64_t)(cz >> 3) * D[1] + (cy >> 3)) * D[0] + (cx >> 3);
  };
  t.gTop.assign(numMC, -1);
  int64_t numBlocks = 0;
  for (const Gridlet &g : m.gridlets)
    for (int z = 0; z < g.dims[2]; ++z)
      for (int y = 0; y < g.dims[1]; ++y)
        for (int x = 0; x < g.dims[0]; ++x) {
          const int64_t mc = mcOf(g, g.lower[0] + x, g.lower[1] + y, g.lower[2] + z);
          if (t.gTop[mc] < 0 && valid(g, x, y, z)) t.gTop[mc] = (int32_t)numBlocks++;
        }
  if (numBlocks * LocateTables::kOwnerBlock >= (1ll << 31)) {
    t.gTop.clear();
    return;
  }
  t.gOwner.assign(numBlocks * LocateTables::kOwnerBlock, -1);
  for (size_t gi = 0; gi < m.gridlets.size(); ++gi) {
    const Gridlet &g = m.gridlets[gi];
    for (int z = 0; z < g.dims[2]; ++z)
      for (int y = 0; y < g.dims[1]; ++y)
        for (int x = 0; x < g.dims[0]; ++x) {
          const int cx = g.lower[0] + x, cy = g.lower[1] + y, cz = g.lower[2] + z;
          const int32_t blk = t.gTop[mcOf(g, cx, cy, cz)];
          if (blk < 0) continue;
          int32_t &o = t.gOwner[(int64_t)blk * LocateTables::kOwnerBlock + ((cz & 7) * 8 + (cy & 7)) * 8 + (cx & 7)];
          if (o < 0 && valid(g, x, y, z)) o = (int32_t)gi;
        }
  }
}

void gridletMargins(const HostModel &m, float &bandMargin, float &fastMargin) {
  // the largest coordinate magnitude: cell bounds cover the gridlet boxes and
  // the element vertices
  float maxAbs = 1.f;
  for (int k = 0; k < 6; ++k) maxAbs = std::max(maxAbs, fabsf(m.cellBounds[k]));
  int e = 1;
  while (e < 120 && ldexpf(1.f, e) <= maxAbs) ++e;
  bandMargin = ldexpf(1.f, e - 21);
  fastMargin = ldexpf(1.f, e - 20);
}

void buildLocateTables(const HostModel &m, LocateTables &t) {
  t = LocateTables();
  // ---- gridlets: per-level macrocell lists --------------------------------
  int maxLevel = -1;
  for (const Gridlet &g : m.gridlets) {
    for (int k = 0; k < 3; ++k)
      if (g.lower[k] < 0) throw std::runtime_error("locate tables need gridlets with lower >= 0");
    maxLevel = std::max(maxLevel, g.level);
  }
  if (maxLevel >= LocateTables::kMaxLevels) throw std::runtime_error("locate tables: too many levels");
  t.numLevels = maxLevel + 1;
  for (const Gridlet &g : m.gridlets)
    for (int k = 0; k < 3; ++k)
      t.levelDims[g.level][k] = std::max(t.levelDims[g.level][k], ((g.lower[k] + g.dims[k]) >> 3) + 1);
  t.levelOffset[0] = 0;
  for (int L = 0; L < t.numLevels; ++L)
    t.levelOffset[L + 1] =
        t.levelOffset[L] + (int64_t)t.levelDims[L][0] * t.levelDims[L][1] * t.levelDims[L][2];
  const int64_t numMC = t.levelOffset[t.numLevels];
  if (numMC >= (1ll << 30)) throw std::runtime_error("locate tables: macrocell tables too large");
  std::vector<uint32_t> cnt(numMC + 1, 0);
  auto forEachMC = [&](const Gridlet &g, auto &&f) {
    const int32_t *D = t.levelDims[g.level];
    const int64_t base = t.levelOffset[g.level];
    for (int z = g.lower[2] >> 3; z <= (g.lower[2] + g.dims[2]) >> 3; ++z)
      for (int y = g.lower[1] >> 3; y <= (g.lower[1] + g.dims[1]) >> 3; ++y)
        for (int x = g.lower[0] >> 3; x <= (g.lower[0] + g.dims[0]) >> 3; ++x)
          f(base + ((int64_t)z * D[1] + y) * D[0] + x);
  };
  for (const Gridlet &g : m.gridlets) forEachMC(g, [&](int64_t i) { cnt[i + 1]++; });
  for (int64_t i = 0; i < numMC; ++i) cnt[i + 1] += cnt[i];
  t.gStart = cnt;
  t.gList.resize(cnt[numMC]);
  {
    std::vector<uint32_t> cur(cnt.begin(), cnt.end() - 1);
    for (size_t g = 0; g < m.gridlets.size(); ++g)
      forEachMC(m.gridlets[g], [&](int64_t i) { t.gList[cur[i]++] = (int32_t)g; });
  }
  gridletMargins(m, t.bandMargin, t.fastMargin);
  buildOwnerTables(m, t);
  // ---- elements: uniform bins -----------------------------------------------
  const int64_t ne = m.numElements();
  if (ne == 0) return;
  std::vector<float> boxes;
  elementBoxes(m.indices, m.vertices, boxes);
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  std::vector<float> ext(ne);
  for (int64_t e = 0; e < ne; ++e) {
    const float *b = &boxes[6 * e];
    float mx = 0.f;
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], b[k]);
      hi[k] = std::max(hi[k], b[3 + k]);
      mx = std::max(mx, b[3 + k] - b[k]);
    }
    ext[e] = mx;
  }
  std::nth_element(ext.begin(), ext.begin() + ne / 2, ext.end());
  // Bin size: the power of two at or below half the median element extent
  // (a few AABB candidates per query), doubled while the bins or their list
  // entries would outgrow their budget.
  float size = 1.f;
  while (size * 2.f <= ext[ne / 2] && size < 1024.f) size *= 2.f;
  size *= 0.5f;
  while (size < 1e-3f * std::max(std::max(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2])) size *= 2.f;
  for (;;) {
    for (int k = 0; k < 3; ++k) t.binDims[k] = (int32_t)floorf((hi[k] - lo[k]) * (1.f / size)) + 1;
    if ((int64_t)t.binDims[0] * t.binDims[1] * t.binDims[2] >
        std::min<int64_t>(kMaxBins, std::max<int64_t>(1ll << 20, kBinsPerElement * ne))) {
      size *= 2.f;
      continue;
    }
    int64_t entries = 0;
    for (int64_t e = 0; e < ne; ++e) {
      int64_t n = 1;
      for (int k = 0; k < 3; ++k)
        n *= binOf(boxes[6 * e + 3 + k], lo[k], 1.f / size, t.binDims[k]) - binOf(boxes[6 * e + k], lo[k], 1.f / size, t.binDims[k]) + 1;
      entries += n;
    }
    if (entries > kMaxCandidates && size < 1024.f) {
      size *= 2.f;
      continue;
    }
    break;
  }
  for (int k = 0; k < 3; ++k) t.binLo[k] = lo[k];
  t.binInv = 1.f / size;
  const int64_t nb = (int64_t)t.binDims[0] * t.binDims[1] * t.binDims[2];
  std::vector<std::atomic<uint32_t>> bc(nb);
  for (auto &c : bc) c.store(0, std::memory_order_relaxed);
  auto range = [&](int64_t e, int32_t a[3], int32_t b[3]) {
    const float *bx = &boxes[6 * e];
    for (int k = 0; k < 3; ++k) {
      a[k] = binOf(bx[k], t.binLo[k], t.binInv, t.binDims[k]);
      b[k] = binOf(bx[3 + k], t.binLo[k], t.binInv, t.binDims[k]);
    }
  };
  const int T = hostThreads();
  {
    std::vector<std::thread> pool;
    for (int th = 0; th < T; ++th)
      pool.emplace_back([&, th] {
        for (int64_t e = th; e < ne; e += T) {
          int32_t a[3], b[3];
          range(e, a, b);
          for (int z = a[2]; z <= b[2]; ++z)
            for (int y = a[1]; y <= b[1]; ++y)
              for (int x = a[0]; x <= b[0]; ++x)
                bc[((int64_t)z * t.binDims[1] + y) * t.binDims[0] + x].fetch_add(1, std::memory_order_relaxed);
        }
      });
    for (auto &p : pool) p.join();
  }
  t.eStart.assign(nb + 1, 0);
  for (int64_t i = 0; i < nb; ++i) t.eStart[i + 1] = t.eStart[i] + bc[i].load(std::memory_order_relaxed);
  if (t.eStart[nb] >= (1u << 31)) throw std::runtime_error("locate tables: element bin lists too large");
  if (ne >= (1ll << 28)) throw std::runtime_error("locate tables: over 2^28 stitching elements");
  // List entries, 3 words: id | vertex count << 28, then the box quantised
  // to the entry's bin, 9 bits per bound in 10-bit slots (x, y, z at bits 0,
  // 10, 20): the lo word with every slot's top bit clear, the hi word with
  // every slot's top bit set.
  t.eCand.resize(3 * (size_t)t.eStart[nb]);
  std::vector<uint32_t> cur(t.eStart.begin(), t.eStart.end() - 1);
  const float inv512 = t.binInv * 512.f;
  for (int64_t e = 0; e < ne; ++e) {
    int32_t a[3], b[3];
    range(e, a, b);
    const int32_t *row = &m.indices[8 * e];
    uint32_t nv = 0;
    while (nv < 8 && row[nv] >= 0) nv++;
    const uint32_t head = (uint32_t)e | nv << 28;
    const float *bx = &boxes[6 * e];
    for (int z = a[2]; z <= b[2]; ++z) {
      const uint32_t lz = subCellOf(bx[2], t.binLo[2], inv512, z) << 20;
      const uint32_t hz = subCellOf(bx[5], t.binLo[2], inv512, z) << 20 | 0x20080200u;
      for (int y = a[1]; y <= b[1]; ++y) {
        const uint32_t ly = lz | subCellOf(bx[1], t.binLo[1], inv512, y) << 10;
        const uint32_t hy = hz | subCellOf(bx[4], t.binLo[1], inv512, y) << 10;
        for (int x = a[0]; x <= b[0]; ++x) {
          const size_t k = cur[((int64_t)z * t.binDims[1] + y) * t.binDims[0] + x]++;
          t.eCand[3 * k] = head;
          t.eCand[3 * k + 1] = ly | subCellOf(bx[0], t.binLo[0], inv512, x);
          t.eCand[3 * k + 2] = hy | subCellOf(bx[3], t.binLo[0], inv512, x);
        }
      }
    }
  }
  t.eVerts.assign(32 * (size_t)ne, 0.f);
  for (int64_t e = 0; e < ne; ++e)
    for (int i = 0; i < 8; ++i) {
      const int32_t v = m.indices[8 * e + i];
      if (v < 0) break;
      memcpy(&t.eVerts[32 * (size_t)e + 4 * i], &m.vertices[4 * (size_t)v], 16);
    }
}

}  // namespace exa
