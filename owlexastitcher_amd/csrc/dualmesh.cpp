// dualmesh.cpp -- multi-threaded dual-mesh and gridlet generation, the
// in-library replacement for the reference's offline tools
// (tools/amrMakeDualMesh/makeDual.cpp:596-890 and makeGrids.cpp:46-246).
//
// Cells are visited in the reference's sorted order (makeDual.cpp:878:
// LogicalCell compared as two little-endian uint64 words, i.e. by y, x,
// level, z); the work is split into contiguous chunks whose outputs are
// concatenated in chunk order, so the result equals a serial run:
//   - elements grouped tets, pyramids, wedges, hexes
//     (model/ExaStitchModel.cpp:105-108), each in emission order;
//   - vertices = element-referenced cells in ascending sorted-cell order;
//   - gridlets by level, then macrocell id (x, y, z), as makeGrids.cpp's
//     std::map iterates.
// Cell lookup (Exa::find, makeDual.cpp:172-195) uses a hash table keyed on
// (level, position) instead of binary search in the sorted list.
#include <math.h>
#include <string.h>
#include <algorithm>
#include <atomic>
#include <memory>
#include <stdexcept>
#include <thread>
#include "exa_host.h"

namespace exa {

template <typename F>
static void parallelChunks(int64_t numChunks, F &&f) {
  const int T = std::min<int64_t>(hostThreads(), std::max<int64_t>(numChunks, 1));
  std::atomic<int64_t> next{0};
  auto work = [&] {
    for (;;) {
      const int64_t c = next.fetch_add(1);
      if (c >= numChunks) break;
      f(c);
    }
  };
  if (T <= 1) {
    work();
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < T; ++t) pool.emplace_back(work);
  for (auto &th : pool) th.join();
}

// LSD radix sort of (key, payload) pairs, 16-bit digits; stable.
void radixSort(std::vector<uint64_t> &keys, std::vector<int64_t> &vals) {
  const size_t n = keys.size();
  std::vector<uint64_t> k2(n);
  std::vector<int64_t> v2(n);
  std::vector<size_t> cnt(65537);
  for (int pass = 0; pass < 4; ++pass) {
    const int sh = 16 * pass;
    std::fill(cnt.begin(), cnt.end(), 0);
    for (size_t i = 0; i < n; ++i) cnt[((keys[i] >> sh) & 0xffff) + 1]++;
    bool trivial = false;
    for (int d = 0; d < 65536; ++d)
      if (cnt[d + 1] == n) trivial = true;
    if (trivial) continue;
    for (int d = 0; d < 65536; ++d) cnt[d + 1] += cnt[d];
    for (size_t i = 0; i < n; ++i) {
      const size_t o = cnt[(keys[i] >> sh) & 0xffff]++;
      k2[o] = keys[i];
      v2[o] = vals[i];
    }
    keys.swap(k2);
    vals.swap(v2);
  }
}

namespace {

constexpr int kCoordBits = 20;
constexpr int32_t kCoordLimit = 1 << kCoordBits;

struct SortedCell {
  int32_t pos[3];
  int32_t level;
  int32_t scalarID;
};

inline uint64_t hashKey(int level, int x, int y, int z) {
  return ((uint64_t)level << 60) | ((uint64_t)x << 40) | ((uint64_t)y << 20) | (uint64_t)z;
}
inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

struct CellHash {
  static constexpr uint64_t kEmpty = ~0ull;
  uint64_t mask = 0;
  std::unique_ptr<std::atomic<uint64_t>[]> keys;
  std::unique_ptr<int64_t[]> vals;
  void init(int64_t n) {
    uint64_t cap = 1024;
    while (cap < (uint64_t)(2 * n)) cap <<= 1;
    mask = cap - 1;
    keys.reset(new std::atomic<uint64_t>[cap]);
    vals.reset(new int64_t[cap]);
    for (uint64_t i = 0; i < cap; ++i) keys[i].store(kEmpty, std::memory_order_relaxed);
  }
  bool insert(uint64_t key, int64_t val) {
    uint64_t h = mix64(key) & mask;
    for (;;) {
      uint64_t expected = kEmpty;
      if (keys[h].compare_exchange_strong(expected, key, std::memory_order_relaxed)) {
        vals[h] = val;
        return true;
      }
      if (expected == key) return false;  // duplicate cell
      h = (h + 1) & mask;
    }
  }
  int64_t find(uint64_t key) const {
    uint64_t h = mix64(key) & mask;
    for (;;) {
      const uint64_t k = keys[h].load(std::memory_order_relaxed);
      if (k == key) return vals[h];
      if (k == kEmpty) return -1;
      h = (h + 1) & mask;
    }
  }
};

struct ChunkOut {
  std::vector<int64_t> elem[4];  // flattened sorted-cell indices per type
  std::vector<CubeRec> cubes;
  int64_t degenerate = 0, errors = 0;
};

struct Exa {
  std::vector<SortedCell> cells;
  CellHash hash;
  int minLevel = 100, maxLevel = 0;

  // makeDual.cpp:165-169
  static int lowerOnLevel(float f, int level) {
    f = floorf(f / (float)(1 << level));
    return (int)(f * (float)(1 << level));
  }
  // makeDual.cpp:172-195 Exa::find, restricted to levels >= fromLevel: a
  // corner found on a finer level makes doCell skip the octant just like a
  // missing corner, so finer levels need not be probed.
  int64_t find(const float where[3], int fromLevel) const {
    for (int level = fromLevel; level <= maxLevel; level++) {
      const int x = lowerOnLevel(where[0], level), y = lowerOnLevel(where[1], level),
                z = lowerOnLevel(where[2], level);
      if (x < 0 || y < 0 || z < 0 || x >= kCoordLimit || y >= kCoordLimit || z >= kCoordLimit)
        continue;
      const int64_t r = hash.find(hashKey(level, x, y, z));
      if (r >= 0) return r;
    }
    return -1;
  }
};

inline bool cellLess(const SortedCell &a, const SortedCell &b) {
  const uint64_t a0 = (uint64_t)(uint32_t)a.pos[0] | ((uint64_t)(uint32_t)a.pos[1] << 32);
  const uint64_t b0 = (uint64_t)(uint32_t)b.pos[0] | ((uint64_t)(uint32_t)b.pos[1] << 32);
  if (a0 != b0) return a0 < b0;
  const uint64_t a1 = (uint64_t)(uint32_t)a.pos[2] | ((uint64_t)(uint32_t)a.level << 32);
  const uint64_t b1 = (uint64_t)(uint32_t)b.pos[2] | ((uint64_t)(uint32_t)b.level << 32);
  if (a1 != b1) return a1 < b1;
  return a.scalarID < b.scalarID;
}

// makeDual.cpp:596-872 doCell, with early outs that cannot change the result:
// an octant is dropped as soon as a corner is missing, finer than the cell,
// or a same-level corner precedes the cell in sorted order (each of these
// makes the reference skip the octant).
void doCell(const Exa &exa, int64_t ci, ChunkOut &out) {
  const SortedCell &cell = exa.cells[ci];
  const int cw = 1 << cell.level;
  for (int dz = -1; dz <= 1; dz += 2)
    for (int dy = -1; dy <= 1; dy += 2)
      for (int dx = -1; dx <= 1; dx += 2) {
        int64_t corner[2][2][2];
        int minLevel = cell.level, maxLevel = cell.level;
        bool skip = false;
        for (int iz = 0; iz < 2 && !skip; iz++)
          for (int iy = 0; iy < 2 && !skip; iy++)
            for (int ix = 0; ix < 2 && !skip; ix++) {
              if (!ix && !iy && !iz) {
                corner[0][0][0] = ci;
                continue;
              }
              const float c[3] = {(float)(cell.pos[0] + dx * ix * cw) + 0.5f * (float)cw,
                                  (float)(cell.pos[1] + dy * iy * cw) + 0.5f * (float)cw,
                                  (float)(cell.pos[2] + dz * iz * cw) + 0.5f * (float)cw};
              const int64_t f = exa.find(c, cell.level);
              if (f < 0) { skip = true; break; }
              const SortedCell &fc = exa.cells[f];
              if (fc.level == cell.level && f < ci) { skip = true; break; }
              corner[iz][iy][ix] = f;
              minLevel = std::min(minLevel, fc.level);
              maxLevel = std::max(maxLevel, fc.level);
            }
        if (skip) continue;
        int64_t(*vx)[2][2] = corner;
        if (dx < 0) {
          std::swap(vx[0][0][0], vx[0][0][1]); std::swap(vx[0][1][0], vx[0][1][1]);
          std::swap(vx[1][0][0], vx[1][0][1]); std::swap(vx[1][1][0], vx[1][1][1]);
        }
        if (dy < 0) {
          std::swap(vx[0][0][0], vx[0][1][0]); std::swap(vx[0][0][1], vx[0][1][1]);
          std::swap(vx[1][0][0], vx[1][1][0]); std::swap(vx[1][0][1], vx[1][1][1]);
        }
        if (dz < 0) {
          std::swap(vx[0][0][0], vx[1][0][0]); std::swap(vx[0][0][1], vx[1][0][1]);
          std::swap(vx[0][1][0], vx[1][1][0]); std::swap(vx[0][1][1], vx[1][1][1]);
        }
        const int64_t v[8] = {vx[0][0][0], vx[0][0][1], vx[0][1][1], vx[0][1][0],
                              vx[1][0][0], vx[1][0][1], vx[1][1][1], vx[1][1][0]};
        if (minLevel == maxLevel) {
          // perfect cube (makeDual.cpp:473-480): v[0] is the lowest corner
          CubeRec r;
          r.level = minLevel;
          const SortedCell &c0 = exa.cells[v[0]];
          for (int k = 0; k < 3; ++k) r.cid[k] = c0.pos[k] >> minLevel;
          for (int i = 0; i < 8; ++i) r.scalarIDs[i] = exa.cells[v[i]].scalarID;
          out.cubes.push_back(r);
          continue;
        }
        int numUnique = 0;
        for (int i = 0; i < 8; ++i) {
          bool dup = false;
          for (int j = 0; j < i; ++j) dup |= (v[j] == v[i]);
          numUnique += !dup;
        }
        auto emit = [&](int type, std::initializer_list<int64_t> ids) {
          for (int64_t id : ids) out.elem[type].push_back(id);
        };
        if (numUnique == 8) {
          emit(3, {v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]});
          continue;
        }
        if (numUnique < 4) {
          out.degenerate++;
          continue;
        }
        // makeDual.cpp:501-547 tryPyramid
        auto tryPyramid = [&](int64_t b0, int64_t b1, int64_t b2, int64_t b3, int64_t top) {
          if (numUnique == 5) {
            emit(1, {b0, b1, b2, b3, top});
          } else if (numUnique == 4) {
            if (b0 == b1) emit(0, {b1, b2, b3, top});
            else if (b1 == b2) emit(0, {b2, b3, b0, top});
            else if (b2 == b3) emit(0, {b3, b0, b1, top});
            else if (b3 == b0) emit(0, {b0, b1, b2, top});
            else if (b0 == b2 || b1 == b3) out.degenerate++;
            else out.errors++;
          } else {
            out.errors++;
          }
        };
        auto same = [&](int a, int b) { return v[a] == v[b]; };
        auto allSame = [&](int a, int b, int c, int d) { return same(a, b) && same(a, c) && same(a, d); };
        auto wedge = [&](int f0, int f1, int f2, int b0, int b1, int b2) {
          emit(2, {v[f0], v[f1], v[f2], v[b0], v[b1], v[b2]});
        };
        if (allSame(0, 1, 2, 3)) { tryPyramid(v[4], v[7], v[6], v[5], v[0]); continue; }
        if (allSame(4, 5, 6, 7)) { tryPyramid(v[0], v[1], v[2], v[3], v[4]); continue; }
        if (allSame(0, 1, 4, 5)) { tryPyramid(v[2], v[6], v[7], v[3], v[0]); continue; }
        if (allSame(2, 3, 6, 7)) { tryPyramid(v[0], v[4], v[5], v[1], v[2]); continue; }
        if (allSame(0, 3, 4, 7)) { tryPyramid(v[1], v[5], v[6], v[2], v[0]); continue; }
        if (allSame(1, 2, 5, 6)) { tryPyramid(v[0], v[3], v[7], v[4], v[1]); continue; }
        if (same(0, 1) && same(4, 5)) { wedge(3, 2, 0, 7, 6, 4); continue; }
        if (same(0, 4) && same(1, 5)) { wedge(2, 6, 5, 3, 7, 4); continue; }
        if (same(3, 7) && same(2, 6)) { wedge(5, 1, 2, 4, 0, 3); continue; }
        if (same(2, 3) && same(6, 7)) { wedge(1, 0, 3, 5, 4, 7); continue; }
        if (same(4, 7) && same(5, 6)) { wedge(3, 0, 4, 2, 1, 6); continue; }
        if (same(4, 5) && same(6, 7)) { wedge(0, 1, 4, 3, 2, 7); continue; }
        if (same(0, 1) && same(3, 2)) { wedge(5, 4, 0, 6, 7, 3); continue; }
        if (same(0, 3) && same(1, 2)) { wedge(4, 7, 3, 5, 6, 2); continue; }
        if (same(0, 3) && same(4, 7)) { wedge(5, 6, 7, 1, 2, 3); continue; }
        if (same(0, 4) && same(3, 7)) { wedge(1, 5, 4, 2, 6, 7); continue; }
        if (same(1, 2) && same(5, 6)) { wedge(7, 4, 5, 3, 0, 1); continue; }
        if (same(1, 5) && same(2, 6)) { wedge(4, 0, 1, 7, 3, 2); continue; }
        emit(3, {v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]});
      }
}

}  // namespace

// makeGrids.cpp:46-209: bricks of <= 8^3 same-level cubes per macrocell
// (cid / 8, rounded down), each brick the bounding box of its cubes; a brick
// vertex takes the scalar ID of the cube corner there (vtkOrder), -1 where no
// cube reaches.  Bricks come in (level, mc.x, mc.y, mc.z) order, the order of
// makeGrids's per-level std::map when levels ascend.
void makeBricks(const std::vector<CubeRec> &cubes, std::vector<Gridlet> &gridlets, const BrickSink &sink,
                int64_t &conflicts) {
  const int64_t numCubes = (int64_t)cubes.size();
  auto mcOf = [](int32_t c) { return (c < 0 ? c - 7 : c) / 8; };  // makeGrids.cpp:79-86
  std::vector<uint64_t> ckeys(numCubes);
  std::vector<int64_t> corder(numCubes);
  for (int64_t i = 0; i < numCubes; ++i) {
    const CubeRec &c = cubes[i];
    if (c.level < 0 || c.level > 15) throw std::runtime_error("makeBricks: cube level out of range [0,15]");
    uint64_t key = (uint64_t)c.level << 60;
    for (int k = 0; k < 3; ++k) {
      if (c.cid[k] < -(1 << 22) || c.cid[k] >= (1 << 22))
        throw std::runtime_error("makeBricks: cube position out of range");
      key |= (uint64_t)(mcOf(c.cid[k]) + (1 << 19)) << (40 - 20 * k);
    }
    ckeys[i] = key;
    corder[i] = i;
  }
  radixSort(ckeys, corder);
  std::vector<int64_t> brickStart;
  for (int64_t i = 0; i < numCubes; ++i)
    if (i == 0 || ckeys[i] != ckeys[i - 1]) brickStart.push_back(i);
  brickStart.push_back(numCubes);
  ckeys.clear();
  ckeys.shrink_to_fit();
  const int64_t numBricks = (int64_t)brickStart.size() - 1;
  gridlets.resize(numBricks);
  std::vector<int64_t> scalarStart(numBricks + 1, 0);
  for (int64_t b = 0; b < numBricks; ++b) {
    int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    for (int64_t i = brickStart[b]; i < brickStart[b + 1]; ++i) {
      const CubeRec &c = cubes[corder[i]];
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], c.cid[k]);
        hi[k] = std::max(hi[k], c.cid[k] + 1);
      }
    }
    Gridlet &g = gridlets[b];
    for (int k = 0; k < 3; ++k) {
      g.lower[k] = lo[k];
      g.dims[k] = hi[k] - lo[k];
    }
    g.level = cubes[corder[brickStart[b]]].level;
    scalarStart[b + 1] = scalarStart[b] + (int64_t)(g.dims[0] + 1) * (g.dims[1] + 1) * (g.dims[2] + 1);
    if (scalarStart[b + 1] >= ((int64_t)1 << 32))
      throw std::runtime_error("makeBricks: gridlet scalars exceed 32-bit offsets");
    g.begin = (uint32_t)scalarStart[b];
  }
  sink.allocate(scalarStart[numBricks]);
  std::atomic<int64_t> bad{0};
  parallelChunks(numBricks, [&](int64_t b) {
    const Gridlet &g = gridlets[b];
    const int nx = g.dims[0], ny = g.dims[1];
    const int64_t ns = scalarStart[b + 1] - scalarStart[b];
    std::vector<int32_t> ids(ns, -1);
    static const int vtkOrder[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    int64_t nbad = 0;
    for (int64_t i = brickStart[b]; i < brickStart[b + 1]; ++i) {
      const CubeRec &c = cubes[corder[i]];
      const int bx = c.cid[0] - g.lower[0], by = c.cid[1] - g.lower[1], bz = c.cid[2] - g.lower[2];
      for (int iz = 0; iz < 2; iz++)
        for (int iy = 0; iy < 2; iy++)
          for (int ix = 0; ix < 2; ix++) {
            const int64_t idx = (bx + ix) + (int64_t)(nx + 1) * ((by + iy) + (int64_t)(ny + 1) * (bz + iz));
            const int32_t sid = c.scalarIDs[vtkOrder[4 * iz + 2 * iy + ix]];
            if (ids[idx] != -1 && ids[idx] != sid) nbad++;
            ids[idx] = sid;
          }
    }
    sink.brick(b, scalarStart[b], ids.data(), ns);
    bad += nbad;
  });
  conflicts += bad.load();
}

void buildDualMesh(const AmrCell *in, const float *scalars, int64_t n, DualMesh &out, bool keepIds) {
  if (n <= 0) throw std::runtime_error("buildDualMesh: no cells");
  if (n >= (int64_t)1 << 31) throw std::runtime_error("buildDualMesh: too many cells");
  Exa exa;
  // validate + sort (makeDual.cpp:878)
  std::vector<uint64_t> keys(n);
  std::vector<int64_t> order(n);
  for (int64_t i = 0; i < n; ++i) {
    const AmrCell &c = in[i];
    if (c.level < 0 || c.level > 15)
      throw std::runtime_error("buildDualMesh: cell level out of range [0,15]");
    for (int k = 0; k < 3; ++k) {
      if (c.pos[k] < 0 || c.pos[k] >= kCoordLimit)
        throw std::runtime_error("buildDualMesh: cell position out of range [0,2^20)");
      if (c.pos[k] & ((1 << c.level) - 1))
        throw std::runtime_error("buildDualMesh: cell position not aligned to its level");
    }
    keys[i] = ((uint64_t)c.pos[1] << 44) | ((uint64_t)c.pos[0] << 24) | ((uint64_t)c.level << 20) |
              (uint64_t)c.pos[2];
    order[i] = i;
    exa.minLevel = std::min(exa.minLevel, c.level);
    exa.maxLevel = std::max(exa.maxLevel, c.level);
  }
  radixSort(keys, order);
  keys.clear();
  keys.shrink_to_fit();
  exa.cells.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    const AmrCell &c = in[order[i]];
    exa.cells[i] = SortedCell{{c.pos[0], c.pos[1], c.pos[2]}, c.level, (int32_t)order[i]};
  }
  order.clear();
  order.shrink_to_fit();
  exa.hash.init(n);
  {
    std::atomic<int64_t> dup{0};
    const int64_t chunk = 1 << 16, nc = (n + chunk - 1) / chunk;
    parallelChunks(nc, [&](int64_t c) {
      const int64_t b = c * chunk, e = std::min(n, b + chunk);
      for (int64_t i = b; i < e; ++i) {
        const SortedCell &sc = exa.cells[i];
        if (!exa.hash.insert(hashKey(sc.level, sc.pos[0], sc.pos[1], sc.pos[2]), i)) dup++;
      }
    });
    if (dup.load()) throw std::runtime_error("buildDualMesh: duplicate cells");
  }
  // dual cells, chunked in sorted order
  const int64_t chunk = 1 << 14, numChunks = (n + chunk - 1) / chunk;
  std::vector<ChunkOut> outs(numChunks);
  parallelChunks(numChunks, [&](int64_t c) {
    const int64_t b = c * chunk, e = std::min(n, b + chunk);
    for (int64_t i = b; i < e; ++i) doCell(exa, i, outs[c]);
  });
  // vertex compaction (ExaStitchModel.cpp:120-156)
  std::vector<int32_t> vid(n, -1);
  int64_t totalElem[4] = {0, 0, 0, 0};
  static const int nv[4] = {4, 5, 6, 8};
  for (auto &o : outs)
    for (int t = 0; t < 4; ++t) {
      for (int64_t id : o.elem[t]) vid[id] = 0;
      totalElem[t] += (int64_t)o.elem[t].size() / nv[t];
    }
  int64_t numVerts = 0;
  for (int64_t i = 0; i < n; ++i)
    if (vid[i] >= 0) vid[i] = (int32_t)numVerts++;
  out.vertices.assign(4 * numVerts, 0.f);
  if (keepIds) out.vertexTags.assign(numVerts, -1);
  for (int64_t i = 0; i < n; ++i)
    if (vid[i] >= 0) {
      const SortedCell &c = exa.cells[i];
      if (keepIds) out.vertexTags[vid[i]] = c.scalarID;
      float *v = &out.vertices[4 * (int64_t)vid[i]];
      for (int k = 0; k < 3; ++k) v[k] = (float)c.pos[k] + 0.5f * (float)(1 << c.level);
      v[3] = scalars[c.scalarID];
    }
  const int64_t numElems = totalElem[0] + totalElem[1] + totalElem[2] + totalElem[3];
  out.indices.assign(8 * numElems, -1);
  int64_t e = 0;
  out.degenerate = 0;
  out.conflicts = 0;
  for (int t = 0; t < 4; ++t) {
    out.counts[t] = totalElem[t];
    for (auto &o : outs) {
      const auto &L = o.elem[t];
      for (size_t i = 0; i < L.size(); i += nv[t], ++e)
        for (int j = 0; j < nv[t]; ++j) out.indices[8 * e + j] = vid[L[i + j]];
    }
  }
  std::vector<CubeRec> cubes;
  {
    size_t nc = 0;
    for (auto &o : outs) nc += o.cubes.size();
    cubes.reserve(nc);
    for (auto &o : outs) {
      out.degenerate += o.degenerate;
      out.conflicts += o.errors;
      for (int t = 0; t < 4; ++t) std::vector<int64_t>().swap(o.elem[t]);
      cubes.insert(cubes.end(), o.cubes.begin(), o.cubes.end());
      std::vector<CubeRec>().swap(o.cubes);
    }
  }
  outs.clear();
  out.numCubes = (int64_t)cubes.size();
  out.gridletScalarIDs.clear();
  BrickSink sink;
  sink.allocate = [&](int64_t total) {
    out.gridletScalars.assign(total, 0.f);
    if (keepIds) out.gridletScalarIDs.assign(total, -1);
  };
  sink.brick = [&](int64_t, int64_t begin, const int32_t *ids, int64_t ns) {
    float *dst = &out.gridletScalars[begin];
    for (int64_t i = 0; i < ns; ++i) dst[i] = ((uint32_t)ids[i] < (uint64_t)n) ? scalars[ids[i]] : NAN;
    if (keepIds) memcpy(&out.gridletScalarIDs[begin], ids, ns * sizeof(int32_t));
  };
  int64_t conflicts = 0;
  makeBricks(cubes, out.gridlets, sink, conflicts);
  out.conflicts += conflicts;
  if (keepIds) out.cubes.swap(cubes);
}

}  // namespace exa
