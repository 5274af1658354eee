// bricks.cpp -- host side of the ExaBricks comparison sampler (bricks.h).
//
// The active-brick-region build restates ABRs.cpp (buildRec 88-201, addLeaf
// 40-86, computeValueRange 204-270, buildFrom 272-352) operation for
// operation, so the regions, their IDs and leaf lists are the reference's;
// while recursing it records the split planes as a kd-tree, which the
// device traverses instead of a BVH over the regions.
#include "bricks.h"
#include <math.h>
#include <string.h>
#include <algorithm>
#include <atomic>
#include <fstream>
#include <stdexcept>
#include <thread>
#include "exa_host.h"

namespace exa {

namespace {

template <typename F>
void parallelItems(int64_t n, F &&f) {
  const int T = (int)std::min<int64_t>(hostThreads(), std::max<int64_t>(n, 1));
  std::atomic<int64_t> next{0};
  auto work = [&] {
    for (;;) {
      const int64_t i = next.fetch_add(1);
      if (i >= n) break;
      f(i);
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

int64_t brickCells(const Brick &b) { return (int64_t)b.size[0] * b.size[1] * b.size[2]; }

void checkBrick(const Brick &b) {
  if (b.size[0] <= 0 || b.size[1] <= 0 || b.size[2] <= 0) throw std::runtime_error("brick with an empty size");
  if (b.level < 0 || b.level > 30) throw std::runtime_error("brick level out of range");
  for (int k = 0; k < 3; ++k) {
    const int64_t hi = (int64_t)b.lower[k] + ((int64_t)b.size[k] << b.level);
    if (hi > INT32_MAX) throw std::runtime_error("brick extends beyond the int32 coordinate range");
  }
  // the device reads a brick's scalar offset (begin) and its in-brick cell
  // index as int32: reject before the product can overflow
  const int64_t xy = (int64_t)b.size[0] * b.size[1];
  if (xy > (int64_t)INT32_MAX || xy * b.size[2] > (int64_t)INT32_MAX)
    throw std::runtime_error("brick holds more than 2^31-1 cells");
}

// A brick's domain clipped to a region (ABRs.cpp buildPrims)
struct Prim {
  float lo[3], hi[3];
  int32_t id;
};

// owl arg_max of a vec3f: the first index of the largest component
int argMax3(const float v[3]) {
  int k = 0;
  if (v[1] > v[k]) k = 1;
  if (v[2] > v[k]) k = 2;
  return k;
}

struct AbrBuilder {
  BrickModel &m;
  int32_t maxDepth = 0;

  static int32_t leafInfo(int32_t abr) { return ((abr + 1) << 2) | 3; }

  // ABRs.cpp:40-86 addLeaf: the sorted set of brick IDs of the region;
  // returns the new ABR's ID or -1
  int32_t addLeaf(const std::vector<Prim> &prims, const float dom[6]) {
    for (int k = 0; k < 3; ++k)
      if (dom[k] >= dom[3 + k]) return -1;
    std::vector<int32_t> ids;
    ids.reserve(prims.size());
    for (const Prim &p : prims) ids.push_back(p.id);
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    if (ids.empty()) return -1;
    Abr a;
    memset(&a, 0, sizeof(a));
    memcpy(a.domain, dom, sizeof(a.domain));
    a.leafListSize = (int32_t)ids.size();
    a.leafListBegin = (int32_t)m.leafList.size();
    m.leafList.insert(m.leafList.end(), ids.begin(), ids.end());
    m.abrs.push_back(a);
    return (int32_t)m.abrs.size() - 1;
  }

  // ABRs.cpp:88-201 buildRec; appends this region's kd node, then the upper
  // side's subtree, then the lower side's (the reference recurses into the
  // upper side first, which fixes the ABR IDs)
  void build(std::vector<Prim> &prims, const float dom[6], int32_t depth) {
    // every split separates two distinct face planes of the bricks, so the
    // depth stays far below this; the bound only keeps the host stack safe
    if (depth > 4096) throw std::runtime_error("active brick regions nested too deeply");
    const int32_t self = (int32_t)m.nodes.size();
    m.nodes.push_back(KdNode{0.f, leafInfo(-1)});
    maxDepth = std::max(maxDepth, depth);
    if (prims.empty()) return;
    for (int k = 0; k < 3; ++k)
      if (dom[3 + k] == dom[k]) return;  // "EMPTY DOMAIN!?"
    float tgt[3], bestPos[3], bestDist[3], span[3];
    for (int k = 0; k < 3; ++k) {
      tgt[k] = (dom[k] + dom[3 + k]) / 2.f;
      bestPos[k] = dom[k];
      bestDist[k] = dom[3 + k] - dom[k];
      span[k] = dom[3 + k] - dom[k];
    }
    // the candidate plane nearest the centre per axis: each clipped brick's
    // upper then lower face, first strict improvement wins
    for (const Prim &p : prims)
      for (int dim = 0; dim < 3; ++dim)
        for (int side = 0; side < 2; ++side) {
          const float pos = side ? p.lo[dim] : p.hi[dim];
          if (pos <= dom[dim] || pos >= dom[3 + dim]) continue;
          const float dist = fabsf(tgt[dim] - pos);
          if (dist >= bestDist[dim]) continue;
          bestPos[dim] = pos;
          bestDist[dim] = dist;
        }
    int splitDim = -1;
    float splitPos = 0.f;
    const int widest = argMax3(span);
    for (int i = 0; i < 3; ++i) {
      const int dim = (widest + i) % 3;
      if (bestPos[dim] <= dom[dim] || bestPos[dim] >= dom[3 + dim]) continue;
      splitDim = dim;
      splitPos = bestPos[dim];
      break;
    }
    if (splitDim < 0) {
      m.nodes[self].info = leafInfo(addLeaf(prims, dom));
      return;
    }
    float domL[6], domR[6];
    memcpy(domL, dom, sizeof(domL));
    memcpy(domR, dom, sizeof(domR));
    domL[3 + splitDim] = splitPos;
    domR[splitDim] = splitPos;
    std::vector<Prim> primsL, primsR;
    auto clipInto = [](const Prim &p, const float d[6], std::vector<Prim> &out) {
      Prim c;
      c.id = p.id;
      for (int k = 0; k < 3; ++k) {
        c.lo[k] = std::max(p.lo[k], d[k]);
        c.hi[k] = std::min(p.hi[k], d[3 + k]);
      }
      if (c.lo[0] < c.hi[0] && c.lo[1] < c.hi[1] && c.lo[2] < c.hi[2]) out.push_back(c);
    };
    for (const Prim &p : prims) {
      clipInto(p, domL, primsL);
      clipInto(p, domR, primsR);
    }
    std::vector<Prim>().swap(prims);
    m.nodes[self].split = splitPos;
    build(primsR, domR, depth + 1);
    const int32_t lower = (int32_t)m.nodes.size();
    build(primsL, domL, depth + 1);
    m.nodes[self].info = (lower << 2) | splitDim;
  }
};

// ABRs.cpp:204-270 computeValueRange: every cell of every brick of the
// region whose centre lies within one cell width of the region (closed)
void abrValueRange(const BrickModel &m, Abr &a) {
  float lo = INFINITY, hi = -INFINITY;
  std::vector<char> vx, vy, vz;
  for (int32_t i = 0; i < a.leafListSize; ++i) {
    const Brick &b = m.bricks[m.leafList[a.leafListBegin + i]];
    const float cw = (float)(1 << b.level);
    std::vector<char> *valid[3] = {&vx, &vy, &vz};
    for (int k = 0; k < 3; ++k) {
      valid[k]->assign(b.size[k], 0);
      for (int j = 0; j < b.size[k]; ++j) {
        const float pos = (float)b.lower[k] + ((float)j + .5f) * cw;
        const float l = pos - cw, h = pos + cw;
        (*valid[k])[j] = (l <= a.domain[3 + k]) && (h >= a.domain[k]);
      }
    }
    for (int iz = 0; iz < b.size[2]; ++iz) {
      if (!vz[iz]) continue;
      for (int iy = 0; iy < b.size[1]; ++iy) {
        if (!vy[iy]) continue;
        for (int ix = 0; ix < b.size[0]; ++ix) {
          if (!vx[ix]) continue;
          const float s = m.scalars[(size_t)b.begin + ix + (size_t)b.size[0] * iy +
                                    (size_t)b.size[0] * b.size[1] * iz];
          lo = std::min(lo, s);
          hi = std::max(hi, s);
        }
      }
    }
  }
  a.valueRange[0] = lo;
  a.valueRange[1] = hi;
}

}  // namespace

void brickBounds(const Brick &b, float o[6]) {
  const int32_t w = 1 << b.level;
  for (int k = 0; k < 3; ++k) {
    o[k] = (float)b.lower[k];
    o[3 + k] = (float)(b.lower[k] + b.size[k] * w);
  }
}

void brickDomain(const Brick &b, float o[6]) {
  const float cw = (float)(1 << b.level);
  for (int k = 0; k < 3; ++k) {
    o[k] = (float)b.lower[k] - 0.5f * cw;
    o[3 + k] = (float)b.lower[k] + ((float)b.size[k] + 0.5f) * cw;
  }
}

void finalizeBrickModel(BrickModel &m) {
  m.abrs.clear();
  m.leafList.clear();
  m.nodes.clear();
  m.kdDepth = 0;
  // ABRs::buildFrom (ABRs.cpp:286-318): the brick domains in brick order,
  // the root region is their union
  std::vector<Prim> prims(m.bricks.size());
  float root[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (size_t i = 0; i < m.bricks.size(); ++i) {
    float d[6];
    brickDomain(m.bricks[i], d);
    for (int k = 0; k < 3; ++k) {
      prims[i].lo[k] = d[k];
      prims[i].hi[k] = d[3 + k];
      root[k] = std::min(root[k], d[k]);
      root[3 + k] = std::max(root[3 + k], d[3 + k]);
    }
    prims[i].id = (int32_t)i;
  }
  AbrBuilder builder{m};
  if (!prims.empty()) builder.build(prims, root, 0);
  else m.nodes.push_back(KdNode{0.f, AbrBuilder::leafInfo(-1)});
  m.kdDepth = builder.maxDepth;
  // ABRs.cpp:320-337: finest level and value range per region
  parallelItems((int64_t)m.abrs.size(), [&](int64_t r) {
    Abr &a = m.abrs[r];
    int finest = 1 << 30;
    for (int32_t i = 0; i < a.leafListSize; ++i) finest = std::min(finest, m.bricks[m.leafList[a.leafListBegin + i]].level);
    a.finestLevelCellWidth = (float)(1 << finest);
    abrValueRange(m, a);
  });
  // ExaBrickModel::init (ExaBrickModel.cpp:177-187)
  for (int k = 0; k < 3; ++k) {
    m.cellBounds[k] = INFINITY;
    m.cellBounds[3 + k] = -INFINITY;
  }
  for (const Brick &b : m.bricks) {
    float bb[6];
    brickBounds(b, bb);
    for (int k = 0; k < 3; ++k) {
      m.cellBounds[k] = std::min(m.cellBounds[k], bb[k]);
      m.cellBounds[3 + k] = std::max(m.cellBounds[3 + k], bb[3 + k]);
    }
  }
  m.valueRange[0] = INFINITY;
  m.valueRange[1] = -INFINITY;
  for (const Abr &a : m.abrs) {
    m.valueRange[0] = std::min(m.valueRange[0], a.valueRange[0]);
    m.valueRange[1] = std::max(m.valueRange[1], a.valueRange[1]);
  }
}

void brickModelFromArrays(const Brick *bricks, const float *scalars, size_t numBricks, BrickModel &m) {
  if (numBricks == 0 || !bricks || !scalars) throw std::runtime_error("no bricks");
  for (size_t i = 0; i < numBricks; ++i) checkBrick(bricks[i]);
  const int64_t numCells = (int64_t)bricks[numBricks - 1].begin + brickCells(bricks[numBricks - 1]);
  if (numCells > (int64_t)INT32_MAX) throw std::runtime_error("too many brick cells (more than 2^31-1)");
  for (size_t i = 0; i < numBricks; ++i)
    if ((int64_t)bricks[i].begin + brickCells(bricks[i]) > numCells)
      throw std::runtime_error("a brick reaches beyond the last brick's cells");
  m.bricks.assign(bricks, bricks + numBricks);
  m.scalars.assign(scalars, scalars + numCells);
  finalizeBrickModel(m);
}

void readBrickFile(const std::string &path, std::vector<Brick> &bricks, std::vector<int32_t> &cellIDs) {
  std::ifstream in(path, std::ios::binary);
  if (!in.good()) throw std::runtime_error("cannot open " + path);
  bricks.clear();
  cellIDs.clear();
  for (;;) {
    Brick b;
    in.read((char *)b.size, sizeof(b.size));
    in.read((char *)b.lower, sizeof(b.lower));
    in.read((char *)&b.level, sizeof(b.level));
    if (!in.good()) break;
    checkBrick(b);
    b.begin = (uint32_t)cellIDs.size();
    const int64_t n = brickCells(b);
    if ((int64_t)cellIDs.size() + n > (int64_t)INT32_MAX)
      throw std::runtime_error("too many brick cells (more than 2^31-1)");
    cellIDs.resize(cellIDs.size() + n);
    in.read((char *)(cellIDs.data() + b.begin), n * sizeof(int32_t));
    if (!in.good()) throw std::runtime_error("truncated brick file " + path);
    bricks.push_back(b);
  }
}

void writeBrickFile(const std::string &path, const std::vector<Brick> &bricks, const std::vector<int32_t> &cellIDs) {
  std::ofstream out(path, std::ios::binary);
  if (!out.good()) throw std::runtime_error("cannot write " + path);
  for (const Brick &b : bricks) {
    const int64_t n = brickCells(b);
    if ((int64_t)b.begin + n > (int64_t)cellIDs.size()) throw std::runtime_error("brick beyond its cell IDs");
    out.write((const char *)b.size, sizeof(b.size));
    out.write((const char *)b.lower, sizeof(b.lower));
    out.write((const char *)&b.level, sizeof(b.level));
    out.write((const char *)(cellIDs.data() + b.begin), n * sizeof(int32_t));
  }
  if (!out.good()) throw std::runtime_error("write error on " + path);
}

void brickModelFromCellIDs(const std::vector<Brick> &bricks, const std::vector<int32_t> &cellIDs,
                           const float *orderedScalars, int64_t numScalars, BrickModel &m) {
  if (bricks.empty()) throw std::runtime_error("no bricks");
  if (cellIDs.size() > (size_t)INT32_MAX) throw std::runtime_error("too many brick cells (more than 2^31-1)");
  m.bricks = bricks;
  m.scalars.resize(cellIDs.size());
  // ExaBrickModel.cpp:76-95
  for (size_t i = 0; i < cellIDs.size(); ++i) {
    const int32_t id = cellIDs[i];
    if (id < 0) throw std::runtime_error("negative cell ID");
    if (id >= numScalars) throw std::runtime_error("invalid cell ID");
    m.scalars[i] = orderedScalars[id];
  }
  for (const Brick &b : m.bricks)
    if ((int64_t)b.begin + brickCells(b) > (int64_t)m.scalars.size()) throw std::runtime_error("brick beyond its cells");
  finalizeBrickModel(m);
}

void loadBrickModel(const std::string &bricksPath, const std::string &scalarsPath, BrickModel &m) {
  std::vector<float> ordered;
  {
    std::ifstream in(scalarsPath, std::ios::binary | std::ios::ate);
    if (!in.good()) throw std::runtime_error("cannot open " + scalarsPath);
    const std::streamoff bytes = in.tellg();
    in.seekg(0);
    ordered.resize((size_t)bytes / sizeof(float));
    in.read((char *)ordered.data(), ordered.size() * sizeof(float));
  }
  std::vector<Brick> bricks;
  std::vector<int32_t> ids;
  readBrickFile(bricksPath, bricks, ids);
  brickModelFromCellIDs(bricks, ids, ordered.data(), (int64_t)ordered.size(), m);
}

void bricksFromCells(const AmrCell *cells, int64_t n, int blockCells, std::vector<Brick> &bricks,
                     std::vector<int32_t> &cellIDs) {
  bricks.clear();
  cellIDs.clear();
  if (blockCells < 1 || blockCells > 16) throw std::runtime_error("blockCells must be in [1, 16]");
  if (n <= 0) return;
  if (n > INT32_MAX) throw std::runtime_error("too many cells");
  // shift positions to non-negative block-aligned coordinates
  int32_t maxLevel = 0;
  int64_t minPos[3] = {INT64_MAX, INT64_MAX, INT64_MAX};
  for (int64_t i = 0; i < n; ++i) {
    if (cells[i].level < 0 || cells[i].level > 15) throw std::runtime_error("cell level out of range");
    maxLevel = std::max(maxLevel, cells[i].level);
    for (int k = 0; k < 3; ++k) minPos[k] = std::min<int64_t>(minPos[k], cells[i].pos[k]);
  }
  const int64_t align = (int64_t)blockCells << maxLevel;
  int64_t shift[3];
  for (int k = 0; k < 3; ++k) shift[k] = (minPos[k] >= 0 ? minPos[k] / align : -((-minPos[k] + align - 1) / align)) * align;
  // key: level | block z, y, x | cell within the block (z, y, x)
  constexpr int kBlockBits = 13;
  std::vector<uint64_t> keys(n);
  std::vector<int64_t> order(n);
  for (int64_t i = 0; i < n; ++i) {
    const AmrCell &c = cells[i];
    uint64_t bc[3], lc[3];
    for (int k = 0; k < 3; ++k) {
      const int64_t p = (int64_t)c.pos[k] - shift[k];
      if (p & ((1ll << c.level) - 1)) throw std::runtime_error("cell position not a multiple of its width");
      const int64_t ci = p >> c.level;
      bc[k] = (uint64_t)(ci / blockCells);
      lc[k] = (uint64_t)(ci % blockCells);
      if (bc[k] >= (1u << kBlockBits)) throw std::runtime_error("cell coordinates too large for the brick maker");
    }
    keys[i] = ((uint64_t)c.level << 51) | (bc[2] << 38) | (bc[1] << 25) | (bc[0] << 12) | (lc[2] << 8) |
              (lc[1] << 4) | lc[0];
    order[i] = i;
  }
  radixSort(keys, order);
  for (int64_t i = 1; i < n; ++i)
    if (keys[i] == keys[i - 1]) throw std::runtime_error("duplicate AMR cell");
  // groups of one level and block
  std::vector<int64_t> groupStart;
  for (int64_t i = 0; i < n; ++i)
    if (i == 0 || (keys[i] >> 12) != (keys[i - 1] >> 12)) groupStart.push_back(i);
  groupStart.push_back(n);
  const int64_t numGroups = (int64_t)groupStart.size() - 1;
  struct Out {
    std::vector<Brick> bricks;
    std::vector<int32_t> ids;
  };
  const int64_t chunk = 256;
  const int64_t numChunks = (numGroups + chunk - 1) / chunk;
  std::vector<Out> outs(numChunks);
  parallelItems(numChunks, [&](int64_t ch) {
    Out &o = outs[ch];
    const int B = blockCells;
    std::vector<int32_t> grid(B * B * B);
    for (int64_t g = ch * chunk; g < std::min(numGroups, (ch + 1) * chunk); ++g) {
      std::fill(grid.begin(), grid.end(), -1);
      const uint64_t k0 = keys[groupStart[g]];
      const int level = (int)(k0 >> 51);
      const int64_t bcz = (int64_t)((k0 >> 38) & 0x1fff), bcy = (int64_t)((k0 >> 25) & 0x1fff),
                    bcx = (int64_t)((k0 >> 12) & 0x1fff);
      for (int64_t i = groupStart[g]; i < groupStart[g + 1]; ++i) {
        const int lx = (int)(keys[i] & 15), ly = (int)((keys[i] >> 4) & 15), lz = (int)((keys[i] >> 8) & 15);
        grid[(lz * B + ly) * B + lx] = (int32_t)order[i];
      }
      auto free_ = [&](int x, int y, int z) { return grid[(z * B + y) * B + x] >= 0; };
      for (int z = 0; z < B; ++z)
        for (int y = 0; y < B; ++y)
          for (int x = 0; x < B; ++x) {
            if (!free_(x, y, z)) continue;
            int nx = 1, ny = 1, nz = 1;
            while (x + nx < B && free_(x + nx, y, z)) ++nx;
            auto rowFree = [&](int yy, int zz) {
              for (int i = 0; i < nx; ++i)
                if (!free_(x + i, yy, zz)) return false;
              return true;
            };
            while (y + ny < B && rowFree(y + ny, z)) ++ny;
            auto sliceFree = [&](int zz) {
              for (int j = 0; j < ny; ++j)
                if (!rowFree(y + j, zz)) return false;
              return true;
            };
            while (z + nz < B && sliceFree(z + nz)) ++nz;
            Brick b;
            const int64_t base[3] = {bcx * B + x, bcy * B + y, bcz * B + z};
            // the position of the brick's first cell, an input cell (no overflow)
            for (int k = 0; k < 3; ++k) b.lower[k] = (int32_t)((base[k] << level) + shift[k]);
            b.size[0] = nx;
            b.size[1] = ny;
            b.size[2] = nz;
            b.level = level;
            b.begin = (uint32_t)o.ids.size();  // relative; made absolute below
            for (int k = 0; k < nz; ++k)
              for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                  int32_t &c = grid[((z + k) * B + y + j) * B + x + i];
                  o.ids.push_back(c);
                  c = -1;
                }
            o.bricks.push_back(b);
          }
    }
  });
  for (Out &o : outs) {
    if ((int64_t)cellIDs.size() + (int64_t)o.ids.size() > (int64_t)INT32_MAX)
      throw std::runtime_error("too many brick cells (more than 2^31-1)");
    const uint32_t off = (uint32_t)cellIDs.size();
    for (Brick b : o.bricks) {
      b.begin += off;
      bricks.push_back(b);
    }
    cellIDs.insert(cellIDs.end(), o.ids.begin(), o.ids.end());
  }
}

}  // namespace exa
