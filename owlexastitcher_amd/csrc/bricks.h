// bricks.h -- host side of the ExaBricks comparison sampler: bricks, active
// brick regions (ABRs), the kd-tree that the ABR build's recursion describes,
// the .bricks file format and a brick maker for AMR leaf cells.
#pragma once
#include <stdint.h>
#include <string>
#include <vector>
#include "exa_types.h"

namespace exa {

// model/ABRs.h:28-63 ExaBrick: a dense box of size[0]*size[1]*size[2] cells
// of width 2^level whose lowest cell starts at lower (finest-cell units); the
// scalar of cell (x, y, z) is scalars[begin + x + size[0]*(y + size[1]*z)].
struct Brick {
  int32_t lower[3];
  int32_t size[3];
  int32_t level;
  uint32_t begin;
};
static_assert(sizeof(Brick) == 32, "Brick must stay 32 bytes");

// model/ABRs.h:65-75 ABR: a box in which one fixed set of bricks overlaps
// (leafList[leafListBegin .. +leafListSize), ascending brick IDs).
struct Abr {
  float domain[6];  // lower xyz, upper xyz
  float valueRange[2];
  int32_t leafListBegin;
  int32_t leafListSize;
  float finestLevelCellWidth;
};
static_assert(sizeof(Abr) == 44, "Abr must stay 44 bytes");

// One node of the kd-tree over the ABRs, in the order the build visits them
// (a node, then its upper-side subtree, then its lower-side subtree).
//   inner node: info = (lower child << 2) | dim, upper child = this + 1,
//               split = the plane's coordinate along dim;
//   leaf:       info = ((ABR id + 1) << 2) | 3 (ABR id -1: no ABR there).
struct KdNode {
  float split;
  int32_t info;
};
static_assert(sizeof(KdNode) == 8, "KdNode must stay 8 bytes");

struct BrickModel {
  std::vector<Brick> bricks;
  std::vector<float> scalars;  // flattened, brick after brick
  std::vector<Abr> abrs;
  std::vector<int32_t> leafList;
  std::vector<KdNode> nodes;
  int32_t kdDepth = 0;
  float cellBounds[6];  // union of brick bounds (ExaBrickModel.cpp:179-182)
  float valueRange[2];  // union of ABR value ranges (184-187)
};

// ABRs.h:39-52 getBounds / getDomain
void brickBounds(const Brick &b, float out[6]);
void brickDomain(const Brick &b, float out[6]);

// ExaBrickModel::init (ExaBrickModel.cpp:148-213) without the adjacency list:
// ABRs (ABRs.cpp:286-352), kd nodes, cellBounds and valueRange.
void finalizeBrickModel(BrickModel &m);

// ExaBrickModel::load(bricks, scalars, numBricks) (ExaBrickModel.cpp:119-145):
// the scalars of the cells of all bricks, numCells = last brick's begin +
// its cell count.  Throws std::runtime_error when a brick is invalid or
// reaches beyond numCells.
void brickModelFromArrays(const Brick *bricks, const float *scalars, size_t numBricks, BrickModel &m);

// .bricks file (ExaBrickModel.cpp:53-70, tools/makeTestData.cpp:23-31): per
// brick size (3 int32), lower (3 int32), level (int32), then the
// size[0]*size[1]*size[2] cell IDs (int32, x fastest) into the scalar file.
void readBrickFile(const std::string &path, std::vector<Brick> &bricks, std::vector<int32_t> &cellIDs);
void writeBrickFile(const std::string &path, const std::vector<Brick> &bricks, const std::vector<int32_t> &cellIDs);
// ExaBrickModel::load(brickFile, scalarFile, "") (ExaBrickModel.cpp:27-117):
// scalars[i] = orderedScalars[cellIDs[i]]; an ID outside the scalar file
// throws, as in the reference.
void loadBrickModel(const std::string &bricksPath, const std::string &scalarsPath, BrickModel &m);
// the same from cell IDs already in memory
void brickModelFromCellIDs(const std::vector<Brick> &bricks, const std::vector<int32_t> &cellIDs,
                           const float *orderedScalars, int64_t numScalars, BrickModel &m);

// Brick maker (no reference counterpart; the reference reads bricks made by
// external tools): the leaf cells of each level, grouped into blocks of
// blockCells^3 cells of that level, are covered by dense boxes, greedily
// (first free cell in x-fastest order, grown along x, then y, then z).
// cellIDs holds the input index of every brick cell, brick after brick.
void bricksFromCells(const AmrCell *cells, int64_t n, int blockCells, std::vector<Brick> &bricks,
                     std::vector<int32_t> &cellIDs);

}  // namespace exa
