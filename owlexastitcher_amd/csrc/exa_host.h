// exa_host.h -- host-side builders of the exaStitch path (no HIP runtime):
// synthetic AMR scenes, the dual-mesh/gridlet generator that replaces the
// reference's offline amrMakeDualMesh + amrMakeGrids tools, the in-memory
// ExaStitchModel, and the point-location BVH.
#pragma once
#include <stdint.h>
#include <functional>
#include <string>
#include <vector>
#include "exa_types.h"

namespace exa {

int hostThreads();
// LSD radix sort of (key, payload) pairs by key, stable (dualmesh.cpp)
void radixSort(std::vector<uint64_t> &keys, std::vector<int64_t> &vals);

// ---------------------------------------------------------------------------
// synthetic AMR scenes
// ---------------------------------------------------------------------------
struct AmrScene {
  std::vector<AmrCell> cells;
  std::vector<float> scalars;  // one per cell, input order == scalar ID
  std::vector<float> params;   // generator parameters (for tests)
};

// Two-level Gaussian scene: coarse cells of width 2^coarseLevel on a
// baseRes^3 grid; coarse cells lying inside [fineLo,fineHi)^3 are split once.
// Scalar at each cell centre = sum of numGaussians isotropic Gaussians
// (centre U[0,domain]^3, sigma U[sigLo,sigHi], amplitude U[ampLo,ampHi]).
void amrGaussians(AmrScene &out, int baseRes, int coarseLevel, int fineLo, int fineHi,
                  int numGaussians, uint64_t seed);

// Three-level fractal-noise scene: base cells of width 2^baseLevel on a
// baseRes^3 grid; a cell is split where the noise at its centre exceeds
// thresh1 (next level) and thresh2 (the level after); a vertex 2:1 balance
// pass follows.  Scalar = noise * radial falloff.
void amrNoise(AmrScene &out, int baseRes, int baseLevel, float thresh1, float thresh2,
              int octaves, float contrast, uint64_t seed);

// Multi-level expanding-shock scene: base cells of width 2^baseLevel, the
// shell (radius, thickness) refined to level 1 with level-0 patches where the
// turbulence exceeds thetaCore, level >= 2 in the wake where it exceeds
// thetaWake; graded so that touching leaves differ by at most one level.
void amrShock(AmrScene &out, int baseRes, int baseLevel, float radius, float thickness, float thetaCore,
              float thetaWake, uint64_t seed);
// leaf counts per level of the same scene, without building it
void amrShockCount(int baseRes, int baseLevel, float radius, float thickness, float thetaCore, float thetaWake,
                   uint64_t seed, int64_t counts[16]);

// ---------------------------------------------------------------------------
// dual mesh (tools/amrMakeDualMesh/makeDual.cpp, makeGrids.cpp)
// ---------------------------------------------------------------------------
// A perfect dual cube: 8 same-level cells; cid = lowest cell position /
// 2^level; scalarIDs in the emitter's corner order (makeDual.cpp:445-449,
// 473-480).
struct CubeRec {
  int32_t level;
  int32_t cid[3];
  int32_t scalarIDs[8];
};

struct DualMesh {
  std::vector<float> vertices;     // float4 per vertex (cell centre, scalar)
  std::vector<int32_t> indices;    // int[8] per element, -1 padded
  std::vector<Gridlet> gridlets;
  std::vector<float> gridletScalars;
  // kept on request (the on-disk outputs need them): per vertex its cell's
  // scalar ID, per gridlet scalar the scalar ID or -1, the perfect cubes
  std::vector<int32_t> vertexTags;
  std::vector<int32_t> gridletScalarIDs;
  std::vector<CubeRec> cubes;
  int64_t counts[4] = {0, 0, 0, 0};  // tets, pyramids, wedges, hexes
  int64_t numCubes = 0;
  int64_t conflicts = 0;   // brick scalar write conflicts + impossible cases
  int64_t degenerate = 0;  // dual cells dropped as degenerate
};
// Throws std::runtime_error on invalid input.
void buildDualMesh(const AmrCell *cells, const float *scalars, int64_t n, DualMesh &out, bool keepIds = false);

// Gridlets from perfect cubes (makeGrids.cpp:46-209).  allocate(total) is
// called once with the number of gridlet scalars, then brick(b, begin, ids,
// n) once per brick, from several threads, with the scalar IDs (-1: hole)
// of brick b, whose scalars start at begin.
struct BrickSink {
  std::function<void(int64_t)> allocate;
  std::function<void(int64_t, int64_t, const int32_t *, int64_t)> brick;
};
void makeBricks(const std::vector<CubeRec> &cubes, std::vector<Gridlet> &gridlets, const BrickSink &sink,
                int64_t &conflicts);

// ---------------------------------------------------------------------------
// model (model/ExaStitchModel.cpp:24-257 in-memory form)
// ---------------------------------------------------------------------------
struct HostModel {
  std::vector<float> vertices;
  std::vector<int32_t> indices;
  std::vector<Gridlet> gridlets;
  std::vector<float> gridletScalars;
  float cellBounds[6];
  float valueRange[2];
  float voxelXfm[12];      // Model::voxelSpaceTransform (identity by default)
  int32_t numMCs[3] = {0, 0, 0};
  int64_t numElements() const { return (int64_t)indices.size() / 8; }
};
// Validates element rows and gridlet records; computes cellBounds and
// valueRange as ExaStitchModel::load does.  Throws on invalid input.
void finalizeModel(HostModel &m);

// ---------------------------------------------------------------------------
// point-location BVH (replaces the OptiX user-geometry BLAS/TLAS)
// ---------------------------------------------------------------------------
// 64-byte binary node holding both children's boxes.  A child ref >= 0 is an
// inner node index; a child with count > 0 is a leaf covering
// primRefs[ref .. ref+count); count == 0 marks an inner node.  An empty child
// has count == -1 and an inverted
// box.
struct BvhNode {
  float lo0[3], hi0[3];
  float lo1[3], hi1[3];
  int32_t ref0, ref1;
  int32_t cnt0, cnt1;
};
static_assert(sizeof(BvhNode) == 64, "BvhNode must stay 64 bytes");

struct Bvh {
  std::vector<BvhNode> nodes;
  std::vector<int32_t> primRefs;  // ascending within each leaf
  int depth = 0;
};
constexpr int kBvhMaxDepth = 32;
// boxes: 6 floats (lo xyz, hi xyz) per primitive.
void buildBvh(const float *boxes, int64_t n, int leafSize, Bvh &out);

void gridletBoxes(const std::vector<Gridlet> &g, std::vector<float> &boxes);

// ---------------------------------------------------------------------------
// direct point-location tables (the default accelerator)
// ---------------------------------------------------------------------------
// Gridlets: per refinement level L, a grid of 8^3-cell macrocells; the list
// of macrocell m holds every level-L gridlet whose box reaches into m (box
// [(lower+.5)2^L, (lower+dims+.5)2^L] <=> t = p/2^L - .5 in [lower,
// lower+dims]), ascending.  A point only needs the list of floor(t/8) per
// level.
// Elements: a uniform bin grid over the element AABBs; the list of a bin
// holds every element whose AABB overlaps it, ascending.  Bin indices are
// computed with the same float expression on host and device, so every
// element whose AABB contains a point is in that point's bin.
struct LocateTables {
  static constexpr int kMaxLevels = 16;
  static constexpr int kOwnerBlock = 512;  // 8^3 dual cells per macrocell
  int32_t numLevels = 0;                 // levels 0..numLevels-1 (may be empty)
  int32_t levelDims[kMaxLevels][3] = {};  // macrocells per axis
  int64_t levelOffset[kMaxLevels + 1] = {};
  std::vector<uint32_t> gStart;  // per macrocell (all levels concatenated) + 1
  std::vector<int32_t> gList;
  // Rounding margins of the gridlet tests, from the largest coordinate
  // magnitude 2^e of the model (any p below 2^e is computed to within
  // 2^(e-24) in level-L cell units): a gridlet hit whose fractional cell
  // coordinates keep bandMargin = 2^(e-21) from every face is the unique
  // accepting gridlet; the owner tables answer points that keep
  // fastMargin = 2^(e-20) from the faces of their level's dual cells.
  float bandMargin = 0.00390625f;
  float fastMargin = 0.0078125f;
  // Dual-cell owner tables (empty when fastMargin would exceed 1/16): per
  // macrocell the index of its owner block or -1; per block, for each of its
  // 8^3 dual cells, the lowest-ID gridlet whose cell there has 8 non-NaN
  // corners, or -1.
  std::vector<int32_t> gTop;
  std::vector<int32_t> gOwner;
  float binLo[3] = {0, 0, 0};
  float binInv = 1.f;  // 1 / bin size (a power of two)
  int32_t binDims[3] = {0, 0, 0};
  std::vector<uint32_t> eStart;  // per bin + 1
  // three words per list entry: element id | vertex count << 28, then the
  // element's box quantised to 512 sub-cells of the entry's bin (9 bits per
  // bound in 10-bit slots at bits 0, 10, 20: lo with the slots' top bits
  // clear, hi with them set)
  std::vector<uint32_t> eCand;
  // per element its vertices (xyz + scalar) in index order, 8 slots
  std::vector<float> eVerts;
};
void buildLocateTables(const HostModel &m, LocateTables &t);
// LocateTables::bandMargin / fastMargin of a model
void gridletMargins(const HostModel &m, float &bandMargin, float &fastMargin);
void elementBoxes(const std::vector<int32_t> &indices, const std::vector<float> &vertices,
                  std::vector<float> &boxes);

}  // namespace exa
