// exa_io.h -- on-disk formats of the ExaStitch pipeline (io.cpp).
#pragma once
#include <stdint.h>
#include <string>
#include <vector>
#include "exa_host.h"

namespace exa {

// element mesh: the content the reference reads from a umesh file
struct ElementMesh {
  std::vector<float> xyz;            // 3 per vertex
  std::vector<int32_t> vertexTag;    // scalar ID per vertex (may be empty)
  std::vector<float> values;         // scalar per vertex (may be empty)
  std::vector<int32_t> elems[4];     // tets (4), pyramids (5), wedges (6), hexes (8)
};

struct XfFile {
  float opacityScale = 100.f;
  float absDomain[2] = {0.f, 1.f};
  float relDomain[2] = {0.f, 99.f};
  std::vector<float> colorMap;  // RGBA
};

void readCells(const std::string &path, std::vector<AmrCell> &cells);
void writeCells(const std::string &path, const std::vector<AmrCell> &cells);
void readScalars(const std::string &path, std::vector<float> &s);
void writeScalars(const std::string &path, const std::vector<float> &s);
void readElementMesh(const std::string &path, ElementMesh &m);
void writeElementMesh(const std::string &path, const ElementMesh &m);
// one file prefix_<level>.cubes per level present
void writeCubes(const std::string &prefix, const std::vector<CubeRec> &cubes);
int cubesFileLevel(const std::string &path);
void readCubes(const std::string &path, std::vector<CubeRec> &cubes);
void writeGrids(const std::string &path, const std::vector<Gridlet> &gridlets, const std::vector<int32_t> &ids,
                bool append);
// amrMakeGrids: bricks of every cubes file, appended to one grids file in
// argument order; *conflicts counts inconsistent brick-vertex writes
void makeGridsFiles(const std::vector<std::string> &cubesPaths, const std::string &gridsPath, int64_t *conflicts);
// ExaStitchModel::load(umesh, grids, scalars); empty paths are skipped
void loadModelFiles(const std::string &elemPath, const std::string &gridsPath, const std::string &scalarsPath,
                    HostModel &m);
void readXf(const std::string &path, XfFile &xf);
void writeXf(const std::string &path, const XfFile &xf);

}  // namespace exa
