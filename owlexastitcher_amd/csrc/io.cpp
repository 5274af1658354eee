// io.cpp -- on-disk formats of the ExaStitch pipeline:
//   .cells    16-byte leaf cells {int32 x, y, z, level}, read until EOF
//             (tools/amrMakeDualMesh/makeDual.cpp:932-940);
//   .scalars  raw float32, one per cell (model/ExaStitchModel.cpp:44-54);
//   _L.cubes  48-byte perfect dual cubes {float lower[3]; int32 level;
//             int32 scalarIDs[8]}, one file per level
//             (makeDual.cpp:445-449, 893-905; makeGrids.cpp:262-282);
//   .grids    per gridlet {int32 lower[3]; int32 level; int32 dims[3];
//             int32 scalarIDs[(dims+1)^3]} (makeGrids.cpp:240-246,
//             ExaStitchModel.cpp:191-213);
//   .xf       transfer function (headless.cpp:258-283);
// and the element mesh, which the reference keeps in the umesh library's
// format (absent from the reference).  This library uses its own file for the
// same content (vertices, per-vertex scalar IDs and/or values, tets,
// pyramids, wedges, hexes), little-endian:
//   char magic[8] = "EXAELEM1";
//   uint64 numVertices, numTets, numPyramids, numWedges, numHexes, flags;
//   float32 xyz[numVertices][3];
//   int32 vertexTag[numVertices]   if flags & 1 (scalar ID per vertex)
//   float32 value[numVertices]     if flags & 2 (scalar per vertex)
//   int32 tets[][4], pyramids[][5], wedges[][6], hexes[][8].
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <stdexcept>
#include "exa_host.h"
#include "exa_io.h"

namespace exa {

namespace {

struct File {
  FILE *f = nullptr;
  std::string path;
  File(const std::string &p, const char *mode) : path(p) {
    f = fopen(p.c_str(), mode);
    if (!f) throw std::runtime_error("cannot open '" + p + "'");
  }
  ~File() {
    if (f) fclose(f);
  }
  int64_t size() {
    fseeko(f, 0, SEEK_END);
    const int64_t n = ftello(f);
    fseeko(f, 0, SEEK_SET);
    return n;
  }
  void read(void *dst, size_t bytes, const char *what) {
    if (bytes && fread(dst, 1, bytes, f) != bytes)
      throw std::runtime_error("'" + path + "': truncated " + what);
  }
  // reads up to bytes; returns how many were read
  size_t readSome(void *dst, size_t bytes) { return fread(dst, 1, bytes, f); }
  void write(const void *src, size_t bytes) {
    if (bytes && fwrite(src, 1, bytes, f) != bytes) throw std::runtime_error("'" + path + "': write failed");
  }
  void close() {
    if (f && fclose(f) != 0) {
      f = nullptr;
      throw std::runtime_error("'" + path + "': close failed");
    }
    f = nullptr;
  }
};

constexpr char kElemMagic[8] = {'E', 'X', 'A', 'E', 'L', 'E', 'M', '1'};
constexpr int kNumVerts[4] = {4, 5, 6, 8};

}  // namespace

// ---- .cells / .scalars -------------------------------------------------------
void readCells(const std::string &path, std::vector<AmrCell> &cells) {
  File in(path, "rb");
  const int64_t n = in.size();
  if (n % (int64_t)sizeof(AmrCell))
    throw std::runtime_error("'" + path + "': size is not a multiple of 16 bytes (not a .cells file?)");
  cells.resize(n / sizeof(AmrCell));
  in.read(cells.data(), n, "cells");
}

void writeCells(const std::string &path, const std::vector<AmrCell> &cells) {
  File out(path, "wb");
  out.write(cells.data(), cells.size() * sizeof(AmrCell));
  out.close();
}

void readScalars(const std::string &path, std::vector<float> &s) {
  File in(path, "rb");
  const int64_t n = in.size();
  s.resize(n / sizeof(float));  // trailing bytes ignored, as the reference does
  in.read(s.data(), s.size() * sizeof(float), "scalars");
}

void writeScalars(const std::string &path, const std::vector<float> &s) {
  File out(path, "wb");
  out.write(s.data(), s.size() * sizeof(float));
  out.close();
}

// ---- element mesh ------------------------------------------------------------
void readElementMesh(const std::string &path, ElementMesh &m) {
  File in(path, "rb");
  char magic[8];
  in.read(magic, 8, "header");
  if (memcmp(magic, kElemMagic, 8) != 0)
    throw std::runtime_error("'" + path +
                             "': not an element-mesh file (EXAELEM1); umesh-library files must be converted");
  uint64_t h[6];
  in.read(h, sizeof(h), "header");
  const int64_t total = in.size();
  fseeko(in.f, 8 + sizeof(h), SEEK_SET);
  // every count below 2^31 before any product is formed, so the size sum
  // cannot wrap around to the file length
  for (int t = 0; t < 5; ++t)
    if (h[t] >= (1ull << 31)) throw std::runtime_error("'" + path + "': a count in the header exceeds 2^31-1");
  uint64_t need = h[0] * 12 + ((h[5] & 1) ? h[0] * 4 : 0) + ((h[5] & 2) ? h[0] * 4 : 0);
  for (int t = 0; t < 4; ++t) need += h[1 + t] * 4 * kNumVerts[t];
  if (need + 8 + sizeof(h) != (uint64_t)total)
    throw std::runtime_error("'" + path + "': sizes in the header do not match the file");
  m.xyz.resize(3 * h[0]);
  in.read(m.xyz.data(), m.xyz.size() * 4, "vertices");
  m.vertexTag.clear();
  m.values.clear();
  if (h[5] & 1) {
    m.vertexTag.resize(h[0]);
    in.read(m.vertexTag.data(), h[0] * 4, "vertex tags");
  }
  if (h[5] & 2) {
    m.values.resize(h[0]);
    in.read(m.values.data(), h[0] * 4, "vertex values");
  }
  for (int t = 0; t < 4; ++t) {
    m.elems[t].resize(h[1 + t] * kNumVerts[t]);
    in.read(m.elems[t].data(), m.elems[t].size() * 4, "elements");
  }
}

void writeElementMesh(const std::string &path, const ElementMesh &m) {
  const uint64_t nv = m.xyz.size() / 3;
  if (!m.vertexTag.empty() && m.vertexTag.size() != nv)
    throw std::runtime_error("element mesh: vertexTag size mismatch");
  if (!m.values.empty() && m.values.size() != nv) throw std::runtime_error("element mesh: values size mismatch");
  uint64_t h[6] = {nv, 0, 0, 0, 0, (uint64_t)(m.vertexTag.empty() ? 0 : 1) | (m.values.empty() ? 0 : 2)};
  for (int t = 0; t < 4; ++t) {
    if (m.elems[t].size() % kNumVerts[t]) throw std::runtime_error("element mesh: ragged element list");
    h[1 + t] = m.elems[t].size() / kNumVerts[t];
  }
  File out(path, "wb");
  out.write(kElemMagic, 8);
  out.write(h, sizeof(h));
  out.write(m.xyz.data(), m.xyz.size() * 4);
  out.write(m.vertexTag.data(), m.vertexTag.size() * 4);
  out.write(m.values.data(), m.values.size() * 4);
  for (int t = 0; t < 4; ++t) out.write(m.elems[t].data(), m.elems[t].size() * 4);
  out.close();
}

// ---- cubes -------------------------------------------------------------------
namespace {
struct CubeFileRec {
  float lower[3];
  int32_t level;
  int32_t scalarIDs[8];
};
static_assert(sizeof(CubeFileRec) == 48, "cube records are 48 bytes");
}  // namespace

void writeCubes(const std::string &prefix, const std::vector<CubeRec> &cubes) {
  int maxLevel = -1;
  for (const CubeRec &c : cubes) maxLevel = std::max(maxLevel, c.level);
  for (int L = 0; L <= maxLevel; ++L) {
    std::vector<CubeFileRec> recs;
    for (const CubeRec &c : cubes) {
      if (c.level != L) continue;
      CubeFileRec r;
      // the lowest corner vertex, a cell centre (makeDual.cpp:474-477)
      const float w = (float)(1 << L);
      for (int k = 0; k < 3; ++k) r.lower[k] = (float)c.cid[k] * w + 0.5f * w;
      r.level = L;
      memcpy(r.scalarIDs, c.scalarIDs, sizeof(r.scalarIDs));
      recs.push_back(r);
    }
    if (recs.empty()) continue;
    File out(prefix + "_" + std::to_string(L) + ".cubes", "wb");
    out.write(recs.data(), recs.size() * sizeof(CubeFileRec));
    out.close();
  }
}

int cubesFileLevel(const std::string &path) {
  // makeGrids.cpp:268-282: the text after the last '_' reads as "_%i.cubes"
  const size_t us = path.rfind('_');
  if (us == std::string::npos) throw std::runtime_error("'" + path + "' is not a cubes file (no _<level>)");
  const char *p = path.c_str() + us + 1;
  char *end = nullptr;
  const long level = strtol(p, &end, 0);
  if (end == p || level < 0 || level > 15)
    throw std::runtime_error("'" + path + "' is not a cubes file (bad level)");
  return (int)level;
}

void readCubes(const std::string &path, std::vector<CubeRec> &cubes) {
  const int level = cubesFileLevel(path);
  File in(path, "rb");
  const int64_t n = in.size();
  if (n % (int64_t)sizeof(CubeFileRec))
    throw std::runtime_error("'" + path + "': size is not a multiple of 48 bytes");
  std::vector<CubeFileRec> recs(n / sizeof(CubeFileRec));
  in.read(recs.data(), n, "cubes");
  cubes.resize(recs.size());
  const int w = 1 << level;
  for (size_t i = 0; i < recs.size(); ++i) {
    CubeRec &c = cubes[i];
    c.level = level;  // the brick level comes from the file name (makeGrids.cpp:292-293)
    for (int k = 0; k < 3; ++k) {
      const float lo = recs[i].lower[k];
      if (!(fabsf(lo) < 2e9f)) throw std::runtime_error("'" + path + "': cube position out of range");
      int32_t cid = (int32_t)lo;  // cellID (makeGrids.cpp:54-66)
      if (lo < 0.f) cid -= w - 1;
      c.cid[k] = cid / w;
    }
    memcpy(c.scalarIDs, recs[i].scalarIDs, sizeof(c.scalarIDs));
  }
}

// ---- grids -------------------------------------------------------------------
void writeGrids(const std::string &path, const std::vector<Gridlet> &gridlets, const std::vector<int32_t> &ids,
                bool append) {
  File out(path, append ? "ab" : "wb");
  for (const Gridlet &g : gridlets) {
    const int64_t ns = (int64_t)(g.dims[0] + 1) * (g.dims[1] + 1) * (g.dims[2] + 1);
    if ((int64_t)g.begin + ns > (int64_t)ids.size()) throw std::runtime_error("writeGrids: scalar IDs out of range");
    out.write(g.lower, 12);
    out.write(&g.level, 4);
    out.write(g.dims, 12);
    out.write(&ids[g.begin], ns * 4);
  }
  out.close();
}

void makeGridsFiles(const std::vector<std::string> &cubesPaths, const std::string &gridsPath, int64_t *conflicts) {
  int64_t bad = 0;
  for (size_t f = 0; f < cubesPaths.size(); ++f) {
    std::vector<CubeRec> cubes;
    readCubes(cubesPaths[f], cubes);
    std::vector<Gridlet> gridlets;
    std::vector<int32_t> ids;
    BrickSink sink;
    sink.allocate = [&](int64_t total) { ids.assign(total, -1); };
    sink.brick = [&](int64_t, int64_t begin, const int32_t *src, int64_t n) {
      memcpy(&ids[begin], src, n * sizeof(int32_t));
    };
    makeBricks(cubes, gridlets, sink, bad);
    writeGrids(gridsPath, gridlets, ids, f > 0);  // the first file truncates (makeGrids.cpp:324-328)
  }
  if (cubesPaths.empty()) writeGrids(gridsPath, {}, {}, false);
  if (conflicts) *conflicts = bad;
}

// ---- model files (ExaStitchModel.cpp:24-257) -----------------------------------
void loadModelFiles(const std::string &elemPath, const std::string &gridsPath, const std::string &scalarsPath,
                    HostModel &m) {
  m.vertices.clear();
  m.indices.clear();
  m.gridlets.clear();
  m.gridletScalars.clear();
  std::vector<float> scalars;
  if (!scalarsPath.empty()) readScalars(scalarsPath, scalars);
  if (!elemPath.empty()) {
    ElementMesh em;
    readElementMesh(elemPath, em);
    const int64_t nv = (int64_t)em.xyz.size() / 3;
    // value: scalars[vertexTag] when both exist, else the per-vertex value,
    // else 0 (ExaStitchModel.cpp:64-76)
    m.vertices.resize(4 * nv);
    for (int64_t i = 0; i < nv; ++i) {
      float value = 0.f;
      if (!scalars.empty() && !em.vertexTag.empty()) {
        const int32_t tag = em.vertexTag[i];
        if ((uint32_t)tag >= scalars.size())
          throw std::runtime_error("'" + elemPath + "': vertex tag " + std::to_string(tag) +
                                   " beyond the scalars file");
        value = scalars[tag];
      } else if (!em.values.empty()) {
        value = em.values[i];
      }
      for (int k = 0; k < 3; ++k) m.vertices[4 * i + k] = em.xyz[3 * i + k];
      m.vertices[4 * i + 3] = value;
    }
    int64_t ne = 0;
    for (int t = 0; t < 4; ++t) ne += (int64_t)em.elems[t].size() / kNumVerts[t];
    m.indices.assign(8 * ne, -1);
    int64_t e = 0;
    for (int t = 0; t < 4; ++t)  // tets, pyramids, wedges, hexes (ExaStitchModel.cpp:105-108)
      for (size_t i = 0; i < em.elems[t].size(); i += kNumVerts[t], ++e)
        for (int j = 0; j < kNumVerts[t]; ++j) {
          const int32_t v = em.elems[t][i + j];
          if (v < 0 || v >= nv) throw std::runtime_error("'" + elemPath + "': vertex index out of range");
          m.indices[8 * e + j] = v;
        }
  }
  if (!gridsPath.empty()) {
    // compact unreferenced vertices, keeping their order (ExaStitchModel.cpp:120-156)
    const int64_t nv = (int64_t)m.vertices.size() / 4;
    std::vector<int32_t> remap(nv, -1);
    for (int32_t v : m.indices)
      if (v >= 0) remap[v] = 0;
    int32_t next = 0;
    for (int64_t i = 0; i < nv; ++i)
      if (remap[i] >= 0) {
        for (int k = 0; k < 4; ++k) m.vertices[4 * (int64_t)next + k] = m.vertices[4 * i + k];
        remap[i] = next++;
      }
    m.vertices.resize(4 * (int64_t)next);
    for (int32_t &v : m.indices)
      if (v >= 0) v = remap[v];
    // gridlets: records until EOF; a scalar ID beyond the scalars is a hole
    // (NaN) (ExaStitchModel.cpp:186-247)
    File in(gridsPath, "rb");
    std::vector<int32_t> ids;
    for (;;) {
      int32_t hdr[7];
      const size_t got = in.readSome(hdr, sizeof(hdr));
      if (got == 0) break;
      if (got != sizeof(hdr)) throw std::runtime_error("'" + gridsPath + "': truncated gridlet header");
      Gridlet g;
      memcpy(g.lower, hdr, 12);
      g.level = hdr[3];
      memcpy(g.dims, hdr + 4, 12);
      if (g.dims[0] < 1 || g.dims[1] < 1 || g.dims[2] < 1 || g.dims[0] > 1024 || g.dims[1] > 1024 ||
          g.dims[2] > 1024 || g.level < 0 || g.level > 20)
        throw std::runtime_error("'" + gridsPath + "': gridlet " + std::to_string(m.gridlets.size()) +
                                 " has invalid level/dims");
      const int64_t ns = (int64_t)(g.dims[0] + 1) * (g.dims[1] + 1) * (g.dims[2] + 1);
      ids.resize(ns);
      in.read(ids.data(), ns * 4, "gridlet scalar IDs");
      if ((int64_t)m.gridletScalars.size() + ns >= ((int64_t)1 << 32))
        throw std::runtime_error("'" + gridsPath + "': gridlet scalars exceed 32-bit offsets");
      g.begin = (uint32_t)m.gridletScalars.size();
      for (int64_t i = 0; i < ns; ++i)
        m.gridletScalars.push_back((uint32_t)ids[i] < scalars.size() ? scalars[ids[i]] : NAN);
      m.gridlets.push_back(g);
    }
  }
  finalizeModel(m);
}

// ---- transfer function (.xf) ---------------------------------------------------
static const uint64_t kXfMagic = 0x1235abc000ull;  // headless.cpp:250

void readXf(const std::string &path, XfFile &xf) {
  File in(path, "rb");
  uint64_t magic = 0;
  in.read(&magic, 8, "header");
  if (magic != kXfMagic) throw std::runtime_error("'" + path + "': not a .xf transfer function file");
  float p[5];
  in.read(p, sizeof(p), "header");
  int32_t n = 0;
  in.read(&n, 4, "header");
  if (n <= 0 || n > (1 << 20)) throw std::runtime_error("'" + path + "': bad color map size");
  xf.opacityScale = p[0];
  xf.absDomain[0] = p[1];
  xf.absDomain[1] = p[2];
  xf.relDomain[0] = p[3];
  xf.relDomain[1] = p[4];
  xf.colorMap.resize(4 * (size_t)n);
  in.read(xf.colorMap.data(), xf.colorMap.size() * 4, "color map");
}

void writeXf(const std::string &path, const XfFile &xf) {
  if (xf.colorMap.size() % 4) throw std::runtime_error("xf: color map is not RGBA");
  File out(path, "wb");
  const uint64_t magic = kXfMagic;
  out.write(&magic, 8);
  const float p[5] = {xf.opacityScale, xf.absDomain[0], xf.absDomain[1], xf.relDomain[0], xf.relDomain[1]};
  out.write(p, sizeof(p));
  const int32_t n = (int32_t)(xf.colorMap.size() / 4);
  out.write(&n, 4);
  out.write(xf.colorMap.data(), xf.colorMap.size() * 4);
  out.close();
}

}  // namespace exa
