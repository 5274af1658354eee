// exa_types.h -- plain data layouts shared by host C++ and HIP code.
//
// These mirror the reference's in-memory records so a model built for the
// reference is byte-compatible:
//   Gridlet      Gridlet.h:22-37               (32 bytes)
//   Sample       sampler/Sampler.h:72-76       (12 bytes)
//   element rows model/ExaStitchModel.cpp:79-108 (int[8], -1 padded,
//                tets, pyramids, wedges, hexes)
//   vertices     float4 xyz + scalar           (ExaStitchModel.cpp:64-76)
#pragma once
#include <stdint.h>

namespace exa {

struct Gridlet {
  int32_t lower[3];
  int32_t level;
  int32_t dims[3];
  uint32_t begin;
};
static_assert(sizeof(Gridlet) == 32, "Gridlet must stay 32 bytes");

struct Sample {
  int32_t primID;
  int32_t cellID;
  float value;
};
static_assert(sizeof(Sample) == 12, "Sample must stay 12 bytes");

// AMR leaf cell as stored in the reference's .cells files
// (tools/amrMakeDualMesh/makeDual.cpp:75-84,932-940): position in finest
// cell units (a multiple of 2^level) and level (0 = finest).
struct AmrCell {
  int32_t pos[3];
  int32_t level;
};
static_assert(sizeof(AmrCell) == 16, "AmrCell must stay 16 bytes");

}  // namespace exa
