// bricks_kernels.h -- device view and kernel launch wrappers of the
// ExaBricks comparison sampler (render_bricks.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "exa_kernels.h"

namespace exa {

// ExaBrickSampler::LP (sampler/ExaBrickSampler.h:38-47) with the kd-tree in
// place of the OptiX BVH over the regions.
struct BrickView {
  const int4 *bricks;       // Brick records, 2 int4 each: lower.xyz size.x | size.yz level begin
  const float *scalars;
  const float4 *abrLo;      // per ABR: domain lower xyz, w = leafListBegin (int bits)
  const float4 *abrHi;      // per ABR: domain upper xyz, w = leafListSize (int bits)
  const float *abrCellWidth;  // finestLevelCellWidth
  const int32_t *leafList;
  const int2 *nodes;        // KdNode: split (float bits), info
  const float *abrMaxOp;    // nullptr: nothing culled
  int32_t numAbrs;
};

// point location: out[i] = {0, -1, value} inside a region that is not
// culled, else {0, 0, 0}
void launchBrickLocate(const BrickView &v, const float *pos, int64_t n, Sample *out, hipStream_t s);
// one frame of the ExaBricks renderer (shade mode ignored: Default).  With
// `work` (one uint32 of device memory, reset by the launch) the path tracer
// and direct lighting run in persistent lanes; the ray marcher, heat-map
// frames and work == nullptr run one thread per pixel; both give the same
// frame.  Counters: kCntGridletHit samples blended, kCntTetHit brick records
// read, kCntPyrHit kd leaves visited, kCntDdaCells regions entered,
// kCntTfFetch, kCntPixels.
void launchRenderBricks(const RenderParams &p, const BrickView &v, uint32_t *work, unsigned long long *counters,
                        hipStream_t s);

}  // namespace exa
