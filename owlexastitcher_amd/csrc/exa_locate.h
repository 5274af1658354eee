// exa_locate.h -- device point location over gridlets and stitching
// elements: the software replacement of the zero-length OptiX ray of
// sampler/ExaStitchSampler.h:128-140 and of testSample (109-126).
//
// Winner among several accepting primitives (OptiX leaves this to traversal
// order): the lowest-ID accepting gridlet, else the lowest-ID accepting
// element whose vertex AABB contains the point, else a miss {-1,-1,0}.
// Primitives whose max opacity is 0 under the current transfer function are
// skipped (programs/ExaStitchSampler.cu:35-37,81-83 give them empty bounds).
#pragma once
#include "exa_kernels.h"

namespace exa {

// Binary-BVH point query with a per-thread stack in LDS (stack[i*stride]).
// leaf(first, count) returns true to end the traversal.  The build caps the
// depth at kStackDepth, and a DFS stack holds at most one entry per level.
template <typename LeafFn>
EXA_DEV void bvhPointQuery(const BvhNode *__restrict__ nodes, const vec3f p, int32_t *stack,
                           const int stride, LeafFn &&leaf) {
  int node = 0, sp = 0;
  for (;;) {
    const float4 *nd = (const float4 *)(nodes + node);
    const float4 a = nd[0], b = nd[1], c = nd[2];
    const int4 d = ((const int4 *)nd)[3];
    const bool h0 = p.x >= a.x && p.y >= a.y && p.z >= a.z && p.x <= a.w && p.y <= b.x && p.z <= b.y;
    const bool h1 = p.x >= b.z && p.y >= b.w && p.z >= c.x && p.x <= c.y && p.y <= c.z && p.z <= c.w;
    int next = -1;
    if (h0) {
      if (d.z > 0) {
        if (leaf(d.x, d.z)) return;
      } else if (d.z == 0) {
        next = d.x;
      }
    }
    if (h1) {
      if (d.w > 0) {
        if (leaf(d.y, d.w)) return;
      } else if (d.w == 0) {
        if (next < 0) next = d.y;
        else stack[(sp++) * stride] = d.y;
      }
    }
    if (next < 0) {
      if (sp == 0) return;
      next = stack[(--sp) * stride];
    }
    node = next;
  }
}

// Element test with the AABB filter of StitchGeomBounds
// (programs/ExaStitchSampler.cu:73-94); returns the element type (0 tet,
// 1 pyramid, 2 wedge, 3 hex) on a hit, -1 otherwise.
EXA_DEV int testElement(float &value, vec3f P, const int32_t *__restrict__ idx8,
                        const float4 *__restrict__ vertices) {
  const int4 i0 = *(const int4 *)idx8;
  const int4 i1 = *(const int4 *)(idx8 + 4);
  const int nv = 4 + (i1.x >= 0) + (i1.y >= 0) + (i1.z >= 0) + (i1.w >= 0);
  float4 v[8];
  v[0] = vertices[i0.x];
  v[1] = vertices[i0.y];
  v[2] = vertices[i0.z];
  v[3] = vertices[i0.w];
  if (nv > 4) v[4] = vertices[i1.x];
  if (nv > 5) v[5] = vertices[i1.y];
  if (nv > 6) { v[6] = vertices[i1.z]; v[7] = vertices[i1.w]; }
  float lo[3] = {v[0].x, v[0].y, v[0].z}, hi[3] = {v[0].x, v[0].y, v[0].z};
#pragma unroll
  for (int i = 1; i < 8; ++i)
    if (i < nv) {
      lo[0] = fminf(lo[0], v[i].x); hi[0] = fmaxf(hi[0], v[i].x);
      lo[1] = fminf(lo[1], v[i].y); hi[1] = fmaxf(hi[1], v[i].y);
      lo[2] = fminf(lo[2], v[i].z); hi[2] = fmaxf(hi[2], v[i].z);
    }
  if (!(P.x >= lo[0] && P.y >= lo[1] && P.z >= lo[2] && P.x <= hi[0] && P.y <= hi[1] && P.z <= hi[2]))
    return -1;
  bool hit;
  if (nv == 4) hit = intersectTet(value, P, v[0], v[1], v[2], v[3]);
  else if (nv == 5) hit = intersectNewton<5>(value, P, v);
  else if (nv == 6) hit = intersectNewton<6>(value, P, v);
  else if (nv == 8) hit = intersectNewton<8>(value, P, v);
  else hit = false;
  if (!hit) return -1;
  return nv == 4 ? 0 : (nv == 5 ? 1 : (nv == 6 ? 2 : 3));
}

// Returns the sample; kind receives -1 miss, 0 gridlet, 1..4 tet/pyr/wedge/hex.
EXA_DEV Sample locatePoint(const SamplerView &sv, const vec3f p, int32_t *stack, const int stride,
                           int &kind) {
  Sample s{-1, -1, 0.f};
  kind = -1;
  if (sv.numGridlets > 0) {
    int32_t best = 0x7fffffff;
    bvhPointQuery(sv.gNodes, p, stack, stride, [&](int32_t first, int32_t count) -> bool {
      for (int32_t k = 0; k < count; ++k) {
        const int32_t g = sv.gRefs[first + k];
        if (g >= best) break;  // refs ascend within a leaf
        if (sv.gridletMaxOp && sv.gridletMaxOp[g] == 0.f) continue;
        float v;
        int cid;
        const int r = intersectGridlet(v, cid, p, sv.gridlets[g], sv.gridletScalars);
        if (r) {
          best = g;
          s.primID = g;
          s.cellID = cid;
          s.value = v;
          if (r == 2) return true;
          break;
        }
      }
      return false;
    });
    if (s.primID >= 0) {
      kind = 0;
      return s;
    }
  }
  if (sv.numElements > 0) {
    int32_t best = 0x7fffffff;
    int bestType = -1;
    bvhPointQuery(sv.eNodes, p, stack, stride, [&](int32_t first, int32_t count) -> bool {
      for (int32_t k = 0; k < count; ++k) {
        const int32_t e = sv.eRefs[first + k];
        if (e >= best) break;
        if (sv.elemMaxOp && sv.elemMaxOp[e] == 0.f) continue;
        float v;
        const int t = testElement(v, p, sv.indices + 8 * (int64_t)e, sv.vertices);
        if (t >= 0) {
          best = e;
          bestType = t;
          s.primID = e;
          s.cellID = -1;
          s.value = v;
          break;
        }
      }
      return false;
    });
    if (s.primID >= 0) kind = 1 + bestType;
  }
  return s;
}

// sampler/ExaStitchSampler.h:109-126 testSample
EXA_DEV Sample testSampleGridlet(const SamplerView &sv, const vec3f p, int32_t primID) {
  Sample s{-1, -1, 0.f};
  if (primID < 0) return s;
  float v;
  int cid;
  if (intersectGridlet(v, cid, p, sv.gridlets[primID], sv.gridletScalars)) {
    s.primID = primID;
    s.cellID = cid;
    s.value = v;
  }
  return s;
}

}  // namespace exa
