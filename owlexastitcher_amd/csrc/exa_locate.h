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

// The per-primitive max-opacity arrays a lookup culls with (nullptr: none).
// They follow from the transfer function, so a batched launch passes each
// lane its own.
struct Culling {
  const float *gridletMaxOp;
  const float *elemMaxOp;
};
EXA_DEV Culling cullingOf(const SamplerView &sv) { return Culling{sv.gridletMaxOp, sv.elemMaxOp}; }

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
// 1 pyramid, 2 wedge, 3 hex) on a hit, -1 otherwise.  kCheckBox = false when
// the caller has already tested the stored AABB (same values).
template <bool kCheckBox = true>
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
  if (kCheckBox) {
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
  }
  bool hit;
  if (nv == 4) hit = intersectTet(value, P, v[0], v[1], v[2], v[3]);
  else if (nv == 5 || nv == 6 || nv == 8) hit = intersectNewtonAny(value, P, v, nv);
  else hit = false;
  if (!hit) return -1;
  return nv == 4 ? 0 : (nv == 5 ? 1 : (nv == 6 ? 2 : 3));
}

// Direct-table point location (LocateTables), gridlet stage: per level, the
// gridlet list of macrocell floor((p/2^L - .5)/8).  A hit strictly inside a
// gridlet cell ends the search (dual cells of one mesh tile space, so no
// lower-ID gridlet can also hold the point); a boundary hit only bounds the
// IDs still worth testing.
EXA_DEV Sample locateTablesGridlets(const SamplerView &sv, const Culling &cu, const vec3f p) {
  Sample s{-1, -1, 0.f};
  int32_t best = 0x7fffffff;
  for (int L = 0; L < sv.numLevels; ++L) {
    const int32_t *D = sv.levelDims[L];
    const float inv = 1.f / (float)(1 << L);
    const float gx = p.x * inv - 0.5f, gy = p.y * inv - 0.5f, gz = p.z * inv - 0.5f;
    const float mxf = floorf(gx * 0.125f);
    const float myf = floorf(gy * 0.125f);
    const float mzf = floorf(gz * 0.125f);
    if (!(mxf >= 0.f && myf >= 0.f && mzf >= 0.f && mxf < (float)D[0] && myf < (float)D[1] &&
          mzf < (float)D[2]))
      continue;
    const int32_t idx = sv.levelOffset[L] + ((int32_t)mzf * D[1] + (int32_t)myf) * D[0] + (int32_t)mxf;
    if (sv.gTop) {
      // Owner-table fast path for points at least fastMargin of a cell away
      // from every dual-cell face of this level (so the gridlet test sees
      // fractions at least bandMargin inside, a band-2 hit).  There the
      // lowest-ID accepting gridlet of the level is the cell's owner.
      const float cx = floorf(gx), cy = floorf(gy), cz = floorf(gz);
      const float fx = gx - cx, fy = gy - cy, fz = gz - cz;
      const float m0 = sv.fastMargin, m1 = 1.f - sv.fastMargin;
      if (fx >= m0 && fx <= m1 && fy >= m0 && fy <= m1 && fz >= m0 && fz <= m1) {
        const int32_t blk = sv.gTop[idx];
        if (blk < 0) continue;
        const int32_t local = (((int32_t)cz & 7) * 8 + ((int32_t)cy & 7)) * 8 + ((int32_t)cx & 7);
        const int32_t g = sv.gOwner[(int64_t)blk * LocateTables::kOwnerBlock + local];
        if (g < 0) continue;
        if (!(cu.gridletMaxOp && cu.gridletMaxOp[g] == 0.f)) {
          float v;
          int cid;
          if (intersectGridlet(v, cid, p, sv.gridlets[g], sv.gridletScalars, sv.bandMargin) == 2) {
            s.primID = g;
            s.cellID = cid;
            s.value = v;
            return s;
          }
        }
        // culled owner: the list scan below finds the next accepting gridlet
      }
    }
    const uint32_t k1 = sv.gStart[idx + 1];
    for (uint32_t k = sv.gStart[idx]; k < k1; ++k) {
      const int32_t g = sv.gList[k];
      if (g >= best) break;
      if (cu.gridletMaxOp && cu.gridletMaxOp[g] == 0.f) continue;
      float v;
      int cid;
      const int r = intersectGridlet(v, cid, p, sv.gridlets[g], sv.gridletScalars, sv.bandMargin);
      if (r) {
        best = g;
        s.primID = g;
        s.cellID = cid;
        s.value = v;
        if (r == 2) return s;
        break;
      }
    }
  }
  return s;
}

// The Woodcock loop's sample lookup with the direct tables: testSample on
// the cached gridlet (ExaStitchSampler.h:109-126) and, when that misses,
// full point location (gridlet stage).  Same result as
// testSampleGridlet() followed by locateTablesGridlets(), but the cached
// gridlet's record and every level's owner tables are loaded together and a
// lane runs one gridlet test (on the cached gridlet or on the owner), so a
// lookup costs about four dependent memory round trips and one test.
// cachedHit tells which of the two answered.
constexpr int kFastLevels = 8;
EXA_DEV Sample lookupTablesGridlets(const SamplerView &sv, const Culling &cu, const vec3f p, int32_t cached,
                                    bool &cachedHit) {
  cachedHit = false;
  if (sv.numGridlets <= 0) return Sample{-1, -1, 0.f};  // no gridlet record to read
  const bool fast = sv.gTop != nullptr && sv.numLevels <= kFastLevels;
  // round 1: cached gridlet record, owner blocks of every level
  const Gridlet cg = sv.gridlets[cached >= 0 ? cached : 0];
  int32_t blk[kFastLevels], local[kFastLevels];
  // first level at which the owner tables cannot decide (point near a
  // dual-cell face): the list scan has to take over from there
  int undecided = kFastLevels;
#pragma unroll
  for (int L = 0; L < kFastLevels; ++L) {
    blk[L] = -1;
    local[L] = 0;
    if (!fast || L >= sv.numLevels) continue;
    const int32_t *D = sv.levelDims[L];
    const float inv = 1.f / (float)(1 << L);
    const float gx = p.x * inv - 0.5f, gy = p.y * inv - 0.5f, gz = p.z * inv - 0.5f;
    const float mxf = floorf(gx * 0.125f);
    const float myf = floorf(gy * 0.125f);
    const float mzf = floorf(gz * 0.125f);
    if (!(mxf >= 0.f && myf >= 0.f && mzf >= 0.f && mxf < (float)D[0] && myf < (float)D[1] &&
          mzf < (float)D[2]))
      continue;
    const float cx = floorf(gx), cy = floorf(gy), cz = floorf(gz);
    const float fx = gx - cx, fy = gy - cy, fz = gz - cz;
    const float m0 = sv.fastMargin, m1 = 1.f - sv.fastMargin;
    if (!(fx >= m0 && fx <= m1 && fy >= m0 && fy <= m1 && fz >= m0 && fz <= m1)) {
      if (undecided == kFastLevels) undecided = L;
      continue;
    }
    const int32_t idx = sv.levelOffset[L] + ((int32_t)mzf * D[1] + (int32_t)myf) * D[0] + (int32_t)mxf;
    blk[L] = sv.gTop[idx];
    local[L] = (((int32_t)cz & 7) * 8 + ((int32_t)cy & 7)) * 8 + ((int32_t)cx & 7);
  }
  // round 2: owners
  int32_t own[kFastLevels];
#pragma unroll
  for (int L = 0; L < kFastLevels; ++L)
    own[L] = blk[L] >= 0 ? sv.gOwner[(int64_t)blk[L] * LocateTables::kOwnerBlock + local[L]] : -1;
  // the lowest level whose owner answers, if no undecided level lies below
  int32_t g = -1;
  if (fast) {
#pragma unroll
    for (int L = kFastLevels - 1; L >= 0; --L)
      if (own[L] >= 0 && L < undecided) g = own[L];
  }
  // The cached gridlet's test accepts only points inside the gridlet's box,
  // the test's own first check.  So each lane runs one gridlet test: on the
  // cached gridlet where p lies in its box, else on the owner (rounds 3 and
  // 4: owner record and culling flag, then the corner scalars).  Only a
  // cached gridlet that rejects a point inside its box (a hole there) leads
  // to a second test, on the owner.
  vec3f clo, chi;
  gridletBounds(cg, clo, chi);
  const bool inCached = cached >= 0 && p.x >= clo.x && p.y >= clo.y && p.z >= clo.z && p.x <= chi.x &&
                        p.y <= chi.y && p.z <= chi.z;
  const bool haveOwner = fast && g >= 0;
  bool ownerTest = false;
  if (!inCached && haveOwner) ownerTest = !(cu.gridletMaxOp && cu.gridletMaxOp[g] == 0.f);
  float v = 0.f;
  int cid = -1;
  int r = 0;
  if (inCached || ownerTest) {
    Gridlet gt = cg;
    if (!inCached) gt = sv.gridlets[g];
    r = intersectGridlet(v, cid, p, gt, sv.gridletScalars, sv.bandMargin);
  }
  if (inCached) {
    if (r) {
      cachedHit = true;
      return Sample{cached, cid, v};
    }
    ownerTest = haveOwner && !(cu.gridletMaxOp && cu.gridletMaxOp[g] == 0.f);
    r = ownerTest ? intersectGridlet(v, cid, p, sv.gridlets[g], sv.gridletScalars, sv.bandMargin) : 0;
  }
  if (!fast) return locateTablesGridlets(sv, cu, p);
  if (g < 0) {
    if (undecided == kFastLevels) return Sample{-1, -1, 0.f};
    return locateTablesGridlets(sv, cu, p);
  }
  if (ownerTest && r == 2) return Sample{g, cid, v};
  return locateTablesGridlets(sv, cu, p);
}

// Element test on an element's own vertex record (8 float4 slots holding the
// same values as vertices[indices[...]]), behind the exact AABB check when
// the caller cannot rule it out (boxCheck): the box of its nv vertices (the
// float box the bins were built from: the same min/max over the same
// floats).  Then the arithmetic of testElement.  -1 or 0..3
// (tet/pyr/wedge/hex).
EXA_DEV int testElementBoxed(float &value, vec3f P, int nv, const float4 *__restrict__ rec, bool boxCheck) {
  if (nv < 4) return -1;
  // all 8 slots, whatever nv (the record is 128 bytes with zeros past nv):
  // one run of independent loads instead of loads behind per-lane branches
  float4 v[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = rec[i];
  if (boxCheck) {
    float lx = v[0].x, ly = v[0].y, lz = v[0].z, hx = lx, hy = ly, hz = lz;
#pragma unroll
    for (int i = 1; i < 8; ++i)
      if (i < nv) {
        lx = fminf(lx, v[i].x);
        ly = fminf(ly, v[i].y);
        lz = fminf(lz, v[i].z);
        hx = fmaxf(hx, v[i].x);
        hy = fmaxf(hy, v[i].y);
        hz = fmaxf(hz, v[i].z);
      }
    if (!(P.x >= lx && P.y >= ly && P.z >= lz && P.x <= hx && P.y <= hy && P.z <= hz)) return -1;
  }
  bool hit;
  if (nv == 4) hit = intersectTet(value, P, v[0], v[1], v[2], v[3]);
  else if (nv == 5 || nv == 6 || nv == 8) hit = intersectNewtonAny(value, P, v, nv);
  else hit = false;
  if (!hit) return -1;
  return nv == 4 ? 0 : (nv == 5 ? 1 : (nv == 6 ? 2 : 3));
}

// Direct-table point location, element stage: the element list of p's bin;
// kind receives 1..4 (tet/pyr/wedge/hex) or -1.  Each 12-byte entry holds the
// element's box quantised to 512 sub-cells of the bin, 9 bits per bound in
// 10-bit slots, lo with the slots' top bits clear and hi with them set.  With
// p's sub-cells in the same slots, one subtraction compares three bounds:
// slot by slot it computes 512 + a - b (1..1023, so no borrow crosses a slot)
// and the slot's top bit survives exactly when a >= b.  An entry whose box
// misses p's sub-cell is rejected without touching the element.  If p's
// sub-cell lies strictly inside the quantised box, p lies inside the float
// box; otherwise the exact box decides.  The quantised box is conservative,
// so the first accepting element in ID order is the one the float boxes give.
EXA_DEV Sample locateTablesElements(const SamplerView &sv, const Culling &cu, const vec3f p, int &kind) {
  constexpr uint32_t kTop = 0x20080200u;  // bit 9 of each 10-bit slot
  constexpr uint32_t kOne = 0x00100401u;  // bit 0 of each slot
  Sample s{-1, -1, 0.f};
  kind = -1;
  if (sv.numElements <= 0) return s;
  int32_t b[3];
  uint32_t pq = 0;
  const float q[3] = {p.x, p.y, p.z};
  const float inv512 = sv.binInv * 512.f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float f = floorf((q[k] - sv.binLo[k]) * sv.binInv);
    b[k] = f != f ? 0 : (f < 0.f ? 0 : (f > (float)(sv.binDims[k] - 1) ? sv.binDims[k] - 1 : (int32_t)f));
    const float g = floorf((q[k] - sv.binLo[k]) * inv512) - (float)b[k] * 512.f;
    pq |= (g != g ? 0u : (g < 0.f ? 0u : (g > 511.f ? 511u : (uint32_t)g))) << (10 * k);
  }
  const uint32_t pTop = pq | kTop;
  const int64_t bin = ((int64_t)b[2] * sv.binDims[1] + b[1]) * sv.binDims[0] + b[0];
  const uint32_t k0 = sv.eStart[bin], k1 = sv.eStart[bin + 1];
  // entries in ID order, four per batch of independent loads; the sub-cell
  // tests of a batch give a mask, then one exact test per set bit (lanes
  // test their own candidates together, whatever the element types)
  for (uint32_t k = k0; k < k1; k += 4) {
    uint3 c[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (k + j < k1) c[j] = sv.eCand[k + j];
    unsigned hits = 0, inner = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t dLo = pTop - c[j].y, dHi = c[j].z - pq;
      if (k + j < k1 && (dLo & dHi & kTop) == kTop) hits |= 1u << j;
      if (((dLo - kOne) & (dHi - kOne) & kTop) == kTop) inner |= 1u << j;
    }
    while (hits) {
      const int j = __builtin_ctz(hits);
      hits &= hits - 1;
      // the candidate's head word (selects, not an indexed array)
      const uint32_t h = j == 0 ? c[0].x : (j == 1 ? c[1].x : (j == 2 ? c[2].x : c[3].x));
      const int32_t e = (int32_t)(h & 0x0fffffffu);
      if (cu.elemMaxOp && cu.elemMaxOp[e] == 0.f) continue;
      float v;
      const int t = testElementBoxed(v, p, (int)(h >> 28), sv.eVerts + 8 * (int64_t)e, !((inner >> j) & 1u));
      if (t >= 0) {
        s.primID = e;
        s.cellID = -1;
        s.value = v;
        kind = 1 + t;
        return s;
      }
    }
  }
  return s;
}

// BVH point location, gridlet stage.
EXA_DEV Sample locateBvhGridlets(const SamplerView &sv, const Culling &cu, const vec3f p, int32_t *stack,
                                 const int stride) {
  Sample s{-1, -1, 0.f};
  if (sv.numGridlets <= 0) return s;
  int32_t best = 0x7fffffff;
  bvhPointQuery(sv.gNodes, p, stack, stride, [&](int32_t first, int32_t count) -> bool {
    for (int32_t k = 0; k < count; ++k) {
      const int32_t g = sv.gRefs[first + k];
      if (g >= best) break;  // refs ascend within a leaf
      if (cu.gridletMaxOp && cu.gridletMaxOp[g] == 0.f) continue;
      float v;
      int cid;
      const int r = intersectGridlet(v, cid, p, sv.gridlets[g], sv.gridletScalars, sv.bandMargin);
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
  return s;
}

// BVH point location, element stage; kind receives 1..4 or -1.
EXA_DEV Sample locateBvhElements(const SamplerView &sv, const Culling &cu, const vec3f p, int32_t *stack,
                                 const int stride,
                                 int &kind) {
  Sample s{-1, -1, 0.f};
  kind = -1;
  if (sv.numElements <= 0) return s;
  int32_t best = 0x7fffffff;
  int bestType = -1;
  bvhPointQuery(sv.eNodes, p, stack, stride, [&](int32_t first, int32_t count) -> bool {
    for (int32_t k = 0; k < count; ++k) {
      const int32_t e = sv.eRefs[first + k];
      if (e >= best) break;
      if (cu.elemMaxOp && cu.elemMaxOp[e] == 0.f) continue;
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
  return s;
}

// The two stages of point location, by accelerator.
template <int kAccel>
EXA_DEV Sample locateGridlets(const SamplerView &sv, const Culling &cu, const vec3f p, int32_t *stack,
                              const int stride) {
  if (kAccel == kAccelTables) return locateTablesGridlets(sv, cu, p);
  return locateBvhGridlets(sv, cu, p, stack, stride);
}
template <int kAccel>
EXA_DEV Sample locateElements(const SamplerView &sv, const Culling &cu, const vec3f p, int32_t *stack,
                              const int stride, int &kind) {
  if (kAccel == kAccelTables) return locateTablesElements(sv, cu, p, kind);
  return locateBvhElements(sv, cu, p, stack, stride, kind);
}

// Full point location; kind receives -1 miss, 0 gridlet, 1..4 tet/pyr/wedge/hex.
template <int kAccel>
EXA_DEV Sample locatePoint(const SamplerView &sv, const vec3f p, int32_t *stack, const int stride, int &kind) {
  const Culling cu = cullingOf(sv);
  const Sample s = locateGridlets<kAccel>(sv, cu, p, stack, stride);
  if (s.primID >= 0) {
    kind = 0;
    return s;
  }
  return locateElements<kAccel>(sv, cu, p, stack, stride, kind);
}

// sampler/ExaStitchSampler.h:109-126 testSample
EXA_DEV Sample testSampleGridlet(const SamplerView &sv, const vec3f p, int32_t primID) {
  Sample s{-1, -1, 0.f};
  if (primID < 0) return s;
  float v;
  int cid;
  if (intersectGridlet(v, cid, p, sv.gridlets[primID], sv.gridletScalars, sv.bandMargin)) {
    s.primID = primID;
    s.cellID = cid;
    s.value = v;
  }
  return s;
}

}  // namespace exa
