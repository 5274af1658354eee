// bricks_device.h -- device side of the ExaBricks comparison sampler: the
// basis-function blend over an active brick region (ABR) and the walk over
// the ABR kd-tree that yields the regions a ray crosses, front to back.
#pragma once
#include "exa_render_common.h"
#include "bricks_kernels.h"

namespace exa {

// A ray walk never visits more regions than this (a bound for degenerate
// input; real rays cross far fewer).
constexpr int kMaxAbrVisits = 1 << 20;

EXA_DEV float comp(vec3f v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// sampler/ExaBrickSampler.h:102-202 addBasisFunctions: the tent basis
// function of every cell of the brick whose support contains pos, weight
// (1-f) or f per axis, products in z, y, x order.
EXA_DEV void addBasisFunctions(const BrickView &v, int32_t brickID, vec3f pos, float &sumWeightedValues,
                               float &sumWeights) {
  const int4 a = v.bricks[2 * brickID];
  const int4 b = v.bricks[2 * brickID + 1];
  const int32_t sx = a.w, sy = b.x, sz = b.y;
  const int32_t begin = b.w;
  const float cw = (float)(1 << b.z);
  const vec3f lp = mk3((pos.x - (float)a.x) / cw - 0.5f, (pos.y - (float)a.y) / cw - 0.5f,
                       (pos.z - (float)a.z) / cw - 0.5f);
  int32_t lx = sat_i32(floorf(lp.x)), ly = sat_i32(floorf(lp.y)), lz = sat_i32(floorf(lp.z));
  lx = lx < -1 ? -1 : lx;
  ly = ly < -1 ? -1 : ly;
  lz = lz < -1 ? -1 : lz;
  const int32_t hx = lx + 1, hy = ly + 1, hz = lz + 1;
  const vec3f f = mk3(lp.x - (float)lx, lp.y - (float)ly, lp.z - (float)lz);
  const vec3f nf = mk3(1.f - f.x, 1.f - f.y, 1.f - f.z);
  const float *S = v.scalars + begin;
  auto add = [&](int32_t x, int32_t y, int32_t z, float w) {
    const float s = S[x + y * sx + z * sx * sy];
    sumWeights += w;
    sumWeightedValues += w * s;
  };
  const bool xl = lx >= 0 && lx < sx, xh = hx < sx;
  if (lz >= 0 && lz < sz) {
    if (ly >= 0 && ly < sy) {
      if (xl) add(lx, ly, lz, nf.z * nf.y * nf.x);
      if (xh) add(hx, ly, lz, nf.z * nf.y * f.x);
    }
    if (hy < sy) {
      if (xl) add(lx, hy, lz, nf.z * f.y * nf.x);
      if (xh) add(hx, hy, lz, nf.z * f.y * f.x);
    }
  }
  if (hz < sz) {
    if (ly >= 0 && ly < sy) {
      if (xl) add(lx, ly, hz, f.z * nf.y * nf.x);
      if (xh) add(hx, ly, hz, f.z * nf.y * f.x);
    }
    if (hy < sy) {
      if (xl) add(lx, hy, hz, f.z * f.y * nf.x);
      if (xh) add(hx, hy, hz, f.z * f.y * f.x);
    }
  }
}

// sample() in ABR mode (sampler/ExaBrickSampler.h:208-226): the blend over
// the region's bricks, ascending brick IDs; nBricks counts the records read.
EXA_DEV float abrValue(const BrickView &v, int32_t abr, vec3f pos, uint32_t &nBricks) {
  const int32_t begin = __float_as_int(v.abrLo[abr].w);
  const int32_t count = __float_as_int(v.abrHi[abr].w);
  float sumWeightedValues = 0.f, sumWeights = 0.f;
  for (int32_t i = 0; i < count; ++i) addBasisFunctions(v, v.leafList[begin + i], pos, sumWeightedValues, sumWeights);
  nBricks += (uint32_t)count;
  return sumWeightedValues / sumWeights;
}

EXA_DEV bool abrCulled(const BrickView &v, int32_t abr) { return v.abrMaxOp && v.abrMaxOp[abr] == 0.f; }

// boxTest (sampler/Sampler.h:103-118) of a ray against a region's domain
EXA_DEV bool abrBoxTest(const BrickView &v, int32_t abr, vec3f org, vec3f dir, float tmin, float tmax, float &t0,
                        float &t1) {
  const float4 lo4 = v.abrLo[abr], hi4 = v.abrHi[abr];
  const vec3f lo = (mk3(lo4.x, lo4.y, lo4.z) - org) / dir;
  const vec3f hi = (mk3(hi4.x, hi4.y, hi4.z) - org) / dir;
  t0 = fmaxf(tmin, reduce_max(vmin(lo, hi)));
  t1 = fminf(tmax, reduce_min(vmax(lo, hi)));
  return t0 < t1;
}

// The closest hit of a ray from tq among the regions not culled
// (ExaBrickABRGeomIsect, sampler/programs/ExaBrickSampler.cu:45-67, under
// deviceCode.cu:527-547 traverse), answered by walking the kd-tree front to
// back: the first region whose box test from tq passes.
//
// Restart form: each step descends from the root to the leaf holding the
// walk parameter ta (toward the lower side where ta is strictly before the
// plane, else the upper side, by the direction's sign), whose kd interval
// ends at tb > ta; the next step starts at tb.  A region whose box test
// passes lies on the near side of every plane at which its entry is later,
// so the walk reaches it before any region entered later.
//
// The walk below gives the same leaves without the restarts and keeps its
// place between queries.  A descent pushes the far child of every plane it
// passes on the near side while the far side is still inside the interval
// (the entry keeps the interval's end); the planes of the entries strictly
// decrease toward the top, and the top's plane is the exit of the leaf
// reached, so popping it continues exactly where a restart from the root at
// that exit would go.  The stack holds kWalkStack entries; when it runs empty
// the walk restarts at the root (needed only after an entry had to be
// dropped).  A query from tq after a region [t0, t1] returned at tq >= t1
// continues at that region's exit: the leaves in between are that region
// (whose box test fails from tq) or lie before tq.
constexpr int kWalkStack = 6;
struct AbrWalk {
  float ta, tb, tmax;
  int32_t node;
  int32_t sp;
  bool dropped;
  int32_t sn[kWalkStack];
  float st[kWalkStack];
};

EXA_DEV void walkBegin(AbrWalk &w, float tmin, float tmax) {
  w.ta = tmin;
  w.tb = tmax;
  w.tmax = tmax;
  w.node = 0;
  w.sp = 0;
  w.dropped = false;
}

// One plane of a descent at node n toward the leaf holding ta: pushes the far
// child with the interval's end when the far side is still inside it (the
// stack's top is entry 0, static register moves) and returns the child to
// descend into.
EXA_DEV int32_t walkDescend(AbrWalk &w, int2 n, int32_t node, vec3f org, vec3f dir, float ta, float &tb) {
  const int dim = n.y & 3;
  const float d = comp(dir, dim);
  const float ts = (__int_as_float(n.x) - comp(org, dim)) / d;
  const int32_t lowerChild = n.y >> 2, upperChild = node + 1;
  const bool neg = signbit(d);
  const int32_t nearC = neg ? upperChild : lowerChild, farC = neg ? lowerChild : upperChild;
  if (!(ta < ts)) return farC;
  if (ts < tb) {
    if (w.sp == kWalkStack) w.dropped = true;
    else ++w.sp;
#pragma unroll
    for (int k = kWalkStack - 1; k > 0; --k) {
      w.sn[k] = w.sn[k - 1];
      w.st[k] = w.st[k - 1];
    }
    w.sn[0] = farC;
    w.st[0] = tb;
    tb = ts;
  }
  return nearC;
}

// The continuation after a leaf whose kd interval ends at tb: the stack's top,
// or the root (an empty stack with nothing dropped leaves tb == tmax, which
// ends the walk).
EXA_DEV void walkLeafDone(AbrWalk &w, float tb) {
  w.ta = tb;
  if (w.sp > 0) {
    w.node = w.sn[0];
    w.tb = w.st[0];
#pragma unroll
    for (int k = 0; k < kWalkStack - 1; ++k) {
      w.sn[k] = w.sn[k + 1];
      w.st[k] = w.st[k + 1];
    }
    --w.sp;
  } else {
    w.node = 0;
    w.tb = w.tmax;
    w.dropped = false;
  }
}

// The first region not culled whose box test from tq passes, or -1.
EXA_DEV int32_t walkNext(const BrickView &v, AbrWalk &w, vec3f org, vec3f dir, float tq, float &t0, float &t1,
                         uint32_t &leaves) {
  for (int guard = 0; guard < kMaxAbrVisits && w.ta < w.tmax; ++guard) {
    int32_t node = w.node;
    float tb = w.tb;
    const float ta = w.ta;
    int2 n = v.nodes[node];
    while ((n.y & 3) != 3) {
      node = walkDescend(w, n, node, org, dir, ta, tb);
      n = v.nodes[node];
    }
    ++leaves;
    walkLeafDone(w, tb);
    const int32_t abr = (n.y >> 2) - 1;
    if (abr >= 0 && !abrCulled(v, abr) && abrBoxTest(v, abr, org, dir, tq, w.tmax, t0, t1)) return abr;
  }
  return -1;
}

// walkNext in bounded pieces: at most `budget` kd nodes (planes descended and
// leaves reached) per call.  Returns kWalkPending when the budget runs out
// before a region is found or the walk ends, with the walk's place kept in w
// (the descent's node and interval end) so that the next call continues it;
// `guard` counts the leaves reached since the last completed call (the
// caller resets it to 0 after a result other than kWalkPending).  A sequence
// of calls does exactly the work of one walkNext and returns its result.
constexpr int32_t kWalkPending = -2;
EXA_DEV int32_t walkNextBounded(const BrickView &v, AbrWalk &w, vec3f org, vec3f dir, float tq, float &t0,
                                float &t1, uint32_t &leaves, int budget, int32_t &guard) {
  while (guard < kMaxAbrVisits && w.ta < w.tmax) {
    int32_t node = w.node;
    float tb = w.tb;
    const float ta = w.ta;
    int2 n = v.nodes[node];
    while ((n.y & 3) != 3) {
      if (budget-- <= 0) {
        w.node = node;
        w.tb = tb;
        return kWalkPending;
      }
      node = walkDescend(w, n, node, org, dir, ta, tb);
      n = v.nodes[node];
    }
    ++guard;
    ++leaves;
    walkLeafDone(w, tb);
    const int32_t abr = (n.y >> 2) - 1;
    if (abr >= 0 && !abrCulled(v, abr) && abrBoxTest(v, abr, org, dir, tq, w.tmax, t0, t1)) return abr;
    if (budget-- <= 0) return kWalkPending;
  }
  return -1;
}

// A region whose closed domain contains p and that is not culled: the leaf
// reached by descending to the side of each plane p lies on, trying the
// other side too where p lies on the plane (kTieStack such branches at most).
constexpr int kTieStack = 8;
EXA_DEV int32_t abrContaining(const BrickView &v, vec3f p) {
  int32_t stack[kTieStack];
  int sp = 0;
  int32_t node = 0;
  for (int guard = 0; guard < kMaxAbrVisits; ++guard) {
    int2 n = v.nodes[node];
    while ((n.y & 3) != 3) {
      const int dim = n.y & 3;
      const float s = __int_as_float(n.x);
      const float c = comp(p, dim);
      const int32_t lowerChild = n.y >> 2, upperChild = node + 1;
      if (c < s) {
        node = lowerChild;
      } else {
        if (c == s && sp < kTieStack) stack[sp++] = lowerChild;
        node = upperChild;
      }
      n = v.nodes[node];
    }
    const int32_t abr = (n.y >> 2) - 1;
    if (abr >= 0 && !abrCulled(v, abr)) {
      const float4 lo = v.abrLo[abr], hi = v.abrHi[abr];
      if (p.x >= lo.x && p.y >= lo.y && p.z >= lo.z && p.x <= hi.x && p.y <= hi.y && p.z <= hi.z) return abr;
    }
    if (sp == 0) return -1;
    node = stack[--sp];
  }
  return -1;
}

}  // namespace exa
