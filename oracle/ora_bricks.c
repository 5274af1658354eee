/*
 * ora_bricks.c -- ExaBricks active brick regions (model/ABRs.cpp), restated
 * in plain C.  TEST INFRASTRUCTURE ONLY (see exa_oracle.h).
 *
 * buildRec (ABRs.cpp:88-201): in a region, per axis the face of a clipped
 * brick domain (upper face, then lower face, brick by brick) nearest the
 * region's centre and strictly inside it; the split axis is the first, from
 * the widest axis on (owl arg_max: first index of the largest span), whose
 * candidate exists; both halves get the clipped domains with positive volume;
 * the upper half is built first.  addLeaf (40-86): the sorted, unique brick
 * IDs.  computeValueRange (204-270) and the finest level (320-337) follow.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "exa_oracle.h"

typedef struct {
  float lo[3], hi[3];
  int32_t id;
} Prim;

struct ora_abrs {
  ora_abr *abrs;
  int64_t n, cap;
  int32_t *leaf;
  int64_t nl, capl;
};

static int cmp_i32(const void *a, const void *b) {
  const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

static void add_leaf(struct ora_abrs *o, const Prim *prims, int64_t np, const float dom[6]) {
  for (int k = 0; k < 3; ++k)
    if (dom[k] >= dom[3 + k]) return;
  if (np == 0) return;
  int32_t *ids = (int32_t *)malloc(np * sizeof(int32_t));
  for (int64_t i = 0; i < np; ++i) ids[i] = prims[i].id;
  qsort(ids, np, sizeof(int32_t), cmp_i32);
  int64_t u = 0;
  for (int64_t i = 0; i < np; ++i)
    if (i == 0 || ids[i] != ids[i - 1]) ids[u++] = ids[i];
  if (o->n == o->cap) {
    o->cap = o->cap ? 2 * o->cap : 64;
    o->abrs = (ora_abr *)realloc(o->abrs, o->cap * sizeof(ora_abr));
  }
  while (o->nl + u > o->capl) {
    o->capl = o->capl ? 2 * o->capl : 256;
    o->leaf = (int32_t *)realloc(o->leaf, o->capl * sizeof(int32_t));
  }
  ora_abr *a = &o->abrs[o->n++];
  memset(a, 0, sizeof(*a));
  memcpy(a->domain, dom, sizeof(a->domain));
  a->leafListBegin = (int32_t)o->nl;
  a->leafListSize = (int32_t)u;
  memcpy(o->leaf + o->nl, ids, u * sizeof(int32_t));
  o->nl += u;
  free(ids);
}

/* takes ownership of prims */
static void build_rec(struct ora_abrs *o, Prim *prims, int64_t np, const float dom[6]) {
  if (np == 0) {
    free(prims);
    return;
  }
  for (int k = 0; k < 3; ++k)
    if (dom[3 + k] == dom[k]) {
      free(prims);
      return;
    }
  float tgt[3], bestPos[3], bestDist[3], span[3];
  for (int k = 0; k < 3; ++k) {
    tgt[k] = (dom[k] + dom[3 + k]) / 2.f;
    bestPos[k] = dom[k];
    bestDist[k] = dom[3 + k] - dom[k];
    span[k] = dom[3 + k] - dom[k];
  }
  for (int64_t i = 0; i < np; ++i)
    for (int dim = 0; dim < 3; ++dim)
      for (int side = 0; side < 2; ++side) {
        const float pos = side ? prims[i].lo[dim] : prims[i].hi[dim];
        if (pos <= dom[dim] || pos >= dom[3 + dim]) continue;
        const float dist = fabsf(tgt[dim] - pos);
        if (dist >= bestDist[dim]) continue;
        bestPos[dim] = pos;
        bestDist[dim] = dist;
      }
  int widest = 0;
  if (span[1] > span[widest]) widest = 1;
  if (span[2] > span[widest]) widest = 2;
  int splitDim = -1;
  float splitPos = 0.f;
  for (int i = 0; i < 3; ++i) {
    const int dim = (widest + i) % 3;
    if (bestPos[dim] <= dom[dim] || bestPos[dim] >= dom[3 + dim]) continue;
    splitDim = dim;
    splitPos = bestPos[dim];
    break;
  }
  if (splitDim < 0) {
    add_leaf(o, prims, np, dom);
    free(prims);
    return;
  }
  float domL[6], domR[6];
  memcpy(domL, dom, sizeof(domL));
  memcpy(domR, dom, sizeof(domR));
  domL[3 + splitDim] = splitPos;
  domR[splitDim] = splitPos;
  Prim *pl = (Prim *)malloc(np * sizeof(Prim)), *pr = (Prim *)malloc(np * sizeof(Prim));
  int64_t nl = 0, nr = 0;
  for (int64_t i = 0; i < np; ++i) {
    Prim cl, cr;
    cl.id = cr.id = prims[i].id;
    for (int k = 0; k < 3; ++k) {
      cl.lo[k] = fmaxf(prims[i].lo[k], domL[k]);
      cl.hi[k] = fminf(prims[i].hi[k], domL[3 + k]);
      cr.lo[k] = fmaxf(prims[i].lo[k], domR[k]);
      cr.hi[k] = fminf(prims[i].hi[k], domR[3 + k]);
    }
    if (cl.lo[0] < cl.hi[0] && cl.lo[1] < cl.hi[1] && cl.lo[2] < cl.hi[2]) pl[nl++] = cl;
    if (cr.lo[0] < cr.hi[0] && cr.lo[1] < cr.hi[1] && cr.lo[2] < cr.hi[2]) pr[nr++] = cr;
  }
  free(prims);
  build_rec(o, pr, nr, domR);
  build_rec(o, pl, nl, domL);
}

static void value_range(const ora_brick *bricks, const float *scalars, const int32_t *leaf, ora_abr *a) {
  float lo = INFINITY, hi = -INFINITY;
  for (int32_t i = 0; i < a->leafListSize; ++i) {
    const ora_brick *b = &bricks[leaf[a->leafListBegin + i]];
    const float cw = (float)(1 << b->level);
    for (int iz = 0; iz < b->size[2]; ++iz) {
      const float pz = (float)b->lower[2] + ((float)iz + .5f) * cw;
      if (!((pz - cw) <= a->domain[5] && (pz + cw) >= a->domain[2])) continue;
      for (int iy = 0; iy < b->size[1]; ++iy) {
        const float py = (float)b->lower[1] + ((float)iy + .5f) * cw;
        if (!((py - cw) <= a->domain[4] && (py + cw) >= a->domain[1])) continue;
        for (int ix = 0; ix < b->size[0]; ++ix) {
          const float px = (float)b->lower[0] + ((float)ix + .5f) * cw;
          if (!((px - cw) <= a->domain[3] && (px + cw) >= a->domain[0])) continue;
          const float s = scalars[(int64_t)b->begin + ix + (int64_t)b->size[0] * iy +
                                  (int64_t)b->size[0] * b->size[1] * iz];
          lo = fminf(lo, s);
          hi = fmaxf(hi, s);
        }
      }
    }
  }
  a->valueRange[0] = lo;
  a->valueRange[1] = hi;
}

ora_abrs *ora_abrs_build(const ora_brick *bricks, int64_t n, const float *scalars) {
  struct ora_abrs *o = (struct ora_abrs *)calloc(1, sizeof(*o));
  if (n <= 0) return o;
  Prim *prims = (Prim *)malloc(n * sizeof(Prim));
  float root[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = 0; i < n; ++i) {
    const ora_brick *b = &bricks[i];
    const float cw = (float)(1 << b->level);
    for (int k = 0; k < 3; ++k) {
      prims[i].lo[k] = (float)b->lower[k] - 0.5f * cw;
      prims[i].hi[k] = (float)b->lower[k] + ((float)b->size[k] + 0.5f) * cw;
      root[k] = fminf(root[k], prims[i].lo[k]);
      root[3 + k] = fmaxf(root[3 + k], prims[i].hi[k]);
    }
    prims[i].id = (int32_t)i;
  }
  build_rec(o, prims, n, root);
  for (int64_t r = 0; r < o->n; ++r) {
    ora_abr *a = &o->abrs[r];
    int finest = 1 << 30;
    for (int32_t i = 0; i < a->leafListSize; ++i) {
      const int lv = bricks[o->leaf[a->leafListBegin + i]].level;
      if (lv < finest) finest = lv;
    }
    a->finestLevelCellWidth = (float)(1 << finest);
    value_range(bricks, scalars, o->leaf, a);
  }
  return o;
}

int64_t ora_abrs_count(const ora_abrs *a, int what) { return what == 0 ? a->n : a->nl; }

void ora_abrs_get(const ora_abrs *a, ora_abr *abrs, int32_t *leafList) {
  if (abrs) memcpy(abrs, a->abrs, a->n * sizeof(ora_abr));
  if (leafList) memcpy(leafList, a->leaf, a->nl * sizeof(int32_t));
}

void ora_abrs_free(ora_abrs *a) {
  if (!a) return;
  free(a->abrs);
  free(a->leaf);
  free(a);
}
