/*
 * ora_dual.c -- CPU restatement of the offline dual-mesh generation
 * (tools/amrMakeDualMesh/makeDual.cpp) and gridlet clustering
 * (tools/amrMakeDualMesh/makeGrids.cpp), run serially.
 * TEST INFRASTRUCTURE ONLY (see exa_oracle.h).
 *
 * Ordering conventions (the reference is thread-order dependent here):
 *   - cells are processed in the reference's sorted order (makeDual.cpp:878);
 *   - elements are emitted in that processing order, octants in the
 *     dz,dy,dx loop order of makeDual.cpp:608-610; the model concatenates
 *     tets, pyramids, wedges, hexes (model/ExaStitchModel.cpp:105-108);
 *   - vertices are the element-referenced cells in ascending sorted-cell
 *     order (the compaction of ExaStitchModel.cpp:120-156 keeps the relative
 *     order of first emission, which the reference leaves to thread timing);
 *   - bricks are ordered by level, then by macrocell id (x, then y, then z)
 *     as a std::map<vec3i,...> iterates (makeGrids.cpp:191-209).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "exa_oracle.h"

typedef struct {
  int32_t pos[3];
  int32_t level;
  int32_t scalarID;
} Cell;

/* makeDual.cpp:115-137: LogicalCell compared as two little-endian uint64s */
static inline uint64_t key0(const Cell *c) {
  return (uint64_t)(uint32_t)c->pos[0] | ((uint64_t)(uint32_t)c->pos[1] << 32);
}
static inline uint64_t key1(const Cell *c) {
  return (uint64_t)(uint32_t)c->pos[2] | ((uint64_t)(uint32_t)c->level << 32);
}
static int logical_cmp(const Cell *a, const Cell *b) {
  uint64_t a0 = key0(a), b0 = key0(b);
  if (a0 != b0) return a0 < b0 ? -1 : 1;
  uint64_t a1 = key1(a), b1 = key1(b);
  if (a1 != b1) return a1 < b1 ? -1 : 1;
  return 0;
}
static int cell_cmp(const void *pa, const void *pb) {
  const Cell *a = (const Cell *)pa, *b = (const Cell *)pb;
  int c = logical_cmp(a, b);
  if (c) return c;
  return a->scalarID < b->scalarID ? -1 : (a->scalarID > b->scalarID ? 1 : 0);
}

typedef struct {
  Cell *cells;
  int64_t n;
  int minLevel, maxLevel;
} Exa;

/* makeDual.cpp:165-169 */
static int lower_on_level(float f, int level) {
  f = floorf(f / (float)(1 << level));
  return (int)(f * (float)(1 << level));
}

/* makeDual.cpp:172-195 Exa::find */
static int64_t exa_find(const Exa *exa, const float where[3]) {
  for (int level = exa->minLevel; level <= exa->maxLevel; level++) {
    Cell q;
    q.pos[0] = lower_on_level(where[0], level);
    q.pos[1] = lower_on_level(where[1], level);
    q.pos[2] = lower_on_level(where[2], level);
    q.level = level;
    q.scalarID = 0;
    int64_t lo = 0, hi = exa->n;
    while (lo < hi) {
      int64_t mid = lo + (hi - lo) / 2;
      if (logical_cmp(&exa->cells[mid], &q) < 0) lo = mid + 1;
      else hi = mid;
    }
    if (lo < exa->n && logical_cmp(&exa->cells[lo], &q) == 0) return lo;
  }
  return -1;
}

static void center_of(const Cell *c, float out[3]) {
  for (int k = 0; k < 3; ++k) out[k] = (float)c->pos[k] + 0.5f * (float)(1 << c->level);
}

/* growable int arrays */
typedef struct { int64_t *v; int64_t n, cap; } Vec;
static void vpush(Vec *a, int64_t x) {
  if (a->n == a->cap) {
    a->cap = a->cap ? 2 * a->cap : 1024;
    a->v = (int64_t *)realloc(a->v, sizeof(int64_t) * a->cap);
  }
  a->v[a->n++] = x;
}

typedef struct {
  int32_t level;
  float lower[3];
  int32_t scalarIDs[8];
} Cube;

struct ora_dual {
  Vec tets, pyrs, wedges, hexes; /* flattened sorted-cell indices */
  Cube *cubes;
  int64_t numCubes, capCubes;
  int64_t errors, degenerate;
  /* outputs */
  int64_t numElements, numVertices, numGridlets, numGridletScalars, conflicts;
  int32_t *indices;
  float *vertices;
  ora_gridlet *gridlets;
  float *gridletScalars;
};

static void push_cube(ora_dual *d, Cube c) {
  if (d->numCubes == d->capCubes) {
    d->capCubes = d->capCubes ? 2 * d->capCubes : 1024;
    d->cubes = (Cube *)realloc(d->cubes, sizeof(Cube) * d->capCubes);
  }
  d->cubes[d->numCubes++] = c;
}

/* makeDual.cpp:501-547 tryPyramid */
static void try_pyramid(ora_dual *d, const int64_t base[4], int64_t top, int numUnique) {
  if (numUnique == 5) {
    for (int i = 0; i < 4; ++i) vpush(&d->pyrs, base[i]);
    vpush(&d->pyrs, top);
    return;
  }
  if (numUnique == 4) {
    int64_t t[4];
    if (base[0] == base[1]) { t[0] = base[1]; t[1] = base[2]; t[2] = base[3]; t[3] = top; }
    else if (base[1] == base[2]) { t[0] = base[2]; t[1] = base[3]; t[2] = base[0]; t[3] = top; }
    else if (base[2] == base[3]) { t[0] = base[3]; t[1] = base[0]; t[2] = base[1]; t[3] = top; }
    else if (base[3] == base[0]) { t[0] = base[0]; t[1] = base[1]; t[2] = base[2]; t[3] = top; }
    else if (base[0] == base[2] || base[1] == base[3]) { d->degenerate++; return; }
    else { d->errors++; return; }
    for (int i = 0; i < 4; ++i) vpush(&d->tets, t[i]);
    return;
  }
  d->errors++;
}

static void try_wedge(ora_dual *d, const int64_t v[8], int f0, int f1, int f2, int b0, int b1, int b2) {
  vpush(&d->wedges, v[f0]); vpush(&d->wedges, v[f1]); vpush(&d->wedges, v[f2]);
  vpush(&d->wedges, v[b0]); vpush(&d->wedges, v[b1]); vpush(&d->wedges, v[b2]);
}

static void emit_hex(ora_dual *d, const Exa *exa, const int64_t v[8], int level) {
  if (level != -1) {
    /* makeDual.cpp:473-480: perfect cube -> gridlet input */
    Cube c;
    c.level = level;
    float p[3];
    center_of(&exa->cells[v[0]], c.lower);
    for (int i = 0; i < 8; ++i) {
      center_of(&exa->cells[v[i]], p);
      for (int k = 0; k < 3; ++k) c.lower[k] = fminf(c.lower[k], p[k]);
      c.scalarIDs[i] = exa->cells[v[i]].scalarID;
    }
    push_cube(d, c);
  } else {
    for (int i = 0; i < 8; ++i) vpush(&d->hexes, v[i]);
  }
}

static inline void swap64(int64_t *a, int64_t *b) { int64_t t = *a; *a = *b; *b = t; }

/* makeDual.cpp:596-872 doCell */
static void do_cell(ora_dual *d, const Exa *exa, int64_t ci) {
  const Cell *cell = &exa->cells[ci];
  const int cw = 1 << cell->level;
  for (int dz = -1; dz <= 1; dz += 2)
    for (int dy = -1; dy <= 1; dy += 2)
      for (int dx = -1; dx <= 1; dx += 2) {
        int64_t corner[2][2][2];
        int minLevel = 1000, maxLevel = -1, numFound = 0;
        for (int iz = 0; iz < 2; iz++)
          for (int iy = 0; iy < 2; iy++)
            for (int ix = 0; ix < 2; ix++) {
              Cell nb = *cell;
              nb.pos[0] += dx * ix * cw;
              nb.pos[1] += dy * iy * cw;
              nb.pos[2] += dz * iz * cw;
              float cc[3];
              center_of(&nb, cc);
              corner[iz][iy][ix] = exa_find(exa, cc);
              if (corner[iz][iy][ix] < 0) continue;
              const int lv = exa->cells[corner[iz][iy][ix]].level;
              if (lv < minLevel) minLevel = lv;
              if (lv > maxLevel) maxLevel = lv;
              ++numFound;
            }
        if (numFound < 8) continue;
        if (minLevel < cell->level) continue;
        int64_t minCell = ci;
        for (int iz = 0; iz < 2; iz++)
          for (int iy = 0; iy < 2; iy++)
            for (int ix = 0; ix < 2; ix++) {
              const int64_t c = corner[iz][iy][ix];
              if (exa->cells[c].level == cell->level &&
                  cell_cmp(&exa->cells[c], &exa->cells[minCell]) < 0)
                minCell = c;
            }
        if (minCell != ci) continue;
        int64_t (*vx)[2][2] = corner;
        if (dx < 0) {
          swap64(&vx[0][0][0], &vx[0][0][1]); swap64(&vx[0][1][0], &vx[0][1][1]);
          swap64(&vx[1][0][0], &vx[1][0][1]); swap64(&vx[1][1][0], &vx[1][1][1]);
        }
        if (dy < 0) {
          swap64(&vx[0][0][0], &vx[0][1][0]); swap64(&vx[0][0][1], &vx[0][1][1]);
          swap64(&vx[1][0][0], &vx[1][1][0]); swap64(&vx[1][0][1], &vx[1][1][1]);
        }
        if (dz < 0) {
          swap64(&vx[0][0][0], &vx[1][0][0]); swap64(&vx[0][0][1], &vx[1][0][1]);
          swap64(&vx[0][1][0], &vx[1][1][0]); swap64(&vx[0][1][1], &vx[1][1][1]);
        }
        const int64_t v[8] = {vx[0][0][0], vx[0][0][1], vx[0][1][1], vx[0][1][0],
                              vx[1][0][0], vx[1][0][1], vx[1][1][1], vx[1][1][0]};
        int numUnique = 0;
        for (int i = 0; i < 8; ++i) {
          int dup = 0;
          for (int j = 0; j < i; ++j) if (v[j] == v[i]) dup = 1;
          if (!dup) numUnique++;
        }
#define SAME(a, b) (v[a] == v[b])
#define ALLSAME(a, b, c, e) (SAME(a, b) && SAME(a, c) && SAME(a, e))
        if (minLevel == maxLevel) { emit_hex(d, exa, v, minLevel); continue; }
        if (numUnique == 8) { emit_hex(d, exa, v, -1); continue; }
        if (numUnique < 4) { d->degenerate++; continue; }
        if (ALLSAME(0, 1, 2, 3)) { int64_t b[4] = {v[4], v[7], v[6], v[5]}; try_pyramid(d, b, v[0], numUnique); continue; }
        if (ALLSAME(4, 5, 6, 7)) { int64_t b[4] = {v[0], v[1], v[2], v[3]}; try_pyramid(d, b, v[4], numUnique); continue; }
        if (ALLSAME(0, 1, 4, 5)) { int64_t b[4] = {v[2], v[6], v[7], v[3]}; try_pyramid(d, b, v[0], numUnique); continue; }
        if (ALLSAME(2, 3, 6, 7)) { int64_t b[4] = {v[0], v[4], v[5], v[1]}; try_pyramid(d, b, v[2], numUnique); continue; }
        if (ALLSAME(0, 3, 4, 7)) { int64_t b[4] = {v[1], v[5], v[6], v[2]}; try_pyramid(d, b, v[0], numUnique); continue; }
        if (ALLSAME(1, 2, 5, 6)) { int64_t b[4] = {v[0], v[3], v[7], v[4]}; try_pyramid(d, b, v[1], numUnique); continue; }
        if (SAME(0, 1) && SAME(4, 5)) { try_wedge(d, v, 3, 2, 0, 7, 6, 4); continue; }
        if (SAME(0, 4) && SAME(1, 5)) { try_wedge(d, v, 2, 6, 5, 3, 7, 4); continue; }
        if (SAME(3, 7) && SAME(2, 6)) { try_wedge(d, v, 5, 1, 2, 4, 0, 3); continue; }
        if (SAME(2, 3) && SAME(6, 7)) { try_wedge(d, v, 1, 0, 3, 5, 4, 7); continue; }
        if (SAME(4, 7) && SAME(5, 6)) { try_wedge(d, v, 3, 0, 4, 2, 1, 6); continue; }
        if (SAME(4, 5) && SAME(6, 7)) { try_wedge(d, v, 0, 1, 4, 3, 2, 7); continue; }
        if (SAME(0, 1) && SAME(3, 2)) { try_wedge(d, v, 5, 4, 0, 6, 7, 3); continue; }
        if (SAME(0, 3) && SAME(1, 2)) { try_wedge(d, v, 4, 7, 3, 5, 6, 2); continue; }
        if (SAME(0, 3) && SAME(4, 7)) { try_wedge(d, v, 5, 6, 7, 1, 2, 3); continue; }
        if (SAME(0, 4) && SAME(3, 7)) { try_wedge(d, v, 1, 5, 4, 2, 6, 7); continue; }
        if (SAME(1, 2) && SAME(5, 6)) { try_wedge(d, v, 7, 4, 5, 3, 0, 1); continue; }
        if (SAME(1, 5) && SAME(2, 6)) { try_wedge(d, v, 4, 0, 1, 7, 3, 2); continue; }
#undef SAME
#undef ALLSAME
        emit_hex(d, exa, v, -1);
      }
}

/* makeGrids.cpp:54-93 */
static void cube_cell_id(const Cube *c, int32_t out[3]) {
  const int w = 1 << c->level;
  for (int k = 0; k < 3; ++k) {
    int cid = (int)c->lower[k];
    if (c->lower[k] < 0.f) cid -= (w - 1);
    out[k] = cid / w;
  }
}
static void cube_mc_id(const Cube *c, int32_t out[3]) {
  cube_cell_id(c, out);
  for (int k = 0; k < 3; ++k) {
    if (out[k] < 0) out[k] -= 7;
    out[k] = out[k] / 8;
  }
}

typedef struct { int32_t mc[3]; int64_t cube; } CubeRef;
static int cuberef_cmp(const void *pa, const void *pb) {
  const CubeRef *a = (const CubeRef *)pa, *b = (const CubeRef *)pb;
  for (int k = 0; k < 3; ++k)
    if (a->mc[k] != b->mc[k]) return a->mc[k] < b->mc[k] ? -1 : 1;
  return a->cube < b->cube ? -1 : (a->cube > b->cube ? 1 : 0);
}

static void make_gridlets(ora_dual *d, const float *scalars, int64_t numScalars) {
  int maxLevel = -1;
  for (int64_t i = 0; i < d->numCubes; ++i)
    if (d->cubes[i].level > maxLevel) maxLevel = d->cubes[i].level;
  int64_t capG = 1024, capS = 65536;
  d->gridlets = (ora_gridlet *)malloc(sizeof(ora_gridlet) * capG);
  d->gridletScalars = (float *)malloc(sizeof(float) * capS);
  CubeRef *refs = (CubeRef *)malloc(sizeof(CubeRef) * (d->numCubes + 1));
  for (int level = 0; level <= maxLevel; ++level) {
    int64_t nr = 0;
    for (int64_t i = 0; i < d->numCubes; ++i) {
      if (d->cubes[i].level != level) continue;
      cube_mc_id(&d->cubes[i], refs[nr].mc);
      refs[nr].cube = i;
      nr++;
    }
    qsort(refs, nr, sizeof(CubeRef), cuberef_cmp);
    for (int64_t a = 0; a < nr;) {
      int64_t b = a;
      int32_t lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
      int32_t hi[3] = {-0x7fffffff, -0x7fffffff, -0x7fffffff};
      while (b < nr && !memcmp(refs[b].mc, refs[a].mc, sizeof(int32_t) * 3)) {
        int32_t cid[3];
        cube_cell_id(&d->cubes[refs[b].cube], cid);
        for (int k = 0; k < 3; ++k) {
          if (cid[k] < lo[k]) lo[k] = cid[k];
          if (cid[k] + 1 > hi[k]) hi[k] = cid[k] + 1;
        }
        b++;
      }
      const int nx = hi[0] - lo[0], ny = hi[1] - lo[1], nz = hi[2] - lo[2];
      const int64_t ns = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
      int32_t *ids = (int32_t *)malloc(sizeof(int32_t) * ns);
      for (int64_t i = 0; i < ns; ++i) ids[i] = -1;
      static const int vtkOrder[8] = {0, 1, 3, 2, 4, 5, 7, 6};
      for (int64_t r = a; r < b; ++r) {
        const Cube *c = &d->cubes[refs[r].cube];
        int32_t cid[3];
        cube_cell_id(c, cid);
        const int bx = cid[0] - lo[0], by = cid[1] - lo[1], bz = cid[2] - lo[2];
        for (int iz = 0; iz < 2; iz++)
          for (int iy = 0; iy < 2; iy++)
            for (int ix = 0; ix < 2; ix++) {
              const int64_t idx = (bx + ix) + (int64_t)(nx + 1) * ((by + iy) + (int64_t)(ny + 1) * (bz + iz));
              const int32_t sid = c->scalarIDs[vtkOrder[4 * iz + 2 * iy + ix]];
              if (ids[idx] != -1 && ids[idx] != sid) d->conflicts++;
              ids[idx] = sid;
            }
      }
      if (d->numGridlets == capG) { capG *= 2; d->gridlets = (ora_gridlet *)realloc(d->gridlets, sizeof(ora_gridlet) * capG); }
      while (d->numGridletScalars + ns > capS) { capS *= 2; d->gridletScalars = (float *)realloc(d->gridletScalars, sizeof(float) * capS); }
      ora_gridlet *g = &d->gridlets[d->numGridlets++];
      g->lower[0] = lo[0]; g->lower[1] = lo[1]; g->lower[2] = lo[2];
      g->level = level;
      g->dims[0] = nx; g->dims[1] = ny; g->dims[2] = nz;
      g->begin = (uint32_t)d->numGridletScalars;
      for (int64_t i = 0; i < ns; ++i)
        d->gridletScalars[d->numGridletScalars++] =
            ((uint32_t)ids[i] < (uint64_t)numScalars) ? scalars[ids[i]] : NAN;
      free(ids);
      a = b;
    }
  }
  free(refs);
}

ora_dual *ora_dual_build(const ora_cell *in, const float *scalars, int64_t n) {
  ora_dual *d = (ora_dual *)calloc(1, sizeof(ora_dual));
  Exa exa;
  exa.n = n;
  exa.cells = (Cell *)malloc(sizeof(Cell) * (n + 1));
  exa.minLevel = 100;
  exa.maxLevel = 0;
  for (int64_t i = 0; i < n; ++i) {
    exa.cells[i].pos[0] = in[i].pos[0];
    exa.cells[i].pos[1] = in[i].pos[1];
    exa.cells[i].pos[2] = in[i].pos[2];
    exa.cells[i].level = in[i].level;
    exa.cells[i].scalarID = (int32_t)i;
    if (in[i].level < exa.minLevel) exa.minLevel = in[i].level;
    if (in[i].level > exa.maxLevel) exa.maxLevel = in[i].level;
  }
  qsort(exa.cells, n, sizeof(Cell), cell_cmp);
  for (int64_t ci = 0; ci < n; ++ci) do_cell(d, &exa, ci);

  /* vertex compaction: element-referenced cells, ascending sorted order */
  int64_t *vid = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
  for (int64_t i = 0; i < n; ++i) vid[i] = -1;
  Vec *lists[4] = {&d->tets, &d->pyrs, &d->wedges, &d->hexes};
  for (int t = 0; t < 4; ++t)
    for (int64_t i = 0; i < lists[t]->n; ++i) vid[lists[t]->v[i]] = 0;
  d->numVertices = 0;
  for (int64_t i = 0; i < n; ++i)
    if (vid[i] >= 0) vid[i] = d->numVertices++;
  d->vertices = (float *)malloc(sizeof(float) * 4 * (d->numVertices + 1));
  for (int64_t i = 0; i < n; ++i)
    if (vid[i] >= 0) {
      float c[3];
      center_of(&exa.cells[i], c);
      float *v = d->vertices + 4 * vid[i];
      v[0] = c[0]; v[1] = c[1]; v[2] = c[2];
      v[3] = scalars[exa.cells[i].scalarID];
    }
  const int nv[4] = {4, 5, 6, 8};
  d->numElements = d->tets.n / 4 + d->pyrs.n / 5 + d->wedges.n / 6 + d->hexes.n / 8;
  d->indices = (int32_t *)malloc(sizeof(int32_t) * 8 * (d->numElements + 1));
  int64_t e = 0;
  for (int t = 0; t < 4; ++t)
    for (int64_t i = 0; i < lists[t]->n; i += nv[t], ++e) {
      for (int j = 0; j < 8; ++j) d->indices[8 * e + j] = j < nv[t] ? (int32_t)vid[lists[t]->v[i + j]] : -1;
    }
  free(vid);
  make_gridlets(d, scalars, n);
  free(exa.cells);
  return d;
}

int64_t ora_dual_count(const ora_dual *d, int what) {
  switch (what) {
  case 0: return d->tets.n / 4;
  case 1: return d->pyrs.n / 5;
  case 2: return d->wedges.n / 6;
  case 3: return d->hexes.n / 8;
  case 4: return d->numElements;
  case 5: return d->numVertices;
  case 6: return d->numGridlets;
  case 7: return d->numGridletScalars;
  case 8: return d->numCubes;
  case 9: return d->conflicts + d->errors;
  case 10: return d->degenerate;
  default: return -1;
  }
}

void ora_dual_get(const ora_dual *d, int32_t *indices, float *vertices, ora_gridlet *gridlets,
                  float *gridletScalars) {
  if (indices) memcpy(indices, d->indices, sizeof(int32_t) * 8 * d->numElements);
  if (vertices) memcpy(vertices, d->vertices, sizeof(float) * 4 * d->numVertices);
  if (gridlets) memcpy(gridlets, d->gridlets, sizeof(ora_gridlet) * d->numGridlets);
  if (gridletScalars) memcpy(gridletScalars, d->gridletScalars, sizeof(float) * d->numGridletScalars);
}

void ora_dual_free(ora_dual *d) {
  if (!d) return;
  free(d->tets.v); free(d->pyrs.v); free(d->wedges.v); free(d->hexes.v);
  free(d->cubes); free(d->indices); free(d->vertices); free(d->gridlets); free(d->gridletScalars);
  free(d);
}
