/*
 * ora_f32.c -- single-precision oracle build plus the precision-independent
 * parts (RNG, math functions, majorant grid, max opacities, binning).
 * TEST INFRASTRUCTURE ONLY (see exa_oracle.h).
 */
#define REAL float
#define FMIN fminf
#define FMAX fmaxf
#define SQRT sqrtf
#define FABS fabsf
#define FMA fmaf
#define FLOOR floorf
#define LOGF ora_logf
#define SINCOS ora_sincosf
#define POWF ora_powf
#define CEIL ceilf
#include "ora_core.inc"

/* ===================================================================== */
/* RNG: owl::common::LCG<4> (deviceCode.cu:41, 1278).  The OWL source is   */
/* not vendored; this restates its published methods: TEA-based seeding    */
/* (Zafar, Olano, Curtis, HPG 2010) with 4 rounds, then the Numerical       */
/* Recipes LCG; a sample keeps the low 24 bits of the state.               */
/* ===================================================================== */
uint32_t ora_lcg_init(uint32_t val0, uint32_t val1) {
  uint32_t v0 = val0, v1 = val1, s0 = 0;
  for (int n = 0; n < 4; n++) {
    s0 += 0x9e3779b9u;
    v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
    v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
  }
  return v0;
}

float ora_lcg_next(uint32_t *state) {
  *state = 1664525u * *state + 1013904223u;
  return (float)(*state & 0x00FFFFFFu) / (float)0x01000000;
}

/* ===================================================================== */
/* Own single-precision log and sin/cos (the same algorithms run in the    */
/* HIP kernels, so the oracle and the device agree bit for bit).            */
/*  logf: Cody & Waite argument reduction x = m*2^e, m in [sqrt(1/2),       */
/*        sqrt(2)) (Cody & Waite, "Software Manual for the Elementary       */
/*        Functions", 1980); log(m) = f + f^2 Q(f), f = m - 1, Q of degree  */
/*        8 fitted to (log1p(f) - f)/f^2 by reweighted least squares        */
/*        (fit_logf.py), Horner by fused multiply-adds, no division.        */
/*        Domain: positive normal floats.                                   */
/*  sincosf: reduction by the nearest multiple of pi/2 in double, Taylor    */
/*        polynomials on [-pi/4, pi/4] (sin to r^9, cos to r^10).  Domain:  */
/*        |x| < 1e6.                                                        */
/* ===================================================================== */
float ora_logf(float x) {
  union { float f; uint32_t u; } b;
  b.f = x;
  int e = (int)((b.u >> 23) & 0xffu) - 127;
  b.u = (b.u & 0x007fffffu) | 0x3f800000u;
  float m = b.f;
  if (m > 1.41421353816986083984375f) {
    m = m * 0.5f;
    e += 1;
  }
  const float f = m - 1.0f;
  float q = -0.07823377847671509f;
  q = fmaf(q, f, 0.13004831969738007f);
  q = fmaf(q, f, -0.13208796083927155f);
  q = fmaf(q, f, 0.14163468778133392f);
  q = fmaf(q, f, -0.16611149907112122f);
  q = fmaf(q, f, 0.2000301629304886f);
  q = fmaf(q, f, -0.2500152289867401f);
  q = fmaf(q, f, 0.3333330750465393f);
  q = fmaf(q, f, -0.49999988079071045f);
  const float logm = fmaf(f * f, q, f);
  const float fe = (float)e;
  return fmaf(fe, 0.693145751953125f, fmaf(fe, 1.428606765330187045e-06f, logm));
}

void ora_sincosf(float x, float *sn, float *cs) {
  const double kd = rint((double)x * 0.63661977236758134308);
  const float r = (float)((double)x - kd * 1.57079632679489661923);
  const int q = ((int)kd) & 3;
  const float r2 = r * r;
  const float sr = r + r * r2 * (-0.16666667163372039794921875f +
                                 r2 * (0.0083333337679505348205566f +
                                       r2 * (-0.00019841270113829523324966f +
                                             r2 * 2.7557319981497130356729e-06f)));
  const float cr = 1.0f + r2 * (-0.5f + r2 * (0.041666667908430099487304688f +
                                             r2 * (-0.0013888889225199818611145f +
                                                   r2 * (2.4801587642286904156208e-05f +
                                                         r2 * -2.7557319981497130356729e-07f))));
  switch (q) {
  case 0: *sn = sr; *cs = cr; break;
  case 1: *sn = cr; *cs = -sr; break;
  case 2: *sn = -sr; *cs = -cr; break;
  default: *sn = -cr; *cs = sr; break;
  }
}

/* Own exp / pow (the same algorithms run in the HIP kernels).
   expf: Cody & Waite reduction x = n ln2 + r (ln2 split as in ora_logf),
   Taylor polynomial of e^r to r^7 on |r| <= ln2/2, times 2^n; results below
   the smallest normal float are 0.  powf: exp(y log x) for x > 0, with
   pow(x, 0) = 1, pow(1, y) = 1, pow(0, y > 0) = 0. */
float ora_expf(float x) {
  if (!(x == x)) return x;
  if (x < -87.3f) return 0.f;
  if (x > 88.72f) return INFINITY;
  const float n = rintf(x * 1.44269502162933349609375f);
  const float r = (x - n * 0.693145751953125f) - n * 1.428606765330187045e-06f;
  const float p =
      1.0f + r * (1.0f + r * (0.5f + r * (0.16666667163372039794921875f +
                                          r * (0.0416666679084300994873046875f +
                                               r * (0.008333333767950534820556640625f +
                                                    r * (0.001388888922519981861114501953125f +
                                                         r * 0.0001984127011382952332496643066406f))))));
  int ni = (int)n;
  float q = p;
  if (ni > 127) {
    q = q * 2.f;
    ni -= 1;
  }
  union { uint32_t u; float f; } sc;
  sc.u = (uint32_t)(ni + 127) << 23;
  return q * sc.f;
}

float ora_powf(float x, float y) {
  if (y == 0.f) return 1.f;
  if (x == 1.f) return 1.f;
  if (x == 0.f) return y > 0.f ? 0.f : INFINITY;
  if (!(x > 0.f)) return NAN;
  return ora_expf(y * ora_logf(x));
}

/* ===================================================================== */
/* public f32 element math                                                 */
/* ===================================================================== */
int ora_intersect_gridlet_f32(float *value, int32_t *cellID, const float pos[3],
                              const ora_gridlet *g, const float *scalars) {
  return isect_gridlet(value, cellID, ld3(pos), g, scalars);
}
int ora_intersect_tet_f32(float *value, const float P[3], const float *v) {
  return isect_tet(value, ld3(P), ld4(v), ld4(v + 4), ld4(v + 8), ld4(v + 12));
}
int ora_intersect_pyr_f32(float *value, const float P[3], const float *v) {
  V4 V[5];
  for (int i = 0; i < 5; ++i) V[i] = ld4(v + 4 * i);
  return isect_newton(value, ld3(P), V, 5, pyr_sf, pyr_df, 0);
}
int ora_intersect_wedge_f32(float *value, const float P[3], const float *v) {
  V4 V[6];
  for (int i = 0; i < 6; ++i) V[i] = ld4(v + 4 * i);
  return isect_newton(value, ld3(P), V, 6, wedge_sf, wedge_df, 1);
}
int ora_intersect_hex_f32(float *value, const float P[3], const float *v) {
  V4 V[8];
  for (int i = 0; i < 8; ++i) V[i] = ld4(v + 4 * i);
  return isect_newton(value, ld3(P), V, 8, hex_sf, hex_df, 0);
}
int ora_intersect_element_f32(float *value, const float P[3], const int32_t *idx8,
                              const float *vertices) {
  return isect_element(value, ld3(P), idx8, vertices);
}

void ora_locate_brute_f32(const ora_model *m, const float *pos, int64_t n, ora_sample *out,
                          int32_t *numContaining) {
  for (int64_t i = 0; i < n; ++i) {
    SampleR s = locate_brute(m, ld3(pos + 3 * i), numContaining ? numContaining + i : NULL);
    out[i].primID = s.primID;
    out[i].cellID = s.cellID;
    out[i].value = s.value;
  }
}

void ora_locate_binned_f32(const ora_bins *b, const ora_model *m, const float *pos, int64_t n,
                           ora_sample *out) {
  for (int64_t i = 0; i < n; ++i) {
    SampleR s = locate_binned(b, m, ld3(pos + 3 * i));
    out[i].primID = s.primID;
    out[i].cellID = s.cellID;
    out[i].value = s.value;
  }
}

/* ===================================================================== */
/* binning                                                                 */
/* ===================================================================== */
static void bin_range(double lo, double hi, double blo, double bhi, int res, int *a, int *b) {
  double ta = (lo - blo) / (bhi - blo) * res - 1e-6;
  double tb = (hi - blo) / (bhi - blo) * res + 1e-6;
  int ia = (int)floor(ta), ib = (int)floor(tb);
  *a = ia < 0 ? 0 : (ia >= res ? res - 1 : ia);
  *b = ib < 0 ? 0 : (ib >= res ? res - 1 : ib);
}

ora_bins *ora_bins_build(const ora_model *m, int res) {
  ora_bins *b = (ora_bins *)calloc(1, sizeof(ora_bins));
  b->res = res;
  for (int k = 0; k < 3; ++k) { b->lo[k] = INFINITY; b->hi[k] = -INFINITY; }
  double (*gbox)[6] = (double(*)[6])malloc(sizeof(double[6]) * (m->numGridlets + 1));
  b->eBox = (float *)malloc(sizeof(float) * 6 * (m->numElements + 1));
  for (int64_t g = 0; g < m->numGridlets; ++g) {
    V3 lo, hi;
    gridlet_bounds(&m->gridlets[g], &lo, &hi);
    gbox[g][0] = lo.x; gbox[g][1] = lo.y; gbox[g][2] = lo.z;
    gbox[g][3] = hi.x; gbox[g][4] = hi.y; gbox[g][5] = hi.z;
    for (int k = 0; k < 3; ++k) {
      if (gbox[g][k] < b->lo[k]) b->lo[k] = gbox[g][k];
      if (gbox[g][3 + k] > b->hi[k]) b->hi[k] = gbox[g][3 + k];
    }
  }
  for (int64_t e = 0; e < m->numElements; ++e) {
    float *bx = b->eBox + 6 * e;
    bx[0] = bx[1] = bx[2] = INFINITY;
    bx[3] = bx[4] = bx[5] = -INFINITY;
    for (int i = 0; i < 8; ++i) {
      int idx = m->indices[8 * e + i];
      if (idx < 0) break;
      for (int k = 0; k < 3; ++k) {
        bx[k] = fminf(bx[k], m->vertices[4 * (int64_t)idx + k]);
        bx[3 + k] = fmaxf(bx[3 + k], m->vertices[4 * (int64_t)idx + k]);
      }
    }
    for (int k = 0; k < 3; ++k) {
      if (bx[k] < b->lo[k]) b->lo[k] = bx[k];
      if (bx[3 + k] > b->hi[k]) b->hi[k] = bx[3 + k];
    }
  }
  for (int k = 0; k < 3; ++k)
    if (!(b->hi[k] > b->lo[k])) { b->lo[k] = 0; b->hi[k] = 1; }
  const int64_t nb = (int64_t)res * res * res;
  b->gStart = (int64_t *)calloc(nb + 1, sizeof(int64_t));
  b->eStart = (int64_t *)calloc(nb + 1, sizeof(int64_t));
  for (int pass = 0; pass < 2; ++pass) {
    int64_t *cnt = (int64_t *)calloc(nb, sizeof(int64_t));
    for (int64_t g = 0; g < m->numGridlets; ++g) {
      int a[3], c[3];
      for (int k = 0; k < 3; ++k) bin_range(gbox[g][k], gbox[g][3 + k], b->lo[k], b->hi[k], res, &a[k], &c[k]);
      for (int z = a[2]; z <= c[2]; ++z)
        for (int y = a[1]; y <= c[1]; ++y)
          for (int x = a[0]; x <= c[0]; ++x) {
            int64_t bin = ((int64_t)z * res + y) * res + x;
            if (pass == 1) b->gList[b->gStart[bin] + cnt[bin]] = (int32_t)g;
            cnt[bin]++;
          }
    }
    if (pass == 0) {
      for (int64_t i = 0; i < nb; ++i) b->gStart[i + 1] = b->gStart[i] + cnt[i];
      b->gList = (int32_t *)malloc(sizeof(int32_t) * (b->gStart[nb] + 1));
    }
    free(cnt);
  }
  for (int pass = 0; pass < 2; ++pass) {
    int64_t *cnt = (int64_t *)calloc(nb, sizeof(int64_t));
    for (int64_t e = 0; e < m->numElements; ++e) {
      const float *bx = b->eBox + 6 * e;
      int a[3], c[3];
      for (int k = 0; k < 3; ++k) bin_range(bx[k], bx[3 + k], b->lo[k], b->hi[k], res, &a[k], &c[k]);
      for (int z = a[2]; z <= c[2]; ++z)
        for (int y = a[1]; y <= c[1]; ++y)
          for (int x = a[0]; x <= c[0]; ++x) {
            int64_t bin = ((int64_t)z * res + y) * res + x;
            if (pass == 1) b->eList[b->eStart[bin] + cnt[bin]] = (int32_t)e;
            cnt[bin]++;
          }
    }
    if (pass == 0) {
      for (int64_t i = 0; i < nb; ++i) b->eStart[i + 1] = b->eStart[i] + cnt[i];
      b->eList = (int32_t *)malloc(sizeof(int32_t) * (b->eStart[nb] + 1));
    }
    free(cnt);
  }
  free(gbox);
  return b;
}

void ora_bins_free(ora_bins *b) {
  if (!b) return;
  free(b->gStart); free(b->eStart); free(b->gList); free(b->eList); free(b->eBox);
  free(b);
}

/* ===================================================================== */
/* majorant grid: Grid.cu:43-246 (build), 687-719 (max opacity)            */
/* ===================================================================== */
void ora_project_on_grid(const float V[3], const int32_t dims[3], const float bounds[6],
                         int32_t out[3]) {
  project_on_grid(ld3(V), dims, bounds, out);
}

/* atomicOp.cuh:4-26: the CAS loops only replace on a strict improvement */
static inline void range_update(float *r, float lo, float hi) {
  if (lo < r[0]) r[0] = lo;
  if (hi > r[1]) r[1] = hi;
}

void ora_grid_build(const ora_model *m, const int32_t dims[3], const float bounds[6],
                    float *ranges) {
  const int64_t nmc = (int64_t)dims[0] * dims[1] * dims[2];
  for (int64_t i = 0; i < nmc; ++i) { ranges[2 * i] = 1e30f; ranges[2 * i + 1] = -1e30f; }
  /* elements, Grid.cu:73-117 */
  for (int64_t e = 0; e < m->numElements; ++e) {
    const int32_t *I = m->indices + 8 * e;
    float blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float vlo = 1e30f, vhi = -1e30f;
    for (int i = 0; i < 8; ++i) {
      if (I[i] < 0) break;
      const float *V = m->vertices + 4 * (int64_t)I[i];
      if (!isnan(V[3])) {
        for (int k = 0; k < 3; ++k) { blo[k] = fminf(blo[k], V[k]); bhi[k] = fmaxf(bhi[k], V[k]); }
        vlo = fminf(vlo, V[3]);
        vhi = fmaxf(vhi, V[3]);
      }
    }
    int32_t lo[3], up[3];
    ora_project_on_grid(blo, dims, bounds, lo);
    ora_project_on_grid(bhi, dims, bounds, up);
    for (int z = lo[2]; z <= up[2]; ++z)
      for (int y = lo[1]; y <= up[1]; ++y)
        for (int x = lo[0]; x <= up[0]; ++x)
          range_update(ranges + 2 * (((int64_t)z * dims[1] + y) * dims[0] + x), vlo, vhi);
  }
  /* gridlets, Grid.cu:134-246 */
  const float mcSize[3] = {(bounds[3] - bounds[0]) / (float)dims[0],
                           (bounds[4] - bounds[1]) / (float)dims[1],
                           (bounds[5] - bounds[2]) / (float)dims[2]};
  for (int64_t gi = 0; gi < m->numGridlets; ++gi) {
    const ora_gridlet *g = &m->gridlets[gi];
    V3 glo, ghi;
    gridlet_bounds(g, &glo, &ghi);
    const float gl[3] = {glo.x, glo.y, glo.z}, gh[3] = {ghi.x, ghi.y, ghi.z};
    int32_t lo[3], up[3];
    ora_project_on_grid(gl, dims, bounds, lo);
    ora_project_on_grid(gh, dims, bounds, up);
    const int nsx = g->dims[0] + 1, nsy = g->dims[1] + 1;
    const float cw = (float)(1 << g->level);
    const float *s = m->gridletScalars + g->begin;
    for (int z = 0; z < g->dims[2]; ++z)
      for (int y = 0; y < g->dims[1]; ++y)
        for (int x = 0; x < g->dims[0]; ++x) {
          const int c[3] = {x, y, z};
          float clo[3], chi[3];
          for (int k = 0; k < 3; ++k) {
            clo[k] = (float)(g->lower[k] + c[k]) * cw + .5f * cw;
            chi[k] = (float)(g->lower[k] + c[k] + 1) * cw + .5f * cw;
          }
          float f[8];
          int n = 0;
          for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
              for (int dx = 0; dx < 2; ++dx)
                f[n++] = s[(int64_t)(z + dz) * nsy * nsx + (int64_t)(y + dy) * nsx + (x + dx)];
          for (int mz = lo[2]; mz <= up[2]; ++mz)
            for (int my = lo[1]; my <= up[1]; ++my)
              for (int mx = lo[0]; mx <= up[0]; ++mx) {
                const int mc[3] = {mx, my, mz};
                float mlo[3], mhi[3];
                for (int k = 0; k < 3; ++k) {
                  mlo[k] = bounds[k] + (float)mc[k] * mcSize[k];
                  mhi[k] = bounds[k] + (float)(mc[k] + 1) * mcSize[k];
                }
                float vlo = 1e30f, vhi = -1e30f;
                /* box3f::overlaps, inclusive on both ends */
                if (!(chi[0] < mlo[0] || chi[1] < mlo[1] || chi[2] < mlo[2] ||
                      clo[0] > mhi[0] || clo[1] > mhi[1] || clo[2] > mhi[2])) {
                  for (int i = 0; i < 8; ++i)
                    if (!isnan(f[i])) { vlo = fminf(vlo, f[i]); vhi = fmaxf(vhi, f[i]); }
                }
                range_update(ranges + 2 * (((int64_t)mz * dims[1] + my) * dims[0] + mx), vlo, vhi);
              }
        }
  }
}

/* ExaStitchSampler.cu:27-142: per-gridlet min/max over its non-NaN corners */
void ora_gridlet_value_ranges(const ora_model *m, float *ranges) {
  for (int64_t gi = 0; gi < m->numGridlets; ++gi) {
    const ora_gridlet *g = &m->gridlets[gi];
    float *r = ranges + 2 * gi;
    r[0] = 1e30f;
    r[1] = -1e30f;
    const int nsx = g->dims[0] + 1, nsy = g->dims[1] + 1;
    const float *s = m->gridletScalars + g->begin;
    for (int z = 0; z < g->dims[2]; ++z)
      for (int y = 0; y < g->dims[1]; ++y)
        for (int x = 0; x < g->dims[0]; ++x) {
          float vlo = 1e30f, vhi = -1e30f;
          for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
              for (int dx = 0; dx < 2; ++dx) {
                const float f = s[(int64_t)(z + dz) * nsy * nsx + (int64_t)(y + dy) * nsx + (x + dx)];
                if (!isnan(f)) { vlo = fminf(vlo, f); vhi = fmaxf(vhi, f); }
              }
          range_update(r, vlo, vhi);
        }
  }
}

/* Grid.cu:687-719 == ExaStitchSampler.cu:144-176 index mapping */
static float map_one(float lo, float hi, const float *rgba, int32_t numColors, const float xf[2]) {
  if (hi < lo) return 0.f;
  lo -= xf[0];
  lo /= xf[1] - xf[0];
  hi -= xf[0];
  hi /= xf[1] - xf[0];
  const float nm1 = (float)(numColors - 1);
  const int32_t ilo = clampi(sat_i32(lo * nm1), 0, numColors - 1);
  const int32_t ihi = clampi((int32_t)((uint32_t)sat_i32(hi * nm1) + 1u), 0, numColors - 1);
  float mo = 0.f;
  for (int i = ilo; i <= ihi; ++i) mo = fmaxf(mo, rgba[4 * i + 3]);
  return mo;
}

void ora_map_max_opacity(const float *ranges, int64_t n, const float *rgba, int32_t numColors,
                         const float xfRange[2], float *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = map_one(ranges[2 * i], ranges[2 * i + 1], rgba, numColors, xfRange);
}

/* ExaStitchSampler.cu:178-227 */
void ora_element_max_opacity(const ora_model *m, const float *rgba, int32_t numColors,
                             const float xfRange[2], float *out) {
  for (int64_t e = 0; e < m->numElements; ++e) {
    float lo = 1e30f, hi = -1e30f;
    for (int i = 0; i < 8; ++i) {
      int idx = m->indices[8 * e + i];
      if (idx < 0) continue;
      const float w = m->vertices[4 * (int64_t)idx + 3];
      lo = fminf(lo, w);
      hi = fmaxf(hi, w);
    }
    out[e] = map_one(lo, hi, rgba, numColors, xfRange);
  }
}

/* ===================================================================== */
/* DDA trace for tests                                                     */
/* ===================================================================== */
typedef struct { int64_t *cells; float *t0s, *t1s; int n, max; } TraceRec;
static int trace_cb(void *user, int64_t cell, float t0, float t1) {
  TraceRec *r = (TraceRec *)user;
  if (r->n < r->max) {
    r->cells[r->n] = cell;
    r->t0s[r->n] = t0;
    r->t1s[r->n] = t1;
  }
  r->n++;
  return 1;
}
int ora_dda_trace_f32(const float org[3], const float dir[3], float tmin, float tmax,
                      const int32_t dims[3], const float bounds[6], int64_t *cells, float *t0s,
                      float *t1s, int maxCells) {
  TraceRec r = {cells, t0s, t1s, 0, maxCells};
  dda3(ld3(org), ld3(dir), tmin, tmax, dims, bounds, trace_cb, &r);
  return r.n;
}

void ora_tf_lookup_f32(const float *rgba, int32_t N, const float domain[2], float v, float out[4]) {
  V4 r = tf_lookup(rgba, N, domain, v);
  out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

void ora_orthonormal_basis_f32(const float w[3], float u[3], float v[3]) {
  V3 uu, vv;
  orthonormal_basis(&uu, &vv, ld3(w));
  u[0] = uu.x; u[1] = uu.y; u[2] = uu.z;
  v[0] = vv.x; v[1] = vv.y; v[2] = vv.z;
}

void ora_hg_sample_f32(const float wo[3], uint32_t v0, uint32_t v1, int64_t n, float *out) {
  Rng rng;
  rng.state = ora_lcg_init(v0, v1);
  for (int64_t i = 0; i < n; ++i) {
    const V3 d = hg_sample(ld3(wo), &rng);
    out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
  }
}

/* ===================================================================== */
/* model/ExaStitchModel.cpp:78-99,224-245                                  */
/* ===================================================================== */
void ora_model_bounds(const ora_model *m, float cb[6], float vr[2]) {
  cb[0] = cb[1] = cb[2] = INFINITY;
  cb[3] = cb[4] = cb[5] = -INFINITY;
  vr[0] = 1e30f;
  vr[1] = -1e30f;
  for (int64_t e = 0; e < m->numElements; ++e)
    for (int i = 0; i < 8; ++i) {
      int idx = m->indices[8 * e + i];
      if (idx < 0) continue;
      const float *v = m->vertices + 4 * (int64_t)idx;
      for (int k = 0; k < 3; ++k) { cb[k] = fminf(cb[k], v[k]); cb[3 + k] = fmaxf(cb[3 + k], v[k]); }
      /* std::min(acc, v) / std::max(acc, v): a NaN v leaves acc unchanged */
      vr[0] = (v[3] < vr[0]) ? v[3] : vr[0];
      vr[1] = (vr[1] < v[3]) ? v[3] : vr[1];
    }
  for (int64_t g = 0; g < m->numGridlets; ++g) {
    V3 lo, hi;
    gridlet_bounds(&m->gridlets[g], &lo, &hi);
    const float l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
    for (int k = 0; k < 3; ++k) {
      cb[k] = fminf(cb[k], l[k]); cb[3 + k] = fmaxf(cb[3 + k], l[k]);
      cb[k] = fminf(cb[k], h[k]); cb[3 + k] = fmaxf(cb[3 + k], h[k]);
    }
    const ora_gridlet *gg = &m->gridlets[g];
    const int64_t ns = (int64_t)(gg->dims[0] + 1) * (gg->dims[1] + 1) * (gg->dims[2] + 1);
    for (int64_t i = 0; i < ns; ++i) {
      const float v = m->gridletScalars[gg->begin + i];
      vr[0] = (v < vr[0]) ? v : vr[0];
      vr[1] = (vr[1] < v) ? v : vr[1];
    }
  }
}

/* ===================================================================== */
/* frame: deviceCode.cu:1268-1345                                          */
/* ===================================================================== */
/* owl make_8bit / make_rgba (owl/common/math/vec.h, not vendored): the byte
   is clamp(int(f*256), 0, 255), channels packed x,y,z,w from the low byte. */
static uint32_t make_8bit(float f) {
  int32_t i = sat_i32(f * 256.f);
  return (uint32_t)(i < 0 ? 0 : (i > 255 ? 255 : i));
}

void ora_render_queries_f32(const ora_render_params *p, const ora_model *m, const ora_bins *b, float *accum,
                            int32_t y0, int32_t y1, float *out, int64_t cap, int64_t *count) {
  QueryRecorder rec = {out, cap, count};
  g_queries = &rec;
  ora_render_f32(p, m, b, accum, NULL, y0, y1);
  g_queries = NULL;
}

void ora_render_f32(const ora_render_params *p, const ora_model *m, const ora_bins *b,
                    float *accum, uint32_t *fb, int32_t y0, int32_t y1) {
  for (int py = y0; py < y1; ++py)
    for (int px = 0; px < p->W; ++px) {
      V4 acc = render_pixel(p, m, b, px, py);
      const int64_t pid = px + (int64_t)p->W * py;
      float *a = accum + 4 * pid;
      if (p->accumID > 0) { acc.x += a[0]; acc.y += a[1]; acc.z += a[2]; acc.w += a[3]; }
      a[0] = acc.x; a[1] = acc.y; a[2] = acc.z; a[3] = acc.w;
      const float sc = 1.f / (float)(p->accumID + 1);
      acc.x *= sc; acc.y *= sc; acc.z *= sc; acc.w *= sc;
      int invert = p->debugCrosshair && (px == p->W / 2 || py == p->H / 2);
      if (p->subImageSelecting) {
        /* deviceCode.cu:134-141 subImageSelectionWin, 1337-1340: the border
           of the selection box */
        const int lx = clampi(sat_i32(p->subImageSel[0] * (float)p->W), 0, p->W - 1);
        const int ly = clampi(sat_i32(p->subImageSel[1] * (float)p->H), 0, p->H - 1);
        const int hx = clampi(sat_i32(p->subImageSel[2] * (float)p->W), 0, p->W - 1);
        const int hy = clampi(sat_i32(p->subImageSel[3] * (float)p->H), 0, p->H - 1);
        invert = invert || (((px == lx || px == hx) && (py >= ly && py <= hy)) ||
                            ((py == ly || py == hy) && (px >= lx && px <= hx)));
      }
      if (invert) {
        acc.x = 1.f - acc.x; acc.y = 1.f - acc.y; acc.z = 1.f - acc.z; acc.w = 1.f - acc.w;
      }
      if (fb) {
        const float is = 1.f / (float)p->spp;
        fb[pid] = (make_8bit(acc.x * is) << 0) + (make_8bit(acc.y * is) << 8) +
                  (make_8bit(acc.z * is) << 16) + (make_8bit(acc.w * is) << 24);
      }
    }
}

/* ===================================================================== */
/* ExaBricks sampling                                                      */
/* ===================================================================== */
static int brick_domain_contains(const ora_brick *b, const float *p) {
  const float cw = (float)(1 << b->level);
  for (int k = 0; k < 3; ++k) {
    const float lo = (float)b->lower[k] - 0.5f * cw;
    const float hi = (float)b->lower[k] + ((float)b->size[k] + 0.5f) * cw;
    if (!(p[k] >= lo && p[k] <= hi)) return 0;
  }
  return 1;
}

void ora_brick_blend_brute_f32(const ora_brick *bricks, int64_t n, const float *scalars, const float *pos,
                               int64_t npts, float *value, float *sumWeights) {
  ora_brick_scene bs;
  memset(&bs, 0, sizeof(bs));
  bs.bricks = bricks;
  bs.numBricks = n;
  bs.scalars = scalars;
  for (int64_t i = 0; i < npts; ++i) {
    const float *p = pos + 3 * i;
    float wv = 0.f, w = 0.f;
    for (int64_t b = 0; b < n; ++b)
      if (brick_domain_contains(&bricks[b], p)) brick_add_basis(&bs, b, v3(p[0], p[1], p[2]), &wv, &w);
    value[i] = wv / w;
    sumWeights[i] = w;
  }
}

void ora_brick_locate_f32(const ora_brick_scene *s, const float *pos, int64_t n, ora_sample *out) {
  for (int64_t i = 0; i < n; ++i) {
    const float *p = pos + 3 * i;
    ora_sample r = {0, 0, 0.f};
    for (int64_t a = 0; a < s->numAbrs; ++a) {
      if (abr_culled(s, a)) continue;
      const float *d = s->abrs[a].domain;
      if (!(p[0] >= d[0] && p[1] >= d[1] && p[2] >= d[2] && p[0] <= d[3] && p[1] <= d[4] && p[2] <= d[5]))
        continue;
      float w;
      const float v = abr_blend(s, a, v3(p[0], p[1], p[2]), &w);
      if (w > 0.f) {
        r.cellID = -1;
        r.value = v;
      }
      break;
    }
    out[i] = r;
  }
}
