/*
 * ora_f64.c -- double-precision build of the oracle core (the FP64 high-spp
 * image oracle).  TEST INFRASTRUCTURE ONLY (see exa_oracle.h).
 */
#define REAL double
#define FMIN fmin
#define FMAX fmax
#define SQRT sqrt
#define FABS fabs
#define FMA fma
#define FLOOR floor
#define LOGF log
#define SINCOS sincos
#define POWF pow
#define CEIL ceil
#define _GNU_SOURCE
#include "ora_core.inc"

int ora_intersect_gridlet_f64(double *value, int32_t *cellID, const double pos[3],
                              const ora_gridlet *g, const float *scalars) {
  return isect_gridlet(value, cellID, v3(pos[0], pos[1], pos[2]), g, scalars);
}

int ora_intersect_element_f64(double *value, const double P[3], const int32_t *idx8,
                              const float *vertices) {
  return isect_element(value, v3(P[0], P[1], P[2]), idx8, vertices);
}

void ora_render_f64(const ora_render_params *p, const ora_model *m, const ora_bins *b,
                    double *accum, int32_t y0, int32_t y1) {
  for (int py = y0; py < y1; ++py)
    for (int px = 0; px < p->W; ++px) {
      V4 acc = render_pixel(p, m, b, px, py);
      double *a = accum + 4 * (px + (int64_t)p->W * py);
      if (p->accumID > 0) { acc.x += a[0]; acc.y += a[1]; acc.z += a[2]; acc.w += a[3]; }
      a[0] = acc.x; a[1] = acc.y; a[2] = acc.z; a[3] = acc.w;
    }
}
