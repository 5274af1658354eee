// exa_device.h -- device-side math, element intersection and point-location
// primitives for the exaStitch hot path (gfx950).
//
// Every routine keeps the reference's operation order; the library is
// compiled with -ffp-contract=off so results are bit-identical to the CPU
// oracle.  Reference citations are given per function.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include "exa_types.h"

namespace exa {

#define EXA_DEV __device__ __forceinline__
#define EXA_HD __host__ __device__ __forceinline__

struct vec3f {
  float x, y, z;
};
EXA_HD vec3f mk3(float x, float y, float z) { return vec3f{x, y, z}; }
EXA_HD vec3f mk3s(float s) { return vec3f{s, s, s}; }
EXA_HD vec3f operator+(vec3f a, vec3f b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
EXA_HD vec3f operator-(vec3f a, vec3f b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
EXA_HD vec3f operator*(vec3f a, vec3f b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
EXA_HD vec3f operator/(vec3f a, vec3f b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
EXA_HD vec3f operator*(vec3f a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
EXA_HD vec3f operator*(float s, vec3f a) { return mk3(s * a.x, s * a.y, s * a.z); }
EXA_HD vec3f operator-(vec3f a) { return mk3(-a.x, -a.y, -a.z); }
EXA_HD float dot(vec3f a, vec3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
EXA_HD vec3f cross(vec3f a, vec3f b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
EXA_HD vec3f vmin(vec3f a, vec3f b) { return mk3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
EXA_HD vec3f vmax(vec3f a, vec3f b) { return mk3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
EXA_HD float reduce_min(vec3f a) { return fminf(fminf(a.x, a.y), a.z); }
EXA_HD float reduce_max(vec3f a) { return fmaxf(fmaxf(a.x, a.y), a.z); }
EXA_HD float length(vec3f a) { return sqrtf(dot(a, a)); }
// owl normalize is v*rsqrt(dot(v,v)); taken here as a correctly rounded 1/sqrt.
EXA_HD vec3f normalize(vec3f a) { return a * (1.f / sqrtf(dot(a, a))); }
EXA_HD vec3f xyz(float4 v) { return mk3(v.x, v.y, v.z); }
// c + a * s with one rounding per component (fused multiply-add)
EXA_HD vec3f fma3(vec3f a, float s, vec3f c) { return mk3(fmaf(a.x, s, c.x), fmaf(a.y, s, c.y), fmaf(a.z, s, c.z)); }

// float -> int32, round toward zero, saturating, NaN -> 0 (the conversion the
// reference's index math relies on, e.g. Grid.cu:711-712, Grid.cuh:62).
EXA_HD int32_t sat_i32(float f) {
  if (f != f) return 0;
  if (f >= 2147483647.f) return 2147483647;
  if (f <= -2147483648.f) return (-2147483647 - 1);
  return (int32_t)f;
}
EXA_HD int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

// common.h:63-67
EXA_HD float lerpf(float a, float b, float x) { return (1.f - x) * a + x * b; }

// ---------------------------------------------------------------------------
// Own single-precision log / sin-cos (identical algorithm in the CPU oracle).
//  logf: Cody & Waite argument reduction x = m*2^e, m in [sqrt(1/2),sqrt(2)),
//        log(m) = f + f^2 Q(f), f = m - 1, Q of degree 8 fitted to
//        (log1p(f) - f)/f^2 by reweighted least squares (oracle/fit_logf.py),
//        Horner by fused multiply-adds: no division.
//  sincos: nearest multiple of pi/2 removed in double, Taylor polynomials on
//        [-pi/4, pi/4] (sin to r^9, cos to r^10).
// ---------------------------------------------------------------------------
EXA_HD float exa_logf(float x) {
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  int e = (int)((u >> 23) & 0xffu) - 127;
  u = (u & 0x007fffffu) | 0x3f800000u;
  float m = __builtin_bit_cast(float, u);
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

EXA_HD void exa_sincosf(float x, float &sn, float &cs) {
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
  if (q == 0) { sn = sr; cs = cr; }
  else if (q == 1) { sn = cr; cs = -sr; }
  else if (q == 2) { sn = -sr; cs = -cr; }
  else { sn = -cr; cs = sr; }
}

// Own single-precision exp / pow (identical algorithm in the CPU oracle).
//  expf: Cody & Waite reduction x = n ln2 + r (ln2 split as in exa_logf),
//        Taylor polynomial of e^r to r^7 on |r| <= ln2/2, scaled by 2^n;
//        results below the smallest normal float return 0.
//  powf: x^y = exp(y log x) for x > 0, with pow(x, 0) = 1, pow(1, y) = 1,
//        pow(0, y > 0) = 0.
EXA_HD float exa_expf(float x) {
  if (!(x == x)) return x;
  if (x < -87.3f) return 0.f;
  if (x > 88.72f) return __builtin_inff();
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
  return q * __builtin_bit_cast(float, (uint32_t)(ni + 127) << 23);
}

EXA_HD float exa_powf(float x, float y) {
  if (y == 0.f) return 1.f;
  if (x == 1.f) return 1.f;
  if (x == 0.f) return y > 0.f ? 0.f : __builtin_inff();
  if (!(x > 0.f)) return __builtin_nanf("");
  return exa_expf(y * exa_logf(x));
}

// owl::common::LCG<4> (deviceCode.cu:41,1278): TEA seeding with 4 rounds
// (Zafar, Olano, Curtis, HPG 2010), then the Numerical Recipes LCG; a sample
// is the low 24 bits of the state scaled by 2^-24.
struct Lcg {
  uint32_t state;
  EXA_HD void init(uint32_t val0, uint32_t val1) {
    uint32_t v0 = val0, v1 = val1, s0 = 0;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      s0 += 0x9e3779b9u;
      v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
      v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    state = v0;
  }
  EXA_HD float operator()() {
    state = 1664525u * state + 1013904223u;
    return (float)(state & 0x00FFFFFFu) / (float)0x01000000;
  }
};

// ---------------------------------------------------------------------------
// Gridlet.h:30-36 getBounds / Gridlet.h:39-98 intersectGridlet
// ---------------------------------------------------------------------------
EXA_HD void gridletBounds(const Gridlet &g, vec3f &lo, vec3f &hi) {
  const float cw = (float)(1 << g.level);
  lo = mk3(((float)g.lower[0] + 0.5f) * cw, ((float)g.lower[1] + 0.5f) * cw,
           ((float)g.lower[2] + 0.5f) * cw);
  hi = mk3(((float)g.lower[0] + (float)g.dims[0] + 0.5f) * cw,
           ((float)g.lower[1] + (float)g.dims[1] + 0.5f) * cw,
           ((float)g.lower[2] + (float)g.dims[2] + 0.5f) * cw);
}

// Returns 0 = miss, 1 = hit on a cell boundary band, 2 = hit strictly inside
// its cell (every fractional coordinate in [margin, 1-margin], margin above
// the rounding of the cell coordinates); a band-2 hit is the unique accepting
// gridlet for that point (dual cells tile space), which lets point location
// stop early.
EXA_HD int intersectGridlet(float &value, int &cellID, vec3f pos, const Gridlet &g,
                            const float *__restrict__ scalars, float margin = 0.00390625f) {
  vec3f lo, hi;
  gridletBounds(g, lo, hi);
  if (!(pos.x >= lo.x && pos.y >= lo.y && pos.z >= lo.z && pos.x <= hi.x && pos.y <= hi.y &&
        pos.z <= hi.z))
    return 0;
  const int nsx = g.dims[0] + 1, nsy = g.dims[1] + 1, nsz = g.dims[2] + 1;
  // (pos - lo) / 2^level, as an exact scaling: a quotient by a power of two
  // and the product with its (exact) reciprocal round identically
  const float rw = ldexpf(1.f, -g.level);
  const vec3f pl = mk3((pos.x - lo.x) * rw, (pos.y - lo.y) * rw, (pos.z - lo.z) * rw);
  const int ix = (int)pl.x, iy = (int)pl.y, iz = (int)pl.z;
  const int ax = ix + 1 < nsx - 1 ? ix + 1 : nsx - 1;
  const int ay = iy + 1 < nsy - 1 ? iy + 1 : nsy - 1;
  const int az = iz + 1 < nsz - 1 ? iz + 1 : nsz - 1;
  const float *s = scalars + g.begin;
  const int sy = nsx, sz = nsy * nsx;
  const float f1 = s[iz * sz + iy * sy + ix];
  const float f2 = s[iz * sz + iy * sy + ax];
  const float f3 = s[iz * sz + ay * sy + ix];
  const float f4 = s[iz * sz + ay * sy + ax];
  const float f5 = s[az * sz + iy * sy + ix];
  const float f6 = s[az * sz + iy * sy + ax];
  const float f7 = s[az * sz + ay * sy + ix];
  const float f8 = s[az * sz + ay * sy + ax];
  if (f1 != f1 || f2 != f2 || f3 != f3 || f4 != f4 || f5 != f5 || f6 != f6 || f7 != f7 || f8 != f8)
    return 0;
  const vec3f frac = mk3(pl.x - (float)ix, pl.y - (float)iy, pl.z - (float)iz);
  const float f12 = lerpf(f1, f2, frac.x);
  const float f56 = lerpf(f5, f6, frac.x);
  const float f34 = lerpf(f3, f4, frac.x);
  const float f78 = lerpf(f7, f8, frac.x);
  const float f1234 = lerpf(f12, f34, frac.y);
  const float f5678 = lerpf(f56, f78, frac.y);
  value = lerpf(f1234, f5678, frac.z);
  cellID = iz * g.dims[1] * g.dims[0] + iy * g.dims[0] + ix;
  const float m0 = margin, m1 = 1.f - margin;
  const bool interior = frac.x >= m0 && frac.x <= m1 && frac.y >= m0 && frac.y <= m1 &&
                        frac.z >= m0 && frac.z <= m1 && ix < g.dims[0] && iy < g.dims[1] &&
                        iz < g.dims[2];
  (void)nsz;
  return interior ? 2 : 1;
}

// ---------------------------------------------------------------------------
// Plane.h:23-54 makePlane/eval and Plane.h:57-99 intersectTet
// ---------------------------------------------------------------------------
struct Plane {
  vec3f N;
  float d;
};
EXA_HD Plane makePlane(vec3f a, vec3f b, vec3f c) {
  Plane p;
  p.N = cross(b - a, c - a);
  p.d = dot(a, p.N);
  return p;
}
EXA_HD float evalPlane(const Plane &p, vec3f v) { return dot(v, p.N) - p.d; }

EXA_HD bool intersectTet(float &value, vec3f P, float4 a, float4 b, float4 c, float4 d) {
  const vec3f va = xyz(a) - P;
  const vec3f vb = xyz(b) - P;
  const vec3f vc = xyz(c) - P;
  const vec3f vd = xyz(d) - P;
  const Plane pa = makePlane(vb, vd, vc);
  const Plane pb = makePlane(va, vc, vd);
  const Plane pc = makePlane(va, vd, vb);
  const Plane pd = makePlane(va, vb, vc);
  const vec3f zero = mk3s(0.f);
  const float fa = evalPlane(pa, zero) / evalPlane(pa, va);
  if (fa < 0.f || fa > 1.f) return false;
  const float fb = evalPlane(pb, zero) / evalPlane(pb, vb);
  if (fb < 0.f || fa > 1.f) return false;
  const float fc = evalPlane(pc, zero) / evalPlane(pc, vc);
  if (fc < 0.f || fa > 1.f) return false;
  const float fd = evalPlane(pd, zero) / evalPlane(pd, vd);
  if (fd < 0.f || fa > 1.f) return false;
  value = fa * a.w + fb * b.w + fc * c.w + fd * d.w;
  return true;
}

// ---------------------------------------------------------------------------
// UElems.h:40-44 det; pyramid 46-177, wedge 180-310, hex 313-464 (Newton on
// the isoparametric map, adapted there from OpenVKL)
// ---------------------------------------------------------------------------
// x . (y x z) of the Newton solve, each cross-product component and the dot
// product's sums by fused multiply-adds (the oracle fuses the same terms)
EXA_HD float det3(vec3f x, vec3f y, vec3f z) {
  const vec3f c = mk3(fmaf(y.y, z.z, -(y.z * z.y)), fmaf(y.z, z.x, -(y.x * z.z)), fmaf(y.x, z.y, -(y.y * z.x)));
  return fmaf(x.z, c.z, fmaf(x.y, c.y, x.x * c.x));
}

template <int N>
struct ShapeFn;

template <>
struct ShapeFn<5> {
  static EXA_HD void sf(const float *pc, float *sf) {
    float rm = 1.f - pc[0], sm = 1.f - pc[1], tm = 1.f - pc[2];
    sf[0] = rm * sm * tm;
    sf[1] = pc[0] * sm * tm;
    sf[2] = pc[0] * pc[1] * tm;
    sf[3] = rm * pc[1] * tm;
    sf[4] = pc[2];
  }
  static EXA_HD void df(const float *pc, float *d) {
    d[0] = -(pc[1] - 1.f) * (pc[2] - 1.f);
    d[1] = (pc[1] - 1.f) * (pc[2] - 1.f);
    d[2] = pc[1] - pc[1] * pc[2];
    d[3] = pc[1] * (pc[2] - 1.f);
    d[4] = 0.f;
    d[5] = -(pc[0] - 1.f) * (pc[2] - 1.f);
    d[6] = pc[0] * (pc[2] - 1.f);
    d[7] = pc[0] - pc[0] * pc[2];
    d[8] = (pc[0] - 1.f) * (pc[2] - 1.f);
    d[9] = 0.f;
    d[10] = -(pc[0] - 1.f) * (pc[1] - 1.f);
    d[11] = pc[0] * (pc[1] - 1.f);
    d[12] = -pc[0] * pc[1];
    d[13] = (pc[0] - 1.f) * pc[1];
    d[14] = 1.f;
  }
};

template <>
struct ShapeFn<6> {
  static EXA_HD void sf(const float *pc, float *sf) {
    sf[0] = (1.f - pc[0] - pc[1]) * (1.f - pc[2]);
    sf[1] = pc[0] * (1.f - pc[2]);
    sf[2] = pc[1] * (1.f - pc[2]);
    sf[3] = (1.f - pc[0] - pc[1]) * pc[2];
    sf[4] = pc[0] * pc[2];
    sf[5] = pc[1] * pc[2];
  }
  static EXA_HD void df(const float *pc, float *d) {
    d[0] = -1.f + pc[2];
    d[1] = 1.f - pc[2];
    d[2] = 0.f;
    d[3] = -pc[2];
    d[4] = pc[2];
    d[5] = 0.f;
    d[6] = -1.f + pc[2];
    d[7] = 0.f;
    d[8] = 1.f - pc[2];
    d[9] = -pc[2];
    d[10] = 0.f;
    d[11] = pc[2];
    d[12] = -1.f + pc[0] + pc[1];
    d[13] = -pc[0];
    d[14] = -pc[1];
    d[15] = 1.f - pc[0] - pc[1];
    d[16] = pc[0];
    d[17] = pc[1];
  }
};

template <>
struct ShapeFn<8> {
  static EXA_HD void sf(const float *pc, float *sf) {
    float rm = 1.f - pc[0], sm = 1.f - pc[1], tm = 1.f - pc[2];
    sf[0] = rm * sm * tm;
    sf[1] = pc[0] * sm * tm;
    sf[2] = pc[0] * pc[1] * tm;
    sf[3] = rm * pc[1] * tm;
    sf[4] = rm * sm * pc[2];
    sf[5] = pc[0] * sm * pc[2];
    sf[6] = pc[0] * pc[1] * pc[2];
    sf[7] = rm * pc[1] * pc[2];
  }
  static EXA_HD void df(const float *pc, float *d) {
    float rm = 1.f - pc[0], sm = 1.f - pc[1], tm = 1.f - pc[2];
    d[0] = -sm * tm;
    d[1] = sm * tm;
    d[2] = pc[1] * tm;
    d[3] = -pc[1] * tm;
    d[4] = -sm * pc[2];
    d[5] = sm * pc[2];
    d[6] = pc[1] * pc[2];
    d[7] = -pc[1] * pc[2];
    d[8] = -rm * tm;
    d[9] = -pc[0] * tm;
    d[10] = pc[0] * tm;
    d[11] = rm * tm;
    d[12] = -rm * pc[2];
    d[13] = -pc[0] * pc[2];
    d[14] = pc[0] * pc[2];
    d[15] = rm * pc[2];
    d[16] = -rm * sm;
    d[17] = -pc[0] * sm;
    d[18] = -pc[0] * pc[1];
    d[19] = -rm * pc[1];
    d[20] = rm * sm;
    d[21] = pc[0] * sm;
    d[22] = pc[0] * pc[1];
    d[23] = rm * pc[1];
  }
};

// UElems.h:88-177 / 219-310 / 372-464.  The position and Jacobian sums and
// the interpolated value accumulate by fused multiply-adds, one rounding per
// term, as nvcc compiles them by default (--fmad=true); the oracle fuses the
// same terms.
template <int N>
EXA_HD bool intersectNewton(float &value, vec3f P, const float4 *V) {
  const float determinantTolerance = 1e-6f;
  float pc[3] = {.5f, .5f, .5f};
  float derivs[3 * N];
  float weights[N];
  bool converged = false;
  for (int iteration = 0; !converged && (iteration < 10); iteration++) {
    ShapeFn<N>::sf(pc, weights);
    ShapeFn<N>::df(pc, derivs);
    vec3f fcol = mk3s(0.f), rcol = mk3s(0.f), scol = mk3s(0.f), tcol = mk3s(0.f);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const vec3f pt = xyz(V[i]);
      fcol = fma3(pt, weights[i], fcol);
      rcol = fma3(pt, derivs[i], rcol);
      scol = fma3(pt, derivs[i + N], scol);
      tcol = fma3(pt, derivs[i + 2 * N], tcol);
    }
    fcol = fcol - P;
    const float d = det3(rcol, scol, tcol);
    if (fabsf(d) < determinantTolerance) return false;
    const float d0 = det3(fcol, scol, tcol) / d;
    const float d1 = det3(rcol, fcol, tcol) / d;
    const float d2 = det3(rcol, scol, fcol) / d;
    pc[0] = pc[0] - d0;
    pc[1] = pc[1] - d1;
    pc[2] = pc[2] - d2;
    // CONVERGED / DIVERGED are double literals in the reference: for a float
    // x, (double)x < 1e-4 holds exactly when x < 0x1.a36e30p-14f (the float
    // just above 1e-4), and 1e6 is a float
    const float kConverged = 0x1.a36e30p-14f, kDiverged = 1e6f;
    if (fabsf(d0) < kConverged && fabsf(d1) < kConverged && fabsf(d2) < kConverged) {
      converged = true;
    } else if (fabsf(pc[0]) > kDiverged || fabsf(pc[1]) > kDiverged || fabsf(pc[2]) > kDiverged) {
      return false;
    }
  }
  if (!converged) return false;
  const float lowerlimit = (float)(0.0 - 1e-6);
  const float upperlimit = (float)(1.0 + 1e-6);
  bool inside = pc[0] >= lowerlimit && pc[0] <= upperlimit && pc[1] >= lowerlimit &&
                pc[1] <= upperlimit && pc[2] >= lowerlimit && pc[2] <= upperlimit;
  if (N == 6) inside = inside && (pc[0] + pc[1] <= upperlimit);
  if (inside) {
    float val = 0.f;
#pragma unroll
    for (int i = 0; i < N; i++) val = fmaf(weights[i], V[i].w, val);
    value = val;
    return true;
  }
  return false;
}

// intersectNewton<N> for a vertex count known only at run time (5, 6 or 8):
// the shape functions are evaluated for that type into 8 slots and the
// iteration is shared, so the lanes of a wave holding different element types
// run one Newton loop together instead of one per type.  The sums stop at
// nv terms, so every value matches intersectNewton<nv> bit for bit.
template <int N>
EXA_HD void shapeSlots(const float *pc, float w[8], float dr[8], float ds[8], float dt[8]) {
  float sw[N], d[3 * N];
  ShapeFn<N>::sf(pc, sw);
  ShapeFn<N>::df(pc, d);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    w[i] = i < N ? sw[i] : 0.f;
    dr[i] = i < N ? d[i] : 0.f;
    ds[i] = i < N ? d[i + N] : 0.f;
    dt[i] = i < N ? d[i + 2 * N] : 0.f;
  }
}
// Pyramids take the hexahedron's evaluation: in slots 0-3 the pyramid's
// shape functions and derivatives are the hexahedron's bit for bit (with
// fl(x - 1) = -fl(1 - x), e.g. -(s-1)*(t-1) = -(sm*tm) = -sm*tm), except
// dr[2] = s - s*t and ds[2] = r - r*t; slot 4 is the apex (w = t, dt = 1).
// So a wave holding pyramids and hexahedra evaluates one set of formulas.
EXA_HD void shapeAny(int nv, const float *pc, float w[8], float dr[8], float ds[8], float dt[8]) {
  if (nv == 6) {
    shapeSlots<6>(pc, w, dr, ds, dt);
    return;
  }
  shapeSlots<8>(pc, w, dr, ds, dt);
  if (nv == 5) {
    dr[2] = pc[1] - pc[1] * pc[2];
    ds[2] = pc[0] - pc[0] * pc[2];
    w[4] = pc[2];
    dr[4] = 0.f;
    ds[4] = 0.f;
    dt[4] = 1.f;
  }
}

EXA_HD bool intersectNewtonAny(float &value, vec3f P, const float4 *V, int nv) {
  const float determinantTolerance = 1e-6f;
  float pc[3] = {.5f, .5f, .5f};
  float w[8], dr[8], ds[8], dt[8];
  bool converged = false;
  for (int iteration = 0; !converged && (iteration < 10); iteration++) {
    shapeAny(nv, pc, w, dr, ds, dt);
    vec3f fcol = mk3s(0.f), rcol = mk3s(0.f), scol = mk3s(0.f), tcol = mk3s(0.f);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (i < nv) {
        const vec3f pt = xyz(V[i]);
        fcol = fma3(pt, w[i], fcol);
        rcol = fma3(pt, dr[i], rcol);
        scol = fma3(pt, ds[i], scol);
        tcol = fma3(pt, dt[i], tcol);
      }
    }
    fcol = fcol - P;
    const float d = det3(rcol, scol, tcol);
    if (fabsf(d) < determinantTolerance) return false;
    const float d0 = det3(fcol, scol, tcol) / d;
    const float d1 = det3(rcol, fcol, tcol) / d;
    const float d2 = det3(rcol, scol, fcol) / d;
    pc[0] = pc[0] - d0;
    pc[1] = pc[1] - d1;
    pc[2] = pc[2] - d2;
    const float kConverged = 0x1.a36e30p-14f, kDiverged = 1e6f;
    if (fabsf(d0) < kConverged && fabsf(d1) < kConverged && fabsf(d2) < kConverged) {
      converged = true;
    } else if (fabsf(pc[0]) > kDiverged || fabsf(pc[1]) > kDiverged || fabsf(pc[2]) > kDiverged) {
      return false;
    }
  }
  if (!converged) return false;
  const float lowerlimit = (float)(0.0 - 1e-6);
  const float upperlimit = (float)(1.0 + 1e-6);
  bool inside = pc[0] >= lowerlimit && pc[0] <= upperlimit && pc[1] >= lowerlimit && pc[1] <= upperlimit &&
                pc[2] >= lowerlimit && pc[2] <= upperlimit;
  if (nv == 6) inside = inside && (pc[0] + pc[1] <= upperlimit);
  if (inside) {
    float val = 0.f;
#pragma unroll
    for (int i = 0; i < 8; i++)
      if (i < nv) val = fmaf(w[i], V[i].w, val);
    value = val;
    return true;
  }
  return false;
}

// programs/ExaStitchSampler.cu:133-161 StitchGeomIsect dispatch on the
// number of valid indices in the element's int[8] row.
EXA_HD bool intersectElement(float &value, vec3f P, const int32_t *__restrict__ idx8,
                             const float4 *__restrict__ vertices) {
  const int4 i0 = *(const int4 *)idx8;
  const int4 i1 = *(const int4 *)(idx8 + 4);
  const int numVerts = (i0.x >= 0) + (i0.y >= 0) + (i0.z >= 0) + (i0.w >= 0) + (i1.x >= 0) +
                       (i1.y >= 0) + (i1.z >= 0) + (i1.w >= 0);
  // rows are validated on the host: indices in range, -1 padding only at the
  // end, 4/5/6/8 valid entries
  if (numVerts < 4) return false;
  float4 v[8];
  v[0] = vertices[i0.x];
  v[1] = vertices[i0.y];
  v[2] = vertices[i0.z];
  v[3] = vertices[i0.w];
  if (numVerts == 4) return intersectTet(value, P, v[0], v[1], v[2], v[3]);
  v[4] = vertices[i1.x];
  if (numVerts == 5) return intersectNewton<5>(value, P, v);
  v[5] = vertices[i1.y];
  if (numVerts == 6) return intersectNewton<6>(value, P, v);
  if (numVerts != 8) return false;
  v[6] = vertices[i1.z];
  v[7] = vertices[i1.w];
  return intersectNewton<8>(value, P, v);
}

}  // namespace exa
