// exa_render_common.h -- device helpers shared by the megakernel
// (render.hip) and the wavefront renderer (wavefront.hip).
#pragma once
#include "exa_locate.h"

namespace exa {

constexpr int kTileX = 16, kTileY = 16, kRenderBlock = kTileX * kTileY;

struct Ray {
  vec3f org, dir;
  float tmin, tmax;
};

EXA_DEV vec3f ld3(const float *a) { return mk3(a[0], a[1], a[2]); }
EXA_DEV vec3f xfmPoint(const float *a, vec3f p) {
  return ld3(a) * p.x + (ld3(a + 3) * p.y + (ld3(a + 6) * p.z + ld3(a + 9)));
}
EXA_DEV vec3f xfmVector(const float *a, vec3f v) { return ld3(a) * v.x + (ld3(a + 3) * v.y + ld3(a + 6) * v.z); }

// sampler/Sampler.h:103-118 boxTest
EXA_DEV bool boxTest(const Ray &ray, const float *box, float &t0, float &t1) {
  const vec3f lo = (ld3(box) - ray.org) / ray.dir;
  const vec3f hi = (ld3(box + 3) - ray.org) / ray.dir;
  const vec3f nr = vmin(lo, hi);
  const vec3f fr = vmax(lo, hi);
  t0 = fmaxf(ray.tmin, reduce_max(nr));
  t1 = fminf(ray.tmax, reduce_min(fr));
  return t0 < t1;
}

// deviceCode.cu:52-60 and the clip-plane blocks at 1010-1021 / 1097-1107
EXA_DEV void clipInterval(const RenderParams &P, const Ray &ray, float &t0, float &t1) {
  if (!P.clipEnabled) return;
  const vec3f N = ld3(P.clipN);
  const float s = dot(N, ray.dir);
  float plane_t;
  bool backFace = false;
  if (s == 0.f) {
    plane_t = 3.402823466e+38f;
  } else {
    backFace = s > 0.f;
    plane_t = (P.clipD - dot(N, ray.org)) / s;
  }
  if (plane_t > t0 && !backFace) t0 = fmaxf(t0, plane_t);
  if (plane_t < t1 && backFace) t1 = fminf(plane_t, t1);
}

// lookupTransferFunction with the texture unit's documented linear filter
// (x = u*N - 0.5, weights with 8 fractional bits, border texels 0).
EXA_DEV float4 lookupTF(const RenderParams &P, float f) {
  float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
  const float x0 = P.tfDomain[0], x1 = P.tfDomain[1];
  const int N = P.tfN;
  if (x0 >= x1 || N <= 0) return r;
  f -= x0;
  f /= (x1 - x0);
  const float x = f * (float)N - 0.5f;
  if (!(x >= -1.f) || !(x < (float)N)) return r;
  const float fl = floorf(x);
  const int i = (int)fl;
  const float alpha = floorf((x - fl) * 256.f + 0.5f) / 256.f;
  float4 t0 = r, t1 = r;
  if (i >= 0 && i < N) t0 = P.tf[i];
  if (i + 1 >= 0 && i + 1 < N) t1 = P.tf[i + 1];
  r.x = (1.f - alpha) * t0.x + alpha * t1.x;
  r.y = (1.f - alpha) * t0.y + alpha * t1.y;
  r.z = (1.f - alpha) * t0.z + alpha * t1.z;
  r.w = (1.f - alpha) * t0.w + alpha * t1.w;
  return r;
}

struct LocalCounters {
  uint32_t c[kNumCounters];
};

enum { kBoundary = 0, kScattering = 1 };

// deviceCode.cu:43-50 backGroundColor
EXA_DEV float4 backgroundColor(int py, int H) {
  const float t = (float)py / (float)H;
  return make_float4((1.f - t) * 1.f + t * 0.5f, (1.f - t) * 1.f + t * 0.7f, (1.f - t) * 1.f + t * 1.f, 1.f);
}

// owl make_8bit: clamp(int(f*256), 0, 255)
EXA_DEV uint32_t make8bit(float f) {
  const int32_t i = sat_i32(f * 256.f);
  return (uint32_t)(i < 0 ? 0 : (i > 255 ? 255 : i));
}


// deviceCode.cu:88-101 generateRay, with the sub-image remap of
// renderFrame_Impl (deviceCode.cu:1285-1296); draws two numbers from rng.
EXA_DEV Ray primaryRay(const RenderParams &P, int px, int py, Lcg &rng) {
  const float rx = rng();
  const float ry = rng();
  float sx = (float)px + rx, sy = (float)py + ry;
  if (P.subImageActive) {
    sx /= (float)P.W;
    sy /= (float)P.H;
    sx = (1.f - sx) * P.subImage[0] + sx * P.subImage[2];
    sy = (1.f - sy) * P.subImage[1] + sy * P.subImage[3];
    sx *= (float)P.W;
    sy *= (float)P.H;
  }
  Ray ray;
  ray.org = ld3(P.camOrg);
  vec3f dir = ld3(P.camDir00) + sx * ld3(P.camDu) + sy * ld3(P.camDv);
  dir = normalize(dir);
  if (fabsf(dir.x) < 1e-5f) dir.x = 1e-5f;
  if (fabsf(dir.y) < 1e-5f) dir.y = 1e-5f;
  if (fabsf(dir.z) < 1e-5f) dir.z = 1e-5f;
  ray.dir = dir;
  ray.tmin = 0.f;
  ray.tmax = 1e10f;
  return ray;
}

// One ray segment through the majorant grid in voxel space: the set-up of
// sampleInteraction (deviceCode.cu:713-719) and the state of dda3
// (DDA.h:26-113).  t0 is the entry distance of the current macrocell.
struct Segment {
  vec3f org, dir;
  float tmax, t0;
  int32_t cell[3];
  vec3f tnext, dist;
};

EXA_DEV void segmentBegin(const RenderParams &P, const GridView &gv, Ray ray, Segment &S) {
  ray.org = xfmPoint(P.voxelXfm, ray.org);
  ray.dir = xfmVector(P.voxelXfm, ray.dir);
  const float dt_scale = length(ray.dir);
  ray.dir = normalize(ray.dir);
  ray.tmin = ray.tmin * dt_scale;
  ray.tmax = ray.tmax * dt_scale;
  const int32_t *dims = gv.dims;
  const float *bounds = gv.bounds;
  const vec3f rcp_dir = mk3(1.f / ray.dir.x, 1.f / ray.dir.y, 1.f / ray.dir.z);
  const vec3f lo = (ld3(bounds) - ray.org) * rcp_dir;
  const vec3f hi = (ld3(bounds + 3) - ray.org) * rcp_dir;
  const vec3f tnear = vmin(lo, hi);
  const vec3f tfar = vmax(lo, hi);
  {
    const vec3f V01 = (ray.org - ld3(bounds)) / (ld3(bounds + 3) - ld3(bounds));
    S.cell[0] = clampi(sat_i32(V01.x * (float)dims[0]), 0, dims[0] - 1);
    S.cell[1] = clampi(sat_i32(V01.y * (float)dims[1]), 0, dims[1] - 1);
    S.cell[2] = clampi(sat_i32(V01.z * (float)dims[2]), 0, dims[2] - 1);
  }
  S.dist = (tfar - tnear) / mk3((float)dims[0], (float)dims[1], (float)dims[2]);
  S.tnext = mk3(ray.dir.x > 0.f ? tnear.x + (float)(S.cell[0] + 1) * S.dist.x
                                : tnear.x + (float)(dims[0] - S.cell[0]) * S.dist.x,
                ray.dir.y > 0.f ? tnear.y + (float)(S.cell[1] + 1) * S.dist.y
                                : tnear.y + (float)(dims[1] - S.cell[1]) * S.dist.y,
                ray.dir.z > 0.f ? tnear.z + (float)(S.cell[2] + 1) * S.dist.z
                                : tnear.z + (float)(dims[2] - S.cell[2]) * S.dist.z);
  S.org = ray.org;
  S.dir = ray.dir;
  S.tmax = ray.tmax;
  S.t0 = fmaxf(ray.tmin, 0.f);
}

// exit distance of the current macrocell, clipped to the segment
EXA_DEV float segmentCellExit(const Segment &S) { return fminf(reduce_min(S.tnext), S.tmax); }

EXA_DEV int64_t segmentCellID(const GridView &gv, const Segment &S) {
  return ((int64_t)S.cell[2] * gv.dims[1] + S.cell[1]) * gv.dims[0] + S.cell[0];
}

// DDA step into the next macrocell entered at t1; false once the ray has
// left the grid.
EXA_DEV bool segmentNextCell(const GridView &gv, Segment &S, float t1) {
  const float t_closest = reduce_min(S.tnext);
  if (S.tnext.x == t_closest) {
    S.tnext.x += S.dist.x;
    S.cell[0] += S.dir.x > 0.f ? 1 : -1;
    if (S.cell[0] == (S.dir.x > 0.f ? gv.dims[0] : -1)) return false;
  }
  if (S.tnext.y == t_closest) {
    S.tnext.y += S.dist.y;
    S.cell[1] += S.dir.y > 0.f ? 1 : -1;
    if (S.cell[1] == (S.dir.y > 0.f ? gv.dims[1] : -1)) return false;
  }
  if (S.tnext.z == t_closest) {
    S.tnext.z += S.dist.z;
    S.cell[2] += S.dir.z > 0.f ? 1 : -1;
    if (S.cell[2] == (S.dir.z > 0.f ? gv.dims[2] : -1)) return false;
  }
  S.t0 = t1;
  return true;
}

// Free-flight step of Woodcock / ratio tracking inside one macrocell
// (deviceCode.cu:734-742): false when the step leaves the cell at t1.
EXA_DEV bool trackingStep(float &t, float majorant, float os, float t1, Lcg &rng) {
  t -= exa_logf(1.f - rng()) / (majorant * os);
  return !(t >= t1);
}

// deviceCode.cu:625-697 classifySample, shade modes 1 (Gridlets) and 2
// (Teaser): the colour of a gridlet sample becomes a checkerboard over the
// gridlet's cells (white / dark blue), an element sample turns red; Teaser
// blends the TF colour into that colour across a diagonal band of the image.
// Alpha stays the TF's, so tracking and the random stream are unchanged.
EXA_DEV float4 shadeSample(const RenderParams &P, const Gridlet *gridlets, const Sample &s, float4 xf, int px,
                           int py) {
  vec3f rgb2;
  if (s.cellID == -1) {
    rgb2 = mk3(1.f, 0.f, 0.f);
  } else {
    const Gridlet &g = gridlets[s.primID];
    const int ix = s.cellID % g.dims[0];
    const int iy = s.cellID / g.dims[0] % g.dims[1];
    const int iz = s.cellID / (g.dims[0] * g.dims[1]);
    int col = (ix % 2 == iy % 2) ? 1 : 0;
    col = iz % 2 == 0 ? col : !col;
    rgb2 = col == 0 ? mk3(1.f, 1.f, 1.f) : mk3(0.f, 0.f, 0.7f);
  }
  if (P.shadeMode == 1) return make_float4(rgb2.x, rgb2.y, rgb2.z, xf.w);
  const float u = (float)px / (float)P.W, v = (float)py / (float)P.H;
  const float ax = 0.4f, ay = 0.6f, bx = 0.6f, by = 0.4f;
  const float A = bx - ax, B = by - ay;
  const float C1 = A * ax + B * ay;
  const float C2 = A * bx + B * by;
  const float C = A * u + B * v;
  if (C <= C1) return xf;
  if (C >= C2) return make_float4(rgb2.x, rgb2.y, rgb2.z, xf.w);
  return make_float4((xf.x * (C2 - C) + rgb2.x * (C - C1)) / (C2 - C1),
                     (xf.y * (C2 - C) + rgb2.y * (C - C1)) / (C2 - C1),
                     (xf.z * (C2 - C) + rgb2.z * (C - C1)) / (C2 - C1), xf.w);
}

// A located sample at a tentative collision (deviceCode.cu:742-789 for
// Woodcock, 800-833 for ratio tracking).  Returns true when the segment
// ends in an interaction; primID caches the last gridlet for testSample.
// (px, py) is the pixel, for the Teaser shade mode.
EXA_DEV bool resolveCollision(const RenderParams &P, const Gridlet *gridlets, int px, int py, const Sample &s,
                              bool shadow, float majorant, Lcg &rng, int32_t &primID, float4 &xf, float &Tr) {
  if (s.primID < 0) return false;
  if (s.cellID != -1) primID = s.primID;
  xf = lookupTF(P, s.value);
  // ratio tracking reads alpha only, so only Woodcock needs the shade mode
  if (P.shadeMode != 0 && !shadow) xf = shadeSample(P, gridlets, s, xf, px, py);
  const float os = P.opacityScale;
  if (!shadow) {
    const float u = rng();
    const float sigmaT = xf.w;
    if (sigmaT * os >= u * majorant) {
      Tr = 0.f;
      return true;
    }
  } else {
    const float sigmaT = xf.w;
    Tr *= 1.f - (sigmaT * os) / majorant;
    if (Tr <= 0.f) return true;
  }
  return false;
}

// deviceCode.cu:151-154 over
EXA_DEV float4 over(float4 a, float4 b) {
  const float k = 1.f - a.w;
  return make_float4(a.x + k * b.x, a.y + k * b.y, a.z + k * b.z, a.w + k * b.w);
}

// deviceCode.cu:104-112 firstSampleT: the first sample at or after t0 on the
// lattice dt * (ils + k)
EXA_DEV float firstSampleT(float t0, float dt, float ils) {
  const float numSegs = floorf((t0 - dt * ils) / dt);
  float t = dt * (ils + numSegs);
  if (t < t0) t += dt;
  return t;
}

// deviceCode.cu:155-185 heatMap = temperature_to_rgb (hue from blue to red,
// value from .5 to 1)
EXA_DEV vec3f heatMapColor(float t) {
  auto sat = [](float x) { return fminf(fmaxf(x, 0.f), 1.f); };
  const float K = 4.0f / 6.0f;
  const float h = K - K * t;
  const float v = .5f + 0.5f * t;
  const float hs = sat(h) * 6.0f;
  const float r = sat(fabsf(hs - 3.f) - 1.0f);
  const float g = sat(2.0f - fabsf(hs - 2.0f));
  const float b = sat(2.0f - fabsf(hs - 4.0f));
  return mk3(v * r, v * g, v * b);
}

// deviceCode.cu:856-926 sampleLight up to the shadow ray (+ directional
// lights, no reference counterpart: unit falloff, ray to the domain exit).
EXA_DEV Ray lightRay(const RenderParams &P, vec3f pos, int numLights, Lcg &rng, int &which, float &ld) {
  which = (int)(rng() * (float)numLights);
  if (which == numLights) which = 0;
  const LightParams &L = P.lights[which];
  Ray ray;
  ld = 1.f;
  ray.org = pos;
  ray.tmin = 0.f;
  if (L.type == 1) {
    ray.dir = normalize(ld3(L.pos));
    ray.tmax = 1e30f;
  } else {
    vec3f lightDir = ld3(L.pos) - pos;
    ld = length(xfmVector(P.lightXfm, lightDir));
    const float ldWorld = length(lightDir);
    lightDir = normalize(lightDir);
    ray.dir = lightDir;
    ray.tmax = ldWorld;
  }
  float t00 = 1e30f, t11 = -1e30f;
  boxTest(ray, P.worldBounds, t00, t11);
  clipInterval(P, ray, t00, t11);
  ray.tmax = t11;
  return ray;
}

// radiance arriving over the shadow ray with transmittance Tr
EXA_DEV vec3f lightResult(const RenderParams &P, int which, float Tr, float ld) {
  const LightParams &L = P.lights[which];
  if (L.type == 1) return mk3s(Tr * L.intensity);
  return mk3s(Tr / (ld * ld) * L.intensity);
}

// deviceCode.cu:144-149, 965-985 (g = 0); draws two numbers
EXA_DEV vec3f hgSample(vec3f wo, Lcg &rng) {
  const float u1 = rng();
  const float u2 = rng();
  const float cost = 1.f - 2.f * u1;
  const float sint = sqrtf(fmaxf(0.f, 1.f - cost * cost));
  const float phi = (float)(2.0f * M_PI * (double)u2);
  const vec3f w = wo;
  const vec3f v = fabsf(w.x) > fabsf(w.y) ? normalize(mk3(-w.z, 0.f, w.x)) : normalize(mk3(0.f, w.z, -w.y));
  const vec3f u = cross(v, w);
  float sp, cp;
  exa_sincosf(phi, sp, cp);
  return (sint * cp) * u + (sint * sp) * v + cost * (-w);
}

// pathTracingIntegrate (deviceCode.cu:1048-1072): bounce cap, albedo and
// Russian roulette at an interaction; false ends the path (throughput 0).
EXA_DEV bool ptInteraction(const RenderParams &P, float4 xf, vec3f &throughput, unsigned &bounce, Lcg &rng) {
  if (bounce++ >= (unsigned)P.maxBounces) {
    throughput = mk3s(0.f);
    return false;
  }
  const vec3f albedo = mk3(xf.x, xf.y, xf.z);
  throughput = throughput * albedo;
  const float Pr = reduce_max(throughput);
  if (Pr < 0.2f) {
    if (rng() > Pr) {
      throughput = mk3s(0.f);
      return false;
    }
    throughput = mk3(throughput.x / Pr, throughput.y / Pr, throughput.z / Pr);
  }
  throughput = throughput + mk3s(0.f);
  return true;
}

// final colour of a path-traced sample (deviceCode.cu:1110-1117)
EXA_DEV float4 ptResult(vec3f Ld, vec3f throughput, unsigned bounce, int numLights, float4 bg) {
  if (numLights == 0) Ld = mk3s(1.f);
  const vec3f L = Ld * throughput;
  return bounce ? make_float4(L.x, L.y, L.z, 1.f) : bg;
}

EXA_DEV int countLights(const RenderParams &P) {
  int n = 0;
  for (int i = 0; i < kLightsMax; ++i)
    if (P.lights[i].on) n++;
  return n;
}

// renderFrame_Impl tail (deviceCode.cu:1320-1344): progressive
// accumulation, crosshair / selection inversion, 8-bit RGBA.
EXA_DEV void finishPixel(const RenderParams &P, int px, int py, float4 acc) {
  const int pixelID = px + P.W * py;
  float4 *a = P.accum + pixelID;
  if (P.accumID > 0) {
    const float4 o = *a;
    acc.x += o.x;
    acc.y += o.y;
    acc.z += o.z;
    acc.w += o.w;
  }
  *a = acc;
  const float sc = 1.f / (float)(P.accumID + 1);
  acc.x *= sc;
  acc.y *= sc;
  acc.z *= sc;
  acc.w *= sc;
  bool invert = P.debugCrosshair && (px == P.W / 2 || py == P.H / 2);
  if (P.subImageSelecting) {
    // deviceCode.cu:134-141 subImageSelectionWin, 1337-1340
    const int lx = clampi(sat_i32(P.subImageSel[0] * (float)P.W), 0, P.W - 1);
    const int ly = clampi(sat_i32(P.subImageSel[1] * (float)P.H), 0, P.H - 1);
    const int hx = clampi(sat_i32(P.subImageSel[2] * (float)P.W), 0, P.W - 1);
    const int hy = clampi(sat_i32(P.subImageSel[3] * (float)P.H), 0, P.H - 1);
    invert = invert || (((px == lx || px == hx) && (py >= ly && py <= hy)) ||
                        ((py == ly || py == hy) && (px >= lx && px <= hx)));
  }
  if (invert) {
    acc.x = 1.f - acc.x;
    acc.y = 1.f - acc.y;
    acc.z = 1.f - acc.z;
    acc.w = 1.f - acc.w;
  }
  if (P.fb) {
    const float is = 1.f / (float)P.spp;
    P.fb[pixelID] = (make8bit(acc.x * is) << 0) + (make8bit(acc.y * is) << 8) + (make8bit(acc.z * is) << 16) +
                    (make8bit(acc.w * is) << 24);
  }
}

}  // namespace exa
