// render.hip -- the raygen megakernel of the exaStitch path tracer for
// gfx950: one thread per pixel, 16x16 pixel tiles per workgroup.
//
// Follows deviceCode.cu: renderFrame_Impl (1268-1345), generateRay (88-101),
// pathTracingIntegrate (991-1120), directLightingIntegrate (1126-1207),
// sampleInteraction with Woodcock / ratio tracking (699-854), sampleLight
// (856-926), henyeyGreensteinSample (965-985), lookupTransferFunction
// (320-341), and dda3 (DDA.h:26-113).
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

template <bool kCount, int kAccel>
struct Ctx {
  const RenderParams &P;
  const SamplerView &sv;
  const GridView &gv;
  Lcg &rng;
  int32_t *stack;
  LocalCounters &cnt;
};

enum { kBoundary = 0, kScattering = 1 };

struct Interaction {
  int type;
  vec3f pos;
  float Tr;
  float4 xf;
};

// deviceCode.cu:699-854: DDA over the majorant grid, Woodcock (shadow=false)
// or ratio tracking (shadow=true) inside each macrocell.
template <bool kCount, int kAccel>
EXA_DEV Interaction sampleInteraction(Ctx<kCount, kAccel> &C, Ray ray, bool shadow) {
  const RenderParams &P = C.P;
  ray.org = xfmPoint(P.voxelXfm, ray.org);
  ray.dir = xfmVector(P.voxelXfm, ray.dir);
  const float dt_scale = length(ray.dir);
  ray.dir = normalize(ray.dir);
  ray.tmin = ray.tmin * dt_scale;
  ray.tmax = ray.tmax * dt_scale;
  Interaction I;
  I.type = kBoundary;
  I.Tr = 1.f;
  I.pos = mk3s(0.f);
  I.xf = make_float4(0.f, 0.f, 0.f, 0.f);
  const float os = P.opacityScale;

  // DDA.h:26-113 dda3
  const int32_t *dims = C.gv.dims;
  const float *bounds = C.gv.bounds;
  const vec3f rcp_dir = mk3(1.f / ray.dir.x, 1.f / ray.dir.y, 1.f / ray.dir.z);
  const vec3f lo = (ld3(bounds) - ray.org) * rcp_dir;
  const vec3f hi = (ld3(bounds + 3) - ray.org) * rcp_dir;
  const vec3f tnear = vmin(lo, hi);
  const vec3f tfar = vmax(lo, hi);
  int32_t cell[3];
  {
    const vec3f V01 = (ray.org - ld3(bounds)) / (ld3(bounds + 3) - ld3(bounds));
    cell[0] = clampi(sat_i32(V01.x * (float)dims[0]), 0, dims[0] - 1);
    cell[1] = clampi(sat_i32(V01.y * (float)dims[1]), 0, dims[1] - 1);
    cell[2] = clampi(sat_i32(V01.z * (float)dims[2]), 0, dims[2] - 1);
  }
  const vec3f dist = (tfar - tnear) / mk3((float)dims[0], (float)dims[1], (float)dims[2]);
  const int step[3] = {ray.dir.x > 0.f ? 1 : -1, ray.dir.y > 0.f ? 1 : -1, ray.dir.z > 0.f ? 1 : -1};
  const int stop[3] = {ray.dir.x > 0.f ? dims[0] : -1, ray.dir.y > 0.f ? dims[1] : -1,
                       ray.dir.z > 0.f ? dims[2] : -1};
  vec3f tnext = mk3(ray.dir.x > 0.f ? tnear.x + (float)(cell[0] + 1) * dist.x
                                    : tnear.x + (float)(dims[0] - cell[0]) * dist.x,
                    ray.dir.y > 0.f ? tnear.y + (float)(cell[1] + 1) * dist.y
                                    : tnear.y + (float)(dims[1] - cell[1]) * dist.y,
                    ray.dir.z > 0.f ? tnear.z + (float)(cell[2] + 1) * dist.z
                                    : tnear.z + (float)(dims[2] - cell[2]) * dist.z);
  float t0 = fmaxf(ray.tmin, 0.f);
  for (;;) {
    const float t1 = fminf(reduce_min(tnext), ray.tmax);
    const int64_t leafID = ((int64_t)cell[2] * dims[1] + cell[1]) * dims[0] + cell[0];
    if (kCount) C.cnt.c[kCntDdaCells]++;
    // woodcockFunc / ratioTrackingFunc (deviceCode.cu:726-839)
    bool goOn = true;
    {
      const float majorant = C.gv.majorants[leafID];
      float t = t0;
      int32_t primID = -1;
      for (;;) {
        if (majorant <= 0.f) break;
        t -= exa_logf(1.f - C.rng()) / (majorant * os);
        if (t >= t1) break;
        I.pos = ray.org + ray.dir * t;
        Sample s = testSampleGridlet(C.sv, I.pos, primID);
        if (s.primID >= 0) {
          if (kCount) C.cnt.c[kCntCachedHit]++;
        } else {
          int kind;
          s = locatePoint<kAccel>(C.sv, I.pos, C.stack, kRenderBlock, kind);
          if (kCount) C.cnt.c[kind < 0 ? kCntMiss : kCntGridletHit + kind]++;
        }
        if (s.primID < 0) continue;
        if (s.cellID != -1) primID = s.primID;
        I.xf = lookupTF(P, s.value);
        if (kCount) C.cnt.c[kCntTfFetch]++;
        if (!shadow) {
          const float u = C.rng();
          const float sigmaT = I.xf.w;
          if (sigmaT * os >= u * majorant) {
            I.Tr = 0.f;
            I.type = kScattering;
            goOn = false;
            break;
          }
        } else {
          const float sigmaT = I.xf.w;
          I.Tr *= 1.f - (sigmaT * os) / majorant;
          if (I.Tr <= 0.f) {
            I.type = kScattering;
            goOn = false;
            break;
          }
        }
      }
    }
    if (!goOn) break;
    const float t_closest = reduce_min(tnext);
    if (tnext.x == t_closest) {
      tnext.x += dist.x;
      cell[0] += step[0];
      if (cell[0] == stop[0]) break;
    }
    if (tnext.y == t_closest) {
      tnext.y += dist.y;
      cell[1] += step[1];
      if (cell[1] == stop[1]) break;
    }
    if (tnext.z == t_closest) {
      tnext.z += dist.z;
      cell[2] += step[2];
      if (cell[2] == stop[2]) break;
    }
    t0 = t1;
  }
  I.pos = xfmPoint(P.voxelXfmInv, I.pos);
  return I;
}

// deviceCode.cu:856-926 sampleLight (+ directional lights, no reference
// counterpart: unit falloff, shadow ray to the domain exit)
template <bool kCount, int kAccel>
EXA_DEV vec3f sampleLight(Ctx<kCount, kAccel> &C, vec3f pos, int numLights) {
  const RenderParams &P = C.P;
  int which = (int)(C.rng() * (float)numLights);
  if (which == numLights) which = 0;
  const LightParams &L = P.lights[which];
  Ray ray;
  float ld = 1.f;
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
  const Interaction I = sampleInteraction(C, ray, true);
  if (L.type == 1) return mk3s(I.Tr * L.intensity);
  return mk3s(I.Tr / (ld * ld) * L.intensity);
}

// deviceCode.cu:144-149, 965-985 (g = 0)
template <bool kCount, int kAccel>
EXA_DEV vec3f hgSample(Ctx<kCount, kAccel> &C, vec3f wo) {
  const float u1 = C.rng();
  const float u2 = C.rng();
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

template <bool kCount, int kAccel>
EXA_DEV float4 pathTrace(Ctx<kCount, kAccel> &C, Ray ray, float4 bg, int numLights) {
  const RenderParams &P = C.P;
  vec3f throughput = mk3s(1.f);
  vec3f Ld = mk3s(0.f);
  unsigned bounce = 0;
  float t0 = 1e30f, t1 = -1e30f;
  if (boxTest(ray, P.worldBounds, t0, t1)) {
    clipInterval(P, ray, t0, t1);
    ray.org = ray.org + ray.dir * t0;
    t1 -= t0;
    for (;;) {
      ray.tmax = t1;
      const Interaction I = sampleInteraction(C, ray, false);
      if (I.type == kBoundary) break;
      if (bounce++ >= (unsigned)P.maxBounces) {
        throughput = mk3s(0.f);
        break;
      }
      const vec3f albedo = mk3(I.xf.x, I.xf.y, I.xf.z);
      throughput = throughput * albedo;
      const float Pr = reduce_max(throughput);
      if (Pr < 0.2f) {
        if (C.rng() > Pr) {
          throughput = mk3s(0.f);
          break;
        }
        throughput = mk3(throughput.x / Pr, throughput.y / Pr, throughput.z / Pr);
      }
      throughput = throughput + mk3s(0.f);
      if (numLights > 0) Ld = Ld + throughput * sampleLight(C, I.pos, numLights);
      const vec3f scatterDir = hgSample(C, -ray.dir);
      ray.org = I.pos;
      ray.dir = scatterDir;
      boxTest(ray, P.worldBounds, t0, t1);
      clipInterval(P, ray, t0, t1);
    }
  }
  if (numLights == 0) Ld = mk3s(1.f);
  const vec3f L = Ld * throughput;
  return bounce ? make_float4(L.x, L.y, L.z, 1.f) : bg;
}

template <bool kCount, int kAccel>
EXA_DEV float4 directLight(Ctx<kCount, kAccel> &C, Ray ray, float4 bg, int numLights) {
  const RenderParams &P = C.P;
  vec3f throughput = mk3s(1.f);
  float t0 = 1e30f, t1 = -1e30f;
  if (!boxTest(ray, P.worldBounds, t0, t1)) return bg;
  clipInterval(P, ray, t0, t1);
  ray.org = ray.org + ray.dir * t0;
  t1 -= t0;
  ray.tmax = t1;
  const Interaction I = sampleInteraction(C, ray, false);
  if (I.type == kBoundary) return bg;
  const vec3f albedo = mk3(I.xf.x, I.xf.y, I.xf.z);
  if (numLights > 0) throughput = throughput * (albedo * sampleLight(C, I.pos, numLights));
  else throughput = throughput * albedo;
  return make_float4(throughput.x, throughput.y, throughput.z, 1.f);
}

// owl make_8bit: clamp(int(f*256), 0, 255)
EXA_DEV uint32_t make8bit(float f) {
  const int32_t i = sat_i32(f * 256.f);
  return (uint32_t)(i < 0 ? 0 : (i > 255 ? 255 : i));
}

template <bool kCount, int kAccel>
__global__ __launch_bounds__(kRenderBlock) void kRender(RenderParams P, SamplerView sv, GridView gv,
                                                        unsigned long long *counters) {
  __shared__ int32_t stack[kAccel == kAccelBvh ? kStackDepth * kRenderBlock : 1];
  const int tid = threadIdx.y * kTileX + threadIdx.x;
  const int px = blockIdx.x * kTileX + threadIdx.x;
  const int py = blockIdx.y * kTileY + threadIdx.y;
  if (px >= P.W || py >= P.H) return;
  LocalCounters cnt;
  if (kCount)
    for (int i = 0; i < kNumCounters; ++i) cnt.c[i] = 0;
  Lcg rng;
  const int pixelID = px + P.W * py;
  rng.init((uint32_t)pixelID, (uint32_t)P.accumID);
  Ctx<kCount, kAccel> C{P, sv, gv, rng, stack + tid, cnt};
  const float t = (float)py / (float)P.H;
  const float4 bg = make_float4((1.f - t) * 1.f + t * 0.5f, (1.f - t) * 1.f + t * 0.7f, (1.f - t) * 1.f + t * 1.f,
                                1.f);
  int numLights = 0;
  for (int i = 0; i < kLightsMax; ++i)
    if (P.lights[i].on) numLights++;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int s = 0; s < P.spp; ++s) {
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
    // deviceCode.cu:88-101 generateRay
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
    const float4 c = P.integrator == 1 ? directLight(C, ray, bg, numLights) : pathTrace(C, ray, bg, numLights);
    acc.x += c.x;
    acc.y += c.y;
    acc.z += c.z;
    acc.w += c.w;
  }
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
  if (kCount) {
    cnt.c[kCntPixels] = 1;
    for (int i = 0; i < kNumCounters; ++i)
      if (cnt.c[i]) atomicAdd(&counters[i], (unsigned long long)cnt.c[i]);
  }
}

void launchRender(const RenderParams &p, const SamplerView &sv, int accel, const GridView &gv,
                  unsigned long long *counters, hipStream_t s) {
  const dim3 grid((p.W + kTileX - 1) / kTileX, (p.H + kTileY - 1) / kTileY);
  const dim3 block(kTileX, kTileY);
  if (accel == kAccelBvh) {
    if (counters) hipLaunchKernelGGL((kRender<true, kAccelBvh>), grid, block, 0, s, p, sv, gv, counters);
    else hipLaunchKernelGGL((kRender<false, kAccelBvh>), grid, block, 0, s, p, sv, gv, counters);
  } else {
    if (counters) hipLaunchKernelGGL((kRender<true, kAccelTables>), grid, block, 0, s, p, sv, gv, counters);
    else hipLaunchKernelGGL((kRender<false, kAccelTables>), grid, block, 0, s, p, sv, gv, counters);
  }
}

}  // namespace exa
