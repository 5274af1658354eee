// exa_integrate.h -- the integrators above sampleInteraction, shared by the
// exaStitch megakernel (render.hip) and the ExaBricks kernel (render_bricks.hip).
// A context type provides P (RenderParams), rng (Lcg) and an overload
// sampleInteraction(Ctx &, const Ray &, bool shadow) found by argument-
// dependent lookup; deviceCode.cu citations per function.
#pragma once
#include "exa_render_common.h"

namespace exa {

struct Interaction {
  int type;
  vec3f pos;
  float Tr;
  float4 xf;
};

template <class Ctx_t>
EXA_DEV vec3f sampleLight(Ctx_t &C, vec3f pos, int numLights) {
  int which;
  float ld;
  const Ray ray = lightRay(C.P, pos, numLights, C.rng, which, ld);
  const Interaction I = sampleInteraction(C, ray, true);
  return lightResult(C.P, which, I.Tr, ld);
}

template <class Ctx_t>
EXA_DEV float4 pathTrace(Ctx_t &C, Ray ray, float4 bg, int numLights) {
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
      if (!ptInteraction(P, I.xf, throughput, bounce, C.rng)) break;
      if (numLights > 0) Ld = Ld + throughput * sampleLight(C, I.pos, numLights);
      const vec3f scatterDir = hgSample(-ray.dir, C.rng);
      ray.org = I.pos;
      ray.dir = scatterDir;
      boxTest(ray, P.worldBounds, t0, t1);
      clipInterval(P, ray, t0, t1);
    }
  }
  return ptResult(Ld, throughput, bounce, numLights, bg);
}

template <class Ctx_t>
EXA_DEV float4 directLight(Ctx_t &C, Ray ray, float4 bg, int numLights) {
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

}  // namespace exa
