// render.hip -- the raygen megakernel of the exaStitch path tracer for
// gfx950: one thread per pixel, 16x16 pixel tiles per workgroup.
//
// Follows deviceCode.cu: renderFrame_Impl (1268-1345), generateRay (88-101),
// pathTracingIntegrate (991-1120), directLightingIntegrate (1126-1207),
// sampleInteraction with Woodcock / ratio tracking (699-854), sampleLight
// (856-926), henyeyGreensteinSample (965-985), lookupTransferFunction
// (320-341), and dda3 (DDA.h:26-113).  The persistent renderer
// (render_persistent.hip) runs the same helpers as a resumable state machine.
#include "exa_integrate.h"

namespace exa {

template <bool kCount, int kAccel>
struct Ctx {
  const RenderParams &P;
  const SamplerView &sv;
  const GridView &gv;
  Lcg &rng;
  int32_t *stack;
  LocalCounters &cnt;
  int px, py;
};

// deviceCode.cu:699-854: DDA over the majorant grid, Woodcock (shadow=false)
// or ratio tracking (shadow=true) inside each macrocell.
template <bool kCount, int kAccel>
EXA_DEV Interaction sampleInteraction(Ctx<kCount, kAccel> &C, const Ray &ray, bool shadow) {
  const RenderParams &P = C.P;
  Interaction I;
  I.type = kBoundary;
  I.Tr = 1.f;
  I.pos = mk3s(0.f);
  I.xf = make_float4(0.f, 0.f, 0.f, 0.f);
  const float os = P.opacityScale;
  Segment S;
  segmentBegin(P, C.gv, ray, S);
  for (;;) {
    const float t1 = segmentCellExit(S);
    if (kCount) C.cnt.c[kCntDdaCells]++;
    bool goOn = true;
    {
      const float majorant = C.gv.majorants[segmentCellID(C.gv, S)];
      float t = S.t0;
      int32_t primID = -1;
      for (;;) {
        if (majorant <= 0.f) break;
        if (!trackingStep(t, majorant, os, t1, C.rng)) break;
        I.pos = S.org + S.dir * t;
        Sample s = testSampleGridlet(C.sv, I.pos, primID);
        if (s.primID >= 0) {
          if (kCount) C.cnt.c[kCntCachedHit]++;
        } else {
          int kind;
          s = locatePoint<kAccel>(C.sv, I.pos, C.stack, kRenderBlock, kind);
          if (kCount) C.cnt.c[kind < 0 ? kCntMiss : kCntGridletHit + kind]++;
        }
        if (kCount && s.primID >= 0) C.cnt.c[kCntTfFetch]++;
        if (resolveCollision(P, C.sv.gridlets, C.px, C.py, s, shadow, majorant, C.rng, primID, I.xf, I.Tr)) {
          I.type = kScattering;
          goOn = false;
          break;
        }
      }
    }
    if (!goOn) break;
    if (!segmentNextCell(C.gv, S, t1)) break;
  }
  I.pos = xfmPoint(P.voxelXfmInv, I.pos);
  return I;
}

// Ray marcher (deviceCode.cu:1213-1260 rayMarchingIntegrate, 348-398
// integrateDVR with Shading): absorption + emission at samples dt apart on a
// lattice jittered by one random number, front-to-back compositing of the
// raw TF colour and alpha (no opacity scale), stopped at alpha 0.99, over the
// background.  The reference reads an uninitialised gradient for lit
// samples; here the gradient is zero, so a lit sample keeps the 0.1 ambient
// term.  At most kMaxMarchSteps samples per ray.
constexpr int kMaxMarchSteps = 1 << 20;

template <bool kCount, int kAccel>
EXA_DEV float4 rayMarch(Ctx<kCount, kAccel> &C, Ray ray, float4 bg, int numLights) {
  const RenderParams &P = C.P;
  float4 color = bg;
  float t0 = 1e30f, t1 = -1e30f;
  if (!boxTest(ray, P.worldBounds, t0, t1)) return color;
  clipInterval(P, ray, t0, t1);
  const float ils = C.rng();
  // integrateDVR: the ray goes to voxel space; t0 and t1 stay as they are
  const vec3f org = xfmPoint(P.voxelXfm, ray.org);
  const vec3f dir = normalize(xfmVector(P.voxelXfm, ray.dir));
  const float dt = P.dt;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  int steps = 0;
  for (float t = firstSampleT(t0, dt, ils); t <= t1 && steps < kMaxMarchSteps; t += dt, ++steps) {
    const vec3f pos = org + dir * t;
    int kind;
    const Sample s = locatePoint<kAccel>(C.sv, pos, C.stack, kRenderBlock, kind);
    if (kCount) C.cnt.c[kind < 0 ? kCntMiss : kCntGridletHit + kind]++;
    if (s.primID < 0) continue;
    if (kCount) C.cnt.c[kCntTfFetch]++;
    const float4 xf = lookupTF(P, s.value);
    const vec3f cd = mk3(xf.x, xf.y, xf.z);
    vec3f shaded = cd * .1f;
    if (numLights == 0) shaded = cd;
    for (int i = 0; i < numLights; ++i) {
      const vec3f lp = ld3(P.lights[i].pos);
      const vec3f L = normalize(lp - pos);
      const float ld = length(lp - pos);
      const float NdotL = 0.f * L.x + 0.f * L.y + 0.f * L.z;
      shaded = shaded + cd * (.9f * NdotL / (ld * ld) * P.lights[i].intensity);
    }
    const float k = (1.f - acc.w) * xf.w;
    acc = make_float4(acc.x + k * shaded.x, acc.y + k * shaded.y, acc.z + k * shaded.z, acc.w + k * 1.f);
    if (acc.w >= 0.99f) break;
  }
  return over(acc, color);
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
  Ctx<kCount, kAccel> C{P, sv, gv, rng, stack + tid, cnt, px, py};
  const float4 bg = backgroundColor(py, P.H);
  const int numLights = countLights(P);
  const uint64_t clockBegin = P.heatMapEnabled ? clock64() : 0;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int s = 0; s < P.spp; ++s) {
    const Ray ray = primaryRay(P, px, py, rng);
    const float4 c = P.integrator == 1   ? directLight(C, ray, bg, numLights)
                     : P.integrator == 2 ? rayMarch(C, ray, bg, numLights)
                                         : pathTrace(C, ray, bg, numLights);
    acc.x += c.x;
    acc.y += c.y;
    acc.z += c.z;
    acc.w += c.w;
  }
  if (P.heatMapEnabled) {
    // deviceCode.cu:1323-1327: cycles spent on this pixel, over the samples
    const float t = (float)(clock64() - clockBegin) * (P.heatMapScale / (float)P.spp);
    const vec3f h = heatMapColor(t);
    acc = over(make_float4(h.x, h.y, h.z, .5f), acc);
  }
  finishPixel(P, px, py, acc);
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
