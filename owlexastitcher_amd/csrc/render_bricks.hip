// render_bricks.hip -- kernels of the ExaBricks comparison sampler for gfx950:
// point location (sample() of sampler/ExaBrickSampler.h:204-241) and the
// raygen megakernel of renderFrame_ExaBrickSampler (deviceCode.cu:1378-1388)
// in its default configuration (EXABRICK_ABR_TRAVERSAL with the ABR sampling
// BVH): majorants per region, the regions a ray crosses walked front to back
// on the kd-tree, the sample at a tentative collision blended over the
// current region's bricks.  One thread per pixel in 16x16 pixel tiles, or
// (path tracer and direct lighting) persistent lanes that take pixels from a
// work counter.
#include <algorithm>
#include "bricks_device.h"
#include "exa_integrate.h"

namespace exa {

__global__ void kBrickLocate(BrickView v, const float *__restrict__ pos, int64_t n, Sample *__restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const vec3f p = mk3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  Sample s{0, 0, 0.f};  // ExaBrickSampler.h:234: no weight -> {0, 0, 0}
  const int32_t abr = abrContaining(v, p);
  if (abr >= 0) {
    const int32_t begin = __float_as_int(v.abrLo[abr].w);
    const int32_t count = __float_as_int(v.abrHi[abr].w);
    float sumWeightedValues = 0.f, sumWeights = 0.f;
    for (int32_t k = 0; k < count; ++k) addBasisFunctions(v, v.leafList[begin + k], p, sumWeightedValues, sumWeights);
    if (sumWeights > 0.f) s = Sample{0, -1, sumWeightedValues / sumWeights};
  }
  out[i] = s;
}

template <bool kCount>
struct BrickCtx {
  const RenderParams &P;
  const BrickView &bv;
  Lcg &rng;
  LocalCounters &cnt;
  int px, py;
};

// deviceCode.cu:699-854 sampleInteraction over the regions (traverse,
// deviceCode.cu:527-547): each region [t0, t1] from the closest-hit query at
// the distance already integrated, Woodcock (shadow=false) or ratio tracking
// (shadow=true) with the region's majorant; the next query starts at
// t1 * 1.0000001.  ExaBricks has no testSample, every sample is a blend
// (cellID -1, so nothing is cached).
template <bool kCount>
EXA_DEV Interaction sampleInteraction(BrickCtx<kCount> &C, const Ray &ray, bool shadow) {
  const RenderParams &P = C.P;
  Interaction I;
  I.type = kBoundary;
  I.Tr = 1.f;
  I.pos = mk3s(0.f);
  I.xf = make_float4(0.f, 0.f, 0.f, 0.f);
  const float os = P.opacityScale;
  const vec3f org = xfmPoint(P.voxelXfm, ray.org);
  vec3f dir = xfmVector(P.voxelXfm, ray.dir);
  const float dt_scale = length(dir);
  dir = normalize(dir);
  const float tmax = ray.tmax * dt_scale;
  float already = ray.tmin * dt_scale;
  uint32_t leaves = 0, bricks = 0;
  AbrWalk walk;
  walkBegin(walk, already, tmax);
  for (int visit = 0; visit < kMaxAbrVisits; ++visit) {
    float t0, t1;
    const int32_t abr = walkNext(C.bv, walk, org, dir, already, t0, t1, leaves);
    if (abr < 0) break;
    if (kCount) C.cnt.c[kCntDdaCells]++;
    const float majorant = C.bv.abrMaxOp[abr];
    float t = t0;
    int32_t primID = -1;
    bool done = false;
    for (;;) {
      if (majorant <= 0.f) break;
      if (!trackingStep(t, majorant, os, t1, C.rng)) break;
      I.pos = org + dir * t;
      const Sample s{0, -1, abrValue(C.bv, abr, I.pos, bricks)};
      if (kCount) {
        C.cnt.c[kCntGridletHit]++;
        C.cnt.c[kCntTfFetch]++;
      }
      if (resolveCollision(P, nullptr, C.px, C.py, s, shadow, majorant, C.rng, primID, I.xf, I.Tr)) {
        I.type = kScattering;
        done = true;
        break;
      }
    }
    if (done) break;
    already = t1 * 1.0000001f;
  }
  if (kCount) {
    C.cnt.c[kCntPyrHit] += leaves;
    C.cnt.c[kCntTetHit] += bricks;
  }
  I.pos = xfmPoint(P.voxelXfmInv, I.pos);
  return I;
}

// Ray marcher with the ExaBricks integrateDVR (deviceCode.cu:1213-1260,
// 401-499): in each region crossed, samples at the midpoints of the steps
// between the lattice points dt_r * (ils + k), dt_r = dt * the region's finest
// cell width, the step cut at the region's ends; alpha corrected to the step
// length, 1 - (1 - a)^len; front-to-back compositing of the raw TF colour,
// stop at alpha 0.99 (colour then premultiplied once more, alpha 1), over the
// background.  The lattice search is bounded to kLatticeSteps moves each way
// and a ray to kMaxMarchSteps samples.
constexpr int kMaxBrickMarchSteps = 1 << 20;
constexpr int kLatticeSteps = 8;

template <bool kCount>
EXA_DEV float4 rayMarchBricks(BrickCtx<kCount> &C, Ray ray, float4 bg) {
  const RenderParams &P = C.P;
  float4 color = bg;
  float t0w = 1e30f, t1w = -1e30f;
  if (!boxTest(ray, P.worldBounds, t0w, t1w)) return color;
  clipInterval(P, ray, t0w, t1w);
  const float ils = C.rng();
  const vec3f org = xfmPoint(P.voxelXfm, ray.org);
  vec3f dir = xfmVector(P.voxelXfm, ray.dir);
  const float dt_scale = length(dir);
  dir = normalize(dir);
  const float tmax = t1w * dt_scale;
  float already = t0w * dt_scale;
  float4 pc = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t leaves = 0, bricks = 0;
  int steps = 0;
  AbrWalk walk;
  walkBegin(walk, already, tmax);
  for (int visit = 0; visit < kMaxAbrVisits && steps < kMaxBrickMarchSteps; ++visit) {
    float t0, t1;
    const int32_t abr = walkNext(C.bv, walk, org, dir, already, t0, t1, leaves);
    if (abr < 0) break;
    if (kCount) C.cnt.c[kCntDdaCells]++;
    const float dt = P.dt * C.bv.abrCellWidth[abr];
    const int32_t i0 = sat_i32(ceilf((t0 - dt * ils) / dt));
    float ti = (ils + (float)i0) * dt;
    for (int k = 0; k < kLatticeSteps && (ti - dt) >= t0; ++k) ti = ti - dt;
    for (int k = 0; k < kLatticeSteps && ti < t0; ++k) ti += dt;
    float tlast = t0;
    for (; steps < kMaxBrickMarchSteps; ti += dt, ++steps) {
      const float tnext = fminf(ti, t1);
      const float ts = 0.5f * (fminf(t1, tnext) + tlast);
      const float adt = tnext - tlast;
      tlast = tnext;
      const vec3f pos = org + ts * dir;
      const float value = abrValue(C.bv, abr, pos, bricks);
      if (kCount) {
        C.cnt.c[kCntGridletHit]++;
        C.cnt.c[kCntTfFetch]++;
      }
      float4 xf = lookupTF(P, value);
      xf.w = 1.f - exa_powf(1.f - xf.w, adt);
      const float k = (1.f - pc.w) * xf.w;
      pc = make_float4(pc.x + k * xf.x, pc.y + k * xf.y, pc.z + k * xf.z, pc.w + k * 1.f);
      if (pc.w >= 0.99f) break;
      if (tnext >= t1) break;
    }
    if (pc.w >= 0.99f) {
      pc = make_float4(pc.x * pc.w, pc.y * pc.w, pc.z * pc.w, 1.f);
      break;
    }
    already = t1 * 1.0000001f;
  }
  if (kCount) {
    C.cnt.c[kCntPyrHit] += leaves;
    C.cnt.c[kCntTetHit] += bricks;
  }
  return over(pc, color);
}

template <bool kCount>
__global__ __launch_bounds__(kRenderBlock) void kRenderBricks(RenderParams P, BrickView bv,
                                                              unsigned long long *counters) {
  const int px = blockIdx.x * kTileX + threadIdx.x;
  const int py = blockIdx.y * kTileY + threadIdx.y;
  if (px >= P.W || py >= P.H) return;
  LocalCounters cnt;
  if (kCount)
    for (int i = 0; i < kNumCounters; ++i) cnt.c[i] = 0;
  Lcg rng;
  rng.init((uint32_t)(px + P.W * py), (uint32_t)P.accumID);
  BrickCtx<kCount> C{P, bv, rng, cnt, px, py};
  const float4 bg = backgroundColor(py, P.H);
  const int numLights = countLights(P);
  const uint64_t clockBegin = P.heatMapEnabled ? clock64() : 0;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int s = 0; s < P.spp; ++s) {
    const Ray ray = primaryRay(P, px, py, rng);
    const float4 c = P.integrator == 1   ? directLight(C, ray, bg, numLights)
                     : P.integrator == 2 ? rayMarchBricks(C, ray, bg)
                                         : pathTrace(C, ray, bg, numLights);
    acc.x += c.x;
    acc.y += c.y;
    acc.z += c.z;
    acc.w += c.w;
  }
  if (P.heatMapEnabled) {
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

// Persistent variant for the path tracer and direct lighting: resident
// waves (4 per SIMD) whose lanes take pixels (8x8 tiles in scan order) from
// a frame-wide counter, one atomic per wave, and run their paths as a state
// machine whose hot step is one pass of sampleInteraction's region loop (the
// next region of the main ray's or the light sample's query and the tracking
// steps inside it), the region walk cut into pieces of kBrickWalkBudget kd
// nodes.  A lane whose query has ended moves on to the next
// segment, sample or pixel while the others still walk regions, so a wave
// waits neither for its longest path nor for its longest region walk.  Each
// lane draws its random numbers and does its arithmetic in the order of
// sampleInteraction and pathTrace / directLight above, so the frames equal
// the one-thread-per-pixel kernel's bit for bit.
constexpr int kBrickPersistBlock = 256;
// kd nodes a lane walks per step before it yields to the lanes in other states
constexpr int kBrickWalkBudget = 8;
constexpr int kBrickPixelTile = 8;
enum : int { kBrGrab = 0, kBrStart = 1, kBrMain = 2, kBrShadow = 3, kBrDone = 4 };

__global__ void kResetBrickWork(uint32_t *work) {
  if (threadIdx.x == 0) *work = 0u;
}

template <bool kCount>
__global__ __launch_bounds__(kBrickPersistBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void kRenderBricksPersistent(RenderParams P, BrickView bv,
                                                                              uint32_t *work, uint32_t numWork,
                                                                              unsigned long long *counters) {
  const int lane = threadIdx.x & 63;
  const int tilesX = (P.W + kBrickPixelTile - 1) / kBrickPixelTile;
  const int numLights = countLights(P);
  const bool direct = P.integrator == 1;
  LocalCounters cnt;
  if (kCount)
    for (int i = 0; i < kNumCounters; ++i) cnt.c[i] = 0;
  Lcg rng;
  BrickCtx<kCount> C{P, bv, rng, cnt, 0, 0};
  int st = kBrGrab, s = 0;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  // path state of pathTrace / directLight; the light sample's ray starts at
  // the scattering point (qray.org), and directLight keeps its albedo in
  // throughput (still 1 there, so 1 * (albedo * L) and albedo * L agree)
  Ray ray, qray;
  vec3f throughput = mk3s(1.f), Ld = mk3s(0.f);
  unsigned bounce = 0;
  float t0 = 1e30f, t1 = -1e30f, ld = 0.f;
  int which = 0;
  // the interaction query in progress (sampleInteraction's locals)
  Interaction I;
  I.type = kBoundary;
  vec3f qo = mk3s(0.f), qd = mk3s(0.f);
  float already = 0.f;
  AbrWalk walk;
  walkBegin(walk, 0.f, 0.f);
  int visits = 0, guard = 0;
  bool inQuery = false;
  uint32_t unused = 0;
  auto endSample = [&](float4 c) {
    acc.x += c.x;
    acc.y += c.y;
    acc.z += c.z;
    acc.w += c.w;
    if (++s < P.spp) {
      st = kBrStart;
      return;
    }
    finishPixel(P, C.px, C.py, acc);
    if (kCount) cnt.c[kCntPixels]++;
    st = kBrGrab;
  };
  auto scatter = [&](vec3f pos) {
    const vec3f scatterDir = hgSample(-ray.dir, rng);
    ray.org = pos;
    ray.dir = scatterDir;
    boxTest(ray, P.worldBounds, t0, t1);
    clipInterval(P, ray, t0, t1);
    ray.tmax = t1;
    st = kBrMain;
  };
  for (;;) {
    const uint64_t want = __ballot(st == kBrGrab);
    if (want) {
      const int leader = __ffsll((unsigned long long)want) - 1;
      uint32_t base = 0;
      if (lane == leader) base = atomicAdd(work, (uint32_t)__popcll(want));
      base = __shfl(base, leader);
      if (st == kBrGrab) {
        const uint32_t w = base + (uint32_t)__popcll(want & ((1ull << lane) - 1ull));
        if (w >= numWork) {
          st = kBrDone;
        } else {
          const uint32_t tile = w >> 6, in = w & 63u;
          const int px = (int)(tile % (uint32_t)tilesX) * kBrickPixelTile + (int)(in & 7u);
          const int py = (int)(tile / (uint32_t)tilesX) * kBrickPixelTile + (int)(in >> 3);
          if (px < P.W && py < P.H) {  // else outside a partial edge tile: grab again
            C.px = px;
            C.py = py;
            rng.init((uint32_t)(px + P.W * py), (uint32_t)P.accumID);
            acc = make_float4(0.f, 0.f, 0.f, 0.f);
            s = 0;
            st = kBrStart;
          }
        }
      }
    }
    if (__ballot(st != kBrDone) == 0) break;
    if (st == kBrStart) {
      ray = primaryRay(P, C.px, C.py, rng);
      throughput = mk3s(1.f);
      Ld = mk3s(0.f);
      bounce = 0;
      t0 = 1e30f;
      t1 = -1e30f;
      if (boxTest(ray, P.worldBounds, t0, t1)) {
        clipInterval(P, ray, t0, t1);
        ray.org = ray.org + ray.dir * t0;
        t1 -= t0;
        ray.tmax = t1;
        st = kBrMain;
      } else {
        const float4 bg = backgroundColor(C.py, P.H);
        endSample(direct ? bg : ptResult(Ld, throughput, bounce, numLights, bg));
      }
    }
    if (st == kBrMain || st == kBrShadow) {
      const bool shadow = st == kBrShadow;
      if (!inQuery) {
        // sampleInteraction's set-up: the query ray in voxel space
        const Ray &q = shadow ? qray : ray;
        I.type = kBoundary;
        I.Tr = 1.f;
        I.pos = mk3s(0.f);
        I.xf = make_float4(0.f, 0.f, 0.f, 0.f);
        qo = xfmPoint(P.voxelXfm, q.org);
        vec3f dir = xfmVector(P.voxelXfm, q.dir);
        const float dt_scale = length(dir);
        qd = normalize(dir);
        already = q.tmin * dt_scale;
        walkBegin(walk, already, q.tmax * dt_scale);
        visits = 0;
        guard = 0;
        inQuery = true;
      }
      // one pass of sampleInteraction's region loop: the next region and the
      // tracking steps inside it
      bool finished = visits >= kMaxAbrVisits;
      if (!finished) {
        float r0, r1;
        const int32_t abr = walkNextBounded(bv, walk, qo, qd, already, r0, r1, kCount ? cnt.c[kCntPyrHit] : unused,
                                            kBrickWalkBudget, guard);
        if (abr != kWalkPending) {
          ++visits;
          guard = 0;
        }
        if (abr == kWalkPending) {
          // the walk continues in the next step
        } else if (abr < 0) {
          finished = true;
        } else {
          if (kCount) cnt.c[kCntDdaCells]++;
          const float majorant = bv.abrMaxOp[abr];
          float t = r0;
          int32_t primID = -1;
          for (;;) {
            if (majorant <= 0.f) break;
            if (!trackingStep(t, majorant, P.opacityScale, r1, rng)) break;
            I.pos = qo + qd * t;
            const Sample smp{0, -1, abrValue(bv, abr, I.pos, kCount ? cnt.c[kCntTetHit] : unused)};
            if (kCount) {
              cnt.c[kCntGridletHit]++;
              cnt.c[kCntTfFetch]++;
            }
            if (resolveCollision(P, nullptr, C.px, C.py, smp, shadow, majorant, rng, primID, I.xf, I.Tr)) {
              I.type = kScattering;
              finished = true;
              break;
            }
          }
          if (!finished) already = r1 * 1.0000001f;
        }
      }
      if (!finished) continue;
      inQuery = false;
      I.pos = xfmPoint(P.voxelXfmInv, I.pos);
      if (direct) {
        if (shadow) {
          throughput = throughput * lightResult(P, which, I.Tr, ld);
          endSample(make_float4(throughput.x, throughput.y, throughput.z, 1.f));
        } else if (I.type == kBoundary) {
          endSample(backgroundColor(C.py, P.H));
        } else {
          throughput = mk3(I.xf.x, I.xf.y, I.xf.z);
          if (numLights > 0) {
            qray = lightRay(P, I.pos, numLights, rng, which, ld);
            st = kBrShadow;
          } else {
            endSample(make_float4(throughput.x, throughput.y, throughput.z, 1.f));
          }
        }
      } else if (shadow) {
        Ld = Ld + throughput * lightResult(P, which, I.Tr, ld);
        scatter(qray.org);
      } else if (I.type == kBoundary || !ptInteraction(P, I.xf, throughput, bounce, rng)) {
        endSample(ptResult(Ld, throughput, bounce, numLights, backgroundColor(C.py, P.H)));
      } else if (numLights > 0) {
        qray = lightRay(P, I.pos, numLights, rng, which, ld);
        st = kBrShadow;
      } else {
        scatter(I.pos);
      }
    }
  }
  if (kCount)
    for (int i = 0; i < kNumCounters; ++i)
      if (cnt.c[i]) atomicAdd(&counters[i], (unsigned long long)cnt.c[i]);
}

template <bool kCount>
static int residentBrickBlocks() {
  static std::atomic<int> cache[kMaxCachedDevices];
  return residentWorkgroups(kRenderBricksPersistent<kCount>, kBrickPersistBlock, cache);
}

void launchBrickLocate(const BrickView &v, const float *pos, int64_t n, Sample *out, hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(kBrickLocate, dim3((unsigned)blocks), dim3(256), 0, s, v, pos, n, out);
}

void launchRenderBricks(const RenderParams &p, const BrickView &v, uint32_t *work, unsigned long long *counters,
                        hipStream_t s) {
  RenderParams q = p;
  q.shadeMode = 0;  // renderFrame_ExaBrickSampler shades with Default only
  if (work && q.integrator != 2 && !q.heatMapEnabled) {
    const uint32_t tilesX = (uint32_t)((q.W + kBrickPixelTile - 1) / kBrickPixelTile);
    const uint32_t tilesY = (uint32_t)((q.H + kBrickPixelTile - 1) / kBrickPixelTile);
    const uint32_t numWork = tilesX * tilesY * 64u;
    const int want = (int)((numWork + kBrickPersistBlock - 1) / kBrickPersistBlock);
    hipLaunchKernelGGL(kResetBrickWork, dim3(1), dim3(64), 0, s, work);
    if (counters) {
      const int blocks = std::max(1, std::min(residentBrickBlocks<true>(), want));
      hipLaunchKernelGGL((kRenderBricksPersistent<true>), dim3(blocks), dim3(kBrickPersistBlock), 0, s, q, v, work,
                         numWork, counters);
    } else {
      const int blocks = std::max(1, std::min(residentBrickBlocks<false>(), want));
      hipLaunchKernelGGL((kRenderBricksPersistent<false>), dim3(blocks), dim3(kBrickPersistBlock), 0, s, q, v, work,
                         numWork, counters);
    }
    return;
  }
  const dim3 grid((q.W + kTileX - 1) / kTileX, (q.H + kTileY - 1) / kTileY);
  const dim3 block(kTileX, kTileY);
  if (counters) hipLaunchKernelGGL((kRenderBricks<true>), grid, block, 0, s, q, v, counters);
  else hipLaunchKernelGGL((kRenderBricks<false>), grid, block, 0, s, q, v, counters);
}

}  // namespace exa
