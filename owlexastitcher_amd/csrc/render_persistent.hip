// render_persistent.hip -- persistent path-regeneration renderer for gfx950.
//
// The megakernel (render.hip) gives every thread one pixel, so a wave runs
// as long as its longest path and executes every branch of the integrator
// that any of its 64 lanes takes.  Here each lane runs its path as a
// resumable state machine, and a wave runs the parts of the machine in
// phases, so that lanes doing the same kind of work run it together:
//  - rare pass: the per-segment work (end of a segment and the integrator
//    step after it, phase-function sampling, camera rays, segment set-up);
//    runs when at least rareBatch/64 of the live lanes wait for it;
//  - tracking pass: free-flight steps and DDA macrocell steps, up to
//    ctrlSteps per lane, until each lane reaches a tentative collision;
//  - gridlet pass: cached-gridlet test, then the gridlet stage of point
//    location;
//  - element pass: the element stage (Newton iterations over pyramids,
//    wedges, hexes) for lanes whose gridlet stage missed; runs when at least
//    elemBatch/64 of the live lanes wait for it;
//  - resolve pass: TF lookup and the Woodcock / ratio-tracking decision.
// When no lane can make progress in the tracking, gridlet or resolve passes,
// both gated passes run regardless of their thresholds, so every wave
// drains.  Waves never wait for each other (pooling waiting lanes over the
// workgroup through LDS was tried: the per-round barriers cost more than the
// fuller element batches gained).  A lane that finishes its pixel takes the next one from a
// frame-wide work counter (one atomic per wave and pass); pixels are handed
// out in 8x8 tiles in scan order.
// Each lane draws random numbers and does its floating-point work in exactly
// the order of the megakernel (both use the helpers of exa_render_common.h),
// so every pixel's result is bit-identical to it; only the interleaving of
// lanes changes.
#include <cstddef>
#include "exa_render_common.h"
#include "neural.h"

namespace exa {

// largest workgroup (the LDS arrays are sized for it); a launch may use 64,
// 128 or 256 threads per workgroup (PersistTuning::blockSize): the waves of
// a workgroup share nothing but the request table and the network weights
constexpr int kPersistBlock = 256;
// minimum waves per SIMD the register allocation must allow (spills beyond)
#ifndef EXA_PERSIST_MIN_WAVES
#define EXA_PERSIST_MIN_WAVES 2
#endif
constexpr int kPixelTile = 8;  // 8x8 pixels per 64-lane grab

enum Stage : int32_t {
  // tracking pass
  kStStep = 0,  // free-flight step inside the macrocell
  kStCellBegin, // enter a macrocell
  // point location and resolution
  kStAtSample,  // collision position ready for point location
  kStPending,   // gridlet stage missed: waits for the element pass
  kStResolve,   // resumed from the element queue: its sample waits in xf
  kStWaitElem,  // element-stage query posted to the workgroup's queue
  // rare pass
  kStSegEnd,       // segment done: continue the integrator
  kStScatter,      // phase-function sample, next path ray
  kStGrab,         // take the next pixel
  kStSampleStart,  // camera ray of the next sample
  kStPtSegment,    // path tracer: begin the segment along the path ray
  kStDone
};
enum SegKind : int32_t { kSegPt = 0, kSegPtShadow, kSegDl, kSegDlShadow };

EXA_DEV bool isTracking(int st) { return st <= kStCellBegin; }
EXA_DEV bool isRare(int st) { return st >= kStSegEnd && st < kStDone; }

struct PathLane {
  int32_t st, kind, px, py, s;
  int32_t req;  // request of a batched launch
  Lcg rng;
  float4 acc;
  // integrator state (pathTracingIntegrate / directLightingIntegrate); the
  // direct-lighting integrator keeps the albedo of its interaction in Ld
  vec3f throughput, Ld;
  unsigned bounce;
  Ray ray;     // path ray; ray.tmax is pathTracingIntegrate's t1
  vec3f spos;  // world-space position of the last interaction
  int32_t which;
  float ld;
  // current segment (sampleInteraction)
  Segment S;
  float t1, majorant, t, Tr;
  int32_t primID, itype;
  vec3f pos;  // voxel-space collision position
  float4 xf;
};

static_assert(sizeof(PathLane) <= kElementQueueVec * sizeof(float4), "a queued path fits its queue entry");
constexpr int kParkVec = (int)((sizeof(PathLane) + sizeof(float4) - 1) / sizeof(float4));

// Element-stage queue.  A lane whose gridlet stage missed writes its whole
// path state into a queue entry and takes other work; a separate launch runs
// the element stage over the entries with every lane busy, and the next
// launch resumes the entries with their samples.  Entry i holds the path's
// float4 k at rec[k * cap + i], so a wave's stores and loads of consecutive
// entries are contiguous.
struct ParkQueue {
  float4 *rec;
  Sample *samples;
  uint32_t *keys;          // per entry: the path's work id (pixel tile order)
  const uint32_t *order;   // entry i of the launch order -> stored entry (image order)
  uint32_t *count;  // entries claimed; those at or beyond cap were not stored
  uint32_t cap;
};

EXA_DEV void parkStore(const PathLane &L, const ParkQueue &q, uint32_t slot) {
  float4 v[kParkVec];
  __builtin_memcpy(v, &L, sizeof(PathLane));
#pragma unroll
  for (int k = 0; k < kParkVec; ++k) q.rec[(size_t)k * q.cap + slot] = v[k];
}

EXA_DEV void parkLoad(PathLane &L, const ParkQueue &q, uint32_t slot) {
  float4 v[kParkVec];
#pragma unroll
  for (int k = 0; k < kParkVec; ++k) v[k] = q.rec[(size_t)k * q.cap + slot];
  __builtin_memcpy(&L, v, sizeof(PathLane));
}

// one 32-bit word of a parked path
EXA_DEV float parkWord(const ParkQueue &q, int word, uint32_t slot) {
  const float4 v = q.rec[(size_t)(word >> 2) * q.cap + slot];
  const int c = word & 3;
  return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w));
}

struct PersistParams {
  NeuralView nv;              // neural field (kNeural launches)
  const uint8_t *netFlags;    // per macrocell: 1 = element stage answered by the network
  uint32_t *work;     // frame-wide pixel-tile counter, zero at launch
  uint32_t numWork;   // 64 per 8x8 pixel tile (of all requests)
  const BatchRequest *reqs;  // batched launch: the requests
  int32_t numReqs;
  int32_t tilesX;
  int32_t ctrlSteps;  // tracking transitions per lane and round
  int32_t elemBatch;  // element pass runs when pending*64 >= live*elemBatch
  int32_t rareBatch;  // rare pass runs when rare*64 >= live*rareBatch
  ParkQueue parkOut;    // kPark: element-stage entries written here
  ParkQueue resumeIn;   // kResume: the work items (instead of pixels)
  uint32_t *resumeNext; // kResume: hand-out counter over resumeIn
  int32_t blockSize;    // threads per workgroup of the launch: 64, 128 or 256
  int32_t sharedQueue;  // 1: element stage through the workgroup's LDS queue
};

EXA_DEV void beginSegment(PathLane &L, const RenderParams &P, const GridView &gv, const Ray &ray, int kind) {
  L.kind = kind;
  segmentBegin(P, gv, ray, L.S);
  L.Tr = 1.f;
  L.itype = kBoundary;
  L.pos = mk3s(0.f);
  L.xf = make_float4(0.f, 0.f, 0.f, 0.f);
  L.st = kStCellBegin;
}

// renderFrame_Impl's sample loop (deviceCode.cu:1283-1316): accumulate the
// sample; after spp samples, write the pixel and ask for the next one.
template <bool kCount>
EXA_DEV void endSample(PathLane &L, const RenderParams &P, float4 c, LocalCounters &cnt) {
  L.acc.x += c.x;
  L.acc.y += c.y;
  L.acc.z += c.z;
  L.acc.w += c.w;
  L.s++;
  if (L.s < P.spp) {
    L.st = kStSampleStart;
    return;
  }
  finishPixel(P, L.px, L.py, L.acc);
  if (kCount) cnt.c[kCntPixels]++;
  L.st = kStGrab;
}

// tracking pass: Woodcock / ratio-tracking free flights and dda3 steps.
// One iteration takes one free flight in the current macrocell and, when it
// leaves the cell, the DDA step into the next one and that cell's set-up, so
// every tracking lane runs the same code each iteration (the per-lane order
// of random draws and arithmetic is that of sampleInteraction).
template <bool kCount>
EXA_DEV void cellBegin(PathLane &L, const GridView &gv, LocalCounters &cnt) {
  L.t1 = segmentCellExit(L.S);
  if (kCount) cnt.c[kCntDdaCells]++;
  L.majorant = gv.majorants[segmentCellID(gv, L.S)];
  L.t = L.S.t0;
  L.primID = -1;
}

template <bool kCount>
EXA_DEV void trackingRun(PathLane &L, const RenderParams &P, const GridView &gv, int steps, LocalCounters &cnt) {
  if (L.st == kStCellBegin) {
    cellBegin<kCount>(L, gv, cnt);
    L.st = kStStep;
  }
  for (int n = 0; n < steps && L.st == kStStep; ++n) {
    if (L.majorant > 0.f && trackingStep(L.t, L.majorant, P.opacityScale, L.t1, L.rng)) {
      L.pos = L.S.org + L.S.dir * L.t;
      L.st = kStAtSample;
    } else if (segmentNextCell(gv, L.S, L.t1)) {
      cellBegin<kCount>(L, gv, cnt);
    } else {
      L.st = kStSegEnd;
    }
  }
}

// end of a segment: the rest of the integrator step that traced it.  The
// branches only decide; the pixel sample is finished, the light ray drawn and
// (by the caller) the next segment begun at one place each, so a wave runs
// that code once whatever mix of cases its lanes are in.
template <bool kCount>
EXA_DEV void segmentEnd(PathLane &L, const RenderParams &P, int numLights, LocalCounters &cnt, Ray &segRay,
                        int &segKind) {
  const vec3f wpos = xfmPoint(P.voxelXfmInv, L.pos);
  bool finish = false, light = false;
  float4 res = make_float4(0.f, 0.f, 0.f, 0.f);
  if (L.kind == kSegPt) {
    if (L.itype == kBoundary || !ptInteraction(P, L.xf, L.throughput, L.bounce, L.rng)) {
      res = ptResult(L.Ld, L.throughput, L.bounce, numLights, backgroundColor(L.py, P.H));
      finish = true;
    } else {
      L.spos = wpos;
      if (numLights > 0) {
        light = true;
        segKind = kSegPtShadow;
      } else {
        L.st = kStScatter;
      }
    }
  } else if (L.kind == kSegPtShadow) {
    L.Ld = L.Ld + L.throughput * lightResult(P, L.which, L.Tr, L.ld);
    L.st = kStScatter;
  } else if (L.kind == kSegDl) {
    if (L.itype == kBoundary) {
      res = backgroundColor(L.py, P.H);
      finish = true;
    } else {
      const vec3f albedo = mk3(L.xf.x, L.xf.y, L.xf.z);
      if (numLights > 0) {
        L.Ld = albedo;
        light = true;
        segKind = kSegDlShadow;
      } else {
        const vec3f th = mk3s(1.f) * albedo;
        res = make_float4(th.x, th.y, th.z, 1.f);
        finish = true;
      }
    }
  } else {
    const vec3f th = mk3s(1.f) * (L.Ld * lightResult(P, L.which, L.Tr, L.ld));
    res = make_float4(th.x, th.y, th.z, 1.f);
    finish = true;
  }
  if (light) {
    segRay = lightRay(P, wpos, numLights, L.rng, L.which, L.ld);
    L.st = kStCellBegin;  // the caller begins the shadow segment
  }
  if (finish) endSample<kCount>(L, P, res, cnt);
}

// pathTracingIntegrate after sampleLight: new direction, new path ray
EXA_DEV void scatter(PathLane &L, const RenderParams &P) {
  const vec3f dir = hgSample(-L.ray.dir, L.rng);
  L.ray.org = L.spos;
  L.ray.dir = dir;
  float t0, t1;
  boxTest(L.ray, P.worldBounds, t0, t1);
  clipInterval(P, L.ray, t0, t1);
  L.ray.tmax = t1;
  L.st = kStPtSegment;
}

// camera ray and the integrators' entry (deviceCode.cu:1000-1027, 1130-1150)
template <bool kCount>
EXA_DEV void sampleStart(PathLane &L, const RenderParams &P, int numLights, LocalCounters &cnt, Ray &segRay,
                         int &segKind) {
  Ray ray = primaryRay(P, L.px, L.py, L.rng);
  float t0 = 1e30f, t1 = -1e30f;
  const bool dl = P.integrator == 1;
  if (!dl) {
    L.throughput = mk3s(1.f);
    L.Ld = mk3s(0.f);
    L.bounce = 0;
  }
  if (!boxTest(ray, P.worldBounds, t0, t1)) {
    const float4 bg = backgroundColor(L.py, P.H);
    endSample<kCount>(L, P, dl ? bg : ptResult(L.Ld, L.throughput, L.bounce, numLights, bg), cnt);
    return;
  }
  clipInterval(P, ray, t0, t1);
  ray.org = ray.org + ray.dir * t0;
  t1 -= t0;
  ray.tmax = t1;
  if (dl) {
    segRay = ray;
    segKind = kSegDl;
    L.st = kStCellBegin;  // the caller begins the segment
  } else {
    L.ray = ray;
    L.st = kStPtSegment;
  }
}

// hand out pixels to the lanes that wait for one: one atomic per wave.  A
// batched launch numbers the tiles of all requests consecutively
// (workBegin); a lane finds its request by binary search in LDS.
template <bool kBatch>
EXA_DEV void grabPixels(PathLane &L, const RenderParams &P, const PersistParams &pp, const uint32_t *workBegin,
                        int lane) {
  const uint64_t want = __ballot(L.st == kStGrab);
  if (!want) return;
  const int leader = __ffsll((unsigned long long)want) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(pp.work, (uint32_t)__popcll(want));
  base = __shfl(base, leader);
  if (L.st != kStGrab) return;
  uint32_t w = base + (uint32_t)__popcll(want & ((1ull << lane) - 1ull));
  if (w >= pp.numWork) {
    L.st = kStDone;
    return;
  }
  int32_t tilesX = (P.W + kPixelTile - 1) / kPixelTile, W = P.W, H = P.H, accumID = P.accumID;
  if (kBatch) {
    int lo = 0, hi = pp.numReqs - 1;  // last request with workBegin <= w
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (workBegin[mid] <= w) lo = mid;
      else hi = mid - 1;
    }
    L.req = lo;
    w -= workBegin[lo];
    const BatchRequest &R = pp.reqs[lo];
    tilesX = R.tilesX;
    W = R.P.W;
    H = R.P.H;
    accumID = R.P.accumID;
  }
  const uint32_t tile = w >> 6, in = w & 63u;
  const int px = (int)(tile % (uint32_t)tilesX) * kPixelTile + (int)(in & 7u);
  const int py = (int)(tile / (uint32_t)tilesX) * kPixelTile + (int)(in >> 3);
  if (px >= W || py >= H) return;  // outside a partial edge tile: grab again
  L.px = px;
  L.py = py;
  L.rng.init((uint32_t)(px + W * py), (uint32_t)accumID);
  L.acc = make_float4(0.f, 0.f, 0.f, 0.f);
  L.s = 0;
  L.st = kStSampleStart;
}

// the work id a path's pixel was handed out under (8x8 tiles in scan order,
// requests one after another): the key that orders the queue by image
// position
template <bool kBatch>
EXA_DEV uint32_t workIdOf(const PathLane &L, const RenderParams &P, const PersistParams &pp,
                          const uint32_t *workBegin) {
  uint32_t tilesX = (uint32_t)((P.W + kPixelTile - 1) / kPixelTile), base = 0;
  if (kBatch) {
    tilesX = (uint32_t)pp.reqs[L.req].tilesX;
    base = workBegin[L.req];
  }
  return base + ((((uint32_t)L.py >> 3) * tilesX + ((uint32_t)L.px >> 3)) << 6) + (((uint32_t)L.py & 7u) << 3) +
         ((uint32_t)L.px & 7u);
}

// hand out queued paths to the lanes that wait for work (resume launches):
// a lane takes the path and its element-stage sample, stashed in xf until
// the resolve pass.  One atomic per wave.
EXA_DEV void grabResumed(PathLane &L, const PersistParams &pp, int lane) {
  const uint64_t want = __ballot(L.st == kStGrab);
  if (!want) return;
  const int leader = __ffsll((unsigned long long)want) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(pp.resumeNext, (uint32_t)__popcll(want));
  base = __shfl(base, leader);
  if (L.st != kStGrab) return;
  const uint32_t w = base + (uint32_t)__popcll(want & ((1ull << lane) - 1ull));
  const uint32_t n = min(*pp.resumeIn.count, pp.resumeIn.cap);
  if (w >= n) {
    L.st = kStDone;
    return;
  }
  const uint32_t slot = pp.resumeIn.order[w];
  parkLoad(L, pp.resumeIn, slot);
  const Sample s = pp.resumeIn.samples[slot];
  L.xf = make_float4(__int_as_float(s.primID), __int_as_float(s.cellID), s.value, 0.f);
  L.st = kStResolve;
}

// Workgroup element queue in LDS (kShared): a lane whose gridlet stage
// missed posts its query (position, request) in its own slot and sets its
// bit in `posted`; a wave of the workgroup claims up to 64 posted queries at
// once (atomic AND on the bit words) and answers them with every lane busy,
// writing each result to the query's slot and setting its bit in `done`;
// the posting lane picks its result up at its next pass.  No barriers: a
// wave whose live lanes all wait answers whatever is posted itself, and
// claimed queries are answered at once, so every query is answered while its
// lane's wave lives.
template <bool kOn>
struct ElementQueueLds {
  static constexpr int N = kOn ? kPersistBlock : 1;
  float4 pos[N];   // x, y, z, request (int bits)
  uint2 res[N];    // primID, value bits
  unsigned long long posted[kPersistBlock / 64], done[kPersistBlock / 64];
  uint16_t claim[kPersistBlock / 64][kOn ? 64 : 1];  // per wave: the slots it claimed
};

template <bool kCount, int kAccel, bool kBatch, bool kNeural, bool kPark = false, bool kResume = false,
          bool kShared = false>
__global__ __launch_bounds__(kPersistBlock) __attribute__((amdgpu_waves_per_eu(EXA_PERSIST_MIN_WAVES, 8))) void
kRenderPersistent(RenderParams P, SamplerView sv, GridView gv, PersistParams pp, unsigned long long *counters) {
  __shared__ ElementQueueLds<kShared> eq;
  __shared__ int32_t stackMem[kAccel == kAccelBvh ? kStackDepth * kPersistBlock : 1];
  __shared__ uint32_t workBegin[kBatch ? kMaxBatchRequests : 1];
  // neural field: bf16 weights, and per wave the encodings and first hidden layer
  __shared__ uint16_t nW1[kNeural ? kNeuralIn * kNeuralHidden : 1], nW2[kNeural ? kNeuralHidden * kNeuralHidden : 1];
  __shared__ uint16_t nX[kNeural ? (kPersistBlock / 64) * 64 * kNeuralIn : 1];
  __shared__ uint16_t nH1[kNeural ? (kPersistBlock / 64) * 64 * kNeuralHidden : 1];
  int32_t *stack = stackMem + threadIdx.x;
  const int lane = (int)__lane_id();
  if (kBatch) {
    for (int i = threadIdx.x; i < pp.numReqs; i += blockDim.x) workBegin[i] = pp.reqs[i].workBegin;
    __syncthreads();
  }
  if (kNeural) {
    for (int i = threadIdx.x; i < kNeuralIn * kNeuralHidden; i += blockDim.x) nW1[i] = pp.nv.W1[i];
    for (int i = threadIdx.x; i < kNeuralHidden * kNeuralHidden; i += blockDim.x) nW2[i] = pp.nv.W2[i];
    __syncthreads();
  }
  const int waveInBlock = (int)threadIdx.x >> 6;
  if (kShared) {
    if (threadIdx.x < kPersistBlock / 64) {
      eq.posted[threadIdx.x] = 0ull;
      eq.done[threadIdx.x] = 0ull;
    }
    __syncthreads();
  }
  const int numWords = (int)(blockDim.x >> 6);
  LocalCounters cnt;
  if (kCount)
    for (int i = 0; i < kNumCounters; ++i) cnt.c[i] = 0;
  const int numLights1 = kBatch ? 0 : countLights(P);
  PathLane L;
  L.st = kStGrab;
  L.req = 0;
  for (;;) {
    const uint64_t live = __ballot(L.st != kStDone);
    if (!live) break;
    const int nLive = __popcll(live);
    // lanes that can progress without the gated passes
    const bool stuck = __ballot(isTracking(L.st) || L.st == kStAtSample || L.st == kStResolve) == 0;
    // rare pass, in the order a lane flows through it
    const uint64_t rare = __ballot(isRare(L.st));
    if (rare && (stuck || __popcll(rare) * 64 >= nLive * pp.rareBatch)) {
      // segRay / segKind: the segment a lane begins at the end of the pass
      Ray segRay;
      int segKind = -1;
      {
        const BatchRequest *R = kBatch ? pp.reqs + L.req : nullptr;
        const RenderParams &Pr = kBatch ? R->P : P;
        if (L.st == kStSegEnd) segmentEnd<kCount>(L, Pr, kBatch ? R->numLights : numLights1, cnt, segRay, segKind);
        if (L.st == kStScatter) scatter(L, Pr);
      }
      if (kResume) grabResumed(L, pp, lane);
      else grabPixels<kBatch>(L, P, pp, workBegin, lane);
      // a grab may have moved the lane to another request (a lane that
      // begins a shadow segment does not grab)
      const BatchRequest *R = kBatch ? pp.reqs + L.req : nullptr;
      const RenderParams &Pr = kBatch ? R->P : P;
      if (L.st == kStSampleStart) sampleStart<kCount>(L, Pr, kBatch ? R->numLights : numLights1, cnt, segRay, segKind);
      if (L.st == kStPtSegment) {
        segRay = L.ray;
        segKind = kSegPt;
      }
      if (segKind >= 0) beginSegment(L, Pr, gv, segRay, segKind);
    }

    // this lane's request (after any grab): parameters, majorants and
    // culling of its transfer function
    const BatchRequest *R = kBatch ? pp.reqs + L.req : nullptr;
    const RenderParams &Pl = kBatch ? R->P : P;
    GridView gvl = gv;
    Culling cu = cullingOf(sv);
    if (kBatch) {
      gvl.majorants = R->majorants;
      cu.gridletMaxOp = R->gridletMaxOp;
      cu.elemMaxOp = R->elemMaxOp;
    }

    // tracking pass
    trackingRun<kCount>(L, Pl, gvl, pp.ctrlSteps, cnt);

    // gridlet pass (cached gridlet first)
    Sample smp{-1, -1, 0.f};
    bool have = false;
    if (L.st == kStAtSample) {
      bool cachedHit;
      if (kAccel == kAccelTables) {
        smp = lookupTablesGridlets(sv, cu, L.pos, L.primID, cachedHit);
      } else {
        smp = testSampleGridlet(sv, L.pos, L.primID);
        cachedHit = smp.primID >= 0;
        if (!cachedHit) smp = locateGridlets<kAccel>(sv, cu, L.pos, stack, kPersistBlock);
      }
      if (smp.primID >= 0) {
        if (kCount) cnt.c[cachedHit ? kCntCachedHit : kCntGridletHit]++;
        have = true;
      } else {
        L.st = kStPending;
      }
    }

    // element-stage queue: lanes whose gridlet stage missed leave their path
    // in the queue and take other work (the lanes beyond its capacity stay
    // for the element pass)
    if (kPark) {
      const uint64_t pm = __ballot(L.st == kStPending);
      if (pm) {
        const int leader = __ffsll((unsigned long long)pm) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(pp.parkOut.count, (uint32_t)__popcll(pm));
        base = __shfl(base, leader);
        const uint32_t slot = base + (uint32_t)__popcll(pm & ((1ull << lane) - 1ull));
        if (L.st == kStPending && slot < pp.parkOut.cap) {
          parkStore(L, pp.parkOut, slot);
          pp.parkOut.keys[slot] = workIdOf<kBatch>(L, Pl, pp, workBegin);
          L.st = kStGrab;
        }
      }
    }

    if (kShared) {
      const int me = (int)threadIdx.x;
      // post: each waiting lane's query in its own slot, one atomic per wave
      const uint64_t pm = __ballot(L.st == kStPending);
      if (pm) {
        if (L.st == kStPending) {
          eq.pos[me] = make_float4(L.pos.x, L.pos.y, L.pos.z, __int_as_float(L.req));
          L.st = kStWaitElem;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane == __ffsll((unsigned long long)pm) - 1) atomicOr(&eq.posted[waveInBlock], (unsigned long long)pm);
      }
      // answer a batch of posted queries: when enough are posted, or when
      // every live lane of this wave waits for an answer
      const bool waiting = __ballot(L.st != kStWaitElem && L.st != kStDone) == 0;
      int total = 0;
      for (int w = 0; w < numWords; ++w) total += __popcll(__hip_atomic_load(&eq.posted[w], __ATOMIC_RELAXED,
                                                                             __HIP_MEMORY_SCOPE_WORKGROUP));
      if (total > 0 && (waiting || total >= (pp.elemBatch > 0 ? pp.elemBatch : 1))) {
        // claim up to 64 slots, word by word (bits another wave took first
        // are not in the returned word)
        int nClaim = 0;
        for (int w = 0; w < numWords && nClaim < 64; ++w) {
          unsigned long long want = 0ull;
          if (lane == 0) {
            const unsigned long long m = __hip_atomic_load(&eq.posted[w], __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_WORKGROUP);
            unsigned long long take = m;
            for (int k = __popcll(m); k > 64 - nClaim; --k) take &= ~(1ull << (63 - __clzll(take)));
            want = take ? (atomicAnd(&eq.posted[w], ~take) & take) : 0ull;
          }
          const unsigned long long got = __shfl(want, 0);
          if (got & (1ull << lane))
            eq.claim[waveInBlock][nClaim + __popcll(got & ((1ull << lane) - 1ull))] = (uint16_t)(w * 64 + lane);
          nClaim += __popcll(got);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        if (lane < nClaim) {
          const int slot = eq.claim[waveInBlock][lane];
          const float4 q = eq.pos[slot];
          Culling cq = cullingOf(sv);
          if (kBatch) {
            const BatchRequest &Rq = pp.reqs[__float_as_int(q.w)];
            cq.gridletMaxOp = Rq.gridletMaxOp;
            cq.elemMaxOp = Rq.elemMaxOp;
          }
          int kind;
          const Sample a = locateTablesElements(sv, cq, mk3(q.x, q.y, q.z), kind);
          if (kCount) cnt.c[kind < 0 ? kCntMiss : kCntGridletHit + kind]++;
          eq.res[slot] = make_uint2((uint32_t)a.primID, __float_as_uint(a.value));
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
          atomicOr(&eq.done[slot >> 6], 1ull << (slot & 63));
        }
      } else if (waiting && total == 0) {
        __builtin_amdgcn_s_sleep(1);  // the answers are being computed by another wave
      }
      // pick up answers
      if (L.st == kStWaitElem) {
        const unsigned long long d = __hip_atomic_load(&eq.done[waveInBlock], __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
        if (d & (1ull << lane)) {
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
          const uint2 r = eq.res[me];
          smp = Sample{(int32_t)r.x, -1, __uint_as_float(r.y)};
          have = true;
          L.st = kStPending;  // resolved below
        }
      }
      const uint64_t taken = __ballot(have && L.st == kStPending);
      if (taken && lane == __ffsll((unsigned long long)taken) - 1) atomicAnd(&eq.done[waveInBlock], ~(unsigned long long)taken);
    }

    // element pass
    const uint64_t pend = __ballot(L.st == kStPending && !have);
    if (!kShared && pend && (stuck || __popcll(pend) * 64 >= nLive * pp.elemBatch)) {
      bool net = false;
      if (kNeural) {
        // validated macrocells: the neural field answers (whole wave, MFMA)
        const int64_t mc = segmentCellID(gvl, L.S);
        net = L.st == kStPending && pp.netFlags[mc];
        if (__ballot(net)) {
          float v = neuralValue(pp.nv, L.pos, nW1, nW2, nX + waveInBlock * 64 * kNeuralIn,
                                nH1 + waveInBlock * 64 * kNeuralHidden);
          if (net) {
            if (pp.nv.widened) {
              const float2 w = pp.nv.widened[mc];
              if (w.x <= w.y) v = fminf(fmaxf(v, w.x), w.y);
            }
            smp = Sample{0x7fffffff, -1, v};
            if (kCount) cnt.c[kCntNeural]++;
            have = true;
          }
        }
      }
      if (L.st == kStPending && !net) {
        int kind;
        smp = locateElements<kAccel>(sv, cu, L.pos, stack, kPersistBlock, kind);
        if (kCount) cnt.c[kind < 0 ? kCntMiss : kCntGridletHit + kind]++;
        have = true;
      }
    }

    // a path resumed from the element queue brings its sample
    if (kResume && L.st == kStResolve) {
      smp = Sample{__float_as_int(L.xf.x), __float_as_int(L.xf.y), L.xf.z};
      have = true;
    }

    // resolve pass (Woodcock accept / ratio-tracking update)
    if (have) {
      if (kCount && smp.primID >= 0) cnt.c[kCntTfFetch]++;
      const bool shadow = L.kind == kSegPtShadow || L.kind == kSegDlShadow;
      if (resolveCollision(Pl, sv.gridlets, L.px, L.py, smp, shadow, L.majorant, L.rng, L.primID, L.xf, L.Tr)) {
        L.itype = kScattering;
        L.st = kStSegEnd;
      } else {
        L.st = kStStep;
      }
    }
  }
  if (kCount)
    for (int i = 0; i < kNumCounters; ++i)
      if (cnt.c[i]) atomicAdd(&counters[i], (unsigned long long)cnt.c[i]);
}

__global__ void kResetWork(uint32_t *work) {
  if (threadIdx.x == 0) *work = 0u;
}

// zero the counters of one round of the element-stage queue (a kernel, not
// a memset, so that a captured graph replays it)
__global__ void kResetQueue(uint32_t *a, uint32_t *b, uint32_t *c) {
  if (threadIdx.x == 0) {
    if (a) *a = 0u;
    if (b) *b = 0u;
    if (c) *c = 0u;
  }
}

// The element stage over the queued paths, one entry per lane: the point
// location of locateTablesElements at the path's collision position, under
// the culling of the path's transfer function.
template <bool kCount, bool kBatch>
__global__ __launch_bounds__(kLocateBlock) void kQueueElements(SamplerView sv, ParkQueue q,
                                                                const BatchRequest *reqs,
                                                                unsigned long long *counters) {
  constexpr int wPos = (int)(offsetof(PathLane, pos) / sizeof(float));
  constexpr int wReq = (int)(offsetof(PathLane, req) / sizeof(float));
  const uint32_t n = min(*q.count, q.cap);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const uint32_t i = q.order[k];
    const vec3f p = mk3(parkWord(q, wPos, i), parkWord(q, wPos + 1, i), parkWord(q, wPos + 2, i));
    Culling cu = cullingOf(sv);
    if (kBatch) {
      const BatchRequest &R = reqs[__float_as_int(parkWord(q, wReq, i))];
      cu.gridletMaxOp = R.gridletMaxOp;
      cu.elemMaxOp = R.elemMaxOp;
    }
    int kind;
    q.samples[i] = locateTablesElements(sv, cu, p, kind);
    if (kCount) atomicAdd(&counters[kind < 0 ? kCntMiss : kCntGridletHit + kind], 1ull);
  }
}

__global__ void kAdvanceBatch(BatchRequest *reqs, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) reqs[i].P.accumID += 1;
}

// Image order of a queue: entries bucketed by work id >> shift (strips of
// 2^shift pixels in tile order), a counting sort in three launches that read
// the entry count on the device.  Order within a bucket is arbitrary.
__global__ void kOrderCount(const uint32_t *keys, const uint32_t *count, uint32_t cap, int shift, uint32_t *hist) {
  const uint32_t n = min(*count, cap);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    atomicAdd(&hist[keys[i] >> shift], 1u);
}

__global__ __launch_bounds__(1024) void kOrderScan(uint32_t *hist, uint32_t nb) {
  // exclusive scan of hist[0..nb) by one workgroup: per-thread chunk sums,
  // a scan of the sums in LDS, then the chunks in place
  __shared__ uint32_t part[1024];
  const uint32_t t = threadIdx.x, per = (nb + 1023u) / 1024u;
  const uint32_t b0 = min(nb, t * per), b1 = min(nb, b0 + per);
  uint32_t sum = 0;
  for (uint32_t b = b0; b < b1; ++b) sum += hist[b];
  part[t] = sum;
  __syncthreads();
  for (uint32_t off = 1; off < 1024u; off <<= 1) {
    const uint32_t v = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = part[t] - sum;
  for (uint32_t b = b0; b < b1; ++b) {
    const uint32_t c = hist[b];
    hist[b] = run;
    run += c;
  }
}

__global__ void kOrderScatter(const uint32_t *keys, const uint32_t *count, uint32_t cap, int shift, uint32_t *offs,
                              uint32_t *order) {
  const uint32_t n = min(*count, cap);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    order[atomicAdd(&offs[keys[i] >> shift], 1u)] = i;
}

__global__ void kZeroU32Array(uint32_t *p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

template <bool kCount, int kAccel, bool kBatch, bool kNeural, bool kPark, bool kResume, bool kShared>
static int residentBlocks() {
  static std::atomic<int> cache[kMaxCachedDevices];
  return residentWorkgroups(kRenderPersistent<kCount, kAccel, kBatch, kNeural, kPark, kResume, kShared>,
                            kPersistBlock, cache);
}

// work items: pixels (64 per 8x8 tile) or, when resuming, at most the
// queue's capacity
template <bool kCount, int kAccel, bool kBatch, bool kNeural = false, bool kPark = false, bool kResume = false,
          bool kShared = false>
static void launchPersistentT(const RenderParams &p, const SamplerView &sv, const GridView &gv,
                              const PersistParams &pp, unsigned long long *counters, hipStream_t s) {
  const uint32_t items = kResume ? pp.resumeIn.cap : pp.numWork;
  const int bs = pp.blockSize;
  // the resident waves (occupancy is set by registers, not by workgroup size)
  const int64_t waves =
      (int64_t)residentBlocks<kCount, kAccel, kBatch, kNeural, kPark, kResume, kShared>() * (kPersistBlock / 64);
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(waves / (bs / 64), ((int64_t)items + bs - 1) / bs));
  hipLaunchKernelGGL((kRenderPersistent<kCount, kAccel, kBatch, kNeural, kPark, kResume, kShared>), dim3(blocks),
                     dim3(bs), 0, s, p, sv, gv, pp, counters);
}

// A frame with the element-stage queue: the first launch renders the pixels
// and queues every element-stage lookup; then, per round, the element stage
// runs over the queue and the next launch resumes the queued paths, queueing
// their next element-stage lookups into the other buffer; the last launch
// resumes without queueing (its lanes run the element pass themselves), so
// every path ends within rounds + 1 launches.  No host synchronisation.
template <bool kCount, bool kBatch>
static void launchQueued(const RenderParams &p, const SamplerView &sv, const GridView &gv, PersistParams pp,
                         unsigned long long *counters, hipStream_t s, const ElementQueueBuffers &qb, int rounds) {
  ParkQueue q[2];
  for (int k = 0; k < 2; ++k) q[k] = ParkQueue{qb.rec[k], qb.samples[k], qb.keys[k], qb.order, qb.count + k, qb.cap};
  uint32_t *resumeNext = qb.count + 2;
  // buckets of 2^shift work ids, at most qb.maxBuckets of them
  int shift = 6;
  while (((uint64_t)pp.numWork >> shift) + 1 > qb.maxBuckets) ++shift;
  const uint32_t nb = (uint32_t)(((uint64_t)pp.numWork >> shift) + 1);
  hipLaunchKernelGGL(kResetQueue, dim3(1), dim3(64), 0, s, pp.work, q[0].count, resumeNext);
  pp.parkOut = q[0];
  pp.resumeIn = q[1];
  pp.resumeNext = resumeNext;
  launchPersistentT<kCount, kAccelTables, kBatch, false, true, false>(p, sv, gv, pp, counters, s);
  const dim3 eg(std::max(1u, std::min(4096u, (qb.cap + kLocateBlock - 1) / kLocateBlock)));
  for (int r = 1; r <= rounds; ++r) {
    const ParkQueue in = q[(r - 1) & 1], out = q[r & 1];
    hipLaunchKernelGGL(kZeroU32Array, dim3(std::min(1024u, (nb + 255) / 256)), dim3(256), 0, s, qb.hist, nb);
    hipLaunchKernelGGL(kOrderCount, eg, dim3(kLocateBlock), 0, s, in.keys, in.count, in.cap, shift, qb.hist);
    hipLaunchKernelGGL(kOrderScan, dim3(1), dim3(1024), 0, s, qb.hist, nb);
    hipLaunchKernelGGL(kOrderScatter, eg, dim3(kLocateBlock), 0, s, in.keys, in.count, in.cap, shift, qb.hist,
                       qb.order);
    hipLaunchKernelGGL((kQueueElements<kCount, kBatch>), eg, dim3(kLocateBlock), 0, s, sv, in, pp.reqs, counters);
    hipLaunchKernelGGL(kResetQueue, dim3(1), dim3(64), 0, s, out.count, resumeNext, nullptr);
    pp.resumeIn = in;
    pp.parkOut = out;
    if (r < rounds) launchPersistentT<kCount, kAccelTables, kBatch, false, true, true>(p, sv, gv, pp, counters, s);
    else launchPersistentT<kCount, kAccelTables, kBatch, false, false, true>(p, sv, gv, pp, counters, s);
  }
}

template <bool kBatch>
static void launchAny(const RenderParams &p, const SamplerView &sv, int accel, const GridView &gv,
                      const PersistParams &pp, unsigned long long *counters, hipStream_t s,
                      const ElementQueueBuffers *qb, int rounds) {
  if (qb && qb->cap > 0 && rounds > 0 && !pp.netFlags && accel == kAccelTables) {
    if (counters) launchQueued<true, kBatch>(p, sv, gv, pp, counters, s, *qb, rounds);
    else launchQueued<false, kBatch>(p, sv, gv, pp, counters, s, *qb, rounds);
    return;
  }
  hipLaunchKernelGGL(kResetWork, dim3(1), dim3(64), 0, s, pp.work);
  if (pp.sharedQueue && !pp.netFlags && accel == kAccelTables) {
    if (counters) launchPersistentT<true, kAccelTables, kBatch, false, false, false, true>(p, sv, gv, pp, counters, s);
    else launchPersistentT<false, kAccelTables, kBatch, false, false, false, true>(p, sv, gv, pp, counters, s);
    return;
  }
  if (pp.netFlags) {  // neural field: tables accelerator
    if (counters) launchPersistentT<true, kAccelTables, kBatch, true>(p, sv, gv, pp, counters, s);
    else launchPersistentT<false, kAccelTables, kBatch, true>(p, sv, gv, pp, counters, s);
    return;
  }
  if (accel == kAccelBvh) {
    if (counters) launchPersistentT<true, kAccelBvh, kBatch>(p, sv, gv, pp, counters, s);
    else launchPersistentT<false, kAccelBvh, kBatch>(p, sv, gv, pp, counters, s);
  } else {
    if (counters) launchPersistentT<true, kAccelTables, kBatch>(p, sv, gv, pp, counters, s);
    else launchPersistentT<false, kAccelTables, kBatch>(p, sv, gv, pp, counters, s);
  }
}

static void setTuning(PersistParams &pp, const PersistTuning &t) {
  pp.ctrlSteps = t.ctrlSteps < 1 ? 1 : t.ctrlSteps;
  pp.elemBatch = t.elemBatch < 0 ? 0 : (t.elemBatch > 64 ? 64 : t.elemBatch);
  pp.rareBatch = t.rareBatch < 0 ? 0 : (t.rareBatch > 64 ? 64 : t.rareBatch);
  pp.parkOut = ParkQueue{};
  pp.resumeIn = ParkQueue{};
  pp.resumeNext = nullptr;
  pp.blockSize = (t.blockSize == 64 || t.blockSize == 128) ? t.blockSize : kPersistBlock;
  pp.sharedQueue = t.elementStage == kElemWorkgroupQueue ? 1 : 0;
}

static int queueRounds(const PersistTuning &t) {
  if (t.elementStage != kElemQueueRounds) return 0;
  return t.queueRounds < 1 ? 1 : (t.queueRounds > kMaxQueueRounds ? kMaxQueueRounds : t.queueRounds);
}

void launchRenderPersistent(const RenderParams &p, const SamplerView &sv, int accel, const GridView &gv,
                            uint32_t *work, const PersistTuning &t, unsigned long long *counters, hipStream_t s,
                            const NeuralView *nv, const uint8_t *netFlags, const ElementQueueBuffers *qb) {
  PersistParams pp;
  pp.nv = NeuralView{};
  if (nv) pp.nv = *nv;
  pp.netFlags = nv ? netFlags : nullptr;
  pp.work = work;
  pp.reqs = nullptr;
  pp.numReqs = 0;
  pp.tilesX = (p.W + kPixelTile - 1) / kPixelTile;
  const uint32_t tilesY = (uint32_t)((p.H + kPixelTile - 1) / kPixelTile);
  pp.numWork = (uint32_t)pp.tilesX * tilesY * 64u;
  setTuning(pp, t);
  launchAny<false>(p, sv, accel, gv, pp, counters, s, qb, queueRounds(t));
}

void launchRenderBatch(const SamplerView &sv, int accel, const GridView &gv, BatchRequest *reqs, int numReqs,
                       uint32_t numWork, uint32_t *work, const PersistTuning &t, unsigned long long *counters,
                       hipStream_t s, const NeuralView *nv, const uint8_t *netFlags, const ElementQueueBuffers *qb) {
  PersistParams pp;
  pp.nv = nv ? *nv : NeuralView{};
  pp.netFlags = nv ? netFlags : nullptr;
  pp.work = work;
  pp.reqs = reqs;
  pp.numReqs = numReqs;
  pp.tilesX = 0;
  pp.numWork = numWork;
  setTuning(pp, t);
  const RenderParams unused{};  // every lane reads its request's parameters
  launchAny<true>(unused, sv, accel, gv, pp, counters, s, qb, queueRounds(t));
  hipLaunchKernelGGL(kAdvanceBatch, dim3((numReqs + 63) / 64), dim3(64), 0, s, reqs, numReqs);
}

}  // namespace exa
