// sampler.hip -- ExaStitch sampler kernels for gfx950:
//   * per-gridlet value ranges   (sampler/ExaStitchSampler.cu:27-142)
//   * majorant-grid build        (Grid.cu:61-246, atomicOp.cuh:4-26)
//   * max-opacity mapping        (Grid.cu:687-719, ExaStitchSampler.cu:144-227)
//   * batched point location     (ExaStitchSampler.h:108-140 + the OptiX
//                                 programs in programs/ExaStitchSampler.cu)
#include <algorithm>
#include "exa_locate.h"

namespace exa {

// Order-preserving float <-> uint32 map so min/max become single integer
// atomics (the reference uses CAS loops, atomicOp.cuh:4-26; both compute the
// exact min/max, independent of order).  -0 is folded onto +0.
EXA_DEV uint32_t encOrd(float f) {
  uint32_t u = __float_as_uint(f);
  if (u == 0x80000000u) u = 0u;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
EXA_DEV float decOrd(uint32_t e) {
  const uint32_t u = (e & 0x80000000u) ? (e & 0x7fffffffu) : ~e;
  return __uint_as_float(u);
}

// Grid.cuh:55-63 projectOnGrid
EXA_DEV void projectOnGrid(vec3f V, const int32_t dims[3], const float bounds[6], int32_t out[3]) {
  const vec3f lo = mk3(bounds[0], bounds[1], bounds[2]);
  const vec3f hi = mk3(bounds[3], bounds[4], bounds[5]);
  const vec3f V01 = (V - lo) / (hi - lo);
  out[0] = clampi(sat_i32(V01.x * (float)dims[0]), 0, dims[0] - 1);
  out[1] = clampi(sat_i32(V01.y * (float)dims[1]), 0, dims[1] - 1);
  out[2] = clampi(sat_i32(V01.z * (float)dims[2]), 0, dims[2] - 1);
}

struct GridParams {
  int32_t dims[3];
  float bounds[6];
};

__global__ void kInitRanges(uint2 *enc, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  enc[i] = make_uint2(encOrd(1e30f), encOrd(-1e30f));
}

__global__ void kDecodeRanges(const uint2 *enc, float2 *ranges, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint2 e = enc[i];
  ranges[i] = make_float2(decOrd(e.x), decOrd(e.y));
}

// Grid.cu:73-117 (UMesh overload): the element's full non-NaN vertex range
// goes to every macrocell of its projected vertex AABB.
__global__ void kGridElements(const float4 *__restrict__ vertices, const int32_t *__restrict__ indices,
                              int64_t numElems, GridParams gp, uint2 *enc) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= numElems) return;
  const int32_t *I = indices + 8 * e;
  vec3f blo = mk3s(INFINITY), bhi = mk3s(-INFINITY);
  float vlo = 1e30f, vhi = -1e30f;
  for (int i = 0; i < 8; ++i) {
    if (I[i] < 0) break;
    const float4 V = vertices[I[i]];
    if (!(V.w != V.w)) {
      blo = vmin(blo, xyz(V));
      bhi = vmax(bhi, xyz(V));
      vlo = fminf(vlo, V.w);
      vhi = fmaxf(vhi, V.w);
    }
  }
  int32_t lo[3], up[3];
  projectOnGrid(blo, gp.dims, gp.bounds, lo);
  projectOnGrid(bhi, gp.dims, gp.bounds, up);
  const uint32_t elo = encOrd(vlo), ehi = encOrd(vhi);
  for (int z = lo[2]; z <= up[2]; ++z)
    for (int y = lo[1]; y <= up[1]; ++y)
      for (int x = lo[0]; x <= up[0]; ++x) {
        uint2 *r = enc + ((int64_t)z * gp.dims[1] + y) * gp.dims[0] + x;
        atomicMin(&r->x, elo);
        atomicMax(&r->y, ehi);
      }
}

// Grid.cu:134-246 (gridlet overload): one block per gridlet, one thread per
// cell; each cell's non-NaN 8-corner range goes to every macrocell (within
// the gridlet's projected box) whose box overlaps the cell box, inclusive.
// The candidate macrocells are narrowed to the cell's projected box +-2
// before the reference's exact overlap test, which selects the same set.
__global__ void kGridGridlets(const Gridlet *__restrict__ gridlets, const float *__restrict__ scalars,
                              int64_t numGridlets, GridParams gp, uint2 *enc) {
  const int64_t gi = blockIdx.x;
  if (gi >= numGridlets) return;
  const Gridlet G = gridlets[gi];
  vec3f glo, ghi;
  gridletBounds(G, glo, ghi);
  int32_t loMC[3], upMC[3];
  projectOnGrid(glo, gp.dims, gp.bounds, loMC);
  projectOnGrid(ghi, gp.dims, gp.bounds, upMC);
  const float mcSize[3] = {(gp.bounds[3] - gp.bounds[0]) / (float)gp.dims[0],
                           (gp.bounds[4] - gp.bounds[1]) / (float)gp.dims[1],
                           (gp.bounds[5] - gp.bounds[2]) / (float)gp.dims[2]};
  const int nx = G.dims[0], ny = G.dims[1], nz = G.dims[2];
  const int nsx = nx + 1, nsy = ny + 1;
  const float cw = (float)(1 << G.level);
  const float *s = scalars + G.begin;
  const int ncell = nx * ny * nz;
  for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
    const int x = c % nx, y = (c / nx) % ny, z = c / (nx * ny);
    float vlo = 1e30f, vhi = -1e30f;
#pragma unroll
    for (int dz = 0; dz < 2; ++dz)
#pragma unroll
      for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
          const float f = s[(z + dz) * nsy * nsx + (y + dy) * nsx + (x + dx)];
          if (!(f != f)) {
            vlo = fminf(vlo, f);
            vhi = fmaxf(vhi, f);
          }
        }
    if (vhi < vlo) continue;  // all corners NaN: the reference's update is a no-op
    const int cc[3] = {x, y, z};
    float clo[3], chi[3];
    for (int k = 0; k < 3; ++k) {
      clo[k] = (float)(G.lower[k] + cc[k]) * cw + .5f * cw;
      chi[k] = (float)(G.lower[k] + cc[k] + 1) * cw + .5f * cw;
    }
    int32_t pl[3], ph[3];
    projectOnGrid(mk3(clo[0], clo[1], clo[2]), gp.dims, gp.bounds, pl);
    projectOnGrid(mk3(chi[0], chi[1], chi[2]), gp.dims, gp.bounds, ph);
    int32_t a[3], b[3];
    for (int k = 0; k < 3; ++k) {
      a[k] = max(loMC[k], pl[k] - 2);
      b[k] = min(upMC[k], ph[k] + 2);
    }
    const uint32_t elo = encOrd(vlo), ehi = encOrd(vhi);
    for (int mz = a[2]; mz <= b[2]; ++mz)
      for (int my = a[1]; my <= b[1]; ++my)
        for (int mx = a[0]; mx <= b[0]; ++mx) {
          const int mc[3] = {mx, my, mz};
          bool ov = true;
          for (int k = 0; k < 3; ++k) {
            const float mlo = gp.bounds[k] + (float)mc[k] * mcSize[k];
            const float mhi = gp.bounds[k] + (float)(mc[k] + 1) * mcSize[k];
            ov = ov && !(chi[k] < mlo || clo[k] > mhi);
          }
          if (!ov) continue;
          uint2 *r = enc + ((int64_t)mz * gp.dims[1] + my) * gp.dims[0] + mx;
          atomicMin(&r->x, elo);
          atomicMax(&r->y, ehi);
        }
  }
}

// ExaStitchSampler.cu:52-142: one block per gridlet, block-wide min/max.
__global__ void kGridletRanges(const Gridlet *__restrict__ gridlets, const float *__restrict__ scalars,
                               int64_t numGridlets, float2 *ranges) {
  const int64_t gi = blockIdx.x;
  if (gi >= numGridlets) return;
  const Gridlet G = gridlets[gi];
  const int nx = G.dims[0], ny = G.dims[1], nz = G.dims[2];
  const int nsx = nx + 1, nsy = ny + 1;
  const float *s = scalars + G.begin;
  float vlo = 1e30f, vhi = -1e30f;
  const int ncell = nx * ny * nz;
  for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
    const int x = c % nx, y = (c / nx) % ny, z = c / (nx * ny);
    for (int dz = 0; dz < 2; ++dz)
      for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
          const float f = s[(z + dz) * nsy * nsx + (y + dy) * nsx + (x + dx)];
          if (!(f != f)) {
            vlo = fminf(vlo, f);
            vhi = fmaxf(vhi, f);
          }
        }
  }
  for (int off = 32; off > 0; off >>= 1) {
    vlo = fminf(vlo, __shfl_xor(vlo, off, 64));
    vhi = fmaxf(vhi, __shfl_xor(vhi, off, 64));
  }
  __shared__ float slo[16], shi[16];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    slo[wave] = vlo;
    shi[wave] = vhi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 1; w < nw; ++w) {
      vlo = fminf(vlo, slo[w]);
      vhi = fmaxf(vhi, shi[w]);
    }
    ranges[gi] = make_float2(vlo, vhi);
  }
}

// Grid.cu:687-719 == ExaStitchSampler.cu:144-176
EXA_DEV float mapMaxOpacity(float lo, float hi, const float4 *__restrict__ cmap, int numColors, float x0,
                            float x1) {
  if (hi < lo) return 0.f;
  lo -= x0;
  lo /= x1 - x0;
  hi -= x0;
  hi /= x1 - x0;
  const float nm1 = (float)(numColors - 1);
  const int32_t ilo = clampi(sat_i32(lo * nm1), 0, numColors - 1);
  const int32_t ihi = clampi((int32_t)((uint32_t)sat_i32(hi * nm1) + 1u), 0, numColors - 1);
  float mo = 0.f;
  for (int i = ilo; i <= ihi; ++i) mo = fmaxf(mo, cmap[i].w);
  return mo;
}

__global__ void kMapMaxOpacity(const float2 *__restrict__ ranges, int64_t n, const float4 *__restrict__ cmap,
                               int numColors, float x0, float x1, float *out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 r = ranges[i];
  out[i] = mapMaxOpacity(r.x, r.y, cmap, numColors, x0, x1);
}

// ExaStitchSampler.cu:178-227
__global__ void kElementMaxOpacity(const float4 *__restrict__ vertices, const int32_t *__restrict__ indices,
                                   int64_t n, const float4 *__restrict__ cmap, int numColors, float x0,
                                   float x1, float *out) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  float lo = 1e30f, hi = -1e30f;
  for (int i = 0; i < 8; ++i) {
    const int32_t idx = indices[8 * e + i];
    if (idx < 0) continue;
    const float w = vertices[idx].w;
    lo = fminf(lo, w);
    hi = fmaxf(hi, w);
  }
  out[e] = mapMaxOpacity(lo, hi, cmap, numColors, x0, x1);
}

template <bool kCount, int kAccel>
__global__ __launch_bounds__(kLocateBlock) void kLocate(SamplerView sv, const float *__restrict__ pos,
                                                        int64_t n, Sample *__restrict__ out,
                                                        unsigned long long *counters) {
  __shared__ int32_t stack[kAccel == kAccelBvh ? kStackDepth * kLocateBlock : 1];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const vec3f p = mk3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  int kind;
  const Sample s = locatePoint<kAccel>(sv, p, stack + threadIdx.x, kLocateBlock, kind);
  out[i] = s;
  if (kCount) atomicAdd(&counters[kind < 0 ? kCntMiss : kCntGridletHit + kind], 1ull);
}

// Batched point location in two stages.  Stage 1 answers every query that a
// gridlet holds (the owner tables' speculative lookup with the tables
// accelerator) and appends the others to a queue, one atomic per wave; stage
// 2 runs the element stage over the queue, so its waves are full of element
// queries instead of carrying the ~20% of them spread over every wave.
// queue[0] counts the entries queue[1..].
template <bool kCount, int kAccel>
__global__ __launch_bounds__(kLocateBlock) void kLocateGridletStage(SamplerView sv, const float *__restrict__ pos,
                                                                    int64_t n, Sample *__restrict__ out,
                                                                    uint32_t *__restrict__ queue,
                                                                    unsigned long long *counters) {
  __shared__ int32_t stack[kAccel == kAccelBvh ? kStackDepth * kLocateBlock : 1];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  bool pending = false;
  if (i < n) {
    const vec3f p = mk3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    const Culling cu = cullingOf(sv);
    Sample s;
    if (kAccel == kAccelTables) {
      bool cachedHit;
      s = lookupTablesGridlets(sv, cu, p, -1, cachedHit);
    } else {
      s = locateBvhGridlets(sv, cu, p, stack + threadIdx.x, kLocateBlock);
    }
    if (s.primID >= 0) {
      out[i] = s;
      if (kCount) atomicAdd(&counters[kCntGridletHit], 1ull);
    } else {
      pending = true;
    }
  }
  const uint64_t mask = __ballot(pending);
  if (mask == 0) return;
  const int lane = (int)__lane_id();
  const int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(&queue[0], (uint32_t)__popcll(mask));
  base = __shfl(base, leader);
  if (pending) queue[1 + base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = (uint32_t)i;
}

template <bool kCount, int kAccel>
__global__ __launch_bounds__(kLocateBlock) void kLocateElementStage(SamplerView sv, const float *__restrict__ pos,
                                                                     Sample *__restrict__ out,
                                                                     const uint32_t *__restrict__ queue,
                                                                     unsigned long long *counters) {
  __shared__ int32_t stack[kAccel == kAccelBvh ? kStackDepth * kLocateBlock : 1];
  const uint32_t count = queue[0];
  const Culling cu = cullingOf(sv);
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < count; q += gridDim.x * blockDim.x) {
    const uint32_t i = queue[1 + q];
    const vec3f p = mk3(pos[3 * (int64_t)i], pos[3 * (int64_t)i + 1], pos[3 * (int64_t)i + 2]);
    int kind;
    const Sample s = locateElements<kAccel>(sv, cu, p, stack + threadIdx.x, kLocateBlock, kind);
    out[i] = s;
    if (kCount) atomicAdd(&counters[kind < 0 ? kCntMiss : kCntGridletHit + kind], 1ull);
  }
}

__global__ void kZeroU32(uint32_t *p) { *p = 0u; }

__global__ void kTestSample(SamplerView sv, const float *__restrict__ pos, const int32_t *__restrict__ prim,
                            int64_t n, Sample *__restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const vec3f p = mk3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  const int32_t g = prim[i];
  out[i] = (g >= 0 && g < sv.numGridlets) ? testSampleGridlet(sv, p, g) : Sample{-1, -1, 0.f};
}

static inline unsigned blocksFor(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void launchGridletValueRanges(const Gridlet *g, const float *scalars, int64_t n, float2 *ranges,
                              hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(kGridletRanges, dim3((unsigned)n), dim3(256), 0, s, g, scalars, n, ranges);
}

void launchGridBuild(const SamplerView &sv, const int32_t dims[3], const float bounds[6], float2 *ranges,
                     uint2 *scratch, hipStream_t s) {
  GridParams gp;
  for (int k = 0; k < 3; ++k) gp.dims[k] = dims[k];
  for (int k = 0; k < 6; ++k) gp.bounds[k] = bounds[k];
  const int64_t nmc = (int64_t)dims[0] * dims[1] * dims[2];
  hipLaunchKernelGGL(kInitRanges, dim3(blocksFor(nmc, 256)), dim3(256), 0, s, scratch, nmc);
  if (sv.numElements > 0)
    hipLaunchKernelGGL(kGridElements, dim3(blocksFor(sv.numElements, 256)), dim3(256), 0, s, sv.vertices,
                       sv.indices, (int64_t)sv.numElements, gp, scratch);
  if (sv.numGridlets > 0)
    hipLaunchKernelGGL(kGridGridlets, dim3((unsigned)sv.numGridlets), dim3(256), 0, s, sv.gridlets,
                       sv.gridletScalars, (int64_t)sv.numGridlets, gp, scratch);
  hipLaunchKernelGGL(kDecodeRanges, dim3(blocksFor(nmc, 256)), dim3(256), 0, s, scratch, ranges, nmc);
}

void launchMapMaxOpacity(const float2 *ranges, int64_t n, const float4 *cmap, int numColors, float x0,
                         float x1, float *out, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(kMapMaxOpacity, dim3(blocksFor(n, 256)), dim3(256), 0, s, ranges, n, cmap, numColors,
                     x0, x1, out);
}

void launchElementMaxOpacity(const float4 *vertices, const int32_t *indices, int64_t n, const float4 *cmap,
                             int numColors, float x0, float x1, float *out, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(kElementMaxOpacity, dim3(blocksFor(n, 256)), dim3(256), 0, s, vertices, indices, n,
                     cmap, numColors, x0, x1, out);
}

void launchLocate(const SamplerView &sv, int accel, const float *pos, int64_t n, Sample *out,
                  unsigned long long *counters, hipStream_t s, uint32_t *queue) {
  if (n <= 0) return;
  const dim3 g(blocksFor(n, kLocateBlock)), b(kLocateBlock);
  if (!queue) {  // one pass: gridlet and element stage per thread
    if (accel == kAccelBvh) {
      if (counters) hipLaunchKernelGGL((kLocate<true, kAccelBvh>), g, b, 0, s, sv, pos, n, out, counters);
      else hipLaunchKernelGGL((kLocate<false, kAccelBvh>), g, b, 0, s, sv, pos, n, out, counters);
    } else {
      if (counters) hipLaunchKernelGGL((kLocate<true, kAccelTables>), g, b, 0, s, sv, pos, n, out, counters);
      else hipLaunchKernelGGL((kLocate<false, kAccelTables>), g, b, 0, s, sv, pos, n, out, counters);
    }
    return;
  }
  // two stages over a queue of n + 1 entries; stage 2 walks the queue with at
  // most 4096 blocks
  const dim3 g2(std::min<unsigned>(g.x, 4096u));
  hipLaunchKernelGGL(kZeroU32, dim3(1), dim3(1), 0, s, queue);
#define EXA_STAGES(C, A)                                                                                  \
  hipLaunchKernelGGL((kLocateGridletStage<C, A>), g, b, 0, s, sv, pos, n, out, queue, counters);          \
  hipLaunchKernelGGL((kLocateElementStage<C, A>), g2, b, 0, s, sv, pos, out, queue, counters)
  if (accel == kAccelBvh) {
    if (counters) { EXA_STAGES(true, kAccelBvh); } else { EXA_STAGES(false, kAccelBvh); }
  } else {
    if (counters) { EXA_STAGES(true, kAccelTables); } else { EXA_STAGES(false, kAccelTables); }
  }
#undef EXA_STAGES
}

void launchTestSample(const SamplerView &sv, const float *pos, const int32_t *primIDs, int64_t n, Sample *out,
                      hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(kTestSample, dim3(blocksFor(n, 256)), dim3(256), 0, s, sv, pos, primIDs, n, out);
}

}  // namespace exa
