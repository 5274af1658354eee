// neural.hip -- training and inference of the neural field sampler on gfx950:
// hash-grid encoding, bf16-MFMA MLP forward/backward with float32
// accumulation, hash-table gradient scatter, Adam, training-pool generation
// with the exact sampler, and per-macrocell validation.
#include "neural_kernels.h"

namespace exa {

// ---- random numbers: the PCG hash (Jarzynski & Olano, "Hash Functions for
// GPU Rendering", JCGT 2020) ----------------------------------------------
EXA_DEV uint32_t pcgHash(uint32_t v) {
  const uint32_t state = v * 747796405u + 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}
EXA_DEV float hashU01(uint32_t h) { return (float)(h >> 8) * (1.f / 16777216.f); }
EXA_DEV uint32_t hash2(uint32_t a, uint32_t b) { return pcgHash(a ^ pcgHash(b)); }

__global__ void kNeuralUniform(float *p, int64_t n, float lo, float hi, uint32_t seed) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = lo + (hi - lo) * hashU01(hash2(seed, (uint32_t)i ^ (uint32_t)(i >> 32) * 0x9e3779b9u));
}

__global__ void kPackTable(const float *master, uint32_t *packed, int64_t nEntries) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nEntries)
    packed[i] = (uint32_t)toBf16(master[2 * i]) | ((uint32_t)toBf16(master[2 * i + 1]) << 16);
}

__global__ void kToBf16(const float *master, uint16_t *out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = toBf16(master[i]);
}

// Adam with bias correction: c1 = 1 - b1^t, c2 = 1 - b2^t
__global__ void kAdam(float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2,
                      float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void kZero(float *p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// ---- training pool ---------------------------------------------------------
// elemFrac of the points lie in random stitching elements (a random convex
// combination of the element's vertices, Dirichlet(1) weights), the rest
// uniformly in the box of a random gridlet.
__global__ void kPoolPositions(const int32_t *indices, const float4 *vertices, int64_t numElements,
                               const Gridlet *gridlets, int64_t numGridlets, int64_t n, float elemFrac,
                               uint32_t seed, float *pos) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t h = hash2(seed, (uint32_t)i);
  const bool inElem = numElements > 0 && (numGridlets == 0 || hashU01(h) < elemFrac);
  float p[3] = {0.f, 0.f, 0.f};
  if (inElem) {
    h = pcgHash(h);
    const int64_t e = (int64_t)(((uint64_t)h * (uint64_t)numElements) >> 32);
    const int32_t *row = indices + 8 * e;
    float wsum = 0.f, w[8];
    int nv = 0;
    for (int k = 0; k < 8; ++k) {
      if (row[k] < 0) break;
      h = pcgHash(h);
      w[k] = -logf(1.f - hashU01(h));
      wsum += w[k];
      nv++;
    }
    for (int k = 0; k < nv; ++k) {
      const float4 v = vertices[row[k]];
      p[0] += w[k] / wsum * v.x;
      p[1] += w[k] / wsum * v.y;
      p[2] += w[k] / wsum * v.z;
    }
  } else if (numGridlets > 0) {
    h = pcgHash(h);
    const int64_t gi = (int64_t)(((uint64_t)h * (uint64_t)numGridlets) >> 32);
    vec3f lo, hi;
    gridletBounds(gridlets[gi], lo, hi);
    const float lo3[3] = {lo.x, lo.y, lo.z}, hi3[3] = {hi.x, hi.y, hi.z};
    for (int k = 0; k < 3; ++k) {
      h = pcgHash(h);
      p[k] = lo3[k] + (hi3[k] - lo3[k]) * hashU01(h);
    }
  }
  pos[3 * i] = p[0];
  pos[3 * i + 1] = p[1];
  pos[3 * i + 2] = p[2];
}

// pool entry: x in [0,1]^3 and the value normalised to the value range; a
// point the exact sampler does not find gets a NaN target (never sampled)
__global__ void kPoolFinish(const float *pos, const Sample *smp, int64_t n, NeuralView nv, float4 *pool) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Sample s = smp[i];
  const float x = (pos[3 * i] - nv.lo[0]) * nv.invExtent[0];
  const float y = (pos[3 * i + 1] - nv.lo[1]) * nv.invExtent[1];
  const float z = (pos[3 * i + 2] - nv.lo[2]) * nv.invExtent[2];
  const float t = s.primID >= 0 ? (s.value - nv.vmin) / nv.vrange : __int_as_float(0x7fc00000);
  pool[i] = make_float4(x, y, z, t);
}

// ---- training step ----------------------------------------------------------
// One wave per workgroup.  Each wave runs tilesPerWave tiles of 64 batch
// samples through the forward and backward passes, accumulates the MLP weight
// gradients in MFMA registers across its tiles and scatters the hash-table
// gradients with float atomics.
__global__ __launch_bounds__(64) void kNeuralTrain(TrainArgs a) {
  __shared__ uint16_t sW1[kNeuralIn * kNeuralHidden], sW2[kNeuralHidden * kNeuralHidden];
  __shared__ uint16_t sX[64 * kNeuralIn], sH1[64 * kNeuralHidden], sH2[64 * kNeuralHidden],
      sD1[64 * kNeuralHidden];
  __shared__ float sGy[64];
  const NeuralView &nv = a.nv;
  const int lane = (int)__lane_id();
  const int r = lane & 15, q = lane >> 4;
  for (int i = lane; i < kNeuralIn * kNeuralHidden; i += 64) sW1[i] = nv.W1[i];
  for (int i = lane; i < kNeuralHidden * kNeuralHidden; i += 64) sW2[i] = nv.W2[i];
  floatx4 gW2[16], gW1[8];
  for (int t = 0; t < 16; ++t) gW2[t] = floatx4{0, 0, 0, 0};
  for (int t = 0; t < 8; ++t) gW1[t] = floatx4{0, 0, 0, 0};
  float gw3 = 0.f, lossAcc = 0.f;
  const float invB = 2.f / (float)a.batch;
  for (int tile = 0; tile < a.tilesPerWave; ++tile) {
    const int64_t row = ((int64_t)blockIdx.x * a.tilesPerWave + tile) * 64 + lane;
    // batch sample row -> pool entry
    const uint32_t pick = hash2(a.stepSeed, (uint32_t)row);
    const int64_t src = a.identity ? (row < a.poolSize ? row : 0) : (int64_t)(((uint64_t)pick * (uint64_t)a.poolSize) >> 32);
    const float4 s = a.pool[src];
    const bool valid = row < a.batch && s.w == s.w;
    const vec3f x = mk3(s.x, s.y, s.z);
    float enc[kNeuralIn];
    neuralEncode(nv, x, enc);
#pragma unroll
    for (int k = 0; k < kNeuralIn; ++k) sX[lane * kNeuralIn + k] = valid ? toBf16(enc[k]) : (uint16_t)0;
    const float y = mlpForward64(sW1, sW2, nv.w3, sX, sH1, sH2);
    const float res = valid ? y - s.w : 0.f;
    lossAcc += res * res;
    const float gy = invB * res;
    sGy[lane] = gy;
    waveSync();
    // dw3[j] = sum_rows H2[row][j] gy[row]  (lane = j)
    for (int k = 0; k < 64; ++k) gw3 += fromBf16(sH2[k * kNeuralHidden + lane]) * sGy[k];
    waveSync();
    // dH2 = [H2 > 0] gy w3, in place (lane = row)
    for (int j = 0; j < kNeuralHidden; ++j) {
      const float h = fromBf16(sH2[lane * kNeuralHidden + j]);
      sH2[lane * kNeuralHidden + j] = h > 0.f ? toBf16(gy * fromBf16(nv.w3[j])) : (uint16_t)0;
    }
    waveSync();
    // dW2 += H1^T dH2
    for (int mt = 0; mt < 4; ++mt)
      for (int nt = 0; nt < 4; ++nt)
        gW2[mt * 4 + nt] = mfmaTile<true, false>(sH1, kNeuralHidden, sH2, kNeuralHidden, 16 * mt, 16 * nt, 64,
                                                 gW2[mt * 4 + nt]);
    // dH1 = [H1 > 0] (dH2 W2^T)
    for (int nt = 0; nt < 4; ++nt)
      for (int mt = 0; mt < 4; ++mt) {
        const floatx4 acc = mfmaTile<false, true>(sH2, kNeuralHidden, sW2, kNeuralHidden, 16 * mt, 16 * nt,
                                                  kNeuralHidden, floatx4{0, 0, 0, 0});
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int idx = (16 * mt + 4 * q + j) * kNeuralHidden + 16 * nt + r;
          sD1[idx] = fromBf16(sH1[idx]) > 0.f ? toBf16(acc[j]) : (uint16_t)0;
        }
      }
    waveSync();
    // dW1 += X^T dH1
    for (int mt = 0; mt < 2; ++mt)
      for (int nt = 0; nt < 4; ++nt)
        gW1[mt * 4 + nt] =
            mfmaTile<true, false>(sX, kNeuralIn, sD1, kNeuralHidden, 16 * mt, 16 * nt, 64, gW1[mt * 4 + nt]);
    // dX = dH1 W1^T (float, into the H1 buffer, which is no longer needed)
    float *sDX = (float *)sH1;
    waveSync();
    for (int nt = 0; nt < 2; ++nt)
      for (int mt = 0; mt < 4; ++mt) {
        const floatx4 acc = mfmaTile<false, true>(sD1, kNeuralHidden, sW1, kNeuralHidden, 16 * mt, 16 * nt,
                                                  kNeuralHidden, floatx4{0, 0, 0, 0});
#pragma unroll
        for (int j = 0; j < 4; ++j) sDX[(16 * mt + 4 * q + j) * kNeuralIn + 16 * nt + r] = acc[j];
      }
    waveSync();
    // hash-table gradients
    if (valid) {
      for (int l = 0; l < nv.levels; ++l) {
        uint32_t g[3];
        float w[3];
        neuralCell(nv, l, x, g, w);
        const float d0 = sDX[lane * kNeuralIn + 2 * l], d1 = sDX[lane * kNeuralIn + 2 * l + 1];
        float *gt = a.gTable + ((size_t)l << nv.log2T) * 2;
        for (int c = 0; c < 8; ++c) {
          const uint32_t idx = neuralIndex(nv, l, g[0] + (c & 1), g[1] + ((c >> 1) & 1), g[2] + ((c >> 2) & 1));
          const float wc = neuralCornerWeight(w, c);
          atomicAdd(gt + 2 * (size_t)idx, wc * d0);
          atomicAdd(gt + 2 * (size_t)idx + 1, wc * d1);
        }
      }
    }
    waveSync();
  }
  for (int mt = 0; mt < 4; ++mt)
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        atomicAdd(a.gW2 + (16 * mt + 4 * q + j) * kNeuralHidden + 16 * nt + r, gW2[mt * 4 + nt][j]);
  for (int mt = 0; mt < 2; ++mt)
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        atomicAdd(a.gW1 + (16 * mt + 4 * q + j) * kNeuralHidden + 16 * nt + r, gW1[mt * 4 + nt][j]);
  atomicAdd(a.gw3 + lane, gw3);
  for (int off = 32; off > 0; off >>= 1) lossAcc += __shfl_xor(lossAcc, off);
  if (lane == 0) atomicAdd(a.lossSum, lossAcc);
}

// ---- inference -----------------------------------------------------------------
// values at n positions (voxel space), 64 per wave-tile
__global__ __launch_bounds__(64) void kNeuralEval(NeuralView nv, const float *pos, int64_t n, float *out) {
  __shared__ uint16_t sW1[kNeuralIn * kNeuralHidden], sW2[kNeuralHidden * kNeuralHidden];
  __shared__ uint16_t sX[64 * kNeuralIn], sH1[64 * kNeuralHidden];
  const int lane = (int)__lane_id();
  for (int i = lane; i < kNeuralIn * kNeuralHidden; i += 64) sW1[i] = nv.W1[i];
  for (int i = lane; i < kNeuralHidden * kNeuralHidden; i += 64) sW2[i] = nv.W2[i];
  for (int64_t base = (int64_t)blockIdx.x * 64; base < n; base += (int64_t)gridDim.x * 64) {
    const int64_t i = base + lane;
    vec3f p = mk3s(0.f);
    if (i < n) p = mk3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    const float y = neuralValue(nv, p, sW1, sW2, sX, sH1);
    if (i < n) out[i] = y;
  }
}

// ---- per-macrocell validation --------------------------------------------------
// k-th validation point of macrocell mc: uniform in the macrocell's box
__global__ void kValPositions(GridView gv, int pointsPerMC, uint32_t seed, float *pos, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t mc = i / pointsPerMC;
  const int cx = (int)(mc % gv.dims[0]);
  const int cy = (int)((mc / gv.dims[0]) % gv.dims[1]);
  const int cz = (int)(mc / ((int64_t)gv.dims[0] * gv.dims[1]));
  const int c[3] = {cx, cy, cz};
  uint32_t h = hash2(seed, (uint32_t)i);
  for (int k = 0; k < 3; ++k) {
    h = pcgHash(h);
    const float span = (gv.bounds[3 + k] - gv.bounds[k]) / (float)gv.dims[k];
    pos[3 * i + k] = gv.bounds[k] + ((float)c[k] + hashU01(h)) * span;
  }
}

// per macrocell: number of element-stage points, exact misses, max |error|
__global__ void kValReduce(const Sample *exact, const float *net, int64_t n, int pointsPerMC, uint32_t *count,
                           uint32_t *miss, uint32_t *maxErrOrd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t mc = i / pointsPerMC;
  const Sample s = exact[i];
  if (s.primID >= 0 && s.cellID != -1) return;  // the gridlet stage answers here
  if (s.primID < 0) {
    atomicOr(&miss[mc], 1u);
    return;
  }
  atomicAdd(&count[mc], 1u);
  const float err = fabsf(net[i] - s.value);
  atomicMax(&maxErrOrd[mc], __float_as_uint(err));  // non-negative floats order as integers
}

// flags and widened majorants: a macrocell uses the network when it had
// element-stage validation points, no exact miss, and max error <= bound
__global__ void kValFlags(const uint32_t *count, const uint32_t *miss, const uint32_t *maxErrOrd, int64_t nmc,
                          float bound, uint8_t *flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nmc) return;
  flags[i] = (count[i] > 0 && miss[i] == 0 && __uint_as_float(maxErrOrd[i]) <= bound) ? 1 : 0;
}

__global__ void kWidenRanges(const float2 *ranges, const uint8_t *flags, int64_t nmc, float bound, float2 *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nmc) return;
  float2 r = ranges[i];
  if (flags[i] && r.x <= r.y) {
    r.x -= bound;
    r.y += bound;
  }
  out[i] = r;
}

// ---- host launchers -------------------------------------------------------------
static dim3 grid1(int64_t n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }

void neuralUniform(float *p, int64_t n, float lo, float hi, uint32_t seed, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(kNeuralUniform, grid1(n), dim3(256), 0, s, p, n, lo, hi, seed);
}
void neuralPack(const float *tableMaster, uint32_t *table, int64_t entries, const float *mlpMaster, uint16_t *mlp,
                int64_t mlpCount, hipStream_t s) {
  hipLaunchKernelGGL(kPackTable, grid1(entries), dim3(256), 0, s, tableMaster, table, entries);
  hipLaunchKernelGGL(kToBf16, grid1(mlpCount), dim3(256), 0, s, mlpMaster, mlp, mlpCount);
}
void neuralAdam(float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2, float eps,
                int step, hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(kAdam, grid1(n), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
}
void neuralZero(float *p, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(kZero, grid1(n), dim3(256), 0, s, p, n);
}
void neuralPoolPositions(const int32_t *indices, const float4 *vertices, int64_t numElements,
                         const Gridlet *gridlets, int64_t numGridlets, int64_t n, float elemFrac, uint32_t seed,
                         float *pos, hipStream_t s) {
  hipLaunchKernelGGL(kPoolPositions, grid1(n), dim3(256), 0, s, indices, vertices, numElements, gridlets,
                     numGridlets, n, elemFrac, seed, pos);
}
void neuralPoolFinish(const float *pos, const Sample *smp, int64_t n, const NeuralView &nv, float4 *pool,
                      hipStream_t s) {
  hipLaunchKernelGGL(kPoolFinish, grid1(n), dim3(256), 0, s, pos, smp, n, nv, pool);
}
void neuralTrain(const TrainArgs &a, hipStream_t s) {
  const int waves = (int)((a.batch + 64LL * a.tilesPerWave - 1) / (64LL * a.tilesPerWave));
  hipLaunchKernelGGL(kNeuralTrain, dim3(waves), dim3(64), 0, s, a);
}
void neuralEval(const NeuralView &nv, const float *pos, int64_t n, float *out, hipStream_t s) {
  if (n <= 0) return;
  const int64_t tiles = (n + 63) / 64;
  const unsigned blocks = (unsigned)std::min<int64_t>(tiles, 8192);
  hipLaunchKernelGGL(kNeuralEval, dim3(blocks), dim3(64), 0, s, nv, pos, n, out);
}
void neuralValidate(const GridView &gv, int pointsPerMC, uint32_t seed, float *pos, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(kValPositions, grid1(n), dim3(256), 0, s, gv, pointsPerMC, seed, pos, n);
}
void neuralValReduce(const Sample *exact, const float *net, int64_t n, int pointsPerMC, uint32_t *count,
                     uint32_t *miss, uint32_t *maxErrOrd, hipStream_t s) {
  hipLaunchKernelGGL(kValReduce, grid1(n), dim3(256), 0, s, exact, net, n, pointsPerMC, count, miss, maxErrOrd);
}
void neuralValFlags(const uint32_t *count, const uint32_t *miss, const uint32_t *maxErrOrd, int64_t nmc,
                    float bound, uint8_t *flags, hipStream_t s) {
  hipLaunchKernelGGL(kValFlags, grid1(nmc), dim3(256), 0, s, count, miss, maxErrOrd, nmc, bound, flags);
}
void neuralWidenRanges(const float2 *ranges, const uint8_t *flags, int64_t nmc, float bound, float2 *out,
                       hipStream_t s) {
  hipLaunchKernelGGL(kWidenRanges, grid1(nmc), dim3(256), 0, s, ranges, flags, nmc, bound, out);
}

}  // namespace exa
