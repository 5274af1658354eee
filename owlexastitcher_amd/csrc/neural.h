// neural.h -- neural field sampler: multiresolution hash encoding + a small
// MLP (32 -> 64 -> 64 -> 1, no biases, ReLU), evaluated with bf16 MFMA on
// gfx950.  No reference counterpart (the learned part stands beside the
// exact ExaStitch sampler, for the stitching elements).
//
// Encoding: the published multiresolution hash encoding (Mueller et al.,
// SIGGRAPH 2022): level l has resolution N_l = floor(N_min * b^l),
// b = exp((ln N_max - ln N_min) / (L - 1)); a position x in [0,1]^3 is scaled
// to x * N_l; the 8 grid vertices around it index a table of T entries of 2
// features (dense index when (N_l + 1)^3 <= T, else the spatial hash
// (x * 1) ^ (y * 2654435761) ^ (z * 805459861) mod T); features are
// trilinearly interpolated.  Table entries are two bf16 packed in 32 bits.
#pragma once
#include "exa_kernels.h"

namespace exa {

constexpr int kNeuralMaxLevels = 16;
constexpr int kNeuralFeatures = 2;
constexpr int kNeuralIn = kNeuralMaxLevels * kNeuralFeatures;  // 32
constexpr int kNeuralHidden = 64;

struct NeuralView {
  int32_t levels;
  int32_t log2T;
  int32_t res[kNeuralMaxLevels];
  uint32_t dense[kNeuralMaxLevels];  // 1: dense index at this level
  const uint32_t *table;              // levels * 2^log2T packed bf16 pairs
  const uint16_t *W1, *W2, *w3;       // bf16, row-major [in][out]
  float lo[3], invExtent[3];          // x01 = (p - lo) * invExtent
  float vmin, vrange;                 // value = vmin + y * vrange
  // renderer: per macrocell the exact value range widened by the validation
  // bound; the network's value is clamped into its macrocell's range, so the
  // majorant built over that range bounds it (nullptr: raw values)
  const float2 *widened;
};

// ---- bf16 ----------------------------------------------------------------
// round to nearest even (finite inputs)
EXA_HD uint16_t toBf16(float f) {
  const uint32_t u = __builtin_bit_cast(uint32_t, f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
EXA_HD float fromBf16(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

// grid vertex -> table index at a level
EXA_DEV uint32_t neuralIndex(const NeuralView &nv, int l, uint32_t x, uint32_t y, uint32_t z) {
  const uint32_t n1 = (uint32_t)nv.res[l] + 1u;
  if (nv.dense[l]) return x + y * n1 + z * n1 * n1;
  const uint32_t h = (x * 1u) ^ (y * 2654435761u) ^ (z * 805459861u);
  return h & ((1u << nv.log2T) - 1u);
}

// Cell of x01 at level l: base vertex and fractional position.  The base is
// clamped to N_l - 1 so that x = 1 stays inside the grid.
EXA_DEV void neuralCell(const NeuralView &nv, int l, vec3f x01, uint32_t g[3], float w[3]) {
  const float n = (float)nv.res[l];
  const float p[3] = {x01.x * n, x01.y * n, x01.z * n};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float f = floorf(p[k]);
    f = f < 0.f ? 0.f : (f > n - 1.f ? n - 1.f : f);
    g[k] = (uint32_t)f;
    w[k] = p[k] - f;
  }
}

// corner c (bit 0 x, bit 1 y, bit 2 z) weight, as ((wx * wy) * wz)
EXA_DEV float neuralCornerWeight(const float w[3], int c) {
  const float wx = (c & 1) ? w[0] : 1.f - w[0];
  const float wy = (c & 2) ? w[1] : 1.f - w[1];
  const float wz = (c & 4) ? w[2] : 1.f - w[2];
  return wx * wy * wz;
}

// 32 encoded features of one position (float32, corners accumulated in order)
EXA_DEV void neuralEncode(const NeuralView &nv, vec3f x01, float enc[kNeuralIn]) {
#pragma unroll
  for (int i = 0; i < kNeuralIn; ++i) enc[i] = 0.f;
  for (int l = 0; l < nv.levels; ++l) {
    uint32_t g[3];
    float w[3];
    neuralCell(nv, l, x01, g, w);
    const uint32_t *tab = nv.table + ((size_t)l << nv.log2T);
    uint32_t e[8];
#pragma unroll
    for (int c = 0; c < 8; ++c)
      e[c] = tab[neuralIndex(nv, l, g[0] + (c & 1), g[1] + ((c >> 1) & 1), g[2] + ((c >> 2) & 1))];
    float f0 = 0.f, f1 = 0.f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const float wc = neuralCornerWeight(w, c);
      f0 += wc * fromBf16((uint16_t)(e[c] & 0xFFFFu));
      f1 += wc * fromBf16((uint16_t)(e[c] >> 16));
    }
    enc[2 * l] = f0;
    enc[2 * l + 1] = f1;
  }
}

// ---- wave-level bf16 MFMA tiles --------------------------------------------
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// orders LDS accesses of the lanes of one wave
EXA_DEV void waveSync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// 16x16 tile of C = A * B accumulated over K (multiple of 16) with
// v_mfma_f32_16x16x16_bf16.  A is M x K and B is K x N, bf16 in LDS, row-major
// with leading dimensions lda / ldb; kTransA / kTransB read the stored matrix
// transposed.  Lane l supplies A[m0 + (l & 15)][k + 4 (l >> 4) + j] and
// B[k + 4 (l >> 4) + j][n0 + (l & 15)], j = 0..3, and receives
// C[m0 + 4 (l >> 4) + j][n0 + (l & 15)].
template <bool kTransA, bool kTransB>
EXA_DEV floatx4 mfmaTile(const uint16_t *A, int lda, const uint16_t *B, int ldb, int m0, int n0, int K,
                         floatx4 acc) {
  const int lane = (int)__lane_id();
  const int r = lane & 15, q = lane >> 4;
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + 4 * q + j;
      a[j] = (short)(kTransA ? A[k * lda + m0 + r] : A[(m0 + r) * lda + k]);
      b[j] = (short)(kTransB ? B[(n0 + r) * ldb + k] : B[k * ldb + n0 + r]);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  return acc;
}

// Forward pass of 64 rows (one per lane) through the MLP: X (64 x 32 bf16,
// LDS) -> H1 (64 x 64) -> H2 (64 x 64) -> y.  W1/W2 are bf16 in LDS; w3 is
// read from global memory.  Returns this lane's y (row = lane).  All lanes of
// the wave must execute it.
EXA_DEV float mlpForward64(const uint16_t *W1, const uint16_t *W2, const uint16_t *w3, const uint16_t *X,
                           uint16_t *H1, uint16_t *H2) {
  const int lane = (int)__lane_id();
  const int r = lane & 15, q = lane >> 4;
  waveSync();
  for (int n0 = 0; n0 < kNeuralHidden; n0 += 16)
    for (int m0 = 0; m0 < 64; m0 += 16) {
      const floatx4 acc = mfmaTile<false, false>(X, kNeuralIn, W1, kNeuralHidden, m0, n0, kNeuralIn, floatx4{0, 0, 0, 0});
#pragma unroll
      for (int j = 0; j < 4; ++j) H1[(m0 + 4 * q + j) * kNeuralHidden + n0 + r] = toBf16(fmaxf(acc[j], 0.f));
    }
  waveSync();
  for (int n0 = 0; n0 < kNeuralHidden; n0 += 16)
    for (int m0 = 0; m0 < 64; m0 += 16) {
      const floatx4 acc =
          mfmaTile<false, false>(H1, kNeuralHidden, W2, kNeuralHidden, m0, n0, kNeuralHidden, floatx4{0, 0, 0, 0});
#pragma unroll
      for (int j = 0; j < 4; ++j) H2[(m0 + 4 * q + j) * kNeuralHidden + n0 + r] = toBf16(fmaxf(acc[j], 0.f));
    }
  waveSync();
  float y = 0.f;
  for (int k = 0; k < kNeuralHidden; ++k) y += fromBf16(H2[lane * kNeuralHidden + k]) * fromBf16(w3[k]);
  return y;
}

// Value at one position per lane (voxel space) for the whole wave; all lanes
// must execute it.  X (64 x 32) and H1 (64 x 64) are bf16 scratch in LDS, W1
// and W2 bf16 in LDS.  The second hidden layer is reduced against w3 as it is
// produced (no 64 x 64 buffer for it): each lane sums its columns, then the 16
// lanes of a row group add up by shuffles.
EXA_DEV float neuralValue(const NeuralView &nv, vec3f p, const uint16_t *W1, const uint16_t *W2, uint16_t *X,
                          uint16_t *H1) {
  const int lane = (int)__lane_id();
  const int r = lane & 15, q = lane >> 4;
  vec3f x = mk3((p.x - nv.lo[0]) * nv.invExtent[0], (p.y - nv.lo[1]) * nv.invExtent[1],
                (p.z - nv.lo[2]) * nv.invExtent[2]);
  x = mk3(fminf(fmaxf(x.x, 0.f), 1.f), fminf(fmaxf(x.y, 0.f), 1.f), fminf(fmaxf(x.z, 0.f), 1.f));
  float enc[kNeuralIn];
  neuralEncode(nv, x, enc);
  waveSync();
#pragma unroll
  for (int k = 0; k < kNeuralIn; ++k) X[lane * kNeuralIn + k] = toBf16(enc[k]);
  waveSync();
  for (int n0 = 0; n0 < kNeuralHidden; n0 += 16)
    for (int m0 = 0; m0 < 64; m0 += 16) {
      const floatx4 acc = mfmaTile<false, false>(X, kNeuralIn, W1, kNeuralHidden, m0, n0, kNeuralIn, floatx4{0, 0, 0, 0});
#pragma unroll
      for (int j = 0; j < 4; ++j) H1[(m0 + 4 * q + j) * kNeuralHidden + n0 + r] = toBf16(fmaxf(acc[j], 0.f));
    }
  waveSync();
  float part[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int j = 0; j < 4; ++j) part[mt][j] = 0.f;
  for (int n0 = 0; n0 < kNeuralHidden; n0 += 16) {
    const float w = fromBf16(nv.w3[n0 + r]);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const floatx4 acc =
          mfmaTile<false, false>(H1, kNeuralHidden, W2, kNeuralHidden, 16 * mt, n0, kNeuralHidden, floatx4{0, 0, 0, 0});
#pragma unroll
      for (int j = 0; j < 4; ++j) part[mt][j] += fromBf16(toBf16(fmaxf(acc[j], 0.f))) * w;
    }
  }
  float y = 0.f;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float v = part[mt][j];
      v += __shfl_xor(v, 1);
      v += __shfl_xor(v, 2);
      v += __shfl_xor(v, 4);
      v += __shfl_xor(v, 8);
      // row 16 mt + 4 q' + j lives in lane group q'; lane `row` takes it
      const float mine = __shfl(v, 16 * ((lane & 15) >> 2));
      if (mt == (lane >> 4) && j == (lane & 3)) y = mine;
    }
  waveSync();
  return nv.vmin + y * nv.vrange;
}

}  // namespace exa
