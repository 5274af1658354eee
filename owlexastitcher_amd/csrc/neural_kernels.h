// neural_kernels.h -- launch wrappers of the neural field sampler (neural.hip)
#pragma once
#include "neural.h"

namespace exa {

struct TrainArgs {
  NeuralView nv;
  const float4 *pool;  // x in [0,1]^3, normalised target (NaN: no sample)
  int64_t poolSize;
  int32_t batch;
  int32_t tilesPerWave;
  int32_t identity;  // 1: batch row i is pool entry i (else a hashed pick)
  uint32_t stepSeed;
  float *gTable;  // float gradients, 2 per table entry
  float *gW1, *gW2, *gw3;
  float *lossSum;
};

void neuralUniform(float *p, int64_t n, float lo, float hi, uint32_t seed, hipStream_t s);
void neuralPack(const float *tableMaster, uint32_t *table, int64_t entries, const float *mlpMaster, uint16_t *mlp,
                int64_t mlpCount, hipStream_t s);
void neuralAdam(float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2, float eps,
                int step, hipStream_t s);
void neuralZero(float *p, int64_t n, hipStream_t s);
void neuralPoolPositions(const int32_t *indices, const float4 *vertices, int64_t numElements,
                         const Gridlet *gridlets, int64_t numGridlets, int64_t n, float elemFrac, uint32_t seed,
                         float *pos, hipStream_t s);
void neuralPoolFinish(const float *pos, const Sample *smp, int64_t n, const NeuralView &nv, float4 *pool,
                      hipStream_t s);
void neuralTrain(const TrainArgs &a, hipStream_t s);
void neuralEval(const NeuralView &nv, const float *pos, int64_t n, float *out, hipStream_t s);
void neuralValidate(const GridView &gv, int pointsPerMC, uint32_t seed, float *pos, int64_t n, hipStream_t s);
void neuralValReduce(const Sample *exact, const float *net, int64_t n, int pointsPerMC, uint32_t *count,
                     uint32_t *miss, uint32_t *maxErrOrd, hipStream_t s);
void neuralValFlags(const uint32_t *count, const uint32_t *miss, const uint32_t *maxErrOrd, int64_t nmc,
                    float bound, uint8_t *flags, hipStream_t s);
void neuralWidenRanges(const float2 *ranges, const uint8_t *flags, int64_t nmc, float bound, float2 *out,
                       hipStream_t s);

}  // namespace exa
