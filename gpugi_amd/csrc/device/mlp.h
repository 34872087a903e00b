// Radiance-cache MLP: layout, input encoding and launch wrappers.
//
// Network: 96 encoded inputs -> 4 hidden layers of 64 (ReLU) -> 3 outputs
// (padded to 16), no biases.  Weights are BF16 on the matrix cores with FP32
// accumulation; FP32 master weights and Adam moments live beside them.
//
// Raw vertex features (16 floats per sample, the cache's query record):
//   [0..2] position mapped to [0,1]^3 by the scene bounds, [3..5] direction,
//   [6..8] shading normal, [9..11] diffuse albedo, [12] log2(1 + Phong
//   exponent) / 16, [13..15] unused.
// Encoding (96 BF16): e[a*12+k] = sin(2^k pi p_a), e[36+a*12+k] = cos(2^k pi p_a)
// for axes a < 3 and frequencies k < 12; e[72..74] direction, e[75..77]
// normal, e[78..80] albedo, e[81] exponent feature, e[82..95] = 0.
//
// Weight layout: the five matrices W_l[out][in] row-major, concatenated
// (offsets below).  The transposed copy WT_l[in][out32] (output dimension
// padded to 32) feeds the backward pass.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

namespace gpugi {

constexpr int kMlpRaw = 16;
constexpr int kMlpIn = 96;
constexpr int kMlpInUsed = 82;
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 16;
constexpr int kMlpOutUsed = 3;
constexpr int kMlpLayers = 5;
constexpr int kMlpFreqs = 12;

// per layer: inputs, outputs, offset into W (elements), offset into WT (elements)
constexpr int kMlpLayerIn[kMlpLayers] = {kMlpIn, kMlpHidden, kMlpHidden, kMlpHidden, kMlpHidden};
constexpr int kMlpLayerOut[kMlpLayers] = {kMlpHidden, kMlpHidden, kMlpHidden, kMlpHidden, kMlpOut};
constexpr int kMlpWOff[kMlpLayers + 1] = {0, 6144, 10240, 14336, 18432, 19456};
constexpr int kMlpParams = 19456;
// WT_l is [in][32 * ceil(out / 32)]
constexpr int kMlpWTOff[kMlpLayers + 1] = {0, 6144, 10240, 14336, 18432, 20480};
constexpr int kMlpWTSize = 20480;
// The inference copy: W_l rows padded to K_l + 8 elements (so the 16 rows an
// A-operand read touches start 4 LDS banks apart) with the columns permuted
// to the order in which the inference kernel holds a layer's inputs in
// registers (see mlp.hip); per-layer offsets in elements.
constexpr int kMlpInferOff[kMlpLayers + 1] = {0, 6656, 11264, 15872, 20480, 21632};
constexpr int kMlpInferSize = 21632;

struct AdamParams {
  float lr, beta1, beta2, eps;
  uint32_t step;  // 1-based step for bias correction
};

#if defined(__HIPCC__)
// Device buffers of one cache.
struct MlpDev {
  float* w;          // kMlpParams FP32 master weights
  float* m;          // Adam first moments
  float* v;          // Adam second moments
  uint16_t* wb;      // kMlpParams BF16 weights [out][in]
  uint16_t* wtb;     // kMlpWTSize BF16 transposed weights [in][out32]
  uint16_t* wbi;     // kMlpInferSize BF16 inference copy (padded rows, permuted columns)
  float* partials;   // per-tile gradient partials (grad kernel)
  uint32_t partialTiles;
  float* grad;       // kMlpParams reduced gradient
  float* loss;       // per-tile loss partials
};

// Tiles of 32 samples the gradient kernel processes.
inline uint32_t MlpTiles(size_t n) { return (uint32_t)((n + 31) / 32); }

// Samples of one training step.  With count == nullptr: samples [0, n).  With
// a device counter: total = min(*count, n) samples are split into `steps`
// equal chunks (the remainder is dropped) and the step uses chunk `step`, so
// the host never waits for the count.
struct MlpBatch {
  const uint32_t* count;
  uint32_t n;
  uint32_t step;
  uint32_t steps;
};

__host__ __device__ inline void MlpBatchRange(const MlpBatch& b, uint32_t* first, uint32_t* size) {
  if (!b.count) {
    *first = 0;
    *size = b.n;
    return;
  }
  const uint32_t c = *b.count;
  const uint32_t total = c < b.n ? c : b.n;
  *size = total / b.steps;
  *first = b.step * *size;
}

hipError_t LaunchMlpEncode(const float* raw, uint16_t* enc, size_t n, hipStream_t stream);
// radiance: n x 4 floats (rgb, 0).
hipError_t LaunchMlpInfer(const MlpDev& m, const float* raw, float* radiance, size_t n, hipStream_t stream);
// Gradient of the relative L2 loss mean_{s,c<3} (y - t)^2 / (y^2 + 0.01) (y treated as a
// constant in the denominator) over the samples of `batch` into m.grad, and the loss
// into *lossOut (device scalar); both are 0 for an empty batch.  raw / targets:
// n x 16 / n x 4 floats.  Needs m.partialTiles >= MlpTiles(batch.n / batch.steps).
// Two launches: the per-tile partial gradients, then their fixed-order sum.
hipError_t LaunchMlpGradient(const MlpDev& m, const float* raw, const float* targets, const MlpBatch& batch,
                             float* lossOut, hipStream_t stream);
// The same two launches separately; with `adam` the sum also applies the Adam
// step to the master weights and refreshes their BF16 copies (skipped for an
// empty batch), so a caller can order other work between the two.
hipError_t LaunchMlpGradientTiles(const MlpDev& m, const float* raw, const float* targets, const MlpBatch& batch,
                                  hipStream_t stream);
hipError_t LaunchMlpReduce(const MlpDev& m, const MlpBatch& batch, float* lossOut, const AdamParams* adam,
                           hipStream_t stream);
// Re-derives the BF16 copies from the master weights (after an upload).
hipError_t LaunchMlpRefreshBf16(const MlpDev& m, hipStream_t stream);
#endif

}  // namespace gpugi
