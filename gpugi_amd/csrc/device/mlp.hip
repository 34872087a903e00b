// Radiance-cache MLP on the CDNA4 matrix cores (v_mfma_f32_16x16x32_bf16).
//
// Every layer is computed transposed, Y^T = W * X^T, so the weight matrix is
// the A operand and 16 samples form the N dimension:
//   A (16 x 32): lane l holds W[out = 16*ob + l%16][in = 32*ks + 8*(l/16) + j], j < 8
//   B (32 x 16): lane l holds X[sample = l%16][in = 32*ks + 8*(l/16) + j]
//   D (16 x 16): lane l holds Y[sample = l%16][out = 16*ob + 4*(l/16) + v], v < 4
// so each lane reads 16 contiguous bytes of a weight row and of an activation
// row (LDS, sample-major) per instruction and writes back 4 outputs of one
// sample.  The backward pass uses the same instruction for
//   dW_l = dY_l^T X_l      (K = the 32 samples of a tile; feature-major copies)
//   dX_l = W_l^T dY_l      (A = the transposed weights WT_l[in][out32])
// Gradients are written per 32-sample tile and summed in a fixed order (tile
// classes t mod 4 in tile order, then the four classes), so a training step
// is deterministic.
#include <hip/hip_runtime.h>

#include <cmath>

#include "mlp.h"

namespace gpugi {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

// Round to nearest even (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ uint32_t pack_bf2(float lo, float hi) {
  return (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
}

// sin / cos of 2^k pi p for k < 12: sinpi / cospi at k = 0, 3, 6, 9 and two
// double-angle steps from each (sin 2t = 2 sin t cos t, cos 2t = (cos t - sin
// t)(cos t + sin t)); for p in [0, 1] the error stays below 4e-7, far inside
// the BF16 rounding of the features.
__device__ __forceinline__ void encode_axis(float p, float* sn, float* cs) {
#pragma unroll
  for (int a = 0; a < kMlpFreqs; a += 3) {
    const float x = ldexpf(p, a);  // exact
    float s = sinpif(x), c = cospif(x);
    sn[a] = s;
    cs[a] = c;
#pragma unroll
    for (int j = 1; j < 3; ++j) {
      const float s2 = (s + s) * c;
      const float c2 = (c - s) * (c + s);
      s = s2;
      c = c2;
      sn[a + j] = s;
      cs[a + j] = c;
    }
  }
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Encodes one quarter of a sample's 96 features: q < 3 the 24 sin/cos
// features of position axis q, q == 3 features 72..95.  dst(f) receives them.
template <typename Dst>
__device__ __forceinline__ void encode_part(const float* __restrict__ raw, int q, Dst dst) {
  if (q < 3) {
    float sn[kMlpFreqs], cs[kMlpFreqs];
    encode_axis(raw[q], sn, cs);
#pragma unroll
    for (int k = 0; k < kMlpFreqs; ++k) {
      dst(q * kMlpFreqs + k, sn[k]);
      dst(36 + q * kMlpFreqs + k, cs[k]);
    }
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) dst(72 + i, raw[3 + i]);
    dst(81, raw[12]);
#pragma unroll
    for (int i = 82; i < 96; ++i) dst(i, 0.0f);
  }
}

constexpr int kActStride = 104;  // sample-major rows: 96 features + 8 pad (bf16 elements)
constexpr int kHidStride = 72;   // 64 + 8
constexpr int kSmpStride = 40;   // feature-major rows: 32 samples + 8

// ---------------------------------------------------------------- inference

constexpr int kInferBlock = 512;  // 8 waves sharing one LDS copy of the weights

// Orders a wave's LDS accesses across lanes for the compiler (a wavefront-
// scope fence emits no barrier instruction): LDS runs one wave's operations in
// issue order, so a wave-private buffer needs nothing more.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// One hidden layer for the 16 samples of a wave: act rows (stride `stride`)
// are read as K_IN inputs and overwritten with the 64 ReLU outputs.  The rows
// belong to this wave alone, so a wave-level sync orders the reads before the
// overwrite.
template <int K_IN>
__device__ __forceinline__ void infer_hidden(const uint16_t* __restrict__ W, uint16_t* act, int stride, int lane) {
  const int s = lane & 15, q = lane >> 4;
  f32x4 acc[4];
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) acc[ob] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int ks = 0; ks < K_IN / 32; ++ks) {
    const bf16x8 b = ld8(act + s * stride + 32 * ks + 8 * q);
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) acc[ob] = mfma(ld8(W + (16 * ob + s) * K_IN + 32 * ks + 8 * q), b, acc[ob]);
  }
  wave_lds_sync();
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    uint2 pk;
    pk.x = (uint32_t)f2bf(fmaxf(acc[ob][0], 0.0f)) | ((uint32_t)f2bf(fmaxf(acc[ob][1], 0.0f)) << 16);
    pk.y = (uint32_t)f2bf(fmaxf(acc[ob][2], 0.0f)) | ((uint32_t)f2bf(fmaxf(acc[ob][3], 0.0f)) << 16);
    *reinterpret_cast<uint2*>(act + s * stride + 16 * ob + 4 * q) = pk;
  }
  wave_lds_sync();
}

__global__ __launch_bounds__(kInferBlock) void k_mlp_infer(const uint16_t* __restrict__ wb,
                                                           const float* __restrict__ raw, float* __restrict__ out,
                                                           uint32_t n) {
  __shared__ __align__(16) uint16_t sW[kMlpParams];
  __shared__ __align__(16) uint16_t sAct[(kInferBlock / 64) * 16 * kActStride];
  for (int i = threadIdx.x; i < kMlpParams / 8; i += kInferBlock)
    reinterpret_cast<uint4*>(sW)[i] = reinterpret_cast<const uint4*>(wb)[i];
  __syncthreads();  // the weights are shared and read-only from here on
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = lane & 15, q = lane >> 4;
  uint16_t* act = sAct + wave * 16 * kActStride;
  const uint32_t groups = (n + 15) / 16;
  const uint32_t wavesTotal = gridDim.x * (kInferBlock / 64);
  // each wave walks its own sequence of 16-sample groups
  for (uint32_t g = blockIdx.x * (kInferBlock / 64) + wave; g < groups; g += wavesTotal) {
    const uint32_t idx = g * 16 + s;
    float f[kMlpRaw];
#pragma unroll
    for (int i = 0; i < kMlpRaw; ++i) f[i] = 0.0f;
    if (idx < n) {
#pragma unroll
      for (int i = 0; i < kMlpRaw / 4; ++i) {
        const float4 v = reinterpret_cast<const float4*>(raw)[(size_t)idx * 4 + i];
        f[4 * i + 0] = v.x;
        f[4 * i + 1] = v.y;
        f[4 * i + 2] = v.z;
        f[4 * i + 3] = v.w;
      }
    }
    // the sample's row, 8-byte stores: lane q < 3 writes features [12q, 12q + 12)
    // and [36 + 12q, 48 + 12q), lane q == 3 features [72, 96)
    uint16_t* row = act + s * kActStride;
    if (q < 3) {
      float sn[kMlpFreqs], cs[kMlpFreqs];
      encode_axis(f[q], sn, cs);
      uint2* ds = reinterpret_cast<uint2*>(row + q * kMlpFreqs);
      uint2* dc = reinterpret_cast<uint2*>(row + 36 + q * kMlpFreqs);
#pragma unroll
      for (int j = 0; j < kMlpFreqs / 4; ++j) {
        ds[j] = make_uint2(pack_bf2(sn[4 * j], sn[4 * j + 1]), pack_bf2(sn[4 * j + 2], sn[4 * j + 3]));
        dc[j] = make_uint2(pack_bf2(cs[4 * j], cs[4 * j + 1]), pack_bf2(cs[4 * j + 2], cs[4 * j + 3]));
      }
    } else {
      uint2* d = reinterpret_cast<uint2*>(row + 72);
      d[0] = make_uint2(pack_bf2(f[3], f[4]), pack_bf2(f[5], f[6]));
      d[1] = make_uint2(pack_bf2(f[7], f[8]), pack_bf2(f[9], f[10]));
      d[2] = make_uint2(pack_bf2(f[11], f[12]), 0u);
      d[3] = d[4] = d[5] = make_uint2(0u, 0u);
    }
    wave_lds_sync();  // the four lanes of a row wrote it; MFMA lanes read across rows
    infer_hidden<kMlpIn>(sW + kMlpWOff[0], act, kActStride, lane);
    infer_hidden<kMlpHidden>(sW + kMlpWOff[1], act, kActStride, lane);
    infer_hidden<kMlpHidden>(sW + kMlpWOff[2], act, kActStride, lane);
    infer_hidden<kMlpHidden>(sW + kMlpWOff[3], act, kActStride, lane);
    f32x4 y = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ks = 0; ks < kMlpHidden / 32; ++ks)
      y = mfma(ld8(sW + kMlpWOff[4] + s * kMlpHidden + 32 * ks + 8 * q), ld8(act + s * kActStride + 32 * ks + 8 * q), y);
    if (q == 0 && idx < n)
      reinterpret_cast<float4*>(out)[idx] = make_float4(y[0], y[1], y[2], 0.0f);
    wave_lds_sync();  // rows are rewritten by the next group's encoding
  }
}

// ---------------------------------------------------------------- gradient

constexpr int kGradBlock = 64;  // one wave, one 32-sample tile per step

struct GradLds {
  uint16_t a0[32 * kActStride];            // encoded inputs, sample-major
  uint16_t a[4][32 * kHidStride];          // hidden outputs 1..4, sample-major
  uint16_t t0[kMlpIn * kSmpStride];        // encoded inputs, feature-major
  uint16_t t[4][kMlpHidden * kSmpStride];  // hidden outputs, feature-major
  uint16_t dys[2][32 * kHidStride];        // dY, sample-major (ping-pong)
  uint16_t dyt[2][kMlpHidden * kSmpStride];  // dY, feature-major
};

// Forward of layer l for half h (samples 16h..16h+15) of the tile, K_IN inputs
// from the sample-major `src` rows; ReLU outputs go to `dst` (sample-major)
// and `dstT` (feature-major).
template <int K_IN>
__device__ __forceinline__ void grad_fwd_hidden(const uint16_t* __restrict__ W, const uint16_t* src, int srcStride,
                                                uint16_t* dst, uint16_t* dstT, int h, int lane) {
  const int s = lane & 15, q = lane >> 4;
  const int smp = 16 * h + s;
  f32x4 acc[4];
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) acc[ob] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int ks = 0; ks < K_IN / 32; ++ks) {
    const bf16x8 b = ld8(src + smp * srcStride + 32 * ks + 8 * q);
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) acc[ob] = mfma(ld8(W + (16 * ob + s) * K_IN + 32 * ks + 8 * q), b, acc[ob]);
  }
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int o = 16 * ob + 4 * q + v;
      const uint16_t y = f2bf(fmaxf(acc[ob][v], 0.0f));
      dst[smp * kHidStride + o] = y;
      dstT[o * kSmpStride + smp] = y;
    }
  }
}

// dW_l (rows OUT_L, K_IN columns) += dY^T X over the tile's 32 samples.
template <int OUT_L, int K_IN>
__device__ __forceinline__ void grad_dw(const uint16_t* dyt, const uint16_t* xt, float* __restrict__ part, int lane) {
  const int s = lane & 15, q = lane >> 4;
#pragma unroll 1
  for (int ob = 0; ob < OUT_L / 16; ++ob) {
    const bf16x8 a = ld8(dyt + (16 * ob + s) * kSmpStride + 8 * q);
#pragma unroll 2
    for (int ib = 0; ib < K_IN / 16; ++ib) {
      const f32x4 d = mfma(a, ld8(xt + (16 * ib + s) * kSmpStride + 8 * q), f32x4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
      for (int v = 0; v < 4; ++v) part[(16 * ob + 4 * q + v) * K_IN + 16 * ib + s] = d[v];
    }
  }
}

// dX_l = W_l^T dY_l (K = OUT32 outputs), gated by the ReLU of the layer input
// x (sample-major, 64 features), into the next dY buffers.
template <int OUT32>
__device__ __forceinline__ void grad_dx(const uint16_t* __restrict__ WT, const uint16_t* dys, const uint16_t* x,
                                        uint16_t* ndys, uint16_t* ndyt, int lane) {
  const int s = lane & 15, q = lane >> 4;
#pragma unroll 1
  for (int h = 0; h < 2; ++h) {
    const int smp = 16 * h + s;
#pragma unroll 1
    for (int ib = 0; ib < kMlpHidden / 16; ++ib) {
      f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < OUT32 / 32; ++ks)
        acc = mfma(ld8(WT + (16 * ib + s) * OUT32 + 32 * ks + 8 * q), ld8(dys + smp * kHidStride + 32 * ks + 8 * q), acc);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * ib + 4 * q + v;
        const float g = bf2f(x[smp * kHidStride + i]) > 0.0f ? acc[v] : 0.0f;
        const uint16_t gb = f2bf(g);
        ndys[smp * kHidStride + i] = gb;
        ndyt[i * kSmpStride + smp] = gb;
      }
    }
  }
}

__global__ __launch_bounds__(kGradBlock) void k_mlp_grad(const uint16_t* __restrict__ wb, const uint16_t* __restrict__ wtb,
                                                         const float* __restrict__ rawAll, const float* __restrict__ targetsAll,
                                                         MlpBatch batch, float* __restrict__ partials,
                                                         float* __restrict__ lossPart) {
  __shared__ __align__(16) GradLds L;
  const int lane = threadIdx.x;
  const int s = lane & 15, q = lane >> 4;
  uint32_t first, n;
  MlpBatchRange(batch, &first, &n);
  const float* raw = rawAll + (size_t)first * kMlpRaw;
  const float* targets = targetsAll + (size_t)first * 4;
  const float invCount = n ? (float)(1.0 / (3.0 * (double)n)) : 0.0f;
  const uint32_t tiles = (n + 31) / 32;
  for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    // encode both halves
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
      const int smp = 16 * h + s;
      const uint32_t idx = tile * 32 + smp;
      float f[kMlpRaw];
#pragma unroll
      for (int i = 0; i < kMlpRaw; ++i) f[i] = 0.0f;
      if (idx < n) {
#pragma unroll
        for (int i = 0; i < kMlpRaw / 4; ++i) {
          const float4 v = reinterpret_cast<const float4*>(raw)[(size_t)idx * 4 + i];
          f[4 * i + 0] = v.x;
          f[4 * i + 1] = v.y;
          f[4 * i + 2] = v.z;
          f[4 * i + 3] = v.w;
        }
      }
      encode_part(f, q, [&](int k, float x) {
        const uint16_t b = f2bf(x);
        L.a0[smp * kActStride + k] = b;
        L.t0[k * kSmpStride + smp] = b;
      });
    }
    __syncthreads();
    // forward, keeping every layer's activations
#pragma unroll 1
    for (int h = 0; h < 2; ++h) grad_fwd_hidden<kMlpIn>(wb + kMlpWOff[0], L.a0, kActStride, L.a[0], L.t[0], h, lane);
    __syncthreads();
#pragma unroll 1
    for (int l = 1; l < 4; ++l) {
#pragma unroll 1
      for (int h = 0; h < 2; ++h)
        grad_fwd_hidden<kMlpHidden>(wb + kMlpWOff[l], L.a[l - 1], kHidStride, L.a[l], L.t[l], h, lane);
      __syncthreads();
    }
    // output layer and the loss gradient dY4 (outputs 0..2; the rest stay 0)
    float loss = 0.0f;
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
      const int smp = 16 * h + s;
      const uint32_t idx = tile * 32 + smp;
      f32x4 y = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < kMlpHidden / 32; ++ks)
        y = mfma(ld8(wb + kMlpWOff[4] + s * kMlpHidden + 32 * ks + 8 * q), ld8(L.a[3] + smp * kHidStride + 32 * ks + 8 * q), y);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int o = 4 * q + v;  // output 0..15
        float g = 0.0f;
        if (o < kMlpOutUsed && idx < n) {
          const float t = targets[(size_t)idx * 4 + o];
          const float d = y[v] - t;
          const float den = y[v] * y[v] + 0.01f;
          loss += d * d / den * invCount;
          g = 2.0f * d / den * invCount;
        }
        const uint16_t gb = f2bf(g);
        L.dys[0][smp * kHidStride + o] = gb;
        L.dys[0][smp * kHidStride + 16 + o] = 0;  // K padding to 32 outputs
        L.dyt[0][o * kSmpStride + smp] = gb;
      }
    }
    __syncthreads();
    float* part = partials + (size_t)tile * kMlpParams;
    // layer 4: dW (16 x 64) and dX through WT_4 (64 x 32)
    grad_dw<kMlpOut, kMlpHidden>(L.dyt[0], L.t[3], part + kMlpWOff[4], lane);
    grad_dx<32>(wtb + kMlpWTOff[4], L.dys[0], L.a[3], L.dys[1], L.dyt[1], lane);
    __syncthreads();
    int cur = 1;
#pragma unroll 1
    for (int l = 3; l >= 1; --l) {
      grad_dw<kMlpHidden, kMlpHidden>(L.dyt[cur], L.t[l - 1], part + kMlpWOff[l], lane);
      grad_dx<kMlpHidden>(wtb + kMlpWTOff[l], L.dys[cur], L.a[l - 1], L.dys[cur ^ 1], L.dyt[cur ^ 1], lane);
      __syncthreads();
      cur ^= 1;
    }
    grad_dw<kMlpHidden, kMlpIn>(L.dyt[cur], L.t0, part + kMlpWOff[0], lane);
    // tile loss
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
    if (lane == 0) lossPart[tile] = loss;
    __syncthreads();
  }
}

// Sum of the per-tile partials: block = 4 waves x 64 parameters; wave c sums
// tiles c, c + 4, c + 8, ... in order, then the four class sums are added in
// class order.  Block 0 also sums the tile losses (lane l: tiles l, l + 64, ...,
// then a fixed butterfly).
constexpr int kReduceParams = 64;
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ partials, MlpBatch batch,
                                                    const float* __restrict__ lossPart, float* __restrict__ grad,
                                                    float* __restrict__ lossOut) {
  __shared__ float cls[4][kReduceParams];
  uint32_t first, n;
  MlpBatchRange(batch, &first, &n);
  const uint32_t tiles = (n + 31) / 32;
  const int lane = threadIdx.x & 63, c = threadIdx.x >> 6;
  const uint32_t p = blockIdx.x * kReduceParams + lane;
  float g = 0.0f;
  if (p < (uint32_t)kMlpParams)
    for (uint32_t t = (uint32_t)c; t < tiles; t += 4) g += partials[(size_t)t * kMlpParams + p];
  cls[c][lane] = g;
  __syncthreads();
  if (c == 0 && p < (uint32_t)kMlpParams) grad[p] = ((cls[0][lane] + cls[1][lane]) + cls[2][lane]) + cls[3][lane];
  if (blockIdx.x == 0 && c == 0 && lossOut) {
    float l = 0.0f;
    for (uint32_t t = (uint32_t)lane; t < tiles; t += 64) l += lossPart[t];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    if (lane == 0) *lossOut = l;
  }
}

// Index into WT of parameter p = W_l[o][i].
__device__ __forceinline__ int wt_index(int p) {
  int l = 0;
  while (p >= kMlpWOff[l + 1]) ++l;
  const int r = p - kMlpWOff[l];
  const int in = kMlpLayerIn[l];
  const int o = r / in, i = r % in;
  const int out32 = (kMlpLayerOut[l] + 31) / 32 * 32;
  return kMlpWTOff[l] + i * out32 + o;
}

__global__ void k_mlp_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ grad, uint16_t* __restrict__ wb, uint16_t* __restrict__ wtb,
                           float lr, float beta1, float beta2, float eps, float bc1, float bc2, MlpBatch batch) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpParams) return;
  uint32_t first, n;
  MlpBatchRange(batch, &first, &n);
  if (n == 0) return;  // no samples this step
  const float g = grad[p];
  const float mm = beta1 * m[p] + (1.0f - beta1) * g;
  const float vv = beta2 * v[p] + (1.0f - beta2) * (g * g);
  m[p] = mm;
  v[p] = vv;
  const float mhat = mm / bc1;
  const float vhat = vv / bc2;
  const float ww = w[p] - lr * mhat / (sqrtf(vhat) + eps);
  w[p] = ww;
  const uint16_t b = f2bf(ww);
  wb[p] = b;
  wtb[wt_index(p)] = b;
}

__global__ void k_mlp_refresh(const float* __restrict__ w, uint16_t* __restrict__ wb, uint16_t* __restrict__ wtb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpParams) return;
  const uint16_t b = f2bf(w[p]);
  wb[p] = b;
  wtb[wt_index(p)] = b;
}

__global__ void k_mlp_encode(const float* __restrict__ raw, uint16_t* __restrict__ enc, uint32_t n) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t idx = t >> 2;
  const int q = t & 3;
  if (idx >= n) return;
  float f[kMlpRaw];
#pragma unroll
  for (int i = 0; i < kMlpRaw; ++i) f[i] = raw[(size_t)idx * kMlpRaw + i];
  uint16_t* row = enc + (size_t)idx * kMlpIn;
  encode_part(f, q, [&](int k, float x) { row[k] = f2bf(x); });
}

int GridFor(size_t units, int perBlock, int cap) {
  const size_t b = (units + perBlock - 1) / perBlock;
  return (int)(b < (size_t)cap ? (b ? b : 1) : (size_t)cap);
}

}  // namespace

hipError_t LaunchMlpEncode(const float* raw, uint16_t* enc, size_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  if (n > 0x3fffffffull) return hipErrorInvalidValue;
  const size_t threads = n * 4;
  hipLaunchKernelGGL(k_mlp_encode, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, stream, raw, enc,
                     (uint32_t)n);
  return hipGetLastError();
}

hipError_t LaunchMlpInfer(const MlpDev& m, const float* raw, float* radiance, size_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  if (n > 0xffffff00ull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mlp_infer, dim3(GridFor(n, 16 * (kInferBlock / 64) * 8, 512)), dim3(kInferBlock), 0, stream, m.wb, raw, radiance,
                     (uint32_t)n);
  return hipGetLastError();
}

hipError_t LaunchMlpGradient(const MlpDev& m, const float* raw, const float* targets, const MlpBatch& batch,
                             float* lossOut, hipStream_t stream) {
  if (batch.n == 0 || batch.n > 0xffffff00u || batch.steps == 0 || batch.step >= batch.steps)
    return hipErrorInvalidValue;
  if (!batch.count && (batch.step != 0 || batch.steps != 1)) return hipErrorInvalidValue;
  // the most tiles the batch can hold (its count is read on the device)
  const uint32_t tiles = MlpTiles(batch.n / batch.steps);
  if (tiles == 0 || tiles > m.partialTiles) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mlp_grad, dim3(GridFor(tiles, 1, 4096)), dim3(kGradBlock), 0, stream, m.wb, m.wtb, raw,
                     targets, batch, m.partials, m.loss);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + kReduceParams - 1) / kReduceParams), dim3(256), 0, stream,
                     m.partials, batch, m.loss, m.grad, lossOut);
  return hipGetLastError();
}

hipError_t LaunchMlpAdam(const MlpDev& m, const AdamParams& p, const MlpBatch& batch, hipStream_t stream) {
  const double bc1 = 1.0 - std::pow((double)p.beta1, (double)p.step);
  const double bc2 = 1.0 - std::pow((double)p.beta2, (double)p.step);
  hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, stream, m.w, m.m, m.v, m.grad, m.wb,
                     m.wtb, p.lr, p.beta1, p.beta2, p.eps, (float)bc1, (float)bc2, batch);
  return hipGetLastError();
}

hipError_t LaunchMlpRefreshBf16(const MlpDev& m, hipStream_t stream) {
  hipLaunchKernelGGL(k_mlp_refresh, dim3((kMlpParams + 255) / 256), dim3(256), 0, stream, m.w, m.wb, m.wtb);
  return hipGetLastError();
}

}  // namespace gpugi
