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
// Gradients are written per 32-sample tile (four waves per tile) and summed in
// a fixed order (tile classes t mod 16 in tile order, then a pairwise tree over
// the classes), so a training step is deterministic; the Adam step runs in the
// same launch as that sum.
#include <hip/hip_runtime.h>

#include <cmath>

#include "mlp.h"

namespace gpugi {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

// Round to nearest even (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
// Two values rounded in one v_cvt_pk_bf16_f32, lo in the low half.
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack_bf2(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
}
// ReLU of two packed BF16 values: a BF16 bit pattern read as int16 is
// negative exactly when the value is negative, so max(x, 0) as int16 (one
// v_pk_max_i16) equals rounding max(v, 0) for every value but NaN.
typedef short s16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t relu_bf2(uint32_t x) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2, x), (s16x2){0, 0}));
}

// sin(pi x) and cos(pi x) for |x| < 2^31.  x = k + t with k = rint(x) and
// |t| <= 1/2, both exact in FP32; sin(pi t) = pi t + t^3 Q(t^2) and cos(pi t) =
// 1 + t^2 R(t^2), where Q (degree 3) and R (degree 4) are minimax fits on
// t^2 in [0, 1/4] (absolute error 8e-9 and 3e-10 before rounding); the
// parity of k flips both signs.  The FP32 result is within 1.2e-7 of the
// exact value.  18 VALU instructions against the 40 of sinpif + cospif,
// which also handle infinities and the sign of zero.
__device__ __forceinline__ void sincospi_enc(float x, float& sn, float& cs) {
  const float k = rintf(x);
  const float t = x - k;
  const float z = t * t;
  float q = fmaf(z, 0.0780123249f, -0.59851259f);
  q = fmaf(z, q, 2.55012512f);
  q = fmaf(z, q, -5.16771412f);
  float r = fmaf(z, -0.0244206991f, 0.234953925f);
  r = fmaf(z, r, -1.33521605f);
  r = fmaf(z, r, 4.05870962f);
  r = fmaf(z, r, -4.93480206f);
  const float s = fmaf(t, 3.14159274f, (t * z) * q);
  const float c = fmaf(z, r, 1.0f);
  const uint32_t flip = (uint32_t)(int)k << 31;
  sn = __uint_as_float(__float_as_uint(s) ^ flip);
  cs = __uint_as_float(__float_as_uint(c) ^ flip);
}

// sin / cos of 2^k pi p for k < 12: sincospi_enc at k = 0, 3, 6, 9 and two
// double-angle steps from each (sin 2t = 2 sin t cos t, cos 2t = (cos t - sin
// t)(cos t + sin t)), each of which doubles the error; for p in [0, 1] the
// error stays below 7e-7, inside the BF16 rounding of the features.
__device__ __forceinline__ void encode_axis(float p, float* sn, float* cs) {
#pragma unroll
  for (int a = 0; a < kMlpFreqs; a += 3) {
    const float x = ldexpf(p, a);  // exact
    float s, c;
    sincospi_enc(x, s, c);
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
//
// A wave keeps a layer's inputs in registers as the B operand: lane (s, q)
// holds, for each K-step ks, the 8 inputs in "slots" 32*ks + 8*q + j, j < 8.
// The output of a hidden layer lands in the D layout, lane (s, q) holding
// neuron 16*ob + 4*q + v of output block ob; packing D[2*ks'] and
// D[2*ks' + 1] gives exactly that lane's B operand of K-step ks' of the next
// layer, if the next layer's weight column for slot p is the one of neuron
//   16*ob + 4*q + v  where p = 32*(ob >> 1) + 8*q + 4*(ob & 1) + v.
// The encoding fills layer 0's slots in the order the lane computes the
// features (slot 32*(t/8) + 8*q + t%8 holds feature t of lane q, see
// encode_b).  The inference copy of the weights (MlpDev::wbi) stores every
// layer with its columns in slot order, so hidden activations never leave
// the registers.  The A fragments (the weights) do not depend on the samples:
// each wave loads all 38 of them from the block's LDS copy once (152 VGPRs)
// and keeps them for its whole sequence of samples, so the sample loop reads
// no LDS; two groups per step give each weight fragment two independent
// accumulation chains.

// Slot (physical column) of logical input i of layer l.
__host__ __device__ constexpr int infer_col(int l, int i) {
  if (l == 0) {
    int q = 0, t = 0;
    if (i < 72) {  // sin (i < 36) or cos of axis a at frequency k: pair P = 4a + k / 3
      const int a = (i % 36) / 12, k = i % 12, P = 4 * a + k / 3;
      q = P / 3;
      t = 6 * (P % 3) + k % 3 + (i >= 36 ? 3 : 0);
    } else {
      q = (i - 72) / 6;
      t = 18 + (i - 72) % 6;
    }
    return 32 * (t / 8) + 8 * q + t % 8;
  }
  const int ob = i / 16, q = (i % 16) / 4, v = i % 4;
  return 32 * (ob >> 1) + 8 * q + 4 * (ob & 1) + v;
}

constexpr bool infer_cols_are_permutations() {
  for (int l = 0; l < 2; ++l) {
    const int k = l == 0 ? kMlpIn : kMlpHidden;
    bool seen[kMlpIn] = {};
    for (int i = 0; i < k; ++i) {
      const int c = infer_col(l, i);
      if (c < 0 || c >= k || seen[c]) return false;
      seen[c] = true;
    }
  }
  return true;
}
static_assert(infer_cols_are_permutations(), "every layer's inputs map one to one onto its slots");
static_assert(kMlpInferOff[1] == 64 * (kMlpIn + 8) && kMlpInferOff[2] - kMlpInferOff[1] == 64 * (kMlpHidden + 8) &&
                  kMlpInferOff[5] - kMlpInferOff[4] == kMlpOut * (kMlpHidden + 8),
              "inference copy layout");

// Index into the inference copy of parameter p = W_l[o][i].
__device__ __forceinline__ int wi_index(int p) {
  int l = 0;
  while (p >= kMlpWOff[l + 1]) ++l;
  const int r = p - kMlpWOff[l];
  const int in = kMlpLayerIn[l];
  return kMlpInferOff[l] + (r / in) * (in + 8) + infer_col(l, r % in);
}

constexpr int kInferBlock = 512;  // 8 waves sharing one LDS copy of the weights
constexpr int kInferGroups = 2;   // 16-sample groups per wave step

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Lane q's share of one sample's raw features: the position components of
// its three (axis, frequency) pairs below (lo for the first, hi for the last)
// and the 6 values of slots 18..23 (lanes 0 and 1; zeroed later where unused).
struct InferRaw {
  float lo, hi;
  float e[6];
};

// Samples past the end (n >= 1) read the last one: each sample is its own
// column of the products, and only the outputs of samples below n are stored.
__device__ __forceinline__ InferRaw load_raw(const float* __restrict__ raw, uint32_t idx, uint32_t n, int q) {
  const float* row = raw + (size_t)(idx < n ? idx : n - 1) * 16;
  InferRaw r;
  r.lo = row[(3 * q) >> 2];
  r.hi = row[(3 * q + 2) >> 2];
  const int off = q == 1 ? 9 : 3;  // features 3..8 (lane 0) or 9..14 (lane 1)
#pragma unroll
  for (int k = 0; k < 6; ++k) r.e[k] = row[off + k];
  return r;
}

// Layer 0's B operands of one sample (lane q's 24 features, slot order).
// The 72 sin / cos features come from 12 (axis a, base frequency 3b) pairs
// P = 4a + b, each a sincospi_enc and two double-angle steps (the arithmetic
// of encode_axis); lane q evaluates pairs 3q .. 3q + 2, pair m of the lane
// filling slots 6m .. 6m + 2 (sin at frequencies 3b .. 3b + 2) and
// 6m + 3 .. 6m + 5 (cos).  The axis of pair 3q + m is that of pair 3q (lo)
// for m = 0, of pair 3q + 2 (hi) for m = 2, and for m = 1 hi on lane 1, lo
// elsewhere.  Slots 18..23 of lane q hold features 72 + 6q .. 77 + 6q
// (direction, normal, albedo, exponent feature, zeros).
__device__ __forceinline__ void encode_b(const InferRaw& r, int q, bf16x8* b) {
  float v[24];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int P = 3 * q + m;
    const float p = m == 0 ? r.lo : (m == 2 || q == 1 ? r.hi : r.lo);
    const float x = ldexpf(p, 3 * (P & 3));  // exact
    float sn, cs;
    sincospi_enc(x, sn, cs);
    v[6 * m] = sn;
    v[6 * m + 3] = cs;
#pragma unroll
    for (int j = 1; j < 3; ++j) {
      const float s2 = (sn + sn) * cs;
      const float c2 = (cs - sn) * (cs + sn);
      sn = s2;
      cs = c2;
      v[6 * m + j] = sn;
      v[6 * m + 3 + j] = cs;
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) v[18 + k] = (q == 0 || (q == 1 && k < 4)) ? r.e[k] : 0.0f;
#pragma unroll
  for (int ks = 0; ks < 3; ++ks)
    b[ks] = __builtin_bit_cast(bf16x8, (u32x4){pack_bf2(v[8 * ks], v[8 * ks + 1]), pack_bf2(v[8 * ks + 2], v[8 * ks + 3]),
                                               pack_bf2(v[8 * ks + 4], v[8 * ks + 5]),
                                               pack_bf2(v[8 * ks + 6], v[8 * ks + 7])});
}

// The A fragments of a layer with 64 outputs (lane (s, q): 8 inputs of row
// 16 ob + s), read once from the LDS copy and kept in registers.
template <int K_IN>
__device__ __forceinline__ void load_layer(const uint16_t* W, bf16x8 (&a)[K_IN / 32][4], int s, int q) {
#pragma unroll
  for (int ks = 0; ks < K_IN / 32; ++ks)
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) a[ks][ob] = ld8(W + (16 * ob + s) * (K_IN + 8) + 32 * ks + 8 * q);
}

// One layer with 64 outputs for kInferGroups groups: K_IN / 32 K-steps of
// B operands in, ReLU'd BF16 B operands of the next layer out.
template <int K_IN>
__device__ __forceinline__ void infer_layer(const bf16x8 (&A)[K_IN / 32][4], const bf16x8 (&x)[kInferGroups][K_IN / 32],
                                            bf16x8 (&h)[kInferGroups][2]) {
  f32x4 acc[kInferGroups][4];
#pragma unroll
  for (int g = 0; g < kInferGroups; ++g)
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) acc[g][ob] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int ks = 0; ks < K_IN / 32; ++ks)
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
#pragma unroll
      for (int g = 0; g < kInferGroups; ++g) acc[g][ob] = mfma(A[ks][ob], x[g][ks], acc[g][ob]);
#pragma unroll
  for (int g = 0; g < kInferGroups; ++g)
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const f32x4& lo = acc[g][2 * ks];
      const f32x4& hi = acc[g][2 * ks + 1];
      h[g][ks] = __builtin_bit_cast(bf16x8, (u32x4){relu_bf2(pack_bf2(lo[0], lo[1])), relu_bf2(pack_bf2(lo[2], lo[3])),
                                                    relu_bf2(pack_bf2(hi[0], hi[1])), relu_bf2(pack_bf2(hi[2], hi[3]))});
    }
}

// Lane (s, 0) of each group writes its sample's rgb (and a 0).
__device__ __forceinline__ void store_out(float* __restrict__ out, const f32x4 (&y)[kInferGroups], uint32_t u,
                                          uint32_t n, int s, int q) {
#pragma unroll
  for (int g = 0; g < kInferGroups; ++g) {
    const uint32_t idx = (u * kInferGroups + g) * 16 + s;
    if (q == 0 && idx < n) reinterpret_cast<float4*>(out)[idx] = make_float4(y[g][0], y[g][1], y[g][2], 0.0f);
  }
}

__global__ __launch_bounds__(kInferBlock) void k_mlp_infer(const uint16_t* __restrict__ wbi,
                                                           const float* __restrict__ raw, float* __restrict__ out,
                                                           uint32_t n) {
  __shared__ __align__(16) uint16_t sW[kMlpInferSize];
  for (int i = threadIdx.x; i < kMlpInferSize / 8; i += kInferBlock)
    reinterpret_cast<uint4*>(sW)[i] = reinterpret_cast<const uint4*>(wbi)[i];
  __syncthreads();  // the weights are shared and read-only from here on
  // wave-uniform, so the sample loop runs on scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int s = lane & 15, q = lane >> 4;
  const uint32_t steps = (n + 16 * kInferGroups - 1) / (16 * kInferGroups);
  const uint32_t wavesTotal = gridDim.x * (kInferBlock / 64);
  // The A fragments do not depend on the samples: 38 of them (152 VGPRs) stay
  // in registers for the whole loop.
  bf16x8 A0[kMlpIn / 32][4], A1[2][4], A2[2][4], A3[2][4], A4[2];
  load_layer<kMlpIn>(sW + kMlpInferOff[0], A0, s, q);
  load_layer<kMlpHidden>(sW + kMlpInferOff[1], A1, s, q);
  load_layer<kMlpHidden>(sW + kMlpInferOff[2], A2, s, q);
  load_layer<kMlpHidden>(sW + kMlpInferOff[3], A3, s, q);
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) A4[ks] = ld8(sW + kMlpInferOff[4] + s * (kMlpHidden + 8) + 32 * ks + 8 * q);
  // Each wave walks its own sequence of kInferGroups x 16 samples; the raw
  // features of its next step are loaded while it computes this one.
  // A step's outputs are stored at the start of the next one, behind that
  // step's loads, so the wait for those loads never waits for fresh stores.
  const uint32_t first = blockIdx.x * (kInferBlock / 64) + wave;
  uint32_t u = first;
  InferRaw next[kInferGroups];
#pragma unroll
  for (int g = 0; g < kInferGroups; ++g) {
    next[g] = load_raw(raw, (u * kInferGroups + g) * 16 + s, n, q);
    // The first step's features are waited for here, before the loop: the
    // wait at the top of a step then never counts that step's own prefetch.
    asm volatile("" ::"v"(next[g].lo), "v"(next[g].hi), "v"(next[g].e[0]), "v"(next[g].e[1]), "v"(next[g].e[2]),
                 "v"(next[g].e[3]), "v"(next[g].e[4]), "v"(next[g].e[5]));
  }
  f32x4 y[kInferGroups];
  for (; u < steps; u += wavesTotal) {
    InferRaw cur[kInferGroups];
    const uint32_t un = u + wavesTotal;  // past the last step: clamped reads, never used
#pragma unroll
    for (int g = 0; g < kInferGroups; ++g) {
      cur[g] = next[g];
      next[g] = load_raw(raw, (un * kInferGroups + g) * 16 + s, n, q);
    }
    if (u != first) store_out(out, y, u - wavesTotal, n, s, q);
    // keep the loads and stores here, a whole step ahead of the next wait
    __builtin_amdgcn_sched_barrier(0);
    bf16x8 x0[kInferGroups][3];
#pragma unroll
    for (int g = 0; g < kInferGroups; ++g) encode_b(cur[g], q, x0[g]);
    bf16x8 h[kInferGroups][2], h2[kInferGroups][2];
    infer_layer<kMlpIn>(A0, x0, h);
    infer_layer<kMlpHidden>(A1, h, h2);
    infer_layer<kMlpHidden>(A2, h2, h);
    infer_layer<kMlpHidden>(A3, h, h2);
#pragma unroll
    for (int g = 0; g < kInferGroups; ++g) y[g] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
      for (int g = 0; g < kInferGroups; ++g) y[g] = mfma(A4[ks], h2[g][ks], y[g]);
  }
  if (u != first) store_out(out, y, u - wavesTotal, n, s, q);
}

// ---------------------------------------------------------------- gradient

// Four waves share one 32-sample tile; every phase splits its 16 x 16 output
// blocks over the waves (each block is computed by one wave with the same
// instruction sequence whatever the split, so the split changes no bit).
constexpr int kGradWaves = 4;
constexpr int kGradBlock = 64 * kGradWaves;

struct GradLds {
  uint16_t a0[32 * kActStride];            // encoded inputs, sample-major
  uint16_t a[4][32 * kHidStride];          // hidden outputs 1..4, sample-major
  uint16_t t0[kMlpIn * kSmpStride];        // encoded inputs, feature-major
  uint16_t t[4][kMlpHidden * kSmpStride];  // hidden outputs, feature-major
  uint16_t dys[2][32 * kHidStride];        // dY, sample-major (ping-pong)
  uint16_t dyt[2][kMlpHidden * kSmpStride];  // dY, feature-major
  float loss[2];                           // the two halves' loss sums
};

// Forward of layer l for half h (samples 16h..16h+15) of the tile and output
// blocks ob0, ob0 + 1, K_IN inputs from the sample-major `src` rows; ReLU
// outputs go to `dst` (sample-major) and `dstT` (feature-major).
template <int K_IN>
__device__ __forceinline__ void grad_fwd_hidden(const uint16_t* __restrict__ W, const uint16_t* src, int srcStride,
                                                uint16_t* dst, uint16_t* dstT, int h, int ob0, int lane) {
  const int s = lane & 15, q = lane >> 4;
  const int smp = 16 * h + s;
  f32x4 acc[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) acc[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int ks = 0; ks < K_IN / 32; ++ks) {
    const bf16x8 b = ld8(src + smp * srcStride + 32 * ks + 8 * q);
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[j] = mfma(ld8(W + (16 * (ob0 + j) + s) * K_IN + 32 * ks + 8 * q), b, acc[j]);
  }
#pragma unroll
  for (int j = 0; j < 2; ++j) {
#pragma unroll
    for (int v = 0; v < 4; v += 2) {
      const int o = 16 * (ob0 + j) + 4 * q + v;  // even
      const uint32_t y = relu_bf2(pack_bf2(acc[j][v], acc[j][v + 1]));
      *reinterpret_cast<uint32_t*>(dst + smp * kHidStride + o) = y;
      dstT[o * kSmpStride + smp] = (uint16_t)y;
      dstT[(o + 1) * kSmpStride + smp] = (uint16_t)(y >> 16);
    }
  }
}

// dW_l (rows OUT_L, K_IN columns) += dY^T X over the tile's 32 samples: the
// (ob, ib) blocks numbered ob * (K_IN / 16) + ib, block b on wave b mod 4.
template <int OUT_L, int K_IN>
__device__ __forceinline__ void grad_dw(const uint16_t* dyt, const uint16_t* xt, float* __restrict__ part, int wave,
                                        int lane) {
  const int s = lane & 15, q = lane >> 4;
  constexpr int kIb = K_IN / 16;
#pragma unroll 2
  for (int b = wave; b < (OUT_L / 16) * kIb; b += kGradWaves) {
    const int ob = b / kIb, ib = b % kIb;
    const f32x4 d = mfma(ld8(dyt + (16 * ob + s) * kSmpStride + 8 * q), ld8(xt + (16 * ib + s) * kSmpStride + 8 * q),
                         f32x4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
    for (int v = 0; v < 4; ++v) part[(16 * ob + 4 * q + v) * K_IN + 16 * ib + s] = d[v];
  }
}

// dX_l = W_l^T dY_l (K = OUT32 outputs), gated by the ReLU of the layer input
// x (sample-major, 64 features), into the next dY buffers: input block
// ib = wave, both halves.
template <int OUT32>
__device__ __forceinline__ void grad_dx(const uint16_t* __restrict__ WT, const uint16_t* dys, const uint16_t* x,
                                        uint16_t* ndys, uint16_t* ndyt, int wave, int lane) {
  const int s = lane & 15, q = lane >> 4;
  const int ib = wave;
#pragma unroll 1
  for (int h = 0; h < 2; ++h) {
    const int smp = 16 * h + s;
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
static_assert(kMlpHidden / 16 == kGradWaves, "grad_dx gives each wave one 16-input block");

__global__ __launch_bounds__(kGradBlock) void k_mlp_grad(const uint16_t* __restrict__ wb, const uint16_t* __restrict__ wtb,
                                                         const float* __restrict__ rawAll, const float* __restrict__ targetsAll,
                                                         MlpBatch batch, float* __restrict__ partials,
                                                         float* __restrict__ lossPart) {
  __shared__ __align__(16) GradLds L;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = lane & 15, q = lane >> 4;
  // the forward split: half h, output blocks ob0 and ob0 + 1
  const int h = wave & 1, ob0 = 2 * (wave >> 1);
  uint32_t first, n;
  MlpBatchRange(batch, &first, &n);
  const float* raw = rawAll + (size_t)first * kMlpRaw;
  const float* targets = targetsAll + (size_t)first * 4;
  const float invCount = n ? (float)(1.0 / (3.0 * (double)n)) : 0.0f;
  const uint32_t tiles = (n + 31) / 32;
  for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    // encode: waves 0 and 1, one half each
    if (wave < 2) {
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
    grad_fwd_hidden<kMlpIn>(wb + kMlpWOff[0], L.a0, kActStride, L.a[0], L.t[0], h, ob0, lane);
    __syncthreads();
#pragma unroll 1
    for (int l = 1; l < 4; ++l) {
      grad_fwd_hidden<kMlpHidden>(wb + kMlpWOff[l], L.a[l - 1], kHidStride, L.a[l], L.t[l], h, ob0, lane);
      __syncthreads();
    }
    // output layer and the loss gradient dY4 (outputs 0..2; the rest stay 0): waves 0 and 1
    if (wave < 2) {
      const int smp = 16 * h + s;
      const uint32_t idx = tile * 32 + smp;
      float loss = 0.0f;
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
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
      if (lane == 0) L.loss[h] = loss;
    }
    __syncthreads();
    float* part = partials + (size_t)tile * kMlpParams;
    // layer 4: dW (16 x 64) and dX through WT_4 (64 x 32)
    grad_dw<kMlpOut, kMlpHidden>(L.dyt[0], L.t[3], part + kMlpWOff[4], wave, lane);
    grad_dx<32>(wtb + kMlpWTOff[4], L.dys[0], L.a[3], L.dys[1], L.dyt[1], wave, lane);
    __syncthreads();
    int cur = 1;
#pragma unroll 1
    for (int l = 3; l >= 1; --l) {
      grad_dw<kMlpHidden, kMlpHidden>(L.dyt[cur], L.t[l - 1], part + kMlpWOff[l], wave, lane);
      grad_dx<kMlpHidden>(wtb + kMlpWTOff[l], L.dys[cur], L.a[l - 1], L.dys[cur ^ 1], L.dyt[cur ^ 1], wave, lane);
      __syncthreads();
      cur ^= 1;
    }
    grad_dw<kMlpHidden, kMlpIn>(L.dyt[cur], L.t0, part + kMlpWOff[0], wave, lane);
    if (threadIdx.x == 0) lossPart[tile] = L.loss[0] + L.loss[1];
    __syncthreads();
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

// Adam on one weight (bias corrections bc1 = 1 - beta1^t, bc2 = 1 - beta2^t
// computed on the host), then its two BF16 copies.
struct AdamArgs {
  float lr, beta1, beta2, eps, bc1, bc2;
};
__device__ __forceinline__ void adam_update(const MlpDev& d, int p, float g, const AdamArgs& a) {
  const float mm = a.beta1 * d.m[p] + (1.0f - a.beta1) * g;
  const float vv = a.beta2 * d.v[p] + (1.0f - a.beta2) * (g * g);
  d.m[p] = mm;
  d.v[p] = vv;
  const float mhat = mm / a.bc1;
  const float vhat = vv / a.bc2;
  const float ww = d.w[p] - a.lr * mhat / (sqrtf(vhat) + a.eps);
  d.w[p] = ww;
  const uint16_t b = f2bf(ww);
  d.wb[p] = b;
  d.wtb[wt_index(p)] = b;
  d.wbi[wi_index(p)] = b;
}

// Sum of the per-tile partials: block = 16 waves x 64 parameters; wave c
// sums tiles c, c + 16, c + 32, ... in tile order (the loads of an unrolled
// run are independent of the sum, so eight are in flight per lane), then
// wave 0 adds the sixteen class sums as a fixed pairwise tree.  The order
// depends only on the tile count, so a step is deterministic.  With ADAM the
// same lane then applies the Adam step to its weight (skipped for an empty
// batch).  Block 0 also sums the tile losses (lane l: tiles l, l + 64, ...,
// then a fixed butterfly).
constexpr int kReduceParams = 64;
constexpr int kReduceWaves = 16;
template <bool ADAM>
__global__ __launch_bounds__(64 * kReduceWaves) void k_mlp_reduce(MlpDev d, MlpBatch batch, float* __restrict__ lossOut,
                                                                  AdamArgs adam) {
  __shared__ float cls[kReduceWaves][kReduceParams];
  uint32_t first, n;
  MlpBatchRange(batch, &first, &n);
  const uint32_t tiles = (n + 31) / 32;
  const int lane = threadIdx.x & 63, c = threadIdx.x >> 6;
  const uint32_t p = blockIdx.x * kReduceParams + lane;
  const float* __restrict__ partials = d.partials;
  float g = 0.0f;
  if (p < (uint32_t)kMlpParams) {
#pragma unroll 8
    for (uint32_t t = (uint32_t)c; t < tiles; t += kReduceWaves) g += partials[(size_t)t * kMlpParams + p];
  }
  cls[c][lane] = g;
  __syncthreads();
  if (c == 0 && p < (uint32_t)kMlpParams) {
    float v[kReduceWaves];
#pragma unroll
    for (int i = 0; i < kReduceWaves; ++i) v[i] = cls[i][lane];
#pragma unroll
    for (int w = 1; w < kReduceWaves; w *= 2)
#pragma unroll
      for (int i = 0; i < kReduceWaves; i += 2 * w) v[i] = v[i] + v[i + w];
    d.grad[p] = v[0];
    if (ADAM && n > 0) adam_update(d, (int)p, v[0], adam);
  }
  if (blockIdx.x == 0 && c == 0 && lossOut) {
    float l = 0.0f;
    for (uint32_t t = (uint32_t)lane; t < tiles; t += 64) l += d.loss[t];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    if (lane == 0) *lossOut = l;
  }
}

__global__ void k_mlp_refresh(const float* __restrict__ w, uint16_t* __restrict__ wb, uint16_t* __restrict__ wtb,
                              uint16_t* __restrict__ wbi) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpParams) return;
  const uint16_t b = f2bf(w[p]);
  wb[p] = b;
  wtb[wt_index(p)] = b;
  wbi[wi_index(p)] = b;
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
  // at least 8 steps per wave, at most one block per CU (the kernel keeps the
  // whole network in registers, two waves per SIMD)
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  hipLaunchKernelGGL(k_mlp_infer, dim3(GridFor(n, 16 * kInferGroups * (kInferBlock / 64) * 8, cus)), dim3(kInferBlock), 0,
                     stream, m.wbi, raw, radiance, (uint32_t)n);
  return hipGetLastError();
}

hipError_t LaunchMlpGradientTiles(const MlpDev& m, const float* raw, const float* targets, const MlpBatch& batch,
                                  hipStream_t stream) {
  if (batch.n == 0 || batch.n > 0xffffff00u || batch.steps == 0 || batch.step >= batch.steps)
    return hipErrorInvalidValue;
  if (!batch.count && (batch.step != 0 || batch.steps != 1)) return hipErrorInvalidValue;
  // the most tiles the batch can hold (its count is read on the device)
  const uint32_t tiles = MlpTiles(batch.n / batch.steps);
  if (tiles == 0 || tiles > m.partialTiles) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mlp_grad, dim3(GridFor(tiles, 1, 4096)), dim3(kGradBlock), 0, stream, m.wb, m.wtb, raw,
                     targets, batch, m.partials, m.loss);
  return hipGetLastError();
}

hipError_t LaunchMlpReduce(const MlpDev& m, const MlpBatch& batch, float* lossOut, const AdamParams* adam,
                           hipStream_t stream) {
  const dim3 grid((kMlpParams + kReduceParams - 1) / kReduceParams), block(64 * kReduceWaves);
  if (!adam) {
    hipLaunchKernelGGL(k_mlp_reduce<false>, grid, block, 0, stream, m, batch, lossOut, AdamArgs{});
    return hipGetLastError();
  }
  const double bc1 = 1.0 - std::pow((double)adam->beta1, (double)adam->step);
  const double bc2 = 1.0 - std::pow((double)adam->beta2, (double)adam->step);
  const AdamArgs a{adam->lr, adam->beta1, adam->beta2, adam->eps, (float)bc1, (float)bc2};
  hipLaunchKernelGGL(k_mlp_reduce<true>, grid, block, 0, stream, m, batch, lossOut, a);
  return hipGetLastError();
}

hipError_t LaunchMlpGradient(const MlpDev& m, const float* raw, const float* targets, const MlpBatch& batch,
                             float* lossOut, hipStream_t stream) {
  hipError_t e = LaunchMlpGradientTiles(m, raw, targets, batch, stream);
  if (e != hipSuccess) return e;
  return LaunchMlpReduce(m, batch, lossOut, nullptr, stream);
}

hipError_t LaunchMlpRefreshBf16(const MlpDev& m, hipStream_t stream) {
  hipLaunchKernelGGL(k_mlp_refresh, dim3((kMlpParams + 255) / 256), dim3(256), 0, stream, m.w, m.wb, m.wtb, m.wbi);
  return hipGetLastError();
}

}  // namespace gpugi
