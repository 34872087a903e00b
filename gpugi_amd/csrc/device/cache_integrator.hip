// Path tracer with the radiance-cache MLP (no reference counterpart).
//
// Each pixel traces the reference path tracer's estimator (pathtracer.comp:
// same random-number order, NEE, emission only after a vertex without NEE,
// Russian roulette) until it queries the cache, whose prediction, weighted by
// the path throughput, replaces the rest of the path.  Two query rules:
//
// GPUGI_CACHE_SPREAD: once the path's footprint has spread; with
//   a0 = sqrt(|x1 - x0|^2 / (4 pi |cos theta_1|))           at the primary hit,
//   a  = sum_{i>=2} sqrt(|x_i - x_{i-1}|^2 / (pdf_i |cos theta_i|)),
// the path stops at the first vertex i >= 2 (second hit on) where a > c * a0
// (Mueller et al. 2021, "Real-time Neural Radiance Caching for Path
// Tracing"); the cache predicts the radiance that vertex reflects (its NEE
// included).
// GPUGI_CACHE_PRIMARY_INDIRECT: at the primary hit, after its emission and
// NEE; the cache predicts the radiance the vertex reflects of light arriving
// by its BSDF-sampled bounce (everything the path would collect from the next
// vertex on), so direct light stays an exact estimate and the bounce noise
// goes away.
//
// One pixel in every `trainStride` also traces a training path: the same path
// continued `extraBounces` vertices past the query vertex, where the cache
// stands in for the remainder (self-training; with PRIMARY_INDIRECT the tail
// vertex does its NEE first).  Every vertex i of a training path yields a
// sample whose target is what the cache predicts there:
// (C + T_tail * L_tail - S_i) / T_i, with C the path's collected radiance, T_i
// the throughput at x_i and S_i what was collected before the predicted part
// (up to x_i's emission for SPREAD, up to x_i's NEE for PRIMARY_INDIRECT).
// Records live at fixed per-pixel / per-path slots, so a frame is
// deterministic.
//
// With PRIMARY_INDIRECT every pixel's path is only the primary hit, so it
// runs as wavefront passes over all pixels (ray generation, the persistent
// closest-hit trace, shading with the NEE setup, the persistent shadow trace,
// the sum), with the arithmetic of the one-kernel path; the training paths
// alone run in k_cached_paths, one thread per path.
#include <hip/hip_runtime.h>

#include "integrator_common.h"
#include "mlp.h"

namespace gpugi {

namespace {

constexpr int kTrainVertFloats = 24;  // features[16], T[4], S[4]

__device__ __forceinline__ void write_features(float* dst, const CacheParams& cp, gg::V3 x, gg::V3 wo, gg::V3 n,
                                               const MaterialData& m) {
  float4* d4 = reinterpret_cast<float4*>(dst);
  const float px = gg::saturate((x.x - cp.bboxMin[0]) * cp.bboxInvExtent[0]);
  const float py = gg::saturate((x.y - cp.bboxMin[1]) * cp.bboxInvExtent[1]);
  const float pz = gg::saturate((x.z - cp.bboxMin[2]) * cp.bboxInvExtent[2]);
  d4[0] = make_float4(px, py, pz, wo.x);
  d4[1] = make_float4(wo.y, wo.z, n.x, n.y);
  d4[2] = make_float4(n.z, m.diffuse.x, m.diffuse.y, m.diffuse.z);
  d4[3] = make_float4(log2f(1.0f + m.refl[3]) * (1.0f / 16.0f), 0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ void zero_features(float* dst) {
  float4* d4 = reinterpret_cast<float4*>(dst);
  for (int k = 0; k < 4; ++k) d4[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// TRAIN == false: the pixels' paths up to their query vertex (SPREAD; an 8x8
// tile per block), writing pixelColor and the queries.  TRAIN == true: the
// training paths (a 1-D grid, one thread per path), writing only the training
// records; a training path is its pixel's path continued, so both passes
// compute the same prefix.
// Five waves per SIMD, as the plain path tracer.
template <bool WIDE, bool TRAIN>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_cached_paths(DrawParams p,
                                                                                                    CacheParams cp) {
  __shared__ uint2 stk[WIDE ? kWideLdsStack * kTileThreads : 1];
  const int tid = threadIdx.x;
  const bool atPrimary = cp.query == GPUGI_CACHE_PRIMARY_INDIRECT;
  int gx, gy;
  uint32_t linear, pix;
  if (TRAIN) {
    linear = blockIdx.x * kTileThreads + tid;
    pix = linear * cp.trainStride + cp.trainOffset;
    if (pix >= (uint32_t)p.width * (uint32_t)p.height) return;
    gx = (int)(pix % (uint32_t)p.width);
    gy = (int)(pix / (uint32_t)p.width);
  } else {
    gx = blockIdx.x * kTile + (tid % kTile);
    gy = blockIdx.y * kTile + (tid / kTile);
    linear = (blockIdx.y * gridDim.x + blockIdx.x) * kTileThreads + tid;
    pix = (uint32_t)gy * (uint32_t)p.width + (uint32_t)gx;
  }
  TraceCtx c{&p, stk, tid, WIDE ? p.wide.spill + (size_t)linear * p.wide.spillDepth : nullptr};
  const bool isTrain = TRAIN;
  const uint32_t trainIdx = pix / cp.trainStride;
  float* tv = cp.trainVert + (size_t)trainIdx * p.maxPathLength * kTrainVertFloats;

  const float lightSampleWeight = (float)p.numLights / 1.0f;
  uint32_t seed = gg::init_random_seed(p.frameSeed, pix);
  gg::V3 color = gg::v3s(0.0f), cacc = gg::v3s(0.0f), throughput = gg::v3s(1.0f), tailT = gg::v3s(0.0f);
  bool didNEE = false, term = false, tail = false;
  int termDepth = 0;
  uint32_t nTrain = 0;
  float a0 = 0.0f, spread = 0.0f, pdfPrev = 1.0f;
  const float j0 = gg::random01(seed);
  const float j1 = gg::random01(seed);
  const gg::V3 U = gg::v3(p.camera.u[0], p.camera.u[1], p.camera.u[2]);
  const gg::V3 V = gg::v3(p.camera.v[0], p.camera.v[1], p.camera.v[2]);
  const gg::V3 W = gg::v3(p.camera.w[0], p.camera.w[1], p.camera.w[2]);
  gg::V3 o = gg::v3(p.camera.position[0], p.camera.position[1], p.camera.position[2]);
  gg::V3 d = gg::primary_direction(j0, j1, gx, gy, p.width, p.height, U, V, W);
  for (int i = 0; i < p.maxPathLength; ++i) {
    const ClosestHit h = trace_closest<WIDE>(c, o, d);
    if (!h.hit) {  // the environment, as the path tracer adds it (pathtracer.comp:56-58)
      const gg::V3 e = throughput * (p.env ? env_radiance(p.env, p.envSize, d) : gg::v3s(0.0f));
      cacc = cacc + e;
      if (!term) color = color + e;
      break;
    }
    o = o + h.t * d;
    const gg::V3 n = shading_normal(p.ref.vinfo, h.tri, h.bary);
    const MaterialData m = load_material_hit(p.materials, p.tex, p.ref.vinfo, h.tri, h.bary);
    const float cosIn = fmaxf(fabsf(gg::dot(n, d)), 1e-4f);
    const float dist2 = h.t * h.t;
    if (i == 0) a0 = sqrtf(dist2 / (4.0f * gg::kPI * cosIn));
    else spread += sqrtf(dist2 / (fmaxf(pdfPrev, 1e-8f) * cosIn));
    const gg::V3 te = throughput * (didNEE ? gg::v3s(0.0f) : m.emissivity);
    if (!atPrimary && !term && i >= 1 && spread > cp.spreadC * a0) {
      color = color + te;
      if (!TRAIN) {
        write_features(cp.queryFeat + (size_t)pix * kMlpRaw, cp, o, -d, n, m);
        reinterpret_cast<float4*>(cp.queryWeight)[pix] = make_float4(throughput.x, throughput.y, throughput.z, 1.0f);
      }
      term = true;
      termDepth = i;
      if (!isTrain) break;
    }
    if (isTrain && term && i == termDepth + cp.extraBounces) {
      // the cache stands in for what this vertex reflects (of its bounce, after its NEE)
      cacc = cacc + te;
      if (atPrimary) {
        const bool nee = m.refl[3] < 500.0f;
        const int sampleIndex = (int)(gg::xorshift32(seed) % p.numLights);
        cacc = cacc + estimate_direct_light<WIDE>(c, o, n, sampleIndex, d, m, nee) * throughput * lightSampleWeight;
      }
      write_features(cp.tailFeat + (size_t)trainIdx * kMlpRaw, cp, o, -d, n, m);
      tailT = throughput;
      tail = true;
      break;
    }
    float* r = nullptr;
    if (isTrain) {
      r = tv + (size_t)nTrain * kTrainVertFloats;
      write_features(r, cp, o, -d, n, m);
      const gg::V3 s = cacc + te;
      reinterpret_cast<float4*>(r)[4] = make_float4(throughput.x, throughput.y, throughput.z, 0.0f);
      reinterpret_cast<float4*>(r)[5] = make_float4(s.x, s.y, s.z, 0.0f);
      ++nTrain;
    }
    cacc = cacc + te;
    if (!term) color = color + te;
    didNEE = m.refl[3] < 500.0f;
    const int sampleIndex = (int)(gg::xorshift32(seed) % p.numLights);
    const gg::V3 radiance = estimate_direct_light<WIDE>(c, o, n, sampleIndex, d, m, didNEE);
    const gg::V3 contrib = radiance * throughput * lightSampleWeight;
    cacc = cacc + contrib;
    if (!term) color = color + contrib;
    if (atPrimary) {
      if (r) reinterpret_cast<float4*>(r)[5] = make_float4(cacc.x, cacc.y, cacc.z, 0.0f);
      if (i == 0) {  // the cache stands in for everything the bounce from x1 would add
        term = true;
        termDepth = 0;
      }
    }
    gg::V3 tp = gg::v3s(1.0f);
    float pdf;
    bool diffuse;
    d = sample_bsdf(d, m, seed, n, tp, false, pdf, diffuse);
    pdfPrev = pdf;
    o = o + d * gg::kRayHitEpsilon;
    throughput = throughput * tp;
    if (i >= p.rrStartDepth) {  // Russian roulette (pathtracer.comp:99-104) from rrStartDepth on
      const float q = gg::saturate(gg::luminance(tp));
      if (gg::random01(seed) >= q) break;
      throughput = throughput / q;
    }
  }
  if (!TRAIN) {
    reinterpret_cast<float4*>(cp.pixelColor)[pix] = make_float4(color.x, color.y, color.z, 0.0f);
    if (!term) {
      reinterpret_cast<float4*>(cp.queryWeight)[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      zero_features(cp.queryFeat + (size_t)pix * kMlpRaw);
    }
  } else {
    cp.trainCount[trainIdx] = nTrain;
    float4* tp4 = reinterpret_cast<float4*>(cp.trainPath + (size_t)trainIdx * 8);
    tp4[0] = make_float4(cacc.x, cacc.y, cacc.z, 0.0f);
    tp4[1] = make_float4(tailT.x, tailT.y, tailT.z, tail ? 1.0f : 0.0f);
    if (!tail) zero_features(cp.tailFeat + (size_t)trainIdx * kMlpRaw);
  }
}

// ---- PRIMARY_INDIRECT wavefront passes: one thread per ray r, every buffer but
// the per-pixel ones indexed by r.  Rays go in runs of 64 consecutive pixels
// (coherent rays, coalesced pixel writes); run c of the ray array is pixel run
//   (c mod Q) * 8 + c / Q,  Q = W*H / 512,
// so each eighth of the ray array, which the trace kernel hands to one XCD,
// samples the whole image (every 8th run) and the XCDs get equal work
// (identity when W*H is not a multiple of 512).
__device__ __forceinline__ uint32_t wave_pixel(uint32_t r, uint32_t P) {
  if (P & 511u) return r;
  const uint32_t Q = P / 512u, c = r / 64u;
  return ((c % Q) * 8u + c / Q) * 64u + (r % 64u);
}

// Primary rays {origin, tmax}{dir, 0} and each pixel's random state after them.
__global__ void k_pi_raygen(DrawParams p, float4* __restrict__ rays, uint32_t* __restrict__ seeds) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t P = (uint32_t)p.width * (uint32_t)p.height;
  if (r >= P) return;
  const uint32_t pix = wave_pixel(r, P);
  const int gx = (int)(pix % (uint32_t)p.width), gy = (int)(pix / (uint32_t)p.width);
  uint32_t seed = gg::init_random_seed(p.frameSeed, pix);
  const float j0 = gg::random01(seed);
  const float j1 = gg::random01(seed);
  const gg::V3 U = gg::v3(p.camera.u[0], p.camera.u[1], p.camera.u[2]);
  const gg::V3 V = gg::v3(p.camera.v[0], p.camera.v[1], p.camera.v[2]);
  const gg::V3 W = gg::v3(p.camera.w[0], p.camera.w[1], p.camera.w[2]);
  const gg::V3 d = gg::primary_direction(j0, j1, gx, gy, p.width, p.height, U, V, W);
  rays[2 * (size_t)r] = make_float4(p.camera.position[0], p.camera.position[1], p.camera.position[2], gg::kRayMax);
  rays[2 * (size_t)r + 1] = make_float4(d.x, d.y, d.z, 0.0f);
  seeds[r] = seed;
}

// Primary hit: emission into pixelColor, the cache query, the NEE shadow ray
// (in place of the primary ray; tmax 0 when there is nothing to connect) and
// its unoccluded contribution.
__global__ void k_pi_shade(DrawParams p, CacheParams cp, float4* __restrict__ rays, const int4* __restrict__ hits,
                           const uint32_t* __restrict__ seeds, float4* __restrict__ nee) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t P = (uint32_t)p.width * (uint32_t)p.height;
  if (r >= P) return;
  const uint32_t pix = wave_pixel(r, P);
  const float4 r0 = rays[2 * (size_t)r], r1 = rays[2 * (size_t)r + 1];
  const int4 rec = hits[r];
  const ClosestHit h = closest_from_record(p, rec.x, __int_as_float(rec.y), __int_as_float(rec.z),
                                           __int_as_float(rec.w));
  const float lightSampleWeight = (float)p.numLights / 1.0f;
  const gg::V3 throughput = gg::v3s(1.0f);
  gg::V3 color = gg::v3s(0.0f), val = gg::v3s(0.0f);
  float test = 0.0f;
  float4 so = make_float4(0.0f, 0.0f, 0.0f, 0.0f), sd = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
  if (h.hit) {
    const gg::V3 d = gg::v3(r1.x, r1.y, r1.z);
    const gg::V3 o = gg::v3(r0.x, r0.y, r0.z) + h.t * d;
    const gg::V3 n = shading_normal(p.ref.vinfo, h.tri, h.bary);
    const MaterialData m = load_material_hit(p.materials, p.tex, p.ref.vinfo, h.tri, h.bary);
    color = color + throughput * m.emissivity;  // the path's first vertex follows the camera, no NEE before it
    const bool didNEE = m.refl[3] < 500.0f;
    uint32_t seed = seeds[r];
    const int sampleIndex = (int)(gg::xorshift32(seed) % p.numLights);
    const NeeSample s = nee_sample(p, o, n, sampleIndex, d, m, didNEE);
    if (s.test) {
      so = make_float4(s.o.x, s.o.y, s.o.z, s.len);
      sd = make_float4(s.dir.x, s.dir.y, s.dir.z, 0.0f);
      val = s.value * throughput * lightSampleWeight;
      test = 1.0f;
    }
    write_features(cp.queryFeat + (size_t)pix * kMlpRaw, cp, o, -d, n, m);
    reinterpret_cast<float4*>(cp.queryWeight)[pix] = make_float4(throughput.x, throughput.y, throughput.z, 1.0f);
  } else {
    const gg::V3 d = gg::v3(r1.x, r1.y, r1.z);
    color = color + throughput * (p.env ? env_radiance(p.env, p.envSize, d) : gg::v3s(0.0f));
    zero_features(cp.queryFeat + (size_t)pix * kMlpRaw);
    reinterpret_cast<float4*>(cp.queryWeight)[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  rays[2 * (size_t)r] = so;
  rays[2 * (size_t)r + 1] = sd;
  nee[r] = make_float4(val.x, val.y, val.z, test);
  reinterpret_cast<float4*>(cp.pixelColor)[pix] = make_float4(color.x, color.y, color.z, 0.0f);
}

// pixelColor += the NEE contribution (zero when there was nothing to connect
// or the shadow ray is occluded, as the one-kernel path adds it).
__global__ void k_pi_finish(DrawParams p, CacheParams cp, const int4* __restrict__ occluded,
                            const float4* __restrict__ nee) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t P = (uint32_t)p.width * (uint32_t)p.height;
  if (r >= P) return;
  float4 v = nee[r];
  if (v.w == 0.0f || occluded[r].x != 0) v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float4* c = reinterpret_cast<float4*>(cp.pixelColor) + wave_pixel(r, P);
  const float4 col = *c;
  *c = make_float4(col.x + v.x, col.y + v.y, col.z + v.z, col.w);
}

__global__ __launch_bounds__(kTileThreads) void k_cache_resolve(DrawParams p, CacheParams cp) {
  const int tid = threadIdx.x;
  const int gx = blockIdx.x * kTile + (tid % kTile);
  const int gy = blockIdx.y * kTile + (tid / kTile);
  const size_t pix = (size_t)gy * p.width + gx;
  float4 col = reinterpret_cast<const float4*>(cp.pixelColor)[pix];
  const float4 w = reinterpret_cast<const float4*>(cp.queryWeight)[pix];
  if (w.w > 0.5f) {
    const float4 y = reinterpret_cast<const float4*>(cp.queryPred)[pix];
    col.x += w.x * fmaxf(y.x, 0.0f);
    col.y += w.y * fmaxf(y.y, 0.0f);
    col.z += w.z * fmaxf(y.z, 0.0f);
  }
  float4* px = &p.accum[pix];
  const float4 v = *px;
  *px = make_float4(col.x + v.x, col.y + v.y, col.z + v.z, 0.0f + v.w);
}

__device__ __forceinline__ float target_channel(float C, float Tend, float pred, float S, float T) {
  if (!(T > 1e-6f)) return 0.0f;
  const float L = (C + Tend * pred - S) / T;
  return fminf(fmaxf(L, 0.0f), 1e4f);
}

// When the frame yields more training vertices (T) than the cache trains on
// (M), vertex `off` of the raster-ordered list is kept iff it is the first to
// map to floor(off * M / T), which spreads the M samples evenly over the
// image.  The kept sample k then goes to (k % steps) * (M / steps) + k / steps,
// so every training step sees the whole image rather than a band of it.
__global__ void k_cache_targets(CacheParams cp, uint32_t numPaths, int maxLen, const uint32_t* __restrict__ offsets,
                                const uint32_t* __restrict__ total, uint32_t maxSamples, uint32_t steps,
                                float* __restrict__ feat, float* __restrict__ targets) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t j = (uint32_t)(t / (uint32_t)maxLen);
  const uint32_t i = (uint32_t)(t % (uint32_t)maxLen);
  if (j >= numPaths || i >= cp.trainCount[j]) return;
  const uint64_t totalV = *total;
  const uint64_t M = maxSamples;
  const uint64_t rank = offsets[j] + i;
  uint64_t k = rank;
  if (totalV > M) {
    if ((rank * M) % totalV >= M) return;
    k = rank * M / totalV;
  }
  const uint64_t n = totalV < M ? totalV : M;
  const uint64_t per = steps > 0 ? n / steps : n;
  if (steps > 1 && per > 0) {
    if (k >= per * steps) return;  // the remainder does not fill a whole step
    k = (k % steps) * per + k / steps;
  }
  const uint32_t off = (uint32_t)k;
  const float* r = cp.trainVert + ((size_t)j * maxLen + i) * kTrainVertFloats;
  for (int k = 0; k < 4; ++k)
    reinterpret_cast<float4*>(feat + (size_t)off * kMlpRaw)[k] = reinterpret_cast<const float4*>(r)[k];
  const float4 T = reinterpret_cast<const float4*>(r)[4];
  const float4 S = reinterpret_cast<const float4*>(r)[5];
  const float4 C = reinterpret_cast<const float4*>(cp.trainPath + (size_t)j * 8)[0];
  const float4 Te = reinterpret_cast<const float4*>(cp.trainPath + (size_t)j * 8)[1];
  float4 pr = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (Te.w > 0.5f) {
    pr = reinterpret_cast<const float4*>(cp.tailPred)[j];
    pr.x = fmaxf(pr.x, 0.0f);
    pr.y = fmaxf(pr.y, 0.0f);
    pr.z = fmaxf(pr.z, 0.0f);
  }
  reinterpret_cast<float4*>(targets)[off] =
      make_float4(target_channel(C.x, Te.x, pr.x, S.x, T.x), target_channel(C.y, Te.y, pr.y, S.y, T.y),
                  target_channel(C.z, Te.z, pr.z, S.z, T.z), 0.0f);
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

hipError_t LaunchFillU32(uint32_t* p, uint32_t value, size_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, value, n);
  return hipGetLastError();
}

hipError_t LaunchCachedPaths(const DrawParams& p, const CacheParams& cp, void* spill, bool train,
                             hipStream_t stream) {
  dim3 grid(p.width / kTile, p.height / kTile);
  if (grid.x == 0 || grid.y == 0) return hipSuccess;
  if (cp.trainStride == 0 || cp.trainOffset >= cp.trainStride) return hipErrorInvalidValue;
  if (!train && cp.query == GPUGI_CACHE_PRIMARY_INDIRECT) return hipErrorInvalidValue;  // wavefront passes instead
  if (train) grid = dim3((CacheTrainPaths(p.width, p.height, cp.trainStride) + kTileThreads - 1) / kTileThreads);
  DrawParams q = p;
  if (p.useWide) q.wide.spill = (uint2*)spill;
  if (p.useWide && train) hipLaunchKernelGGL((k_cached_paths<true, true>), grid, dim3(kTileThreads), 0, stream, q, cp);
  else if (p.useWide) hipLaunchKernelGGL((k_cached_paths<true, false>), grid, dim3(kTileThreads), 0, stream, q, cp);
  else if (train) hipLaunchKernelGGL((k_cached_paths<false, true>), grid, dim3(kTileThreads), 0, stream, q, cp);
  else hipLaunchKernelGGL((k_cached_paths<false, false>), grid, dim3(kTileThreads), 0, stream, q, cp);
  return hipGetLastError();
}

hipError_t LaunchPrimaryRaygen(const DrawParams& p, float4* rays, uint32_t* seeds, hipStream_t stream) {
  const size_t P = (size_t)p.width * p.height;
  if (P == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pi_raygen, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, stream, p, rays, seeds);
  return hipGetLastError();
}

hipError_t LaunchPrimaryShade(const DrawParams& p, const CacheParams& cp, float4* rays, const int4* hits,
                              const uint32_t* seeds, float4* nee, hipStream_t stream) {
  const size_t P = (size_t)p.width * p.height;
  if (P == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pi_shade, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, stream, p, cp, rays, hits, seeds,
                     nee);
  return hipGetLastError();
}

hipError_t LaunchPrimaryFinish(const DrawParams& p, const CacheParams& cp, const int4* occluded, const float4* nee,
                               hipStream_t stream) {
  const size_t P = (size_t)p.width * p.height;
  if (P == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pi_finish, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, stream, p, cp, occluded, nee);
  return hipGetLastError();
}

hipError_t LaunchCacheResolve(const DrawParams& p, const CacheParams& cp, hipStream_t stream) {
  dim3 grid(p.width / kTile, p.height / kTile);
  if (grid.x == 0 || grid.y == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cache_resolve, grid, dim3(kTileThreads), 0, stream, p, cp);
  return hipGetLastError();
}

hipError_t LaunchCacheTargets(const DrawParams& p, const CacheParams& cp, const uint32_t* offsets,
                              const uint32_t* total, uint32_t maxSamples, uint32_t steps, float* feat, float* targets,
                              hipStream_t stream) {
  const uint32_t paths = CacheTrainPaths(p.width, p.height, cp.trainStride);
  const size_t threads = (size_t)paths * p.maxPathLength;
  if (threads == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cache_targets, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, stream, cp, paths,
                     p.maxPathLength, offsets, total, maxSamples, steps, feat, targets);
  return hipGetLastError();
}

}  // namespace gpugi
