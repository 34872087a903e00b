// Light tracer (lighttracer.comp + showlightcaches.comp) and bidirectional path
// tracer (lighttracer.comp with SAVE_LIGHT_CACHE + pathtracer_bidir.comp).
//
// Compiled with -ffp-contract=off; every kernel keeps the random-number draw
// order of the GLSL it restates.  Camera splats use global float atomics
// (the reference's imageAtomicAdd on an r32f alias of the backbuffer,
// screenreproject.glsl:238-241), so per-pixel sums are exact up to the
// order of the additions.
//
// The BPT light cache is appended deterministically: every light path stages
// its stored vertices in a private slice, an exclusive scan over the per-path
// counts assigns each path a contiguous range, and a scatter writes the cache
// in (path, bounce) order behind the R*N initial vertices.  This is one of the
// orders the reference's atomicAdd append (lighttracer.comp:120) can produce,
// fixed so that renders are reproducible and comparable with a CPU
// restatement.
#include <hip/hip_runtime.h>

#include "bpt_eye.h"
#include "integrator_common.h"
#include "scan.h"

namespace gpugi {

constexpr int kLightBlock = 64;  // local size of lighttracer.comp:48
static_assert(kLightBlock == kTileThreads, "the wide traversal stack is laid out for kTileThreads lanes");

// Float add into the accumulation buffer (device memory): a global atomic, so
// the address needs no test for the LDS or scratch aperture a flat atomic does.
__device__ __forceinline__ void splat_add(float* p, float v) {
  (void)__builtin_amdgcn_global_atomic_fadd_f32((__attribute__((address_space(1))) float*)p, v);
}

// ProjectToScreen without DEPTH_BUFFER_OCCLUSION_TEST / CHAINED_MIS and with
// the box reconstruction filter (screenreproject.glsl:55-227).
template <bool WIDE>
__device__ __forceinline__ void project_to_screen(const TraceCtx& c, gg::V3 rayO, gg::V3 rayD, gg::V3 gN,
                                                  gg::V3 sN, const MaterialData& m, gg::V3 tp, bool computeMIS) {
  const DrawParams& p = *c.p;
  const gg::V3 U = gg::v3(p.camera.u[0], p.camera.u[1], p.camera.u[2]);
  const gg::V3 V = gg::v3(p.camera.v[0], p.camera.v[1], p.camera.v[2]);
  const gg::V3 W = gg::v3(p.camera.w[0], p.camera.w[1], p.camera.w[2]);
  const gg::V3 cp = gg::v3(p.camera.position[0], p.camera.position[1], p.camera.position[2]);
  gg::V3 crd = cp - rayO;
  const float camViewDotRay = -gg::dot(W, crd);
  if (!(camViewDotRay > 0.0f)) return;
  float sx = -gg::dot(U, crd) / camViewDotRay;
  float sy = -gg::dot(V, crd) / camViewDotRay;
  sx = sx / gg::dot(U, U);
  sy = sy / gg::dot(V, V);
  if (!(sx >= -1.0f && sx < 1.0f && sy > -1.0f && sy <= 1.0f)) return;
  const float cameraDistSq = gg::dot(crd, crd);
  const float cameraDist = sqrtf(cameraDistSq);
  crd = crd / cameraDist;
  const gg::V3 co = crd * gg::kRayHitEpsilon + rayO;
  const float ppx = (sx * 0.5f + 0.5f) * (float)p.width;
  const float ppy = (sy * 0.5f + 0.5f) * (float)p.height;
  const int ix = (int)roundf(ppx), iy = (int)roundf(ppy);
  const float cosAtCamera = gg::saturate(camViewDotRay / cameraDist);
  const float f2i = adjoint_corrected_out_dot_n(rayD, crd, gN, sN) /
                    (cosAtCamera * cosAtCamera * cosAtCamera * p.camera.pixel_area * cameraDistSq);
  float pdf;
  gg::V3 bsdf = bsdf_adjoint(rayD, crd, m, sN, pdf);
  if (computeMIS) {
    const float eyeToLight = 1.0f;  // InitialMIS (lightcache.glsl:76-79)
    const float lightToEye = mis_heuristic(pdf);
    const float mis = eyeToLight / fmaxf(lightToEye + eyeToLight, gg::kDivisorEpsilon);
    bsdf = bsdf * mis;
  }
  const gg::V3 color = tp * f2i * bsdf;
  // Splats outside the image are dropped, as image atomics outside the image
  // are, and a splat of exactly zero changes nothing (the accumulated sums
  // are never -0): neither needs the camera ray, which is traced last.
  if (ix < 0 || ix >= p.width || iy < 0 || iy >= p.height) return;
  if (color.x == 0.0f && color.y == 0.0f && color.z == 0.0f) return;
  if (trace_any<WIDE>(c, co, crd, cameraDist)) return;
  float* px = reinterpret_cast<float*>(&p.accum[(size_t)iy * p.width + ix]);
  splat_add(px + 0, color.x);
  splat_add(px + 1, color.y);
  splat_add(px + 2, color.z);
}

enum LightMode { kModeLT = 0, kModeBptLight = 1, kModeBptWarmup = 2 };

__device__ __forceinline__ LightCacheEntryDev make_entry(gg::V3 flux, int mat, gg::V3 pos, float n0, float n1,
                                                         float u, float v, gg::V3 incident, int len) {
  LightCacheEntryDev e;
  e.flux[0] = flux.x; e.flux[1] = flux.y; e.flux[2] = flux.z;
  e.materialIndex = mat;
  e.position[0] = pos.x; e.position[1] = pos.y; e.position[2] = pos.z;
  e.normal0 = n0;
  e.texcoord[0] = u; e.texcoord[1] = v;
  e.normal1 = n1;
  e.pathLength = len;
  gg::pack_normal(incident, e.incidentPacked[0], e.incidentPacked[1]);
  e.chainedMisN = 0.0f;
  e.chainedMisT = 0.0f;
  return e;
}

__device__ __forceinline__ void store_entry(LightCacheEntryDev* dst, const LightCacheEntryDev& e) {
  const float4* s = reinterpret_cast<const float4*>(&e);
  float4* d = reinterpret_cast<float4*>(dst);
  d[0] = s[0]; d[1] = s[1]; d[2] = s[2]; d[3] = s[3];
}

// lighttracer.comp:49-193.  MODE selects LT, BPT light pass or BPT warm-up.
// One light path as a state that advances one bounce at a time.
struct LtPath {
  gg::V3 tp, o, d;
  uint32_t seed, gid, stored;
  int bounce;
};

template <int MODE>
__device__ __forceinline__ void lt_begin(const DrawParams& p, LtPath& s, uint32_t gid, int raysPerSample,
                                         LightCacheEntryDev* __restrict__ cache) {
  s.gid = gid;
  s.seed = gg::init_random_seed(p.frameSeed, gid);
  const int lsIndex = (int)(gid / (uint32_t)raysPerSample);
  const gpugi_light_sample& ls = p.lights[lsIndex];
  const gg::V3 lpos = gg::v3(ls.position[0], ls.position[1], ls.position[2]);
  const gg::V3 lint = gg::v3(ls.intensity[0], ls.intensity[1], ls.intensity[2]);
  if (MODE == kModeBptLight && (int)gid < p.bptCapacity) {
    // Initial light vertex of this path (lighttracer.comp:56-61).
    const gg::V3 flux = (lint / gg::kPI_2) / (float)raysPerSample;
    store_entry(&cache[gid], make_entry(flux, -1, lpos, ls.normal_phi, ls.normal_theta_cos, 0.0f, 0.0f,
                                        gg::v3s(0.0f), 0));
  }
  s.tp = lint * (gg::kPI / (float)raysPerSample);
  s.o = lpos;
  if (ls.normal_theta_cos > 1.5f) {
    const float r0 = gg::random01(s.seed);
    const float r1 = gg::random01(s.seed);
    s.d = gg::sample_direction(r0, r1);
    s.tp = s.tp * 4.0f;
  } else {
    const gg::V3 n = gg::unpack_normal(ls.normal_phi, ls.normal_theta_cos);
    gg::V3 U, V;
    gg::create_onb(n, U, V);
    const float r0 = gg::random01(s.seed);
    const float r1 = gg::random01(s.seed);
    s.d = gg::sample_hemisphere_cosine(r0, r1, U, V, n);
  }
  s.stored = 0;
  s.bounce = 0;
}

// One bounce; true when the path has ended.
template <bool WIDE, int MODE>
__device__ __forceinline__ bool lt_bounce(const DrawParams& p, const TraceCtx& c, LtPath& s,
                                          LightCacheEntryDev* __restrict__ staging) {
  const int i = s.bounce;
  if (i >= p.maxPathLength) return true;
  const ClosestHit h = trace_closest<WIDE>(c, s.o, s.d);
  if (!h.hit) return true;
  s.o = s.o + h.t * s.d;
  const gg::V3 sN = shading_normal(p.ref.vinfo, h.tri, h.bary);
  const float2 uv = hit_texcoord(p.ref.vinfo, h.tri, h.bary);
  const MaterialData m = load_material_uv(p.materials, p.tex, h.tri.w, uv);
  if (MODE != kModeLT && m.refl[3] < 100.0f) {
    if (MODE == kModeBptLight) {
      float n0, n1;
      gg::pack_normal(sN, n0, n1);
      store_entry(&staging[(size_t)s.gid * p.maxPathLength + s.stored],
                  make_entry(s.tp, h.tri.w, s.o, n0, n1, uv.x, uv.y, s.d, i + 1));
    }
    ++s.stored;
  }
  // the camera connection closes a path of i + 1 bounces (lighttracer.comp:133-157)
  if (MODE != kModeBptWarmup && (p.specificPathLength < 0 || i + 1 == p.specificPathLength))
    project_to_screen<WIDE>(c, s.o, s.d, h.geoNormal, sN, m, s.tp, MODE == kModeBptLight);
  gg::V3 t1 = gg::v3s(1.0f);
  float pdf;
  bool diffuse;
  const gg::V3 nd = sample_bsdf(s.d, m, s.seed, sN, t1, true, pdf, diffuse);
  t1 = t1 * adjoint_correction(s.d, nd, h.geoNormal, sN);
  s.tp = s.tp * t1;
  if (i >= p.rrStartDepth) {  // Russian roulette (lighttracer.comp:183-188) from rrStartDepth on
    const float q = gg::saturate(gg::luminance(t1));
    if (gg::random01(s.seed) >= q) return true;
    s.tp = s.tp / q;
  }
  s.d = nd;
  s.o = s.o + s.d * gg::kRayHitEpsilon;
  ++s.bounce;
  return s.bounce >= p.maxPathLength;
}

// All numPaths light paths with path regeneration (persistent_items).  Five
// waves per SIMD (at most 96 VGPRs): the BPT light pass would otherwise take
// 129 VGPRs and 3 waves; the few registers spilled for it cost less than the
// waves gained (clutter room: LT -2 %, BPT -3.6 % with the eye pass below).
template <bool WIDE, int MODE>
__global__ __launch_bounds__(kLightBlock) __attribute__((amdgpu_waves_per_eu(5))) void k_lighttracer(
    DrawParams p, int raysPerSample, uint32_t numPaths, LightCacheEntryDev* __restrict__ cache,
    LightCacheEntryDev* __restrict__ staging, uint32_t* __restrict__ pathCounts, uint32_t* __restrict__ queue) {
  __shared__ uint2 stk[WIDE ? kWideLdsStack * kLightBlock : 1];
  const int tid = threadIdx.x;
  const uint32_t linear = blockIdx.x * kLightBlock + tid;
  TraceCtx c{&p, stk, tid, WIDE ? p.wide.spill + (size_t)linear * p.wide.spillDepth : nullptr};
  LtPath s;
  persistent_items(
      numPaths, queue, [&](uint32_t gid) { lt_begin<MODE>(p, s, gid, raysPerSample, cache); },
      [&]() { return lt_bounce<WIDE, MODE>(p, c, s, staging); },
      [&]() {
        if (MODE != kModeLT) pathCounts[s.gid] = s.stored;
      });
}

// showlightcaches.comp:9-65 for exactly N light samples.
template <bool WIDE>
__global__ __launch_bounds__(kLightBlock) void k_showlightcaches(DrawParams p) {
  __shared__ uint2 stk[WIDE ? kWideLdsStack * kLightBlock : 1];
  const int tid = threadIdx.x;
  const uint32_t i = blockIdx.x * kLightBlock + tid;
  if (i >= p.numLights) return;
  TraceCtx c{&p, stk, tid, WIDE ? p.wide.spill + (size_t)i * p.wide.spillDepth : nullptr};
  const gpugi_light_sample& ls = p.lights[i];
  const gg::V3 lpos = gg::v3(ls.position[0], ls.position[1], ls.position[2]);
  const gg::V3 U = gg::v3(p.camera.u[0], p.camera.u[1], p.camera.u[2]);
  const gg::V3 V = gg::v3(p.camera.v[0], p.camera.v[1], p.camera.v[2]);
  const gg::V3 W = gg::v3(p.camera.w[0], p.camera.w[1], p.camera.w[2]);
  const gg::V3 cp = gg::v3(p.camera.position[0], p.camera.position[1], p.camera.position[2]);
  gg::V3 crd = cp - lpos;
  const float camViewDotRay = -gg::dot(W, crd);
  if (!(camViewDotRay > 0.0f)) return;
  float sx = -gg::dot(U, crd) / camViewDotRay;
  float sy = -gg::dot(V, crd) / camViewDotRay;
  sx = sx / gg::dot(U, U);
  sy = sy / gg::dot(V, V);
  if (!(sx >= -1.0f && sx < 1.0f && sy > -1.0f && sy <= 1.0f)) return;
  const float cameraDistSq = gg::dot(crd, crd);
  const float cameraDist = sqrtf(cameraDistSq);
  crd = crd / cameraDist;
  const gg::V3 co = crd * gg::kRayHitEpsilon + lpos;
  if (trace_any<WIDE>(c, co, crd, cameraDist)) return;
  const gg::V3 n = gg::unpack_normal(ls.normal_phi, ls.normal_theta_cos);
  const float cosTheta = gg::dot(n, crd);
  const float cosAtCamera = gg::saturate(camViewDotRay / cameraDist);
  const float i2r = cosTheta / (gg::kPI_2 * cosAtCamera * cosAtCamera * cosAtCamera * p.camera.pixel_area * cameraDistSq);
  const gg::V3 color = gg::v3(ls.intensity[0], ls.intensity[1], ls.intensity[2]) * i2r;
  const int ix = (int)((sx * 0.5f + 0.5f) * (float)p.width);
  const int iy = (int)((sy * 0.5f + 0.5f) * (float)p.height);
  if (ix < 0 || ix >= p.width || iy < 0 || iy >= p.height) return;
  // The reference does an unsynchronized load/store here (showlightcaches.comp:54-55);
  // atomics give the intended sum.
  float* px = reinterpret_cast<float*>(&p.accum[(size_t)iy * p.width + ix]);
  splat_add(px + 0, color.x);
  splat_add(px + 1, color.y);
  splat_add(px + 2, color.z);
}

// Scatter of the staged BPT vertices behind the initial ones.
__global__ __launch_bounds__(256) void k_bpt_scatter(const LightCacheEntryDev* __restrict__ staging,
                                                     const uint32_t* __restrict__ counts,
                                                     const uint32_t* __restrict__ offsets, uint32_t numPaths,
                                                     int maxLen, int capacity, LightCacheEntryDev* __restrict__ cache) {
  const uint32_t gid = blockIdx.x * 256u + threadIdx.x;
  if (gid >= numPaths) return;
  const uint32_t n = counts[gid];
  const uint64_t base = (uint64_t)numPaths + offsets[gid];
  for (uint32_t k = 0; k < n; ++k) {
    const uint64_t dst = base + k;
    if (dst < (uint64_t)capacity) {
      const float4* s = reinterpret_cast<const float4*>(&staging[(size_t)gid * maxLen + k]);
      float4* d = reinterpret_cast<float4*>(&cache[dst]);
      d[0] = s[0]; d[1] = s[1]; d[2] = s[2]; d[3] = s[3];
    }
  }
}

// One bounce of an eye path (bpt_eye.h); true when the path has ended.
// actualN: light-cache entries in use.
template <bool WIDE>
__device__ __forceinline__ bool bpt_eye_bounce(const DrawParams& p, const TraceCtx& c, BptEyePath& s,
                                               const LightCacheEntryDev* __restrict__ cache, uint32_t actualN) {
  if (s.bounce >= p.maxPathLength) return true;
  const float lightSampleWeight = (float)actualN;
  const ClosestHit h = trace_closest<WIDE>(c, s.o, s.d);
  if (!h.hit) return true;
  BptEyeVertex v;
  bpt_eye_hit(p, s, h, v);
  const BptConnection cn = bpt_connect_setup(p, s, cache, actualN);
  if (cn.test) {  // the value first: a zero one needs no ray (bpt_value_is_zero)
    const gg::V3 add = bpt_connect_value(p, s, v, cn, lightSampleWeight);
    if (!bpt_value_is_zero(add) && !trace_any<WIDE>(c, cn.lo, cn.ld, cn.tmax)) s.color = s.color + add;
  }
  return bpt_eye_continue(p, s, v);
}

// Eye paths of all pixels of the 8x8 tiles with path regeneration
// (persistent_items); each pixel is written by one lane.
template <bool WIDE>
// Five waves per SIMD, as k_lighttracer.
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_bpt_eye(DrawParams p, const LightCacheEntryDev* __restrict__ cache,
                                                          const uint32_t* __restrict__ numEntriesPtr,
                                                          uint32_t* __restrict__ queue) {
  __shared__ uint2 stk[WIDE ? kWideLdsStack * kTileThreads : 1];
  const int tid = threadIdx.x;
  const uint32_t linear = blockIdx.x * kTileThreads + tid;
  TraceCtx c{&p, stk, tid, WIDE ? p.wide.spill + (size_t)linear * p.wide.spillDepth : nullptr};
  const int numEntries = (int)*numEntriesPtr;
  const uint32_t actualN = (uint32_t)(numEntries < p.bptCapacity ? numEntries : p.bptCapacity);
  const uint32_t pixels = ((uint32_t)p.width / kTile) * ((uint32_t)p.height / kTile) * kTileThreads;
  int gx = 0, gy = 0;
  BptEyePath s;
  persistent_items(
      pixels, queue,
      [&](uint32_t item) {
        tile_pixel(item, p.width, gx, gy);
        bpt_eye_begin(p, s, gx, gy);
      },
      [&]() { return bpt_eye_bounce<WIDE>(p, c, s, cache, actualN); },
      [&]() {
        float4* px = &p.accum[(size_t)gy * p.width + gx];
        const float4 v = *px;
        *px = make_float4(s.color.x + v.x, s.color.y + v.y, s.color.z + v.z, 1.0f + v.w);
      });
}

// ---------------------------------------------------------------------------
static hipError_t PrepareSpill(IntegratorState* st, DrawParams* p, size_t threads) {
  if (!p->useWide) return hipSuccess;
  const size_t bytes = std::max<size_t>(8, threads * p->wide.spillDepth * sizeof(uint2));
  hipError_t e = EnsureBuffer(&st->spill, &st->spillBytes, bytes);
  if (e == hipSuccess) p->wide.spill = (uint2*)st->spill;
  return e;
}

// The light paths of whole 64-path work groups (the reference dispatches
// numPaths / 64 groups), one block per group.
template <int MODE>
static hipError_t LaunchLight(IntegratorState* st, const DrawParams& p, int R, uint32_t numPaths,
                              LightCacheEntryDev* cache, LightCacheEntryDev* staging, uint32_t* counts,
                              hipStream_t stream) {
  const uint32_t paths = numPaths / kLightBlock * kLightBlock;
  if (paths == 0) return hipSuccess;
  const void* fn = p.useWide ? reinterpret_cast<const void*>(&k_lighttracer<true, MODE>)
                             : reinterpret_cast<const void*>(&k_lighttracer<false, MODE>);
  uint32_t blocks = 0;
  if (hipError_t e = PersistentGrid(st, fn, paths, &blocks, stream, false)) return e;
  if (p.useWide)
    hipLaunchKernelGGL((k_lighttracer<true, MODE>), dim3(blocks), dim3(kLightBlock), 0, stream, p, R, paths, cache,
                       staging, counts, st->queue);
  else
    hipLaunchKernelGGL((k_lighttracer<false, MODE>), dim3(blocks), dim3(kLightBlock), 0, stream, p, R, paths, cache,
                       staging, counts, st->queue);
  return hipGetLastError();
}

hipError_t DrawLT(IntegratorState* st, const DrawParams& p0, hipStream_t stream) {
  DrawParams p = p0;
  const int R = RaysPerLightSample(p.width, p.height, p.numLights);
  const uint32_t numPaths = p.numLights * (uint32_t)R;
  hipError_t e = PrepareSpill(st, &p, std::max<size_t>(numPaths, p.numLights));
  if (e != hipSuccess) return e;
  e = LaunchLight<kModeLT>(st, p, R, numPaths, nullptr, nullptr, nullptr, stream);
  if (e != hipSuccess) return e;
  // LightPathtracer::Draw then projects the light samples themselves (lightpathtracer.cpp:50).
  dim3 grid((p.numLights + kLightBlock - 1) / kLightBlock);
  if (p.useWide) hipLaunchKernelGGL(k_showlightcaches<true>, grid, dim3(kLightBlock), 0, stream, p);
  else hipLaunchKernelGGL(k_showlightcaches<false>, grid, dim3(kLightBlock), 0, stream, p);
  return hipGetLastError();
}

hipError_t BptWarmup(IntegratorState* st, const DrawParams& p0, hipStream_t stream, BptWarmupCounts* out) {
  DrawParams p = p0;
  const int R = RaysPerLightSample(p.width, p.height, p.numLights);
  const uint32_t numPaths = p.numLights * (uint32_t)R;
  hipError_t e = PrepareSpill(st, &p, numPaths);
  if (e != hipSuccess) return e;
  e = EnsureBuffer(&st->pathCounts, &st->pathCountsBytes, (size_t)numPaths * 4);
  if (e != hipSuccess) return e;
  size_t scanTemp = ExclusiveScanTempBytes(numPaths);
  e = EnsureBuffer(&st->scratch, &st->scratchBytes, scanTemp + 256);
  if (e != hipSuccess) return e;
  uint32_t* counts = (uint32_t*)st->pathCounts;
  uint32_t* total = (uint32_t*)((char*)st->scratch + scanTemp);
  e = LaunchLight<kModeBptWarmup>(st, p, R, numPaths, nullptr, nullptr, counts, stream);
  if (e != hipSuccess) return e;
  e = LaunchExclusiveScanU32(counts, counts, numPaths, st->scratch, total, stream);
  if (e != hipSuccess) return e;
  uint32_t stored = 0;
  e = hipMemcpyAsync(&stored, total, 4, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return e;
  out->lightPathLengthSum = stored;
  out->numLightCacheEntries = (uint64_t)numPaths + stored;
  return hipSuccess;
}

hipError_t BptLightPass(IntegratorState* st, DrawParams* pp, hipStream_t stream, std::string* err,
                        const LightCacheEntryDev** cacheOut, const uint32_t** totalOut) {
  DrawParams& p = *pp;
  *cacheOut = nullptr;
  *totalOut = nullptr;
  if (p.bptCapacity < 0) {
    *err = "the light cache capacity is not set (run the warm-up first)";
    return hipSuccess;
  }
  const int R = RaysPerLightSample(p.width, p.height, p.numLights);
  const uint32_t numPaths = p.numLights * (uint32_t)R;
  const size_t eyeThreads = (size_t)(p.width / kTile) * (p.height / kTile) * kTileThreads;
  hipError_t e = PrepareSpill(st, &p, std::max<size_t>(numPaths, eyeThreads));
  if (e != hipSuccess) return e;
  e = EnsureBuffer(&st->lightCache, &st->lightCacheBytes, std::max<size_t>(64, (size_t)p.bptCapacity * 64));
  if (e != hipSuccess) return e;
  e = EnsureBuffer(&st->staging, &st->stagingBytes, (size_t)numPaths * p.maxPathLength * 64);
  if (e != hipSuccess) return e;
  e = EnsureBuffer(&st->pathCounts, &st->pathCountsBytes, (size_t)numPaths * 8);
  if (e != hipSuccess) return e;
  const size_t scanTemp = ExclusiveScanTempBytes(numPaths);
  e = EnsureBuffer(&st->scratch, &st->scratchBytes, scanTemp + 256);
  if (e != hipSuccess) return e;
  LightCacheEntryDev* cache = (LightCacheEntryDev*)st->lightCache;
  LightCacheEntryDev* staging = (LightCacheEntryDev*)st->staging;
  uint32_t* counts = (uint32_t*)st->pathCounts;
  uint32_t* offsets = counts + numPaths;
  uint32_t* total = (uint32_t*)((char*)st->scratch + scanTemp);
  e = LaunchLight<kModeBptLight>(st, p, R, numPaths, cache, staging, counts, stream);
  if (e != hipSuccess) return e;
  e = LaunchExclusiveScanU32(counts, offsets, numPaths, st->scratch, total, stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bpt_scatter, dim3((numPaths + 255) / 256), dim3(256), 0, stream, staging, counts, offsets,
                     numPaths, p.maxPathLength, p.bptCapacity, cache);
  // NumLightCacheEntries = R*N + appended (resetcounter.comp:9 plus the appends)
  AddConstantU32(total, numPaths, stream);
  st->lastTotal = total;
  *cacheOut = cache;
  *totalOut = total;
  return hipGetLastError();
}

hipError_t DrawBPT(IntegratorState* st, const DrawParams& p0, hipStream_t stream, std::string* err) {
  DrawParams p = p0;
  const LightCacheEntryDev* cache = nullptr;
  const uint32_t* total = nullptr;
  hipError_t e = BptLightPass(st, &p, stream, err, &cache, &total);
  if (e != hipSuccess || !err->empty()) return e;
  const size_t pixels = (size_t)(p.width / kTile) * (p.height / kTile) * kTileThreads;
  if (pixels) {
    const void* fn = p.useWide ? reinterpret_cast<const void*>(&k_bpt_eye<true>)
                               : reinterpret_cast<const void*>(&k_bpt_eye<false>);
    uint32_t blocks = 0;
    if (hipError_t e2 = PersistentGrid(st, fn, pixels, &blocks, stream)) return e2;
    if (p.useWide)
      hipLaunchKernelGGL(k_bpt_eye<true>, dim3(blocks), dim3(kTileThreads), 0, stream, p, cache, total, st->queue);
    else
      hipLaunchKernelGGL(k_bpt_eye<false>, dim3(blocks), dim3(kTileThreads), 0, stream, p, cache, total, st->queue);
  }
  return hipGetLastError();
}

}  // namespace gpugi
