// Wavefront path tracer: the paths of pathtracer.comp:16-110 advanced one
// bounce at a time over the whole image, with the rays of every bounce traced
// by the trace kernels (closest hit, then the NEE shadow rays as any hit) on
// compact device ray queues.  Each path consumes its random numbers and adds
// its terms in the order of the one-kernel path tracer (emission, then NEE,
// then the next vertex), so both schedules give the same image bit for bit.
//
// Per bounce, for the queue of live paths (pixel indices):
//   trace closest hit of rays[i]                 -> hits[i]
//   k_wf_shade:  emission, NEE record i (value, throughput * weight, shadow
//                ray, or a zero-length ray), BSDF sample, roulette; a
//                path that goes on is appended to the next queue with its ray
//   trace any hit of the shadow rays             -> occluded[i]
//   k_wf_resolve: color += (visible ? value : 0) * throughput * weight
// Every pass reads the queue length from device memory: the host enqueues
// maxPathLength bounces without waiting, so a frame can be captured in a
// graph; passes over an empty queue return at once.  The shade and resolve
// passes run a fixed grid of resident blocks that strides over the queue.
// On a second stream, the shadow trace and resolve of one bounce run beside
// the next bounce's closest-hit trace (they touch disjoint buffers); the next
// shade pass waits for the resolve, so each pixel's terms keep their order.
//
// The eye paths of the bidirectional path tracer (pathtracer_bidir.comp) run
// through the same passes: the any-hit ray of a bounce is the connection to a
// light-cache vertex, whose value the shade pass computes with the steps of
// bpt_eye.h and the resolve pass adds when the connection is not blocked.
#include "bpt_eye.h"
#include "integrator_common.h"
#include "launch.h"
#include "wavefront_pt.h"

namespace gpugi {

namespace {

constexpr int kWfBlock = 256;
constexpr int kWfPT = 0, kWfBPT = 1;  // the integrator whose paths the passes advance

// Pixel of path i: 8x8 tiles in row order, row-major inside a tile.
__device__ __forceinline__ void tile_pixel(uint32_t i, uint32_t tilesW, int& gx, int& gy) {
  const uint32_t tile = i / kTileThreads, k = i % kTileThreads;
  gx = (int)((tile % tilesW) * kTile + k % kTile);
  gy = (int)((tile / tilesW) * kTile + k / kTile);
}

template <int MODE>
__global__ __launch_bounds__(kWfBlock) void k_wf_raygen(DrawParams p, WavefrontBuffers b, uint32_t tilesW,
                                                        uint32_t numPaths) {
  const uint32_t i = blockIdx.x * kWfBlock + threadIdx.x;
  if (i >= numPaths) return;
  // the reference dispatches W/8 x H/8 tiles: only those pixels are traced.
  // Queue order is tile by tile, so 64 neighbouring pixels share a wave.
  int gx, gy;
  tile_pixel(i, tilesW, gx, gy);
  const uint32_t pix = (uint32_t)gy * (uint32_t)p.width + (uint32_t)gx;
  gg::V3 o, d;
  if (MODE == kWfPT) {
    PtPath s;
    pt_begin(p, s, gx, gy, p.frameSeed);
    b.colorSeed[pix] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(s.seed));
    b.tpFlags[pix] = make_float4(s.throughput.x, s.throughput.y, s.throughput.z, __uint_as_float(0u));  // bounce 0
    o = s.o;
    d = s.d;
  } else {
    BptEyePath s;
    bpt_eye_begin(p, s, gx, gy);
    b.colorSeed[pix] = make_float4(s.color.x, s.color.y, s.color.z, __uint_as_float(s.seed));
    b.tpFlags[pix] = make_float4(s.tp.x, s.tp.y, s.tp.z, __uint_as_float(0u));  // bounce 0
    b.misState[pix] = make_float2(s.lastConnectionProbability, s.lastCosAbs);
    o = s.o;
    d = s.d;
  }
  b.queue[0][i] = pix;
  b.rays[0][2 * (size_t)i] = make_float4(o.x, o.y, o.z, gg::kRayMax);
  b.rays[0][2 * (size_t)i + 1] = make_float4(d.x, d.y, d.z, 0.0f);
}

// One bounce of pt_bounce for queue entry i (cur: the queue being shaded).
// Returns whether the path goes on; then (o, d) is its next ray.
__device__ __forceinline__ bool wf_shade_one(const DrawParams& p, const WavefrontBuffers& b, int cur, uint32_t i,
                                             uint32_t& pixOut, gg::V3& oOut, gg::V3& dOut) {
  const uint32_t pix = b.queue[cur][i];
  pixOut = pix;
  const float4 r0 = b.rays[cur][2 * (size_t)i], r1 = b.rays[cur][2 * (size_t)i + 1];
  gg::V3 o = gg::v3(r0.x, r0.y, r0.z), d = gg::v3(r1.x, r1.y, r1.z);
  const int4 rec = b.hits[i];
  const ClosestHit h = closest_from_record(p, rec.x, __int_as_float(rec.y), __int_as_float(rec.z),
                                           __int_as_float(rec.w));
  // the path's state: two 16-byte records (color and random state, throughput and flags)
  const float4 c4 = b.colorSeed[pix], t4 = b.tpFlags[pix];
  gg::V3 color = gg::v3(c4.x, c4.y, c4.z), throughput = gg::v3(t4.x, t4.y, t4.z);
  uint32_t seed = __float_as_uint(c4.w);
  const uint32_t flags = __float_as_uint(t4.w);
  const int bounce = (int)(flags >> 1);
  bool didNEE = (flags & 1u) != 0;
  // the shadow ray, written once: tmax 0 where there is nothing to connect (its result is not read)
  float4 sr0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), sr1 = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
  if (!h.hit) {
    // environment on a miss (pathtracer.comp:56-58); the path ends, no NEE term
    color = color + throughput * (p.env ? env_radiance(p.env, p.envSize, d) : gg::v3s(0.0f));
    b.colorSeed[pix] = make_float4(color.x, color.y, color.z, c4.w);
    b.shadowRays[2 * (size_t)i] = sr0;
    b.shadowRays[2 * (size_t)i + 1] = sr1;
    b.neeValue[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return false;
  }
  const float lightSampleWeight = (float)p.numLights / 1.0f;
  o = o + h.t * d;
  const gg::V3 n = shading_normal(p.ref.vinfo, h.tri, h.bary);
  const MaterialData m = load_material_hit(p.materials, p.tex, p.ref.vinfo, h.tri, h.bary);
  // SHOW_SPECIFIC_PATHLENGTH (pathtracer.comp:74-84)
  const int spl = p.specificPathLength;
  if ((spl < 0 || bounce == spl) && !didNEE) color = color + throughput * m.emissivity;
  didNEE = m.refl[3] < 500.0f;
  if (spl < 0 || bounce + 1 == spl) {
    const int sampleIndex = (int)(gg::xorshift32(seed) % p.numLights);
    const NeeSample ns = nee_sample(p, o, n, sampleIndex, d, m, didNEE);
    if (ns.test) {
      sr0 = make_float4(ns.o.x, ns.o.y, ns.o.z, ns.len);
      sr1 = make_float4(ns.dir.x, ns.dir.y, ns.dir.z, 0.0f);
    }
    // w = 1: a term is added; value.xyz counts only when the shadow ray is not blocked
    b.neeValue[i] = make_float4(ns.value.x, ns.value.y, ns.value.z, ns.test ? 1.0f : 0.5f);
    b.neeWeight[i] = make_float4(throughput.x, throughput.y, throughput.z, lightSampleWeight);
  } else {
    b.neeValue[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // no light sample at this bounce
  }
  b.shadowRays[2 * (size_t)i] = sr0;
  b.shadowRays[2 * (size_t)i + 1] = sr1;
  gg::V3 tp = gg::v3s(1.0f);
  float pdf;
  bool diffuse;
  d = sample_bsdf(d, m, seed, n, tp, false, pdf, diffuse);
  o = o + d * gg::kRayHitEpsilon;
  throughput = throughput * tp;
  bool alive = true;
  if (bounce >= p.rrStartDepth) {  // Russian roulette (pathtracer.comp:99-104) from rrStartDepth on
    const float q = gg::saturate(gg::luminance(tp));
    if (gg::random01(seed) >= q) {
      alive = false;
    } else {
      throughput = throughput / q;
    }
  }
  b.colorSeed[pix] = make_float4(color.x, color.y, color.z, __uint_as_float(seed));
  b.tpFlags[pix] = make_float4(throughput.x, throughput.y, throughput.z,
                               __uint_as_float(((uint32_t)(bounce + 1) << 1) | (didNEE ? 1u : 0u)));
  oOut = o;
  dOut = d;
  return alive && bounce + 1 < p.maxPathLength;
}

// The light cache of a BPT draw and its entry count (a device word).
struct WfBpt {
  const LightCacheEntryDev* cache = nullptr;
  const uint32_t* numEntries = nullptr;
};

// One bounce of bpt_eye_bounce (light_integrators.hip) for queue entry i: the
// connection ray becomes the entry's any-hit ray and its value the entry's NEE
// record (w = 1 when a ray is traced).  Returns whether the path goes on; then
// (o, d) is its next ray.  actualN: light-cache entries in use.
__device__ __forceinline__ bool wf_bpt_shade_one(const DrawParams& p, const WavefrontBuffers& b, int cur, uint32_t i,
                                                 const LightCacheEntryDev* __restrict__ cache, uint32_t actualN,
                                                 uint32_t& pixOut, gg::V3& oOut, gg::V3& dOut) {
  const uint32_t pix = b.queue[cur][i];
  pixOut = pix;
  const float4 r0 = b.rays[cur][2 * (size_t)i], r1 = b.rays[cur][2 * (size_t)i + 1];
  const int4 rec = b.hits[i];
  const ClosestHit h = closest_from_record(p, rec.x, __int_as_float(rec.y), __int_as_float(rec.z),
                                           __int_as_float(rec.w));
  float4 sr0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), sr1 = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
  float4 value = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  bool alive = false;
  BptEyePath s;
  s.o = gg::v3(r0.x, r0.y, r0.z);
  s.d = gg::v3(r1.x, r1.y, r1.z);
  if (h.hit) {  // a miss ends the path and adds nothing (pathtracer_bidir.comp:62-63)
    const float4 c4 = b.colorSeed[pix], t4 = b.tpFlags[pix];
    const float2 ms = b.misState[pix];
    s.color = gg::v3(c4.x, c4.y, c4.z);
    s.seed = __float_as_uint(c4.w);
    s.tp = gg::v3(t4.x, t4.y, t4.z);
    s.bounce = (int)(__float_as_uint(t4.w) >> 1);
    s.lastConnectionProbability = ms.x;
    s.lastCosAbs = ms.y;
    s.mis = 0.0f;
    BptEyeVertex v;
    bpt_eye_hit(p, s, h, v);
    const BptConnection cn = bpt_connect_setup(p, s, cache, actualN);
    if (cn.test) {
      const gg::V3 add = bpt_connect_value(p, s, v, cn, (float)actualN);
      if (!bpt_value_is_zero(add)) {  // a zero term needs no ray
        sr0 = make_float4(cn.lo.x, cn.lo.y, cn.lo.z, cn.tmax);
        sr1 = make_float4(cn.ld.x, cn.ld.y, cn.ld.z, 0.0f);
        value = make_float4(add.x, add.y, add.z, 1.0f);
      }
    }
    alive = !bpt_eye_continue(p, s, v);
    b.colorSeed[pix] = make_float4(s.color.x, s.color.y, s.color.z, __uint_as_float(s.seed));
    b.tpFlags[pix] = make_float4(s.tp.x, s.tp.y, s.tp.z, __uint_as_float((uint32_t)s.bounce << 1));
    b.misState[pix] = make_float2(s.lastConnectionProbability, s.lastCosAbs);
  }
  b.shadowRays[2 * (size_t)i] = sr0;
  b.shadowRays[2 * (size_t)i + 1] = sr1;
  b.neeValue[i] = value;
  oOut = s.o;
  dOut = s.d;
  return alive;
}

// The shade pass over the queue's device length; paths that go on are appended
// to the next queue, one atomic per wave and in lane order within the wave.
template <int MODE>
__global__ __launch_bounds__(kWfBlock) void k_wf_shade(DrawParams p, WavefrontBuffers b, int cur, WfBpt bpt) {
  const uint32_t count = b.counters[cur];
  uint32_t actualN = 0;
  if (MODE == kWfBPT) {  // light-cache entries in use, as k_bpt_eye reads them
    const int numEntries = (int)*bpt.numEntries;
    actualN = (uint32_t)(numEntries < p.bptCapacity ? numEntries : p.bptCapacity);
  }
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t below = (1ull << lane) - 1ull;
  // every lane of a wave runs the same trips: i - lane is uniform per wave
  for (uint32_t base = blockIdx.x * kWfBlock + (threadIdx.x & ~63u); base < count; base += gridDim.x * kWfBlock) {
    const uint32_t i = base + lane;
    uint32_t pix = 0;
    gg::V3 o = gg::v3s(0.0f), d = gg::v3s(0.0f);
    const bool push = i < count && (MODE == kWfPT ? wf_shade_one(p, b, cur, i, pix, o, d)
                                                  : wf_bpt_shade_one(p, b, cur, i, bpt.cache, actualN, pix, o, d));
    const uint64_t m = __ballot(push);
    if (m == 0) continue;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t first = 0;
    if ((int)lane == leader) first = atomicAdd(&b.counters[1 - cur], (uint32_t)__popcll(m));
    first = (uint32_t)__shfl((int)first, leader);
    if (push) {
      const uint32_t k = first + (uint32_t)__popcll(m & below);
      b.queue[1 - cur][k] = pix;
      b.rays[1 - cur][2 * (size_t)k] = make_float4(o.x, o.y, o.z, gg::kRayMax);
      b.rays[1 - cur][2 * (size_t)k + 1] = make_float4(d.x, d.y, d.z, 0.0f);
    }
  }
}

// PT: color += radiance * throughput * weight, radiance = 0 when the shadow ray
// was blocked or there was nothing to connect to (as estimate_direct_light
// returns).  BPT: color += the connection's value when its ray was not blocked
// (nothing otherwise, as bpt_eye_bounce).
template <int MODE>
__global__ __launch_bounds__(kWfBlock) void k_wf_resolve(WavefrontBuffers b, int cur) {
  const uint32_t count = b.counters[cur];
  for (uint32_t i = blockIdx.x * kWfBlock + threadIdx.x; i < count; i += gridDim.x * kWfBlock) {
    const float4 v = b.neeValue[i];
    if (v.w == 0.0f) continue;  // a miss: no NEE term
    if (MODE == kWfBPT) {
      if (b.occluded[i].x != 0) continue;
      const uint32_t pix = b.queue[cur][i];
      const float4 c4 = b.colorSeed[pix];
      const gg::V3 color = gg::v3(c4.x, c4.y, c4.z) + gg::v3(v.x, v.y, v.z);
      b.colorSeed[pix] = make_float4(color.x, color.y, color.z, c4.w);
      continue;
    }
    const bool visible = v.w == 1.0f && b.occluded[i].x == 0;
    const gg::V3 radiance = visible ? gg::v3(v.x, v.y, v.z) : gg::v3s(0.0f);
    const float4 w = b.neeWeight[i];
    const uint32_t pix = b.queue[cur][i];
    const float4 c4 = b.colorSeed[pix];
    const gg::V3 color = gg::v3(c4.x, c4.y, c4.z) + radiance * gg::v3(w.x, w.y, w.z) * w.w;
    b.colorSeed[pix] = make_float4(color.x, color.y, color.z, c4.w);
  }
}

// accum += vec4(color, 0) (PT) or vec4(color, 1) (BPT eye pass)
template <int MODE>
__global__ __launch_bounds__(kWfBlock) void k_wf_finish(DrawParams p, WavefrontBuffers b, uint32_t tilesW,
                                                        uint32_t numPaths) {
  const uint32_t i = blockIdx.x * kWfBlock + threadIdx.x;
  if (i >= numPaths) return;
  int gx, gy;
  tile_pixel(i, tilesW, gx, gy);
  const uint32_t pix = (uint32_t)gy * (uint32_t)p.width + (uint32_t)gx;
  const float4 c = b.colorSeed[pix];
  float4* px = &p.accum[pix];
  const float4 a = *px;
  *px = make_float4(c.x + a.x, c.y + a.y, c.z + a.z, (MODE == kWfPT ? 0.0f : 1.0f) + a.w);
}

// counters[0] = the first queue's length, counters[1] = 0
__global__ void k_wf_init_counts(uint32_t* c, uint32_t first) {
  if (threadIdx.x < 2) c[threadIdx.x] = threadIdx.x == 0 ? first : 0u;
}

__global__ void k_wf_zero(uint32_t* c) { c[0] = 0u; }

dim3 Grid(uint32_t n) { return dim3((n + kWfBlock - 1) / kWfBlock); }

}  // namespace

size_t WavefrontBytes(size_t numPixels) {
  // color and random state, throughput and flags (16 B each) and the BPT MIS
  // state (8 B) per pixel; per queue slot: two queues and two ray arrays, hits,
  // shadow rays, occlusion, NEE value and weight
  return numPixels * (16 + 16 + 8) + numPixels * (4 * 2 + 32 * 2 + 16 + 32 + 16 + 16 + 16) +
         16 * 256;  // the counters and each buffer's rounding to 256 bytes
}

WavefrontBuffers CarveWavefront(void* base, size_t numPixels) {
  char* q = static_cast<char*>(base);
  auto take = [&](size_t bytes) {
    char* r = q;
    q += (bytes + 255) & ~(size_t)255;
    return r;
  };
  WavefrontBuffers b;
  b.counters = (uint32_t*)take(256);
  b.colorSeed = (float4*)take(numPixels * 16);
  b.tpFlags = (float4*)take(numPixels * 16);
  b.misState = (float2*)take(numPixels * 8);
  for (int k = 0; k < 2; ++k) {
    b.queue[k] = (uint32_t*)take(numPixels * 4);
    b.rays[k] = (float4*)take(numPixels * 32);
  }
  b.hits = (int4*)take(numPixels * 16);
  b.shadowRays = (float4*)take(numPixels * 32);
  b.occluded = (int4*)take(numPixels * 16);
  b.neeValue = (float4*)take(numPixels * 16);
  b.neeWeight = (float4*)take(numPixels * 16);
  return b;
}

namespace {

template <int MODE>
hipError_t DrawWavefront(const DrawParams& p, const WavefrontTrace& tr, const WavefrontBuffers& b, WfBpt bpt,
                         hipStream_t stream) {
  const uint32_t tilesW = (uint32_t)(p.width / kTile), tilesH = (uint32_t)(p.height / kTile);
  const uint32_t numPaths = tilesW * tilesH * kTileThreads;
  if (numPaths == 0 || p.maxPathLength <= 0) return hipSuccess;
  if (tr.passBlocks == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_wf_raygen<MODE>, Grid(numPaths), dim3(kWfBlock), 0, stream, p, b, tilesW, numPaths);
  hipLaunchKernelGGL(k_wf_init_counts, dim3(1), dim3(64), 0, stream, b.counters, numPaths);
  // resident blocks of the shade / resolve passes, at most one per 256 paths
  const dim3 passGrid(std::min<uint32_t>(tr.passBlocks, (numPaths + kWfBlock - 1) / kWfBlock));
  // with a side stream the shadow trace and resolve pass of bounce k overlap
  // the closest-hit trace of bounce k + 1: the two launches fill each other's
  // drain (the last waves of a persistent trace leave most CUs idle)
  const bool fork = tr.side != nullptr;
  const hipStream_t shadowStream = fork ? tr.side : stream;
  const WideScene& shadowScene = fork ? tr.wsShadow : tr.ws;
  hipError_t e = hipSuccess;
  int cur = 0;
  for (int bounce = 0; bounce < p.maxPathLength; ++bounce) {
    const uint32_t* count = b.counters + cur;  // this bounce's queue length, on the device
    const gpugi_ray* rays = (const gpugi_ray*)b.rays[cur];
    e = tr.wide ? LaunchTraceWide(tr.ws, tr.stack, GPUGI_CLOSEST_HIT, rays, (gpugi_hit*)b.hits, numPaths,
                                  tr.refill, tr.wait, p.watertight, stream, count)
                : LaunchTraceRef(p.ref, GPUGI_CLOSEST_HIT, rays, (gpugi_hit*)b.hits, numPaths, p.watertight,
                                 stream, count);
    if (e != hipSuccess) return e;
    // the previous bounce's resolve pass reads the queue, NEE records and color
    // that this shade pass overwrites (and the counter k_wf_zero clears)
    if (fork && bounce > 0 && (e = hipStreamWaitEvent(stream, tr.evResolve, 0)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_wf_zero, dim3(1), dim3(1), 0, stream, b.counters + (1 - cur));
    hipLaunchKernelGGL(k_wf_shade<MODE>, passGrid, dim3(kWfBlock), 0, stream, p, b, cur, bpt);
    if (fork) {
      if ((e = hipEventRecord(tr.evShade, stream)) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(tr.side, tr.evShade, 0)) != hipSuccess) return e;
    }
    const gpugi_ray* shadow = (const gpugi_ray*)b.shadowRays;
    e = tr.wide ? LaunchTraceWide(shadowScene, tr.stack, GPUGI_ANY_HIT, shadow, (gpugi_hit*)b.occluded, numPaths,
                                  tr.refill, tr.wait, p.watertight, shadowStream, count)
                : LaunchTraceRef(p.ref, GPUGI_ANY_HIT, shadow, (gpugi_hit*)b.occluded, numPaths, p.watertight,
                                 shadowStream, count);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_wf_resolve<MODE>, passGrid, dim3(kWfBlock), 0, shadowStream, b, cur);
    if (fork && (e = hipEventRecord(tr.evResolve, tr.side)) != hipSuccess) return e;
    cur = 1 - cur;
  }
  if (fork && (e = hipStreamWaitEvent(stream, tr.evResolve, 0)) != hipSuccess) return e;  // join
  hipLaunchKernelGGL(k_wf_finish<MODE>, Grid(numPaths), dim3(kWfBlock), 0, stream, p, b, tilesW, numPaths);
  return hipGetLastError();
}

}  // namespace

hipError_t DrawPTWavefront(const DrawParams& p, const WavefrontTrace& tr, const WavefrontBuffers& b,
                           hipStream_t stream) {
  return DrawWavefront<kWfPT>(p, tr, b, WfBpt{}, stream);
}

hipError_t DrawBPTWavefront(const DrawParams& p, const WavefrontTrace& tr, const WavefrontBuffers& b,
                            const LightCacheEntryDev* cache, const uint32_t* numEntries, hipStream_t stream) {
  if (!cache || !numEntries) return hipErrorInvalidValue;
  WfBpt bpt;
  bpt.cache = cache;
  bpt.numEntries = numEntries;
  return DrawWavefront<kWfBPT>(p, tr, b, bpt, stream);
}

}  // namespace gpugi
