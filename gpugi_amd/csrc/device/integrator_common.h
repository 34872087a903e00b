// Traversal and light-sample helpers shared by the integrator kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "integrators.h"
#include "shading.h"
#include "trace_wide.h"

namespace gpugi {

constexpr int kTile = 8;                 // local size 8x8 of pathtracer.comp:14-15
constexpr int kTileThreads = kTile * kTile;

// Closest / any hit through either traversal.  The wide walk needs a per-lane
// LDS stack (S x 64 entries per block) and a global spill slice.
struct TraceCtx {
  const DrawParams* p;
  uint2* stk;
  int tid;
  uint2* spill;
  uint32_t* count = nullptr;  // when set, incremented per ray query
};

// Work items [0, numItems) for a persistent one-wave block, with regeneration:
// the wave takes chunks of `chunk` consecutive items from *queue (zeroed
// before the launch) and gives each lane whose item has finished the next item
// of its chunk.  begin(item) starts an item on the calling lane, step()
// advances it and returns true when it has finished, finish() then completes
// it.  The order in which items run does not change what any item computes.
template <class Begin, class Step, class Finish>
__device__ __forceinline__ void persistent_items(uint32_t numItems, uint32_t* queue, Begin begin, Step step,
                                                 Finish finish, uint32_t chunk = 64) {
  const int lane = threadIdx.x & 63;
  const uint64_t below = (1ull << lane) - 1ull;
  const uint32_t chunks = (numItems + chunk - 1u) / chunk;
  // wave-uniform: the chunk being handed out and how many of its items are gone
  uint32_t cur = 0, pos = chunk;
  bool exhausted = false, live = false;
  for (;;) {
    const uint64_t idle = ~__ballot(live);  // the block is one full wave
    const uint32_t nidle = (uint32_t)__popcll(idle);
    if (!exhausted && nidle > 0u) {
      const uint32_t avail = chunk - pos;
      uint32_t next = chunks;
      if (avail < nidle) {
        uint32_t v = 0;
        if (lane == 0) v = atomicAdd(queue, 1u);
        v = (uint32_t)__shfl((int)v, 0);
        if (v < chunks) next = v;
      }
      if (!live) {
        const uint32_t rank = (uint32_t)__popcll(idle & below);
        uint32_t item = numItems;
        if (rank < avail) item = cur * chunk + pos + rank;
        else if (next < chunks) item = next * chunk + (rank - avail);
        if (item < numItems) {
          begin(item);
          live = true;
        }
      }
      if (avail >= nidle) {
        pos += nidle;
      } else if (next < chunks) {
        cur = next;
        pos = nidle - avail;
      } else {
        exhausted = true;
      }
    }
    if (__ballot(live) == 0) {
      if (exhausted) break;
      continue;
    }
    if (live && step()) {
      finish();
      live = false;
    }
  }
}

// Pixel of item `item` when items run over 8x8 tiles in raster order.
__device__ __forceinline__ void tile_pixel(uint32_t item, int width, int& gx, int& gy) {
  const uint32_t tilesX = (uint32_t)width / kTile, tile = item / kTileThreads, k = item % kTileThreads;
  gx = (int)((tile % tilesX) * kTile + k % kTile);
  gy = (int)((tile / tilesX) * kTile + k / kTile);
}

struct ClosestHit {
  bool hit;
  float t;
  int4 tri;
  gg::V3 bary;
  gg::V3 geoNormal;  // cross(e1, e0), not normalized (TRINORMAL_OUTPUT)
};

// ClosestHit from a hit record {slot, t, b_y, b_z} of the wide walk or a
// trace kernel (b_x = 1 - (b_z + b_y), the reference formula; the geometric
// normal from the vertices with the reference formula cross(p0 - p2, p1 - p0)).
__device__ __forceinline__ ClosestHit closest_from_record(const DrawParams& p, int slot, float t, float by, float bz) {
  ClosestHit h;
  h.hit = slot >= 0;
  h.t = t;
  if (h.hit) {
    h.tri = p.ref.slots[slot];
    h.bary = gg::v3(1.0f - (bz + by), by, bz);
    gg::V3 p0 = load_pos(p.ref.positions, h.tri.x), p1 = load_pos(p.ref.positions, h.tri.y),
           p2 = load_pos(p.ref.positions, h.tri.z);
    h.geoNormal = gg::cross(p0 - p2, p1 - p0);
  }
  return h;
}

// With DrawParams::watertight the watertight triangle test replaces the
// reference one in both traversals (the geometric normal is then rebuilt
// from the vertices with the reference formula cross(p0 - p2, p1 - p0)).
// S: LDS stack entries per lane of the wide walk (the block's stack array
// holds S x kTileThreads entries).
template <bool WIDE, int S = kWideLdsStack>
__device__ __forceinline__ ClosestHit trace_closest(const TraceCtx& c, gg::V3 o, gg::V3 d) {
  ClosestHit h;
  if (c.count) ++*c.count;
  if (WIDE) {
    WideResult r = c.p->watertight
                       ? trace_wide_ray<false, kTileThreads, S, true>(c.p->wide, o, d, gg::kRayMax, c.stk, c.tid,
                                                                      c.spill)
                       : trace_wide_ray<false, kTileThreads, S, false>(c.p->wide, o, d, gg::kRayMax, c.stk, c.tid,
                                                                       c.spill);
    h = closest_from_record(*c.p, r.slot, r.t, r.by, r.bz);
  } else {
    float rayLength = gg::kRayMax;
    RefHit rh;
    uint32_t nb = 0, nt = 0;
    if (c.p->watertight) trace_ref<false, false, true>(c.p->ref, o, d, rayLength, rh, nb, nt);
    else trace_ref<false, false, false>(c.p->ref, o, d, rayLength, rh, nb, nt);
    h.hit = rayLength != gg::kRayMax;  // pathtracer.comp:56 tests rayHit == RAY_MAX
    h.t = rayLength;
    if (h.hit) {
      h.tri = rh.tri;
      h.bary = rh.bary;
      if (c.p->watertight) {
        const gg::V3 p0 = load_pos(c.p->ref.positions, h.tri.x), p1 = load_pos(c.p->ref.positions, h.tri.y),
                     p2 = load_pos(c.p->ref.positions, h.tri.z);
        h.geoNormal = gg::cross(p0 - p2, p1 - p0);
      } else {
        h.geoNormal = rh.triNormal;
      }
    }
  }
  return h;
}

template <bool WIDE, int S = kWideLdsStack>
__device__ __forceinline__ bool trace_any(const TraceCtx& c, gg::V3 o, gg::V3 d, float len) {
  if (c.count) ++*c.count;
  if (WIDE) {
    WideResult r = c.p->watertight
                       ? trace_wide_ray<true, kTileThreads, S, true>(c.p->wide, o, d, len, c.stk, c.tid, c.spill)
                       : trace_wide_ray<true, kTileThreads, S, false>(c.p->wide, o, d, len, c.stk, c.tid, c.spill);
    return r.any;
  }
  float rayLength = len;
  RefHit rh;
  uint32_t nb = 0, nt = 0;
  if (c.p->watertight) return trace_ref<true, false, true>(c.p->ref, o, d, rayLength, rh, nb, nt);
  return trace_ref<true, false, false>(c.p->ref, o, d, rayLength, rh, nb, nt);
}

// The light-sample connection of EstimateDirectLight (lightsource.glsl:3-93,
// without importance break / cache): the shadow ray to test and the radiance
// it carries if unoccluded.  test == false: nothing to connect.
struct NeeSample {
  gg::V3 o, dir;
  float len;
  gg::V3 value;
  bool test;
};

__device__ __forceinline__ NeeSample nee_sample(const DrawParams& p, gg::V3 pos, gg::V3 normal, int idx,
                                                gg::V3 viewDir, const MaterialData& m, bool sampleAreaLights) {
  NeeSample s;
  s.test = false;
  s.value = gg::v3s(0.0f);
  const gpugi_light_sample& ls = p.lights[idx];
  const gg::V3 lpos = gg::v3(ls.position[0], ls.position[1], ls.position[2]);
  const gg::V3 lint = gg::v3(ls.intensity[0], ls.intensity[1], ls.intensity[2]);
  gg::V3 dir = lpos - pos;
  const float lightDistSq = gg::dot(dir, dir) + gg::kDivisorEpsilon;
  const float lightDist = sqrtf(lightDistSq);
  dir = dir / lightDist;
  const float surfaceCos = gg::saturate(gg::dot(dir, normal));
  if (surfaceCos <= 0.0f) return s;
  if (ls.normal_theta_cos > 1.5f) {
    s.o = gg::kRayHitEpsilon * dir + pos;
    s.dir = dir;
    s.len = lightDist - gg::kRayHitEpsilon * 2.0f;
    const gg::V3 irradiance = (surfaceCos / lightDistSq) * lint;
    float pdf;
    const gg::V3 bsdf = bsdf_eye(viewDir, dir, m, normal, pdf);
    s.value = irradiance * bsdf;
    s.test = true;
  } else if (sampleAreaLights) {
    const gg::V3 lightNormal = gg::unpack_normal(ls.normal_phi, ls.normal_theta_cos);
    const float f = gg::saturate(gg::dot(lightNormal, -dir));
    if (f > 0.0f) {
      s.o = gg::kRayHitEpsilon * dir + pos;
      s.dir = dir;
      s.len = lightDist - gg::kRayHitEpsilon * 2.0f;
      float pdf;
      const gg::V3 intensity = lint * f;
      const gg::V3 bsdf = bsdf_eye(viewDir, dir, m, normal, pdf);
      const gg::V3 irradiance = (surfaceCos / lightDistSq) * intensity;
      s.value = irradiance * bsdf;
      s.test = true;
    }
  }
  return s;
}

// EstimateDirectLight with the shadow ray traced here.
template <bool WIDE, int S = kWideLdsStack>
__device__ __forceinline__ gg::V3 estimate_direct_light(const TraceCtx& c, gg::V3 pos, gg::V3 normal, int idx,
                                                        gg::V3 viewDir, const MaterialData& m, bool sampleAreaLights) {
  const NeeSample s = nee_sample(*c.p, pos, normal, idx, viewDir, m, sampleAreaLights);
  if (s.test && !trace_any<WIDE, S>(c, s.o, s.dir, s.len)) return s.value;
  return gg::v3s(0.0f);
}

// One path of pathtracer.comp:16-110 (NUM_LIGHT_SAMPLES 1, RUSSIAN_ROULETTE,
// black environment) as a state that advances one bounce at a time, so a
// persistent kernel can give a finished lane a new pixel.
struct PtPath {
  gg::V3 color, throughput, o, d;
  uint32_t seed;
  int bounce;
  bool didNEE;
};

// Camera ray of pixel (gx, gy) and the random state after it.
__device__ __forceinline__ void pt_begin(const DrawParams& p, PtPath& s, int gx, int gy, uint32_t frameSeed) {
  s.seed = gg::init_random_seed(frameSeed, (uint32_t)(gx + gy * p.width));
  s.color = gg::v3s(0.0f);
  s.throughput = gg::v3s(1.0f);
  s.didNEE = false;
  s.bounce = 0;
  const float j0 = gg::random01(s.seed);
  const float j1 = gg::random01(s.seed);
  const gg::V3 U = gg::v3(p.camera.u[0], p.camera.u[1], p.camera.u[2]);
  const gg::V3 V = gg::v3(p.camera.v[0], p.camera.v[1], p.camera.v[2]);
  const gg::V3 W = gg::v3(p.camera.w[0], p.camera.w[1], p.camera.w[2]);
  s.o = gg::v3(p.camera.position[0], p.camera.position[1], p.camera.position[2]);
  s.d = gg::primary_direction(j0, j1, gx, gy, p.width, p.height, U, V, W);
}

// One bounce of the path (the loop body of pathtracer.comp).  Returns true
// when the path has ended; s.color is then final.
template <bool WIDE, int S = kWideLdsStack>
__device__ __forceinline__ bool pt_bounce(const DrawParams& p, const TraceCtx& c, PtPath& s) {
  if (s.bounce >= p.maxPathLength) return true;
  const float lightSampleWeight = (float)p.numLights / 1.0f;
  const ClosestHit h = trace_closest<WIDE, S>(c, s.o, s.d);
  if (!h.hit) {
    // environment on a miss (pathtracer.comp:56-58); black by default (renderersystem.cpp:39-47)
    s.color = s.color + s.throughput * (p.env ? env_radiance(p.env, p.envSize, s.d) : gg::v3s(0.0f));
    return true;
  }
  s.o = s.o + h.t * s.d;
  const gg::V3 n = shading_normal(p.ref.vinfo, h.tri, h.bary);
  const MaterialData m = load_material_hit(p.materials, p.tex, p.ref.vinfo, h.tri, h.bary);
  // SHOW_SPECIFIC_PATHLENGTH (pathtracer.comp:74-84): the emission at bounce i
  // closes a path of length i, the light sample one of length i + 1
  const int spl = p.specificPathLength;
  if ((spl < 0 || s.bounce == spl) && !s.didNEE) s.color = s.color + s.throughput * m.emissivity;
  s.didNEE = m.refl[3] < 500.0f;
  if (spl < 0 || s.bounce + 1 == spl) {
    const int sampleIndex = (int)(gg::xorshift32(s.seed) % p.numLights);
    const gg::V3 radiance = estimate_direct_light<WIDE, S>(c, s.o, n, sampleIndex, s.d, m, s.didNEE);
    s.color = s.color + radiance * s.throughput * lightSampleWeight;
  }
  gg::V3 tp = gg::v3s(1.0f);
  float pdf;
  bool diffuse;
  s.d = sample_bsdf(s.d, m, s.seed, n, tp, false, pdf, diffuse);
  s.o = s.o + s.d * gg::kRayHitEpsilon;
  s.throughput = s.throughput * tp;
  if (s.bounce >= p.rrStartDepth) {  // Russian roulette (pathtracer.comp:99-104) from rrStartDepth on
    const float q = gg::saturate(gg::luminance(tp));
    if (gg::random01(s.seed) >= q) return true;
    s.throughput = s.throughput / q;
  }
  ++s.bounce;
  return s.bounce >= p.maxPathLength;
}

// Adds vec4(color, 0) to the pixel's accumulation entry.
__device__ __forceinline__ void pt_finish(const DrawParams& p, const PtPath& s, int gx, int gy) {
  float4* px = &p.accum[(size_t)gy * p.width + gx];
  const float4 v = *px;
  *px = make_float4(s.color.x + v.x, s.color.y + v.y, s.color.z + v.z, 0.0f + v.w);
}

// The whole path of pixel (gx, gy) on the calling lane.  rays (optional)
// counts the rays this path traced.
template <bool WIDE, int S = kWideLdsStack>
__device__ __forceinline__ void pt_pixel(const DrawParams& p, TraceCtx c, int gx, int gy, unsigned long long* rays) {
  uint32_t traced = 0;
  if (rays) c.count = &traced;
  PtPath s;
  pt_begin(p, s, gx, gy, p.frameSeed);
  while (!pt_bounce<WIDE, S>(p, c, s)) {
  }
  pt_finish(p, s, gx, gy);
  if (rays) {
    // one atomic per wave
    uint32_t sum = traced;
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(rays, (unsigned long long)sum);
  }
}

}  // namespace gpugi
