// Path tracer, light tracer and bidirectional path tracer kernels.
//
// Compiled with -ffp-contract=off.  Each kernel restates the GLSL main() it
// replaces with the same random-number draw order, so that with the same
// FrameSeed and light samples a render matches a CPU restatement pixel for
// pixel when the reference traversal is selected.
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
  uint32_t* stkNode;
  float* stkT;
  int tid;
  uint2* spill;
};

struct ClosestHit {
  bool hit;
  float t;
  int4 tri;
  gg::V3 bary;
  gg::V3 geoNormal;  // cross(e1, e0), not normalized (TRINORMAL_OUTPUT)
};

template <bool WIDE>
__device__ __forceinline__ ClosestHit trace_closest(const TraceCtx& c, gg::V3 o, gg::V3 d) {
  ClosestHit h;
  if (WIDE) {
    WideResult r = trace_wide_ray<false, kTileThreads, kWideLdsStack>(c.p->wide, o, d, gg::kRayMax, c.stkNode, c.stkT,
                                                                      c.tid, c.spill);
    h.hit = r.slot >= 0;
    h.t = r.t;
    if (h.hit) {
      h.tri = c.p->ref.slots[r.slot];
      h.bary = gg::v3(1.0f - (r.bz + r.by), r.by, r.bz);
      gg::V3 p0 = load_pos(c.p->ref.positions, h.tri.x), p1 = load_pos(c.p->ref.positions, h.tri.y),
             p2 = load_pos(c.p->ref.positions, h.tri.z);
      h.geoNormal = gg::cross(p0 - p2, p1 - p0);
    }
  } else {
    float rayLength = gg::kRayMax;
    RefHit rh;
    uint32_t nb = 0, nt = 0;
    trace_ref<false, false>(c.p->ref, o, d, rayLength, rh, nb, nt);
    h.hit = rayLength != gg::kRayMax;  // pathtracer.comp:56 tests rayHit == RAY_MAX
    h.t = rayLength;
    if (h.hit) {
      h.tri = rh.tri;
      h.bary = rh.bary;
      h.geoNormal = rh.triNormal;
    }
  }
  return h;
}

template <bool WIDE>
__device__ __forceinline__ bool trace_any(const TraceCtx& c, gg::V3 o, gg::V3 d, float len) {
  if (WIDE) {
    WideResult r = trace_wide_ray<true, kTileThreads, kWideLdsStack>(c.p->wide, o, d, len, c.stkNode, c.stkT, c.tid,
                                                                     c.spill);
    return r.any;
  }
  float rayLength = len;
  RefHit rh;
  uint32_t nb = 0, nt = 0;
  return trace_ref<true, false>(c.p->ref, o, d, rayLength, rh, nb, nt);
}

// EstimateDirectLight (lightsource.glsl:3-93) without importance break / cache.
template <bool WIDE>
__device__ __forceinline__ gg::V3 estimate_direct_light(const TraceCtx& c, gg::V3 pos, gg::V3 normal, int idx,
                                                        gg::V3 viewDir, const MaterialData& m, bool sampleAreaLights) {
  const gpugi_light_sample& ls = c.p->lights[idx];
  const gg::V3 lpos = gg::v3(ls.position[0], ls.position[1], ls.position[2]);
  const gg::V3 lint = gg::v3(ls.intensity[0], ls.intensity[1], ls.intensity[2]);
  gg::V3 dir = lpos - pos;
  const float lightDistSq = gg::dot(dir, dir) + gg::kDivisorEpsilon;
  const float lightDist = sqrtf(lightDistSq);
  dir = dir / lightDist;
  const float surfaceCos = gg::saturate(gg::dot(dir, normal));
  if (surfaceCos <= 0.0f) return gg::v3s(0.0f);
  if (ls.normal_theta_cos > 1.5f) {
    const gg::V3 o = gg::kRayHitEpsilon * dir + pos;
    if (!trace_any<WIDE>(c, o, dir, lightDist - gg::kRayHitEpsilon * 2.0f)) {
      const gg::V3 irradiance = (surfaceCos / lightDistSq) * lint;
      float pdf;
      const gg::V3 bsdf = bsdf_eye(viewDir, dir, m, normal, pdf);
      return irradiance * bsdf;
    }
  } else if (sampleAreaLights) {
    const gg::V3 lightNormal = gg::unpack_normal(ls.normal_phi, ls.normal_theta_cos);
    const float f = gg::saturate(gg::dot(lightNormal, -dir));
    if (f > 0.0f) {
      const gg::V3 o = gg::kRayHitEpsilon * dir + pos;
      if (!trace_any<WIDE>(c, o, dir, lightDist - gg::kRayHitEpsilon * 2.0f)) {
        float pdf;
        const gg::V3 intensity = lint * f;
        const gg::V3 bsdf = bsdf_eye(viewDir, dir, m, normal, pdf);
        const gg::V3 irradiance = (surfaceCos / lightDistSq) * intensity;
        return irradiance * bsdf;
      }
    }
  }
  return gg::v3s(0.0f);
}

// pathtracer.comp:16-110 (NUM_LIGHT_SAMPLES 1, RUSSIAN_ROULETTE, black environment).
template <bool WIDE>
__global__ __launch_bounds__(kTileThreads) void k_pathtracer(DrawParams p) {
  __shared__ uint32_t stkNode[WIDE ? kWideLdsStack * kTileThreads : 1];
  __shared__ float stkT[WIDE ? kWideLdsStack * kTileThreads : 1];
  const int tid = threadIdx.x;
  const int gx = blockIdx.x * kTile + (tid % kTile);
  const int gy = blockIdx.y * kTile + (tid / kTile);
  const uint32_t linear = (blockIdx.y * gridDim.x + blockIdx.x) * kTileThreads + tid;
  TraceCtx c{&p, stkNode, stkT, tid, WIDE ? p.wide.spill + (size_t)linear * p.wide.spillDepth : nullptr};

  const float lightSampleWeight = (float)p.numLights / 1.0f;
  uint32_t seed = gg::init_random_seed(p.frameSeed, (uint32_t)(gx + gy * p.width));
  gg::V3 color = gg::v3s(0.0f);
  gg::V3 throughput = gg::v3s(1.0f);
  bool didNEE = false;
  const float j0 = gg::random01(seed);
  const float j1 = gg::random01(seed);
  const gg::V3 U = gg::v3(p.camera.u[0], p.camera.u[1], p.camera.u[2]);
  const gg::V3 V = gg::v3(p.camera.v[0], p.camera.v[1], p.camera.v[2]);
  const gg::V3 W = gg::v3(p.camera.w[0], p.camera.w[1], p.camera.w[2]);
  gg::V3 o = gg::v3(p.camera.position[0], p.camera.position[1], p.camera.position[2]);
  gg::V3 d = gg::primary_direction(j0, j1, gx, gy, p.width, p.height, U, V, W);
  for (int i = 0; i < p.maxPathLength; ++i) {
    const ClosestHit h = trace_closest<WIDE>(c, o, d);
    if (!h.hit) {
      color = color + throughput * gg::v3s(0.0f);  // black environment map (renderersystem.cpp:39-47)
      break;
    }
    o = o + h.t * d;
    const gg::V3 n = shading_normal(p.ref.vinfo, h.tri, h.bary);
    const MaterialData m = load_material(p.materials, h.tri.w);
    if (!didNEE) color = color + throughput * m.emissivity;
    didNEE = m.refl[3] < 500.0f;
    const int sampleIndex = (int)(gg::xorshift32(seed) % p.numLights);
    const gg::V3 radiance = estimate_direct_light<WIDE>(c, o, n, sampleIndex, d, m, didNEE);
    color = color + radiance * throughput * lightSampleWeight;
    gg::V3 tp = gg::v3s(1.0f);
    float pdf;
    bool diffuse;
    d = sample_bsdf(d, m, seed, n, tp, false, pdf, diffuse);
    o = o + d * gg::kRayHitEpsilon;
    throughput = throughput * tp;
    const float q = gg::saturate(gg::luminance(tp));
    if (gg::random01(seed) >= q) break;
    throughput = throughput / q;
  }
  float4* px = &p.accum[(size_t)gy * p.width + gx];
  float4 v = *px;
  *px = make_float4(color.x + v.x, color.y + v.y, color.z + v.z, 0.0f + v.w);
}

static hipError_t EnsureBuffer(void** buf, size_t* cap, size_t bytes) {
  if (*cap >= bytes) return hipSuccess;
  if (*buf) hipFree(*buf);
  *buf = nullptr;
  *cap = 0;
  hipError_t e = hipMalloc(buf, bytes);
  if (e == hipSuccess) *cap = bytes;
  return e;
}

void FreeIntegratorState(IntegratorState* s) {
  for (void** b : {&s->spill, &s->lightCache, &s->scratch})
    if (*b) { hipFree(*b); *b = nullptr; }
  s->spillBytes = s->lightCacheBytes = s->scratchBytes = 0;
  s->lightCacheCapacity = -1;
}

void InvalidateIntegratorCaches(IntegratorState* s) { s->lightCacheCapacity = -1; }

static hipError_t DrawPT(IntegratorState* st, const DrawParams& p0, hipStream_t stream) {
  DrawParams p = p0;
  dim3 grid(p.width / kTile, p.height / kTile);  // truncating, like pathtracer.cpp:50
  if (grid.x == 0 || grid.y == 0) return hipSuccess;
  if (p.useWide) {
    const size_t threads = (size_t)grid.x * grid.y * kTileThreads;
    const size_t bytes = std::max<size_t>(8, threads * p.wide.spillDepth * sizeof(uint2));
    hipError_t e = EnsureBuffer(&st->spill, &st->spillBytes, bytes);
    if (e != hipSuccess) return e;
    p.wide.spill = (uint2*)st->spill;
    hipLaunchKernelGGL(k_pathtracer<true>, grid, dim3(kTileThreads), 0, stream, p);
  } else {
    hipLaunchKernelGGL(k_pathtracer<false>, grid, dim3(kTileThreads), 0, stream, p);
  }
  return hipGetLastError();
}

hipError_t DrawIntegrator(IntegratorState* st, int integrator, const DrawParams& p, hipStream_t stream,
                          std::string* err) {
  switch (integrator) {
    case GPUGI_PT:
      return DrawPT(st, p, stream);
    default:
      *err = "integrator not available in this build";
      return hipSuccess;
  }
}

}  // namespace gpugi
