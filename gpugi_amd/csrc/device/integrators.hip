// Path tracer, light tracer and bidirectional path tracer kernels.
//
// Compiled with -ffp-contract=off.  Each kernel restates the GLSL main() it
// replaces with the same random-number draw order, so that with the same
// FrameSeed and light samples a render matches a CPU restatement pixel for
// pixel when the reference traversal is selected.
#include <hip/hip_runtime.h>

#include "integrator_common.h"

namespace gpugi {

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
    if (i >= p.rrStartDepth) {  // Russian roulette (pathtracer.comp:99-104) from rrStartDepth on
      const float q = gg::saturate(gg::luminance(tp));
      if (gg::random01(seed) >= q) break;
      throughput = throughput / q;
    }
  }
  float4* px = &p.accum[(size_t)gy * p.width + gx];
  float4 v = *px;
  *px = make_float4(color.x + v.x, color.y + v.y, color.z + v.z, 0.0f + v.w);
}

void FreeIntegratorState(IntegratorState* s) {
  for (void** b : {&s->spill, &s->lightCache, &s->scratch, &s->staging, &s->pathCounts})
    if (*b) { (void)hipFree(*b); *b = nullptr; }
  s->spillBytes = s->lightCacheBytes = s->scratchBytes = s->stagingBytes = s->pathCountsBytes = 0;
  s->lastTotal = nullptr;
}

void InvalidateIntegratorCaches(IntegratorState* s) { s->lastTotal = nullptr; }

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
    case GPUGI_LT:
      return DrawLT(st, p, stream);
    case GPUGI_BPT:
      return DrawBPT(st, p, stream, err);
    default:
      *err = "integrator not available in this build";
      return hipSuccess;
  }
}

}  // namespace gpugi
