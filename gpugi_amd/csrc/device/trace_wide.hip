// Persistent-wave ray queries on the 4-wide quantized BVH.
//
// Each wave takes 64 rays at a time from a global counter (one atomic per
// wave), so the grid is sized to the machine rather than to the ray count and
// the per-thread overflow stack has a fixed size.  Hit records are written as
// one 16-byte store per lane, 1 KiB contiguous per wave.  The LDS stack depth
// is a template parameter: deeper stacks spill less but cap the waves per CU
// (LDS is 160 KiB per CU); the autotuner picks the depth per scene.
// Compiled with -ffp-contract=off: the triangle test must round exactly like
// the reference formula.
#include <hip/hip_runtime.h>

#include "launch.h"
#include "trace_wide.h"

namespace gpugi {

template <bool ANY, int B, int S>
__global__ __launch_bounds__(B) void k_trace_wide(WideScene s, const float4* __restrict__ rays,
                                                  int4* __restrict__ hits, uint32_t n,
                                                  uint32_t* __restrict__ counter) {
  __shared__ uint32_t stkNode[S * B];
  __shared__ float stkT[S * B];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const uint32_t gtid = blockIdx.x * B + tid;
  uint2* spill = s.spill + (size_t)gtid * s.spillDepth;
  for (;;) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(counter, 64u);
    base = __shfl(base, 0);
    if (base >= n) break;
    const uint32_t ri = base + (uint32_t)lane;
    if (ri >= n) continue;
    const float4 r0 = rays[2 * (size_t)ri], r1 = rays[2 * (size_t)ri + 1];
    const WideResult r = trace_wide_ray<ANY, B, S>(s, gg::v3(r0.x, r0.y, r0.z), gg::v3(r1.x, r1.y, r1.z), r0.w,
                                                   stkNode, stkT, tid, spill);
    if (ANY) hits[ri] = make_int4(r.any ? 1 : 0, 0, 0, 0);
    else hits[ri] = make_int4(r.slot, __float_as_int(r.t), __float_as_int(r.by), __float_as_int(r.bz));
  }
}

__global__ void k_zero_u32(uint32_t* p) { *p = 0u; }

namespace {

struct Variant {
  int stack;
  const void* closest;
  const void* any;
};

template <int S>
Variant MakeVariant() {
  return Variant{S, reinterpret_cast<const void*>(&k_trace_wide<false, kWideBlock, S>),
                 reinterpret_cast<const void*>(&k_trace_wide<true, kWideBlock, S>)};
}

const Variant* FindVariant(int stack) {
  static const Variant table[] = {MakeVariant<8>(), MakeVariant<12>(), MakeVariant<16>(), MakeVariant<24>(),
                                  MakeVariant<32>()};
  for (const Variant& v : table)
    if (v.stack == stack) return &v;
  return nullptr;
}

template <int S>
void LaunchS(const WideScene& s, int query, const gpugi_ray* rays, gpugi_hit* hits, size_t n, uint32_t* counter,
             hipStream_t stream) {
  dim3 grid(s.numThreads / kWideBlock);
  if (query == GPUGI_ANY_HIT)
    hipLaunchKernelGGL((k_trace_wide<true, kWideBlock, S>), grid, dim3(kWideBlock), 0, stream, s,
                       (const float4*)rays, (int4*)hits, (uint32_t)n, counter);
  else
    hipLaunchKernelGGL((k_trace_wide<false, kWideBlock, S>), grid, dim3(kWideBlock), 0, stream, s,
                       (const float4*)rays, (int4*)hits, (uint32_t)n, counter);
}

}  // namespace

const int kWideStackVariants[kNumWideStackVariants] = {8, 12, 16, 24, 32};

hipError_t WideBlocksPerCU(int stack, int* blocks) {
  const Variant* v = FindVariant(stack);
  if (!v) return hipErrorInvalidValue;
  int a = 0, b = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, v->closest, kWideBlock, 0);
  if (e != hipSuccess) return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, v->any, kWideBlock, 0);
  if (e != hipSuccess) return e;
  *blocks = a > b ? a : b;
  return hipSuccess;
}

hipError_t LaunchTraceWide(const WideScene& s, int stack, int query, const gpugi_ray* rays, gpugi_hit* hits,
                           size_t n, uint32_t* rayCounter, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  if (n > 0xffffffffull - 64) return hipErrorInvalidValue;
  if (s.numThreads == 0 || s.numThreads % kWideBlock != 0) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, stream, rayCounter);
  switch (stack) {
    case 8: LaunchS<8>(s, query, rays, hits, n, rayCounter, stream); break;
    case 12: LaunchS<12>(s, query, rays, hits, n, rayCounter, stream); break;
    case 16: LaunchS<16>(s, query, rays, hits, n, rayCounter, stream); break;
    case 24: LaunchS<24>(s, query, rays, hits, n, rayCounter, stream); break;
    case 32: LaunchS<32>(s, query, rays, hits, n, rayCounter, stream); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace gpugi
