// Traversal kernels and ray generators of the microbenchmark.
// Compiled with -ffp-contract=off: the triangle test must round exactly like
// the reference formula.
#include <hip/hip_runtime.h>

#include "launch.h"

namespace gpugi {

// ---------------------------------------------------------------------------
// Reference layout: one thread per ray.
template <bool ANY, bool WT>
__global__ __launch_bounds__(256) void k_trace_ref(RefScene s, const float4* __restrict__ rays,
                                                   int4* __restrict__ hits, uint32_t n,
                                                   const uint32_t* __restrict__ nDev) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= (nDev ? min(*nDev, n) : n)) return;  // nDev: the ray count as the device holds it
  const float4 r0 = rays[2 * (size_t)i], r1 = rays[2 * (size_t)i + 1];
  gg::V3 o = gg::v3(r0.x, r0.y, r0.z), d = gg::v3(r1.x, r1.y, r1.z);
  float rayLength = r0.w;
  RefHit h;
  uint32_t nb = 0, nt = 0;
  bool any = trace_ref<ANY, false, WT>(s, o, d, rayLength, h, nb, nt);
  if (ANY) {
    hits[i] = make_int4(any ? 1 : 0, 0, 0, 0);
  } else {
    hits[i] = make_int4(h.slot, __float_as_int(rayLength), __float_as_int(h.slot >= 0 ? h.bary.y : 0.0f),
                        __float_as_int(h.slot >= 0 ? h.bary.z : 0.0f));
  }
}

// TRACERAY_IMPORTANCE_BREAK walk; nodes (optional) get {hit node, node
// importance bits, node size^2 bits, 0} per ray.
template <bool ANY, bool WT>
__global__ __launch_bounds__(256) void k_trace_ref_ib(RefScene s, const float4* __restrict__ rays,
                                                      int4* __restrict__ hits, int4* __restrict__ nodes, uint32_t n,
                                                      ImportanceBreak ib) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const float4 r0 = rays[2 * (size_t)i], r1 = rays[2 * (size_t)i + 1];
  gg::V3 o = gg::v3(r0.x, r0.y, r0.z), d = gg::v3(r1.x, r1.y, r1.z);
  float rayLength = r0.w;
  RefHit h;
  uint32_t nb = 0, nt = 0;
  const bool any = trace_ref<ANY, false, WT>(s, o, d, rayLength, h, nb, nt, ib);
  if (ANY) {
    hits[i] = make_int4(any ? 1 : 0, 0, 0, 0);
  } else {
    hits[i] = make_int4(h.slot, __float_as_int(rayLength), __float_as_int(h.slot >= 0 ? h.bary.y : 0.0f),
                        __float_as_int(h.slot >= 0 ? h.bary.z : 0.0f));
    if (nodes) nodes[i] = make_int4(h.node, __float_as_int(h.nodeImportance), __float_as_int(h.nodeSizeSq), 0);
  }
}

template <bool ANY>
__global__ __launch_bounds__(256) void k_trace_ref_counts(RefScene s, const float4* __restrict__ rays,
                                                          uint32_t* __restrict__ nbox,
                                                          uint32_t* __restrict__ ntri, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const float4 r0 = rays[2 * (size_t)i], r1 = rays[2 * (size_t)i + 1];
  gg::V3 o = gg::v3(r0.x, r0.y, r0.z), d = gg::v3(r1.x, r1.y, r1.z);
  float rayLength = r0.w;
  RefHit h;
  uint32_t nb = 0, nt = 0;
  trace_ref<ANY, true>(s, o, d, rayLength, h, nb, nt);
  nbox[i] = nb;
  ntri[i] = nt;
}

hipError_t LaunchTraceRefImportance(const RefScene& s, int query, const gpugi_ray* rays, gpugi_hit* hits,
                                    int4* nodes, size_t n, const ImportanceBreak& ib, bool watertight,
                                    hipStream_t stream) {
  if (n == 0) return hipSuccess;
  dim3 grid((unsigned)((n + 255) / 256));
  const float4* r = (const float4*)rays;
  int4* h = (int4*)hits;
  const uint32_t m = (uint32_t)n;
  if (query == GPUGI_ANY_HIT) {
    if (watertight) hipLaunchKernelGGL((k_trace_ref_ib<true, true>), grid, dim3(256), 0, stream, s, r, h, nodes, m, ib);
    else hipLaunchKernelGGL((k_trace_ref_ib<true, false>), grid, dim3(256), 0, stream, s, r, h, nodes, m, ib);
  } else {
    if (watertight) hipLaunchKernelGGL((k_trace_ref_ib<false, true>), grid, dim3(256), 0, stream, s, r, h, nodes, m, ib);
    else hipLaunchKernelGGL((k_trace_ref_ib<false, false>), grid, dim3(256), 0, stream, s, r, h, nodes, m, ib);
  }
  return hipGetLastError();
}

hipError_t LaunchTraceRef(const RefScene& s, int query, const gpugi_ray* rays, gpugi_hit* hits, size_t n,
                          bool watertight, hipStream_t stream, const uint32_t* nDevice) {
  if (n == 0) return hipSuccess;
  if (n > 0xffffffffull - 255) return hipErrorInvalidValue;  // the kernel indexes rays with 32 bits
  dim3 grid((unsigned)((n + 255) / 256));
  const float4* r = (const float4*)rays;
  int4* h = (int4*)hits;
  const uint32_t m = (uint32_t)n;
  if (query == GPUGI_ANY_HIT) {
    if (watertight) hipLaunchKernelGGL((k_trace_ref<true, true>), grid, dim3(256), 0, stream, s, r, h, m, nDevice);
    else hipLaunchKernelGGL((k_trace_ref<true, false>), grid, dim3(256), 0, stream, s, r, h, m, nDevice);
  } else {
    if (watertight) hipLaunchKernelGGL((k_trace_ref<false, true>), grid, dim3(256), 0, stream, s, r, h, m, nDevice);
    else hipLaunchKernelGGL((k_trace_ref<false, false>), grid, dim3(256), 0, stream, s, r, h, m, nDevice);
  }
  return hipGetLastError();
}

hipError_t LaunchTraceRefCounts(const RefScene& s, int query, const gpugi_ray* rays, uint32_t* nbox,
                                uint32_t* ntri, size_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  dim3 grid((unsigned)((n + 255) / 256));
  if (query == GPUGI_ANY_HIT)
    hipLaunchKernelGGL(k_trace_ref_counts<true>, grid, dim3(256), 0, stream, s, (const float4*)rays, nbox,
                       ntri, (uint32_t)n);
  else
    hipLaunchKernelGGL(k_trace_ref_counts<false>, grid, dim3(256), 0, stream, s, (const float4*)rays, nbox,
                       ntri, (uint32_t)n);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Ray generators.
__global__ __launch_bounds__(256) void k_raygen_primary(gpugi_camera cam, int width, int height, uint32_t seed,
                                                        float4* __restrict__ rays, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const uint32_t pix = i % (uint32_t)(width * height);
  const int px = (int)(pix % (uint32_t)width), py = (int)(pix / (uint32_t)width);
  uint32_t rs = gg::init_random_seed(seed, i);
  float j0 = gg::random01(rs);
  float j1 = gg::random01(rs);
  gg::V3 U = gg::v3(cam.u[0], cam.u[1], cam.u[2]);
  gg::V3 V = gg::v3(cam.v[0], cam.v[1], cam.v[2]);
  gg::V3 W = gg::v3(cam.w[0], cam.w[1], cam.w[2]);
  gg::V3 d = gg::primary_direction(j0, j1, px, py, width, height, U, V, W);
  rays[2 * (size_t)i] = make_float4(cam.position[0], cam.position[1], cam.position[2], gg::kRayMax);
  rays[2 * (size_t)i + 1] = make_float4(d.x, d.y, d.z, 0.0f);
}

hipError_t LaunchRaygenPrimary(const gpugi_camera& cam, int width, int height, uint32_t seed, gpugi_ray* rays,
                               size_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_raygen_primary, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, cam, width,
                     height, seed, (float4*)rays, (uint32_t)n);
  return hipGetLastError();
}

// Deterministic stream compaction of the hit indices (index order kept):
// per-block counts, one-block scan of the counts, scatter.
constexpr int kScanBlock = 1024;

__global__ __launch_bounds__(kScanBlock) void k_compact_count(const int4* __restrict__ hits, uint32_t n,
                                                              uint32_t* __restrict__ blockCounts) {
  __shared__ uint32_t warpSums[kScanBlock / 64];
  const uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
  const bool flag = i < n && hits[i].x >= 0;
  uint64_t mask = __ballot(flag);
  if ((threadIdx.x & 63) == 0) warpSums[threadIdx.x >> 6] = (uint32_t)__popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int w = 0; w < kScanBlock / 64; ++w) s += warpSums[w];
    blockCounts[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(kScanBlock) void k_compact_scan(uint32_t* __restrict__ blockCounts, uint32_t numBlocks,
                                                             uint32_t* __restrict__ total) {
  // Single block: each thread scans a contiguous chunk, then a block scan of chunk sums.
  __shared__ uint32_t sums[kScanBlock];
  const uint32_t per = (numBlocks + kScanBlock - 1) / kScanBlock;
  const uint32_t begin = threadIdx.x * per;
  uint32_t local = 0;
  for (uint32_t k = 0; k < per; ++k)
    if (begin + k < numBlocks) local += blockCounts[begin + k];
  sums[threadIdx.x] = local;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    uint32_t v = threadIdx.x >= (uint32_t)off ? sums[threadIdx.x - off] : 0u;
    __syncthreads();
    sums[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = sums[threadIdx.x] - local;  // exclusive prefix of this chunk
  for (uint32_t k = 0; k < per; ++k) {
    if (begin + k < numBlocks) {
      uint32_t c = blockCounts[begin + k];
      blockCounts[begin + k] = run;
      run += c;
    }
  }
  if (threadIdx.x == kScanBlock - 1) *total = sums[kScanBlock - 1];
}

__global__ __launch_bounds__(kScanBlock) void k_compact_scatter(const int4* __restrict__ hits, uint32_t n,
                                                                const uint32_t* __restrict__ blockOffsets,
                                                                uint32_t* __restrict__ list) {
  __shared__ uint32_t warpSums[kScanBlock / 64];
  const uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
  const bool flag = i < n && hits[i].x >= 0;
  uint64_t mask = __ballot(flag);
  const uint32_t lane = threadIdx.x & 63;
  if (lane == 0) warpSums[threadIdx.x >> 6] = (uint32_t)__popcll(mask);
  __syncthreads();
  uint32_t base = blockOffsets[blockIdx.x];
  for (uint32_t w = 0; w < (threadIdx.x >> 6); ++w) base += warpSums[w];
  const uint64_t below = lane ? (mask & (~0ull >> (64 - lane))) : 0ull;
  if (flag) list[base + (uint32_t)__popcll(below)] = i;
}

hipError_t LaunchHitCompaction(const gpugi_hit* hits, size_t n, uint32_t* list, uint32_t* count, void* temp,
                               size_t* tempBytes, hipStream_t stream) {
  const uint32_t numBlocks = (uint32_t)((n + kScanBlock - 1) / kScanBlock);
  const size_t need = (size_t)numBlocks * 4;
  if (!temp) { *tempBytes = need; return hipSuccess; }
  if (*tempBytes < need) return hipErrorInvalidValue;
  uint32_t* bc = (uint32_t*)temp;
  hipLaunchKernelGGL(k_compact_count, dim3(numBlocks), dim3(kScanBlock), 0, stream, (const int4*)hits,
                     (uint32_t)n, bc);
  hipLaunchKernelGGL(k_compact_scan, dim3(1), dim3(kScanBlock), 0, stream, bc, numBlocks, count);
  hipLaunchKernelGGL(k_compact_scatter, dim3(numBlocks), dim3(kScanBlock), 0, stream, (const int4*)hits,
                     (uint32_t)n, bc, list);
  return hipGetLastError();
}

// Hit point and geometric normal of primary ray j, as pathtracer.comp:56-61
// computes them (origin + t * dir; cross(e1, e0) of intersectiontests.glsl:68).
__device__ __forceinline__ void hit_frame(const RefScene& s, const float4* prim, const int4* hits, uint32_t j,
                                          gg::V3& p, gg::V3& din, gg::V3& ng) {
  const float4 r0 = prim[2 * (size_t)j], r1 = prim[2 * (size_t)j + 1];
  const int4 h = hits[j];
  gg::V3 o = gg::v3(r0.x, r0.y, r0.z);
  din = gg::v3(r1.x, r1.y, r1.z);
  p = o + __int_as_float(h.y) * din;
  const int4 tri = s.slots[h.x];
  gg::V3 p0 = load_pos(s.positions, tri.x), p1 = load_pos(s.positions, tri.y), p2 = load_pos(s.positions, tri.z);
  ng = gg::cross(p0 - p2, p1 - p0);
}

__global__ __launch_bounds__(256) void k_raygen_diffuse(RefScene s, const float4* __restrict__ prim,
                                                        const int4* __restrict__ hits,
                                                        const uint32_t* __restrict__ list,
                                                        const uint32_t* __restrict__ count, uint32_t seed,
                                                        float4* __restrict__ rays, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const uint32_t nh = *count;
  if (nh == 0) {  // nothing was hit: emit an inert ray
    rays[2 * (size_t)i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    rays[2 * (size_t)i + 1] = make_float4(0.0f, 1.0f, 0.0f, 0.0f);
    return;
  }
  const uint32_t j = list[i % nh];
  gg::V3 p, din, ng;
  hit_frame(s, prim, hits, j, p, din, ng);
  gg::V3 N = gg::normalize(ng);
  if (gg::dot(N, din) > 0.0f) N = -N;
  gg::V3 U, V;
  gg::create_onb(N, U, V);
  uint32_t rs = gg::init_random_seed(seed, i);
  float r0 = gg::random01(rs);
  float r1 = gg::random01(rs);
  gg::V3 d = gg::sample_hemisphere_cosine(r0, r1, U, V, N);
  gg::V3 o = p + d * gg::kRayHitEpsilon;
  rays[2 * (size_t)i] = make_float4(o.x, o.y, o.z, gg::kRayMax);
  rays[2 * (size_t)i + 1] = make_float4(d.x, d.y, d.z, 0.0f);
}

__global__ __launch_bounds__(256) void k_raygen_shadow(const float4* __restrict__ prim,
                                                       const int4* __restrict__ hits,
                                                       const uint32_t* __restrict__ list,
                                                       const uint32_t* __restrict__ count,
                                                       const float* __restrict__ lights, uint32_t numLights,
                                                       float4* __restrict__ rays, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const uint32_t nh = *count;
  if (nh == 0) {
    rays[2 * (size_t)i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    rays[2 * (size_t)i + 1] = make_float4(0.0f, 1.0f, 0.0f, 0.0f);
    return;
  }
  const uint32_t j = list[i % nh];
  const float4 a = prim[2 * (size_t)j], b = prim[2 * (size_t)j + 1];
  gg::V3 p = gg::v3(a.x, a.y, a.z) + __int_as_float(hits[j].y) * gg::v3(b.x, b.y, b.z);
  const uint32_t l = i % numLights;
  gg::V3 L = gg::v3(lights[3 * l], lights[3 * l + 1], lights[3 * l + 2]);
  // EstimateDirectLight's shadow ray (lightsource.glsl:18-21,39-40).
  gg::V3 dir = L - p;
  float lightDistSq = gg::dot(dir, dir) + gg::kDivisorEpsilon;
  float lightDist = sqrtf(lightDistSq);
  dir = dir / lightDist;
  gg::V3 o = gg::kRayHitEpsilon * dir + p;
  rays[2 * (size_t)i] = make_float4(o.x, o.y, o.z, lightDist - gg::kRayHitEpsilon * 2.0f);
  rays[2 * (size_t)i + 1] = make_float4(dir.x, dir.y, dir.z, 0.0f);
}

hipError_t LaunchRaygenDiffuse(const RefScene& s, const gpugi_ray* prim, const gpugi_hit* hits,
                               const uint32_t* list, const uint32_t* count, uint32_t seed, gpugi_ray* rays,
                               size_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_raygen_diffuse, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, s,
                     (const float4*)prim, (const int4*)hits, list, count, seed, (float4*)rays, (uint32_t)n);
  return hipGetLastError();
}

hipError_t LaunchRaygenShadow(const gpugi_ray* prim, const gpugi_hit* hits, const uint32_t* list,
                              const uint32_t* count, const float* lights, uint32_t numLights, gpugi_ray* rays,
                              size_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_raygen_shadow, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                     (const float4*)prim, (const int4*)hits, list, count, lights, numLights, (float4*)rays,
                     (uint32_t)n);
  return hipGetLastError();
}

}  // namespace gpugi
