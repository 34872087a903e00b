// Three-phase exclusive scan: per-block sums, one-block scan of the sums,
// per-block scan plus the block's offset.  Wave-level steps use ballot-free
// LDS Hillis-Steele scans; every phase is deterministic.
#include "scan.h"

namespace gpugi {

constexpr int kScanThreads = 1024;

__device__ __forceinline__ uint32_t block_inclusive_scan(uint32_t v, uint32_t* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const uint32_t add = t >= off ? sh[t - off] : 0u;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  return sh[t];
}

__global__ __launch_bounds__(kScanThreads) void k_scan_sums(const uint32_t* __restrict__ in, size_t n,
                                                            uint32_t* __restrict__ sums) {
  __shared__ uint32_t sh[kScanThreads];
  const size_t i = (size_t)blockIdx.x * kScanThreads + threadIdx.x;
  const uint32_t v = i < n ? in[i] : 0u;
  const uint32_t incl = block_inclusive_scan(v, sh);
  if (threadIdx.x == kScanThreads - 1) sums[blockIdx.x] = incl;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_spine(uint32_t* __restrict__ sums, uint32_t numBlocks,
                                                             uint32_t* __restrict__ total) {
  __shared__ uint32_t sh[kScanThreads];
  const uint32_t per = (numBlocks + kScanThreads - 1) / kScanThreads;
  const uint32_t begin = threadIdx.x * per;
  uint32_t local = 0;
  for (uint32_t k = 0; k < per; ++k)
    if (begin + k < numBlocks) local += sums[begin + k];
  const uint32_t incl = block_inclusive_scan(local, sh);
  uint32_t run = incl - local;
  for (uint32_t k = 0; k < per; ++k) {
    if (begin + k < numBlocks) {
      const uint32_t c = sums[begin + k];
      sums[begin + k] = run;
      run += c;
    }
  }
  if (threadIdx.x == kScanThreads - 1) *total = incl;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const uint32_t* in, size_t n,
                                                             const uint32_t* __restrict__ sums, uint32_t* out) {
  __shared__ uint32_t sh[kScanThreads];
  const size_t i = (size_t)blockIdx.x * kScanThreads + threadIdx.x;
  const uint32_t v = i < n ? in[i] : 0u;
  const uint32_t incl = block_inclusive_scan(v, sh);
  if (i < n) out[i] = sums[blockIdx.x] + incl - v;
}

__global__ void k_add_u32(uint32_t* p, uint32_t c) { *p += c; }

size_t ExclusiveScanTempBytes(size_t n) {
  const size_t blocks = (n + kScanThreads - 1) / kScanThreads;
  return ((blocks * 4 + 255) / 256) * 256;
}

hipError_t LaunchExclusiveScanU32(const uint32_t* in, uint32_t* out, size_t n, void* temp, uint32_t* total,
                                  hipStream_t stream) {
  const uint32_t blocks = (uint32_t)((n + kScanThreads - 1) / kScanThreads);
  if (blocks == 0) {
    hipLaunchKernelGGL(k_scan_spine, dim3(1), dim3(kScanThreads), 0, stream, (uint32_t*)temp, 0u, total);
    return hipGetLastError();
  }
  uint32_t* sums = (uint32_t*)temp;
  hipLaunchKernelGGL(k_scan_sums, dim3(blocks), dim3(kScanThreads), 0, stream, in, n, sums);
  hipLaunchKernelGGL(k_scan_spine, dim3(1), dim3(kScanThreads), 0, stream, sums, blocks, total);
  hipLaunchKernelGGL(k_scan_apply, dim3(blocks), dim3(kScanThreads), 0, stream, in, n, sums, out);
  return hipGetLastError();
}

void AddConstantU32(uint32_t* p, uint32_t c, hipStream_t stream) {
  hipLaunchKernelGGL(k_add_u32, dim3(1), dim3(1), 0, stream, p, c);
}

}  // namespace gpugi
