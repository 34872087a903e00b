// Mean squared error of the accumulation buffer against a reference image
// (TextureMSE::ComputeCurrentMSE, texturemse.cpp:39-66, with the per-pixel
// error of shader/utils/sqerror.frag).  Two deterministic passes: fixed-size
// blocks sum their pixels' errors in double, then one block sums the block
// sums in index order.
#include "launch.h"

namespace gpugi {

constexpr int kMseThreads = 256;
constexpr int kMseMaxBlocks = 1024;

__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = kMseThreads / 2; off > 0; off >>= 1) {
    if (t < off) sh[t] += sh[t + off];
    __syncthreads();
  }
  return sh[0];
}

// Per reference pixel: the render texel under the pixel's centre (nearest
// sampling at texcoord (x + 0.5) / refW), color / divisor - reference, and
// dot(diff, diff), all in float as the fragment shader.
__global__ __launch_bounds__(kMseThreads) void k_mse_partial(const float4* __restrict__ accum, int width,
                                                             int height, float divisor,
                                                             const float* __restrict__ ref, int refW, int refH,
                                                             double* __restrict__ partial) {
  __shared__ double sh[kMseThreads];
  const size_t n = (size_t)refW * refH;
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * kMseThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kMseThreads) {
    const int x = (int)(i % (size_t)refW), y = (int)(i / (size_t)refW);
    const int rx = (int)(((2 * (int64_t)x + 1) * width) / (2 * (int64_t)refW));
    const int ry = (int)(((2 * (int64_t)y + 1) * height) / (2 * (int64_t)refH));
    const float4 c = accum[(size_t)ry * width + rx];
    const float dx = c.x / divisor - ref[3 * i + 0];
    const float dy = c.y / divisor - ref[3 * i + 1];
    const float dz = c.z / divisor - ref[3 * i + 2];
    acc += (double)(dx * dx + dy * dy + dz * dz);
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kMseThreads) void k_mse_final(const double* __restrict__ partial, int blocks,
                                                           size_t n, float* __restrict__ out) {
  __shared__ double sh[kMseThreads];
  double acc = 0.0;
  for (int b = threadIdx.x; b < blocks; b += kMseThreads) acc += partial[b];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) *out = (float)(s / (double)n);
}

size_t ImageMseTempBytes() { return kMseMaxBlocks * sizeof(double); }

hipError_t LaunchImageMse(const float4* accum, int width, int height, uint32_t divisor, const float* ref,
                          int refW, int refH, void* temp, float* out, hipStream_t stream) {
  if (!accum || !ref || !temp || !out || width <= 0 || height <= 0 || refW <= 0 || refH <= 0)
    return hipErrorInvalidValue;
  const size_t n = (size_t)refW * refH;
  const int blocks = (int)std::min<size_t>(kMseMaxBlocks, (n + kMseThreads - 1) / kMseThreads);
  double* partial = static_cast<double*>(temp);
  hipLaunchKernelGGL(k_mse_partial, dim3(blocks), dim3(kMseThreads), 0, stream, accum, width, height,
                     (float)divisor, ref, refW, refH, partial);
  hipLaunchKernelGGL(k_mse_final, dim3(1), dim3(kMseThreads), 0, stream, partial, blocks, n, out);
  return hipGetLastError();
}

}  // namespace gpugi
