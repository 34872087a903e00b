// Calibration kernel of the VALU roofline: eight independent FP32 FMA chains
// per lane (64 v_fma_f32 per loop trip, no packed-FP32 forms), enough waves
// to fill every SIMD.  Its VALU wave-instructions per CU cycle under
// rocprofv3 (SQ_INSTS_VALU / GRBM_GUI_ACTIVE) measure the issue rate that
// bounds a VALU-heavy kernel such as the traversal.
#include <hip/hip_runtime.h>

#include "launch.h"

namespace gpugi {

constexpr int kProbeBlock = 256;

__global__ __launch_bounds__(kProbeBlock) void k_probe_valu(float* __restrict__ out, uint32_t iters, float b, float c) {
  const uint32_t t = blockIdx.x * kProbeBlock + threadIdx.x;
  float a0 = (float)t, a1 = a0 + 1.0f, a2 = a0 + 2.0f, a3 = a0 + 3.0f;
  float a4 = a0 + 4.0f, a5 = a0 + 5.0f, a6 = a0 + 6.0f, a7 = a0 + 7.0f;
  for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      a0 = fmaf(a0, b, c);
      a1 = fmaf(a1, b, c);
      a2 = fmaf(a2, b, c);
      a3 = fmaf(a3, b, c);
      a4 = fmaf(a4, b, c);
      a5 = fmaf(a5, b, c);
      a6 = fmaf(a6, b, c);
      a7 = fmaf(a7, b, c);
    }
  }
  out[t] = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
}

hipError_t LaunchProbeValu(float* out, uint32_t blocks, uint32_t iters, hipStream_t stream) {
  if (!out || blocks == 0) return hipErrorInvalidValue;
  // b < 1 keeps the chains finite for any trip count
  hipLaunchKernelGGL(k_probe_valu, dim3(blocks), dim3(kProbeBlock), 0, stream, out, iters, 0.999f, 0.001f);
  return hipGetLastError();
}

}  // namespace gpugi
