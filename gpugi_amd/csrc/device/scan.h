// Deterministic exclusive prefix sum of uint32 arrays (device).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace gpugi {

// Temporary bytes LaunchExclusiveScanU32 needs for n elements.
size_t ExclusiveScanTempBytes(size_t n);
// out[i] = in[0] + ... + in[i-1]; *total = sum of all.  out may alias in.
hipError_t LaunchExclusiveScanU32(const uint32_t* in, uint32_t* out, size_t n, void* temp, uint32_t* total,
                                  hipStream_t stream);
// *p += c on the device (keeps a counter on the device without a host round trip).
void AddConstantU32(uint32_t* p, uint32_t c, hipStream_t stream);

}  // namespace gpugi
