// Host-side launch wrappers of the HIP kernels (defined in the .hip files).
#pragma once
#include <hip/hip_runtime.h>

#include "gpugi_rt.h"
#include "trace_ref.h"
#include "wide_bvh.h"

namespace gpugi {

// watertight: the watertight triangle test in place of the reference one.
// nDevice (optional): device word holding the ray count, at most n (the grid is sized for n).
hipError_t LaunchTraceRef(const RefScene& s, int query, const gpugi_ray* rays, gpugi_hit* hits, size_t n,
                          bool watertight, hipStream_t stream, const uint32_t* nDevice = nullptr);
// The TRACERAY_IMPORTANCE_BREAK walk (ib.importance: one float2 per node);
// nodes (optional, closest hit) receive {node, importance bits, size^2 bits, 0}.
hipError_t LaunchTraceRefImportance(const RefScene& s, int query, const gpugi_ray* rays, gpugi_hit* hits,
                                    int4* nodes, size_t n, const ImportanceBreak& ib, bool watertight,
                                    hipStream_t stream);
hipError_t LaunchTraceRefCounts(const RefScene& s, int query, const gpugi_ray* rays, uint32_t* nbox,
                                uint32_t* ntri, size_t n, hipStream_t stream);
// Wide kernel with an LDS stack of `stack` entries per lane (one of kWideStackVariants).
// refill: a wave reloads idle lanes once at least this many (1..64) are idle.
// wait: a wave runs a leaf pass once at least this many (1..64) lanes wait for one.
// watertight: the watertight triangle test (needs the raw-vertex triangle layout).
// nDevice (optional): device word holding the ray count, at most n, read by the kernel.
hipError_t LaunchTraceWide(const WideScene& s, int stack, int query, const gpugi_ray* rays, gpugi_hit* hits,
                           size_t n, int refill, int wait, bool watertight, hipStream_t stream,
                           const uint32_t* nDevice = nullptr);
// Resident blocks per CU of the wide kernel variant (sizes the persistent grid).
hipError_t WideBlocksPerCU(int stack, int* blocks);
constexpr int kNumWideStackVariants = 6;
extern const int kWideStackVariants[kNumWideStackVariants];

// The allocation-free trace's scratch (trace_noalloc.hip): GPUGI_NOALLOC_SCRATCH_BYTES
// per device, the first kNoAllocQueueBytes for the wide kernel's batch counters,
// the rest for its spill stacks.  *out = the array's device address on `device`
// (looked up once per device with that device current).
constexpr size_t kNoAllocQueueBytes = 1024;
hipError_t NoAllocScratchAddress(int device, void** out);

// VALU calibration kernel: blocks x 256 lanes, each 64 * iters FP32 FMAs;
// out receives one float per lane.
hipError_t LaunchProbeValu(float* out, uint32_t blocks, uint32_t iters, hipStream_t stream);

hipError_t LaunchRaygenPrimary(const gpugi_camera& cam, int width, int height, uint32_t seed,
                               gpugi_ray* rays, size_t n, hipStream_t stream);
// Builds the list of primary rays that hit (in index order) and the diffuse / shadow sets.
hipError_t LaunchHitCompaction(const gpugi_hit* hits, size_t n, uint32_t* list, uint32_t* count,
                               void* temp, size_t* tempBytes, hipStream_t stream);
hipError_t LaunchRaygenDiffuse(const RefScene& s, const gpugi_ray* prim, const gpugi_hit* hits,
                               const uint32_t* list, const uint32_t* count, uint32_t seed, gpugi_ray* rays,
                               size_t n, hipStream_t stream);
hipError_t LaunchRaygenShadow(const gpugi_ray* prim, const gpugi_hit* hits, const uint32_t* list,
                              const uint32_t* count, const float* lights, uint32_t numLights,
                              gpugi_ray* rays, size_t n, hipStream_t stream);

// TextureMSE::ComputeCurrentMSE (texturemse.cpp:39-66): *out = mean over the
// refW x refH reference pixels (RGB32F rows) of dot(d, d), d = accum / divisor
// - reference, the accumulation buffer sampled nearest at each pixel centre.
// temp holds ImageMseTempBytes() bytes.
size_t ImageMseTempBytes();
hipError_t LaunchImageMse(const float4* accum, int width, int height, uint32_t divisor, const float* ref,
                          int refW, int refH, void* temp, float* out, hipStream_t stream);
}  // namespace gpugi
