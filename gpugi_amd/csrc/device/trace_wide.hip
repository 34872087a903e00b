// Persistent-wave ray queries on the 4-wide quantized BVH.
//
// Scheduling: workgroups are dispatched round-robin over the 8 XCDs, so block
// b runs on XCD b % 8; XCD x owns the x-th contiguous eighth of the ray array
// (neighbouring rays start close together, so they share that XCD's L2), and
// its waves take that range's 64-ray batches in order from one atomic counter
// per XCD, so no wave sits idle while its XCD has batches left.
//
// Each wave runs its 64 lanes one box or leaf pass at a time (trace_wide.h);
// a leaf pass is chosen when at least `wait` lanes have only a pending leaf.
// When at least `refill` lanes have finished (ballot + popcount), the idle
// lanes take the next rays of the wave's current batch (and of the next one it
// takes) in one go, so a single long ray no longer holds 63 idle lanes;
// refill = 64 reproduces batch-at-a-time processing.
// The LDS stack depth is a template parameter: deeper stacks spill less but
// cap the waves per CU (160 KiB LDS per CU); the autotuner picks it.
// Compiled with -ffp-contract=off: the triangle test must round exactly like
// the reference formula.
#include <hip/hip_runtime.h>

#include "launch.h"
#include "trace_wide.h"

namespace gpugi {

constexpr uint32_t kNumXcd = 8;

__global__ void k_wide_queue_reset(uint32_t* q) {
  if (threadIdx.x < (unsigned)kWideQueueWords) q[threadIdx.x] = 0u;
}

template <bool ANY, int B, int S, bool WT>
__global__ __launch_bounds__(B) void k_trace_wide(WideScene s, const float4* __restrict__ rays,
                                                  int4* __restrict__ hits, uint32_t n, int refill, int wait,
                                                  const uint32_t* __restrict__ nDev) {
  __shared__ uint2 lds[S * B];
  if (nDev) n = min(*nDev, n);  // the ray count as the device holds it (a queue length)
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const uint32_t gtid = blockIdx.x * B + tid;
  // spill laid out [depth][thread] over the whole grid (s.numThreads == gridDim.x * B)
  const WideStack stk{as_lds(lds), tid, s.spill + gtid, gridDim.x * (uint32_t)B};

  const uint32_t xcd = blockIdx.x % kNumXcd;
  const uint32_t batches = (n + 63u) / 64u;
  const uint32_t perX = (batches + kNumXcd - 1) / kNumXcd;
  const uint32_t bBegin = xcd * perX;
  const uint32_t bEnd = min(batches, bBegin + perX);
  uint32_t* counter = s.queue + xcd * kWideQueueStride;
  const uint64_t below = (1ull << lane) - 1ull;

  // wave-uniform: the batch being handed out and how many of its rays are gone
  uint32_t curBatch = 0, curPos = 64;
  bool exhausted = bBegin >= bEnd;
  bool active = false;
  uint32_t ri = 0;
  WideState st;
  for (;;) {
    uint64_t act = __ballot(active);
    const uint64_t idle = ~act;  // every lane of the wave runs this loop
    const uint32_t nidle = (uint32_t)__popcll(idle);
    if (!exhausted && (nidle >= (uint32_t)refill || nidle == 64u)) {
      const uint32_t avail = 64u - curPos;
      uint32_t next = bEnd;
      if (avail < nidle) {  // take the XCD's next batch
        uint32_t v = 0;
        if (lane == 0) v = atomicAdd(counter, 1u);
        v = (uint32_t)__shfl((int)v, 0);
        if (v < bEnd - bBegin) next = bBegin + v;
      }
      if (!active) {
        const uint32_t rank = (uint32_t)__popcll(idle & below);
        uint64_t r = n;
        if (rank < avail) r = (uint64_t)curBatch * 64u + curPos + rank;
        else if (next < bEnd) r = (uint64_t)next * 64u + (rank - avail);
        if (r < n) {
          const float4 r0 = rays[2 * (size_t)r], r1 = rays[2 * (size_t)r + 1];
          wide_init<WT>(st, gg::v3(r0.x, r0.y, r0.z), gg::v3(r1.x, r1.y, r1.z), r0.w);
          ri = (uint32_t)r;
          active = true;
        }
      }
      if (avail >= nidle) {
        curPos += nidle;
      } else if (next < bEnd) {
        curBatch = next;
        curPos = nidle - avail;
      } else {
        exhausted = true;
      }
      act = __ballot(active);
    }
    if (act == 0) {
      if (exhausted) break;
      continue;
    }
    const bool leafPass = wide_choose_leaf_pass(act, st, wait);
    if (active && wide_pass<ANY, B, S, WT>(s, st, leafPass, stk)) {
      if (ANY) hits[ri] = make_int4(st.found ? 1 : 0, 0, 0, 0);
      else hits[ri] = make_int4(st.bestSlot, __float_as_int(st.tbest), __float_as_int(st.bu), __float_as_int(st.bv));
      active = false;
    }
  }
}

namespace {

struct Variant {
  int stack;
  const void* closest;
  const void* any;
};

template <int S>
Variant MakeVariant() {
  return Variant{S, reinterpret_cast<const void*>(&k_trace_wide<false, kWideBlock, S, false>),
                 reinterpret_cast<const void*>(&k_trace_wide<true, kWideBlock, S, false>)};
}

const Variant* FindVariant(int stack) {
  static const Variant table[] = {MakeVariant<4>(),  MakeVariant<8>(),  MakeVariant<10>(),
                                  MakeVariant<12>(), MakeVariant<16>(), MakeVariant<24>()};
  for (const Variant& v : table)
    if (v.stack == stack) return &v;
  return nullptr;
}

template <int S, bool WT>
void LaunchS(const WideScene& s, int query, const gpugi_ray* rays, gpugi_hit* hits, size_t n, int refill,
             int wait, hipStream_t stream, const uint32_t* nDev) {
  dim3 grid(s.numThreads / kWideBlock);
  if (query == GPUGI_ANY_HIT)
    hipLaunchKernelGGL((k_trace_wide<true, kWideBlock, S, WT>), grid, dim3(kWideBlock), 0, stream, s,
                       (const float4*)rays, (int4*)hits, (uint32_t)n, refill, wait, nDev);
  else
    hipLaunchKernelGGL((k_trace_wide<false, kWideBlock, S, WT>), grid, dim3(kWideBlock), 0, stream, s,
                       (const float4*)rays, (int4*)hits, (uint32_t)n, refill, wait, nDev);
}

template <int S>
void LaunchS(const WideScene& s, int query, const gpugi_ray* rays, gpugi_hit* hits, size_t n, int refill,
             int wait, bool watertight, hipStream_t stream, const uint32_t* nDev) {
  if (watertight) LaunchS<S, true>(s, query, rays, hits, n, refill, wait, stream, nDev);
  else LaunchS<S, false>(s, query, rays, hits, n, refill, wait, stream, nDev);
}

}  // namespace

const int kWideStackVariants[kNumWideStackVariants] = {4, 8, 10, 12, 16, 24};

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
                           size_t n, int refill, int wait, bool watertight, hipStream_t stream,
                           const uint32_t* nDevice) {
  if (n == 0) return hipSuccess;
  if (n > 0xffffffffull - 64) return hipErrorInvalidValue;
  // the static schedule needs a whole number of blocks per XCD
  if (s.numThreads == 0 || s.numThreads % (kWideBlock * kNumXcd) != 0) return hipErrorInvalidConfiguration;
  if (refill < 1 || refill > 64 || wait < 1 || wait > 64) return hipErrorInvalidValue;
  if (!s.queue) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_wide_queue_reset, dim3(1), dim3(kWideQueueWords), 0, stream, s.queue);
  switch (stack) {
    case 4: LaunchS<4>(s, query, rays, hits, n, refill, wait, watertight, stream, nDevice); break;
    case 8: LaunchS<8>(s, query, rays, hits, n, refill, wait, watertight, stream, nDevice); break;
    case 10: LaunchS<10>(s, query, rays, hits, n, refill, wait, watertight, stream, nDevice); break;
    case 12: LaunchS<12>(s, query, rays, hits, n, refill, wait, watertight, stream, nDevice); break;
    case 16: LaunchS<16>(s, query, rays, hits, n, refill, wait, watertight, stream, nDevice); break;
    case 24: LaunchS<24>(s, query, rays, hits, n, refill, wait, watertight, stream, nDevice); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace gpugi
