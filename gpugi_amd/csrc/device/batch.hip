// Batched rendering of many scenes in one persistent launch (no reference
// counterpart: the reference renders one scene per GL context).
//
// Every scene keeps its own BVH, camera, light samples, resolution and
// accumulation buffer; their arrays are packed back to back in one set of
// device buffers and each scene's views point into them, so the traversal
// code runs unchanged on scene-local indices.  The work of a frame is the
// list of all 8x8 pixel tiles of all scenes; persistent workgroups take
// `grab` tiles at a time from one shared atomic counter, so small and large
// scenes interleave and the GPU stays full until the last tile.  A frame
// begins with a one-thread kernel that resets the tile counter and derives
// FrameSeed = WangHash(frame index) on the device, so one captured frame can
// be replayed as a hipGraph any number of times.  Each pixel runs the path
// tracer body of pathtracer.comp (pt_pixel), identical to a single-scene draw.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "gpugi_rt.h"
#include "integrator_common.h"
#include "wide_bvh.h"

using namespace gpugi;

namespace gpugi {
namespace {

struct BatchSlot {
  DrawParams p;       // views into the packed buffers; frameSeed is taken from the device counters
  uint32_t tileBase;  // first tile of this scene in the frame's tile list
  uint32_t tilesX, tilesY;
};

// counters: u32[0] tile cursor, u32[1] frame index, u32[2] FrameSeed, u64 at byte 16: ray queries.
struct BatchCounters {
  uint32_t tile, frame, seed, pad;
  unsigned long long rays;
};

__global__ void k_batch_begin(BatchCounters* c) {
  c->tile = 0;
  c->seed = gg::wanghash(c->frame);
  c->frame = c->frame + 1;
}

__global__ void k_batch_set_frame(BatchCounters* c, uint32_t frame) {
  c->tile = 0;
  c->frame = frame;
  c->seed = 0;
  c->rays = 0;
}

__global__ void k_fill_f4(float4* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// LDS stack depths of the wide walk the batch kernels are built for; the
// spill slices are sized for the smallest.
constexpr int kBatchStacks[3] = {16, 24, 32};

template <bool WIDE, int S>
__device__ __forceinline__ void render_tile(const BatchSlot& sl, uint32_t local, uint32_t seed, uint2* stk,
                                            uint2* spill, BatchCounters* c) {
  DrawParams p = sl.p;
  p.frameSeed = seed;
  const int tid = threadIdx.x;
  const int gx = (int)(local % sl.tilesX) * kTile + (tid % kTile);
  const int gy = (int)(local / sl.tilesX) * kTile + (tid / kTile);
  TraceCtx ctx{&p, stk, tid, spill};
  pt_pixel<WIDE, S>(p, ctx, gx, gy, &c->rays);
}

// Five waves per SIMD, as the single-scene path tracer (no spill here).
template <bool WIDE, int S>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_batch_persistent(
    const BatchSlot* __restrict__ slots, uint32_t numScenes, uint32_t totalTiles, BatchCounters* c, uint2* spillBase,
    uint32_t spillDepth, uint32_t grab) {
  __shared__ uint2 stk[WIDE ? S * kTileThreads : 1];
  __shared__ uint32_t sBase;
  const int tid = threadIdx.x;
  uint2* spill = WIDE ? spillBase + ((size_t)blockIdx.x * kTileThreads + tid) * spillDepth : nullptr;
  const uint32_t seed = c->seed;
  for (;;) {
    if (tid == 0) sBase = atomicAdd(&c->tile, grab);
    __syncthreads();
    const uint32_t base = sBase;
    __syncthreads();
    if (base >= totalTiles) break;
    const uint32_t end = min(totalTiles, base + grab);
    for (uint32_t t = base; t < end; ++t) {
      uint32_t s = 0;
      while (s + 1 < numScenes && slots[s + 1].tileBase <= t) ++s;
      render_tile<WIDE, S>(slots[s], t - slots[s].tileBase, seed, stk, spill, c);
    }
  }
}

// One scene, one tile per workgroup (the "separate launches" baseline).
template <bool WIDE, int S>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_batch_scene(
    const BatchSlot* __restrict__ slot, BatchCounters* c, uint2* spillBase, uint32_t spillDepth) {
  __shared__ uint2 stk[WIDE ? S * kTileThreads : 1];
  uint2* spill = WIDE ? spillBase + ((size_t)blockIdx.x * kTileThreads + threadIdx.x) * spillDepth : nullptr;
  render_tile<WIDE, S>(*slot, blockIdx.x, c->seed, stk, spill, c);
}

// The launch of one frame's tiles for one traversal and stack depth.
template <bool WIDE, int S>
void LaunchTiles(bool persistent, dim3 persistentGrid, const std::vector<BatchSlot>& hSlots, BatchSlot* dSlots,
                 uint32_t totalTiles, BatchCounters* c, uint2* spill, uint32_t spillDepth, uint32_t grab,
                 hipStream_t stream) {
  if (persistent) {
    hipLaunchKernelGGL((k_batch_persistent<WIDE, S>), persistentGrid, dim3(kTileThreads), 0, stream, dSlots,
                       (uint32_t)hSlots.size(), totalTiles, c, spill, spillDepth, grab);
    return;
  }
  for (size_t i = 0; i < hSlots.size(); ++i) {
    const dim3 grid(hSlots[i].tilesX * hSlots[i].tilesY);
    if (grid.x == 0) continue;
    hipLaunchKernelGGL((k_batch_scene<WIDE, S>), grid, dim3(kTileThreads), 0, stream, dSlots + i, c, spill,
                       spillDepth);
  }
}

}  // namespace
}  // namespace gpugi

struct BatchSceneHost {
  std::vector<float> hier, pos, vinfo;
  std::vector<int32_t> tris;
  std::vector<gpugi_material> mats;
  std::vector<gpugi_light_sample> lights;
  gpugi_camera camera{};
  int width = 0, height = 0, maxPathLength = 16;
  uint32_t numNodes = 0, numLeaves = 0, numVertices = 0;
  WideBVHHost wide;
};

struct gpugi_batch {
  int device = 0;
  int numCUs = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string error;
  std::vector<BatchSceneHost> scenes;
  bool finalized = false;
  void* dPacked = nullptr;  // all scene arrays
  BatchSlot* dSlots = nullptr;
  BatchCounters* dCounters = nullptr;
  uint2* dSpill = nullptr;
  size_t spillThreads = 0;
  uint32_t spillDepth = 0;
  std::vector<BatchSlot> hSlots;
  std::vector<size_t> accumOffset;  // float4 index of each scene's accumulation buffer
  size_t accumCount = 0;
  float4* dAccum = nullptr;
  uint32_t totalTiles = 0;
  int kernel = GPUGI_KERNEL_WIDE;
  int grab = 2;
  int gridMul = 2;
  int stackDepth = 16;     // one of kBatchStacks
  int residentRef = 1;     // resident blocks per CU of the reference-walk kernel
  int residentWide[3] = {1, 1, 1};  // ... of the wide kernel per kBatchStacks entry
  hipGraph_t graph = nullptr;
  hipGraphExec_t graphExec = nullptr;
};

namespace {

int BFail(gpugi_batch* b, int code, const std::string& msg) {
  if (b) b->error = msg;
  return code;
}

#define BHIP(b, expr)                                                                           \
  do {                                                                                          \
    hipError_t _e = (expr);                                                                     \
    if (_e != hipSuccess) return BFail(b, GPUGI_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

size_t Align256(size_t x) { return (x + 255) & ~(size_t)255; }

void DropGraph(gpugi_batch* b) {
  if (b->graphExec) hipGraphExecDestroy(b->graphExec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->graphExec = nullptr;
  b->graph = nullptr;
}

int StackIndex(int depth) {
  for (int i = 0; i < 3; ++i)
    if (kBatchStacks[i] == depth) return i;
  return -1;
}

int PersistentGrid(gpugi_batch* b) {
  const int resident = b->kernel == GPUGI_KERNEL_WIDE ? b->residentWide[StackIndex(b->stackDepth)] : b->residentRef;
  return std::max(1, resident * b->numCUs * b->gridMul / 4);
}

// Spill slices for every thread any launch of the current configuration uses.
int EnsureSpill(gpugi_batch* b) {
  size_t threads = (size_t)PersistentGrid(b) * kTileThreads;
  for (const BatchSlot& s : b->hSlots) threads = std::max(threads, (size_t)s.tilesX * s.tilesY * kTileThreads);
  if (b->kernel != GPUGI_KERNEL_WIDE || b->spillDepth == 0) return GPUGI_OK;
  if (b->spillThreads >= threads) return GPUGI_OK;
  if (b->dSpill) hipFree(b->dSpill);
  b->dSpill = nullptr;
  b->spillThreads = 0;
  BHIP(b, hipMalloc((void**)&b->dSpill, threads * b->spillDepth * sizeof(uint2)));
  b->spillThreads = threads;
  return GPUGI_OK;
}

int LaunchFrame(gpugi_batch* b, bool persistent) {
  hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(1), 0, b->stream, b->dCounters);
  const dim3 grid(PersistentGrid(b));
  auto launch = [&](auto fn) {
    fn(persistent, grid, b->hSlots, b->dSlots, b->totalTiles, b->dCounters, b->dSpill, b->spillDepth,
       (uint32_t)b->grab, b->stream);
  };
  if (b->kernel != GPUGI_KERNEL_WIDE) launch(LaunchTiles<false, 16>);
  else if (b->stackDepth == 16) launch(LaunchTiles<true, 16>);
  else if (b->stackDepth == 24) launch(LaunchTiles<true, 24>);
  else launch(LaunchTiles<true, 32>);
  BHIP(b, hipGetLastError());
  return GPUGI_OK;
}

int Ready(gpugi_batch* b, const char* fn) {
  if (!b) return GPUGI_ERR_ARG;
  if (!b->finalized) return BFail(b, GPUGI_ERR_STATE, std::string(fn) + ": call gpugi_batch_finalize first");
  BHIP(b, hipSetDevice(b->device));
  return GPUGI_OK;
}

}  // namespace

extern "C" {

int gpugi_batch_create(int device, gpugi_batch** out) {
  if (!out) return GPUGI_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return GPUGI_ERR_NODEVICE;
  auto* b = new (std::nothrow) gpugi_batch();
  if (!b) return GPUGI_ERR_NOMEM;
  b->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
      hipDeviceGetAttribute(&b->numCUs, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      hipMalloc((void**)&b->dCounters, sizeof(BatchCounters)) != hipSuccess) {
    delete b;
    return GPUGI_ERR_HIP;
  }
  *out = b;
  return GPUGI_OK;
}

void gpugi_batch_destroy(gpugi_batch* b) {
  if (!b) return;
  hipSetDevice(b->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  DropGraph(b);
  for (void* p : {(void*)b->dPacked, (void*)b->dSlots, (void*)b->dCounters, (void*)b->dSpill, (void*)b->dAccum})
    if (p) hipFree(p);
  if (b->ev0) hipEventDestroy(b->ev0);
  if (b->ev1) hipEventDestroy(b->ev1);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

const char* gpugi_batch_last_error(const gpugi_batch* b) { return b ? b->error.c_str() : "null batch"; }

int gpugi_batch_add(gpugi_batch* b, const gpugi_scene_desc* d, const gpugi_camera* camera, int width, int height,
                    const gpugi_light_sample* lights, uint32_t num_lights, int max_path_length) {
  if (!b || !d || !camera || !lights) return BFail(b, GPUGI_ERR_ARG, "gpugi_batch_add: null argument");
  if (b->finalized) return BFail(b, GPUGI_ERR_STATE, "gpugi_batch_add: the batch is finalized");
  if (width <= 0 || height <= 0 || (size_t)width * height > (size_t)1 << 26 || num_lights == 0 ||
      max_path_length < 1 || max_path_length > 64 || d->triangles_per_leaf != 2 || d->num_nodes == 0 ||
      d->num_leaves == 0 || d->num_vertices == 0 || d->num_materials == 0 || d->num_materials > 256)
    return BFail(b, GPUGI_ERR_ARG, "gpugi_batch_add: invalid scene, size or light samples");
  // every index the kernels follow must stay inside this scene's arrays
  for (uint32_t i = 0; i < d->num_nodes; ++i) {
    uint32_t fc, esc;
    std::memcpy(&fc, d->hierarchy + (size_t)i * 8 + 3, 4);
    std::memcpy(&esc, d->hierarchy + (size_t)i * 8 + 7, 4);
    const uint32_t child = fc & 0x7fffffffu;
    if ((fc & 0x80000000u) ? child >= d->num_leaves : child >= d->num_nodes) return BFail(b, GPUGI_ERR_FORMAT, "gpugi_batch_add: child index out of range");
    if (esc >= d->num_nodes) return BFail(b, GPUGI_ERR_FORMAT, "gpugi_batch_add: escape index out of range");
  }
  for (size_t i = 0; i < (size_t)d->num_leaves * 2; ++i) {
    const int32_t* t = d->triangles + i * 4;
    if (t[0] == t[1]) continue;
    for (int k = 0; k < 3; ++k)
      if (t[k] < 0 || (uint32_t)t[k] >= d->num_vertices) return BFail(b, GPUGI_ERR_FORMAT, "gpugi_batch_add: vertex index out of range");
    if (t[3] < 0 || (uint32_t)t[3] >= d->num_materials) return BFail(b, GPUGI_ERR_FORMAT, "gpugi_batch_add: material index out of range");
  }
  BatchSceneHost s;
  s.hier.assign(d->hierarchy, d->hierarchy + (size_t)d->num_nodes * 8);
  s.tris.assign(d->triangles, d->triangles + (size_t)d->num_leaves * 8);
  s.pos.assign(d->positions, d->positions + (size_t)d->num_vertices * 3);
  s.vinfo.assign(d->vertex_info, d->vertex_info + (size_t)d->num_vertices * 4);
  s.mats.assign(d->materials, d->materials + d->num_materials);
  s.lights.assign(lights, lights + num_lights);
  s.camera = *camera;
  s.width = width;
  s.height = height;
  s.maxPathLength = max_path_length;
  s.numNodes = d->num_nodes;
  s.numLeaves = d->num_leaves;
  s.numVertices = d->num_vertices;
  const std::string err = BuildWideBVH(*d, 2, &s.wide);
  if (!err.empty()) return BFail(b, GPUGI_ERR_FORMAT, "gpugi_batch_add: " + err);
  b->scenes.push_back(std::move(s));
  return GPUGI_OK;
}

int gpugi_batch_finalize(gpugi_batch* b) {
  if (!b) return GPUGI_ERR_ARG;
  if (b->finalized) return GPUGI_OK;
  if (b->scenes.empty()) return BFail(b, GPUGI_ERR_STATE, "gpugi_batch_finalize: no scenes");
  BHIP(b, hipSetDevice(b->device));
  // packed layout: per scene nodes, slots, positions, vertex info, materials, lights, wide nodes, wide tris
  std::vector<size_t> off;
  size_t total = 0;
  auto put = [&](size_t bytes) {
    off.push_back(total);
    total += Align256(bytes);
  };
  for (const BatchSceneHost& s : b->scenes) {
    put(s.hier.size() * 4);
    put(s.tris.size() * 4);
    put(s.pos.size() * 4);
    put(s.vinfo.size() * 4);
    put(s.mats.size() * sizeof(gpugi_material));
    put(s.lights.size() * sizeof(gpugi_light_sample));
    put(s.wide.nodes.size() * sizeof(WideNodeHost));
    put(s.wide.tris.size() * sizeof(WideTriHost));
  }
  BHIP(b, hipMalloc(&b->dPacked, total));
  char* base = (char*)b->dPacked;
  size_t k = 0, accum = 0;
  uint32_t tiles = 0;
  b->spillDepth = 0;
  for (const BatchSceneHost& s : b->scenes)
    b->spillDepth = std::max(b->spillDepth, s.wide.maxStack > (uint32_t)kBatchStacks[0] ? s.wide.maxStack - kBatchStacks[0] : 0u);
  for (const BatchSceneHost& s : b->scenes) {
    const void* src[8] = {s.hier.data(), s.tris.data(), s.pos.data(), s.vinfo.data(), s.mats.data(), s.lights.data(),
                          s.wide.nodes.data(), s.wide.tris.data()};
    const size_t bytes[8] = {s.hier.size() * 4, s.tris.size() * 4, s.pos.size() * 4, s.vinfo.size() * 4,
                             s.mats.size() * sizeof(gpugi_material), s.lights.size() * sizeof(gpugi_light_sample),
                             s.wide.nodes.size() * sizeof(WideNodeHost), s.wide.tris.size() * sizeof(WideTriHost)};
    char* ptr[8];
    for (int i = 0; i < 8; ++i) {
      ptr[i] = base + off[k + i];
      if (bytes[i]) BHIP(b, hipMemcpy(ptr[i], src[i], bytes[i], hipMemcpyHostToDevice));
    }
    k += 8;
    BatchSlot sl;
    DrawParams& p = sl.p;
    p.ref.nodes = (const float4*)ptr[0];
    p.ref.slots = (const int4*)ptr[1];
    p.ref.positions = (const float*)ptr[2];
    p.ref.vinfo = (const float4*)ptr[3];
    p.materials = (const gpugi_material*)ptr[4];
    p.lights = (const gpugi_light_sample*)ptr[5];
    p.numLights = (uint32_t)s.lights.size();
    p.wide.nodes = (const uint4*)ptr[6];
    p.wide.tris = (const float4*)ptr[7];
    p.wide.spill = nullptr;  // per thread, set in the kernels
    p.wide.spillDepth = b->spillDepth;
    p.wide.numThreads = 0;
    p.wide.errorFlag = &b->dCounters->pad;
    p.camera = s.camera;
    p.width = s.width;
    p.height = s.height;
    p.maxPathLength = s.maxPathLength;
    sl.tilesX = (uint32_t)(s.width / kTile);
    sl.tilesY = (uint32_t)(s.height / kTile);
    sl.tileBase = tiles;
    tiles += sl.tilesX * sl.tilesY;
    b->accumOffset.push_back(accum);
    accum += (size_t)s.width * s.height;
    b->hSlots.push_back(sl);
  }
  b->totalTiles = tiles;
  b->accumCount = accum;
  BHIP(b, hipMalloc((void**)&b->dAccum, accum * sizeof(float4)));
  for (size_t i = 0; i < b->hSlots.size(); ++i) b->hSlots[i].p.accum = b->dAccum + b->accumOffset[i];
  BHIP(b, hipMalloc((void**)&b->dSlots, b->hSlots.size() * sizeof(BatchSlot)));
  BHIP(b, hipMemcpy(b->dSlots, b->hSlots.data(), b->hSlots.size() * sizeof(BatchSlot), hipMemcpyHostToDevice));
  const void* wideKernels[3] = {(const void*)&k_batch_persistent<true, 16>, (const void*)&k_batch_persistent<true, 24>,
                                (const void*)&k_batch_persistent<true, 32>};
  for (int i = 0; i < 3; ++i) {
    int r = 0;
    BHIP(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&r, wideKernels[i], kTileThreads, 0));
    b->residentWide[i] = std::max(1, r);
  }
  int r = 0;
  BHIP(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&r, (const void*)&k_batch_persistent<false, 16>, kTileThreads, 0));
  b->residentRef = std::max(1, r);
  b->finalized = true;
  return gpugi_batch_reset(b, 0);
}

int gpugi_batch_config(gpugi_batch* b, int kernel, int tiles_per_grab, int grid_multiplier, int stack_depth) {
  if (int rc = Ready(b, "gpugi_batch_config")) return rc;
  if ((kernel != GPUGI_KERNEL_WIDE && kernel != GPUGI_KERNEL_REFERENCE) || tiles_per_grab < 1 ||
      tiles_per_grab > 64 || grid_multiplier < 1 || grid_multiplier > 16 || StackIndex(stack_depth) < 0)
    return BFail(b, GPUGI_ERR_ARG,
                 "gpugi_batch_config: kernel REFERENCE or WIDE, grab in [1, 64], multiplier in [1, 16], "
                 "stack depth 16, 24 or 32");
  b->kernel = kernel;
  b->grab = tiles_per_grab;
  b->gridMul = grid_multiplier;
  b->stackDepth = stack_depth;
  DropGraph(b);
  return EnsureSpill(b);
}

int gpugi_batch_reset(gpugi_batch* b, uint32_t frame_index) {
  if (int rc = Ready(b, "gpugi_batch_reset")) return rc;
  hipLaunchKernelGGL(k_batch_set_frame, dim3(1), dim3(1), 0, b->stream, b->dCounters, frame_index);
  hipLaunchKernelGGL(k_fill_f4, dim3((unsigned)((b->accumCount + 255) / 256)), dim3(256), 0, b->stream, b->dAccum,
                     b->accumCount);
  BHIP(b, hipGetLastError());
  BHIP(b, hipStreamSynchronize(b->stream));
  return EnsureSpill(b);
}

int gpugi_batch_draw(gpugi_batch* b, int persistent) {
  if (int rc = Ready(b, "gpugi_batch_draw")) return rc;
  if (int rc = EnsureSpill(b)) return rc;
  return LaunchFrame(b, persistent != 0);
}

int gpugi_batch_capture(gpugi_batch* b) {
  if (int rc = Ready(b, "gpugi_batch_capture")) return rc;
  if (int rc = EnsureSpill(b)) return rc;
  DropGraph(b);
  BHIP(b, hipStreamSynchronize(b->stream));
  BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int rc = LaunchFrame(b, true);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  BHIP(b, e);
  b->graph = g;
  BHIP(b, hipGraphInstantiate(&b->graphExec, b->graph, nullptr, nullptr, 0));
  return GPUGI_OK;
}

int gpugi_batch_replay(gpugi_batch* b, int times) {
  if (int rc = Ready(b, "gpugi_batch_replay")) return rc;
  if (!b->graphExec) return BFail(b, GPUGI_ERR_STATE, "gpugi_batch_replay: no captured frame");
  for (int i = 0; i < times; ++i) BHIP(b, hipGraphLaunch(b->graphExec, b->stream));
  return GPUGI_OK;
}

int gpugi_batch_run(gpugi_batch* b, int mode, int frames, float* ms) {
  if (int rc = Ready(b, "gpugi_batch_run")) return rc;
  if (mode < 0 || mode > 2 || frames < 1) return BFail(b, GPUGI_ERR_ARG, "gpugi_batch_run: mode 0..2, frames >= 1");
  if (int rc = EnsureSpill(b)) return rc;
  if (mode == 2 && !b->graphExec)
    if (int rc = gpugi_batch_capture(b)) return rc;
  BHIP(b, hipEventRecord(b->ev0, b->stream));
  for (int f = 0; f < frames; ++f) {
    if (mode == 2) BHIP(b, hipGraphLaunch(b->graphExec, b->stream));
    else if (int rc = LaunchFrame(b, mode == 1)) return rc;
  }
  BHIP(b, hipEventRecord(b->ev1, b->stream));
  BHIP(b, hipEventSynchronize(b->ev1));
  float t = 0.0f;
  BHIP(b, hipEventElapsedTime(&t, b->ev0, b->ev1));
  if (ms) *ms = t;
  return GPUGI_OK;
}

int gpugi_batch_readback(gpugi_batch* b, int scene, float* rgba, size_t bytes) {
  if (int rc = Ready(b, "gpugi_batch_readback")) return rc;
  if (scene < 0 || scene >= (int)b->hSlots.size() || !rgba) return BFail(b, GPUGI_ERR_ARG, "gpugi_batch_readback: bad scene");
  const size_t need = (size_t)b->scenes[scene].width * b->scenes[scene].height * sizeof(float4);
  if (bytes < need) return BFail(b, GPUGI_ERR_ARG, "gpugi_batch_readback: buffer too small");
  BHIP(b, hipMemcpyAsync(rgba, b->dAccum + b->accumOffset[scene], need, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return GPUGI_OK;
}

int gpugi_batch_rays(gpugi_batch* b, uint64_t* rays) {
  if (int rc = Ready(b, "gpugi_batch_rays")) return rc;
  if (!rays) return BFail(b, GPUGI_ERR_ARG, "gpugi_batch_rays: null argument");
  BatchCounters c;
  BHIP(b, hipMemcpyAsync(&c, b->dCounters, sizeof(c), hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  *rays = c.rays;
  return GPUGI_OK;
}

int gpugi_batch_autotune(gpugi_batch* b, int frames, int* kernel, int* tiles_per_grab, int* grid_multiplier,
                         int* stack_depth, float* best_ms) {
  if (int rc = Ready(b, "gpugi_batch_autotune")) return rc;
  if (frames < 1) return BFail(b, GPUGI_ERR_ARG, "gpugi_batch_autotune: frames >= 1");
  float best = 1e30f;
  int bk = b->kernel, bg = b->grab, bm = b->gridMul, bs = b->stackDepth;
  for (int k : {GPUGI_KERNEL_REFERENCE, GPUGI_KERNEL_WIDE})
    for (int st : kBatchStacks) {
      if (k == GPUGI_KERNEL_REFERENCE && st != kBatchStacks[0]) continue;  // the escape walk keeps no stack
      for (int g : {1, 2, 4, 8})
        for (int m : {1, 2, 4}) {
          if (int rc = gpugi_batch_config(b, k, g, m, st)) return rc;
          float ms = 0.0f;
          if (int rc = gpugi_batch_run(b, 1, 1, &ms)) return rc;  // warm-up
          if (int rc = gpugi_batch_run(b, 1, frames, &ms)) return rc;
          if (ms < best) {
            best = ms;
            bk = k;
            bg = g;
            bm = m;
            bs = st;
          }
        }
    }
  if (int rc = gpugi_batch_config(b, bk, bg, bm, bs)) return rc;
  if (kernel) *kernel = bk;
  if (tiles_per_grab) *tiles_per_grab = bg;
  if (grid_multiplier) *grid_multiplier = bm;
  if (stack_depth) *stack_depth = bs;
  if (best_ms) *best_ms = best / frames;
  return GPUGI_OK;
}

}  // extern "C"
