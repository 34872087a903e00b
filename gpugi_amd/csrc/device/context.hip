// C ABI of the device context: scene upload and re-pack, traces, timers,
// device buffers and the microbenchmark ray generators.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <set>
#include <string>
#include <vector>

#include "../host/host_api.hpp"
#include "integrators.h"
#include "launch.h"

using namespace gpugi;

struct gpugi_ctx {
  int device = 0;
  int numCUs = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timerStarted = false;
  std::string error;
  // reference layout
  bool hasScene = false;
  void* dNodes = nullptr;
  void* dSlots = nullptr;
  void* dPos = nullptr;
  void* dVinfo = nullptr;
  void* dMats = nullptr;
  uint32_t numNodes = 0, numSlots = 0, numVertices = 0, numMaterials = 0;
  // wide layout
  bool hasWide = false;
  void* dWideNodes = nullptr;
  void* dWideTris = nullptr;
  void* dSpill = nullptr;
  void* dCounters = nullptr;  // [0] ray counter, [1] error flag, [2] hit count
  uint32_t wideThreads = 0, spillDepth = 0, wideMaxStack = 0;
  // image state
  int width = 0, height = 0;
  void* dAccum = nullptr;
  uint32_t iterations = 0;
  gpugi_camera camera{};
  bool hasCamera = false;
  void* dLights = nullptr;
  uint32_t numLights = 0, lightCap = 0;
  int maxPathLength = 16;  // default pathLength (application.cpp:123)
  int integratorKernel = GPUGI_KERNEL_AUTO;
  IntegratorState integ;
  // scratch
  void* dScratch = nullptr;
  size_t scratchCap = 0;
  std::set<void*> userAllocs;
};

namespace {

int Fail(gpugi_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->error = msg;
  else SetGlobalError(msg);
  return code;
}

#define HIP_TRY(ctx, expr)                                                                     \
  do {                                                                                         \
    hipError_t _e = (expr);                                                                    \
    if (_e != hipSuccess)                                                                      \
      return Fail(ctx, GPUGI_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));     \
  } while (0)

void FreeDev(void*& p) {
  if (p) hipFree(p);
  p = nullptr;
}

RefScene MakeRef(const gpugi_ctx* c) {
  RefScene s;
  s.nodes = (const float4*)c->dNodes;
  s.slots = (const int4*)c->dSlots;
  s.positions = (const float*)c->dPos;
  s.vinfo = (const float4*)c->dVinfo;
  return s;
}

WideScene MakeWide(const gpugi_ctx* c) {
  WideScene s;
  s.nodes = (const uint4*)c->dWideNodes;
  s.tris = (const float4*)c->dWideTris;
  s.spill = (uint2*)c->dSpill;
  s.spillDepth = c->spillDepth;
  s.numThreads = c->wideThreads;
  s.errorFlag = (uint32_t*)c->dCounters + 1;
  return s;
}

int EnsureScratch(gpugi_ctx* ctx, size_t bytes) {
  if (ctx->scratchCap >= bytes) return GPUGI_OK;
  FreeDev(ctx->dScratch);
  ctx->scratchCap = 0;
  HIP_TRY(ctx, hipMalloc(&ctx->dScratch, bytes));
  ctx->scratchCap = bytes;
  return GPUGI_OK;
}

int SetDev(gpugi_ctx* ctx) {
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  return GPUGI_OK;
}

}  // namespace

const RefScene* gpugi_internal_refscene(const gpugi_ctx* c, RefScene* tmp) { *tmp = MakeRef(c); return tmp; }

extern "C" {

int gpugi_create(int device, gpugi_ctx** out) {
  if (!out) return Fail(nullptr, GPUGI_ERR_ARG, "gpugi_create: null output");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return Fail(nullptr, GPUGI_ERR_NODEVICE, "gpugi_create: no HIP device available");
  if (device < 0 || device >= count) return Fail(nullptr, GPUGI_ERR_ARG, "gpugi_create: device index out of range");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess)
    return Fail(nullptr, GPUGI_ERR_HIP, "gpugi_create: hipGetDeviceProperties failed");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return Fail(nullptr, GPUGI_ERR_UNSUPPORTED,
                std::string("gpugi_create: this build targets gfx950, device is ") + prop.gcnArchName);
  gpugi_ctx* ctx = new (std::nothrow) gpugi_ctx();
  if (!ctx) return Fail(nullptr, GPUGI_ERR_NOMEM, "gpugi_create: out of host memory");
  ctx->device = device;
  ctx->numCUs = prop.multiProcessorCount;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
      hipMalloc(&ctx->dCounters, 64) != hipSuccess || hipMemset(ctx->dCounters, 0, 64) != hipSuccess) {
    gpugi_destroy(ctx);
    return Fail(nullptr, GPUGI_ERR_HIP, "gpugi_create: HIP stream/event/counter setup failed");
  }
  *out = ctx;
  return GPUGI_OK;
}

void gpugi_destroy(gpugi_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  for (void** p : {&ctx->dNodes, &ctx->dSlots, &ctx->dPos, &ctx->dVinfo, &ctx->dMats, &ctx->dWideNodes,
                   &ctx->dWideTris, &ctx->dSpill, &ctx->dCounters, &ctx->dAccum, &ctx->dLights, &ctx->dScratch})
    FreeDev(*p);
  FreeIntegratorState(&ctx->integ);
  for (void* p : ctx->userAllocs) hipFree(p);
  if (ctx->ev0) hipEventDestroy(ctx->ev0);
  if (ctx->ev1) hipEventDestroy(ctx->ev1);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* gpugi_last_error(const gpugi_ctx* ctx) { return ctx ? ctx->error.c_str() : GlobalError(); }

int gpugi_load_scene(gpugi_ctx* ctx, const gpugi_scene_desc* d) {
  if (!ctx || !d) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_load_scene: null argument");
  if (!d->hierarchy || !d->triangles || !d->positions || !d->vertex_info || d->num_nodes == 0 ||
      d->num_leaves == 0 || d->num_vertices == 0)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_load_scene: missing scene arrays");
  if (d->triangles_per_leaf != 2)
    return Fail(ctx, GPUGI_ERR_UNSUPPORTED, "gpugi_load_scene: TRIANGLES_PER_LEAF must be 2 (scenedata.glsl:5)");
  if (d->num_materials == 0 || d->num_materials > 256 || !d->materials)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_load_scene: need 1..256 materials");
  // Validate every index the kernels follow, so that no kernel can read out of bounds.
  const uint32_t nn = d->num_nodes, nl = d->num_leaves, nv = d->num_vertices;
  for (uint32_t i = 0; i < nn; ++i) {
    uint32_t fc, esc;
    std::memcpy(&fc, d->hierarchy + (size_t)i * 8 + 3, 4);
    std::memcpy(&esc, d->hierarchy + (size_t)i * 8 + 7, 4);
    if ((fc & 0x80000000u) ? (fc & 0x7fffffffu) >= nl : (fc >= nn || fc <= i))
      return Fail(ctx, GPUGI_ERR_FORMAT, "gpugi_load_scene: node " + std::to_string(i) + " has a bad firstChild");
    if (esc >= nn || (esc != 0 && esc <= i))
      return Fail(ctx, GPUGI_ERR_FORMAT, "gpugi_load_scene: node " + std::to_string(i) + " has a bad escape");
  }
  for (size_t s = 0; s < (size_t)nl * 2; ++s) {
    const int32_t* t = d->triangles + s * 4;
    if (t[0] == t[1]) continue;
    for (int k = 0; k < 3; ++k)
      if (t[k] < 0 || (uint32_t)t[k] >= nv) return Fail(ctx, GPUGI_ERR_FORMAT, "gpugi_load_scene: vertex index out of range");
    if (t[3] < 0 || (uint32_t)t[3] >= d->num_materials)
      return Fail(ctx, GPUGI_ERR_FORMAT, "gpugi_load_scene: material index out of range");
  }
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  WideBVHHost wide;
  std::string err = BuildWideBVH(*d, 4, &wide);
  if (!err.empty()) return Fail(ctx, GPUGI_ERR_FORMAT, "gpugi_load_scene: " + err);

  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  for (void** p : {&ctx->dNodes, &ctx->dSlots, &ctx->dPos, &ctx->dVinfo, &ctx->dMats, &ctx->dWideNodes,
                   &ctx->dWideTris, &ctx->dSpill})
    FreeDev(*p);
  ctx->hasScene = ctx->hasWide = false;
  const size_t bNodes = (size_t)nn * 32, bSlots = (size_t)nl * 2 * 16, bPos = (size_t)nv * 12 + 4,
               bInfo = (size_t)nv * 16, bMats = (size_t)d->num_materials * sizeof(gpugi_material);
  HIP_TRY(ctx, hipMalloc(&ctx->dNodes, bNodes));
  HIP_TRY(ctx, hipMalloc(&ctx->dSlots, bSlots));
  HIP_TRY(ctx, hipMalloc(&ctx->dPos, bPos));
  HIP_TRY(ctx, hipMalloc(&ctx->dVinfo, bInfo));
  HIP_TRY(ctx, hipMalloc(&ctx->dMats, bMats));
  HIP_TRY(ctx, hipMemcpy(ctx->dNodes, d->hierarchy, bNodes, hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(ctx->dSlots, d->triangles, bSlots, hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(ctx->dPos, d->positions, (size_t)nv * 12, hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(ctx->dVinfo, d->vertex_info, bInfo, hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(ctx->dMats, d->materials, bMats, hipMemcpyHostToDevice));
  ctx->numNodes = nn;
  ctx->numSlots = nl * 2;
  ctx->numVertices = nv;
  ctx->numMaterials = d->num_materials;
  ctx->hasScene = true;

  // Wide layout + persistent-grid spill stack.
  int blocksPerCU = 0;
  HIP_TRY(ctx, WideBlocksPerCU(&blocksPerCU));
  if (blocksPerCU < 1) blocksPerCU = 1;
  ctx->wideThreads = (uint32_t)(ctx->numCUs * blocksPerCU * kWideBlock);
  ctx->wideMaxStack = wide.maxStack;
  ctx->spillDepth = wide.maxStack > (uint32_t)kWideLdsStack ? wide.maxStack - (uint32_t)kWideLdsStack : 0;
  HIP_TRY(ctx, hipMalloc(&ctx->dWideNodes, wide.nodes.size() * sizeof(WideNodeHost)));
  HIP_TRY(ctx, hipMalloc(&ctx->dWideTris, std::max<size_t>(1, wide.tris.size()) * sizeof(WideTriHost)));
  HIP_TRY(ctx, hipMemcpy(ctx->dWideNodes, wide.nodes.data(), wide.nodes.size() * sizeof(WideNodeHost),
                         hipMemcpyHostToDevice));
  if (!wide.tris.empty())
    HIP_TRY(ctx, hipMemcpy(ctx->dWideTris, wide.tris.data(), wide.tris.size() * sizeof(WideTriHost),
                           hipMemcpyHostToDevice));
  if (ctx->spillDepth > 0)
    HIP_TRY(ctx, hipMalloc(&ctx->dSpill, (size_t)ctx->wideThreads * ctx->spillDepth * sizeof(uint2)));
  ctx->hasWide = true;
  HIP_TRY(ctx, hipMemset(ctx->dCounters, 0, 64));
  if (ctx->dAccum) {
    HIP_TRY(ctx, hipMemset(ctx->dAccum, 0, (size_t)ctx->width * ctx->height * 16));
    ctx->iterations = 0;
  }
  InvalidateIntegratorCaches(&ctx->integ);
  return GPUGI_OK;
}

int gpugi_set_camera(gpugi_ctx* ctx, const gpugi_camera* camera) {
  if (!ctx || !camera) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_camera: null argument");
  ctx->camera = *camera;
  ctx->hasCamera = true;
  if (ctx->dAccum) {
    if (SetDev(ctx)) return GPUGI_ERR_HIP;
    HIP_TRY(ctx, hipMemsetAsync(ctx->dAccum, 0, (size_t)ctx->width * ctx->height * 16, ctx->stream));
    ctx->iterations = 0;
  }
  return GPUGI_OK;
}

int gpugi_set_screen(gpugi_ctx* ctx, int width, int height) {
  if (!ctx || width <= 0 || height <= 0 || (size_t)width * height > (size_t)1 << 28)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_screen: invalid size");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  FreeDev(ctx->dAccum);
  HIP_TRY(ctx, hipMalloc(&ctx->dAccum, (size_t)width * height * 16));
  HIP_TRY(ctx, hipMemset(ctx->dAccum, 0, (size_t)width * height * 16));
  ctx->width = width;
  ctx->height = height;
  ctx->iterations = 0;
  InvalidateIntegratorCaches(&ctx->integ);
  return GPUGI_OK;
}

int gpugi_set_light_samples(gpugi_ctx* ctx, const gpugi_light_sample* samples, uint32_t n) {
  if (!ctx || !samples || n == 0) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_light_samples: need n > 0 samples");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  if (n > ctx->lightCap) {
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    FreeDev(ctx->dLights);
    HIP_TRY(ctx, hipMalloc(&ctx->dLights, (size_t)n * sizeof(gpugi_light_sample)));
    ctx->lightCap = n;
  }
  // Synchronous copy: the caller may reuse its buffer right away (the
  // reference writes a persistently mapped buffer, renderersystem.cpp:109-112).
  HIP_TRY(ctx, hipMemcpyAsync(ctx->dLights, samples, (size_t)n * sizeof(gpugi_light_sample), hipMemcpyHostToDevice,
                              ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  ctx->numLights = n;
  return GPUGI_OK;
}

int gpugi_set_max_path_length(gpugi_ctx* ctx, int max_path_length) {
  if (!ctx || max_path_length < 1 || max_path_length > 64)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_max_path_length: must be in [1, 64]");
  ctx->maxPathLength = max_path_length;
  InvalidateIntegratorCaches(&ctx->integ);
  return GPUGI_OK;
}

int gpugi_set_integrator_kernel(gpugi_ctx* ctx, int kernel) {
  if (!ctx || kernel < GPUGI_KERNEL_AUTO || kernel > GPUGI_KERNEL_WIDE)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_integrator_kernel: bad kernel");
  ctx->integratorKernel = kernel;
  return GPUGI_OK;
}

int gpugi_draw(gpugi_ctx* ctx, int integrator, uint32_t frame_seed) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_draw: null context");
  if (!ctx->hasScene || !ctx->dAccum || !ctx->hasCamera || ctx->numLights == 0)
    return Fail(ctx, GPUGI_ERR_STATE, "gpugi_draw: scene, screen, camera and light samples must be set first");
  if (integrator < GPUGI_PT || integrator > GPUGI_BPT) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_draw: unknown integrator");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  DrawParams p;
  p.ref = MakeRef(ctx);
  p.wide = MakeWide(ctx);
  p.useWide = ctx->hasWide && ctx->integratorKernel != GPUGI_KERNEL_REFERENCE;
  p.materials = (const gpugi_material*)ctx->dMats;
  p.lights = (const gpugi_light_sample*)ctx->dLights;
  p.numLights = ctx->numLights;
  p.camera = ctx->camera;
  p.width = ctx->width;
  p.height = ctx->height;
  p.accum = (float4*)ctx->dAccum;
  p.frameSeed = frame_seed;
  p.maxPathLength = ctx->maxPathLength;
  p.counters = (uint32_t*)ctx->dCounters;
  std::string err;
  hipError_t e = DrawIntegrator(&ctx->integ, integrator, p, ctx->stream, &err);
  if (e != hipSuccess) return Fail(ctx, GPUGI_ERR_HIP, "gpugi_draw: " + err + hipGetErrorString(e));
  if (!err.empty()) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_draw: " + err);
  ++ctx->iterations;
  return GPUGI_OK;
}

int gpugi_reset_accum(gpugi_ctx* ctx) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_reset_accum: null context");
  if (!ctx->dAccum) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_reset_accum: no screen set");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  HIP_TRY(ctx, hipMemsetAsync(ctx->dAccum, 0, (size_t)ctx->width * ctx->height * 16, ctx->stream));
  ctx->iterations = 0;
  return GPUGI_OK;
}

int gpugi_readback(gpugi_ctx* ctx, float* rgba, size_t bytes, uint32_t* iterations) {
  if (!ctx || !rgba) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_readback: null argument");
  if (!ctx->dAccum) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_readback: no screen set");
  const size_t need = (size_t)ctx->width * ctx->height * 16;
  if (bytes < need) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_readback: buffer too small");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  HIP_TRY(ctx, hipMemcpyAsync(rgba, ctx->dAccum, need, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (iterations) *iterations = ctx->iterations;
  return GPUGI_OK;
}

int gpugi_trace(gpugi_ctx* ctx, int query, int kernel, const gpugi_ray* rays, gpugi_hit* hits, size_t n,
                int rays_on_device) {
  if (!ctx || (n > 0 && (!rays || !hits))) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_trace: null argument");
  if (query != GPUGI_CLOSEST_HIT && query != GPUGI_ANY_HIT) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_trace: bad query");
  if (kernel < GPUGI_KERNEL_AUTO || kernel > GPUGI_KERNEL_WIDE) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_trace: bad kernel");
  if (!ctx->hasScene) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_trace: no scene loaded");
  if (n > 0xffffff00ull) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_trace: at most 2^32 - 256 rays per call");
  if (n == 0) return GPUGI_OK;
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  const bool wide = kernel == GPUGI_KERNEL_WIDE || (kernel == GPUGI_KERNEL_AUTO && ctx->hasWide);
  const gpugi_ray* dr = rays;
  gpugi_hit* dh = hits;
  if (!rays_on_device) {
    if (int rc = EnsureScratch(ctx, n * (sizeof(gpugi_ray) + sizeof(gpugi_hit)))) return rc;
    dr = (const gpugi_ray*)ctx->dScratch;
    dh = (gpugi_hit*)((char*)ctx->dScratch + n * sizeof(gpugi_ray));
    HIP_TRY(ctx, hipMemcpyAsync((void*)dr, rays, n * sizeof(gpugi_ray), hipMemcpyHostToDevice, ctx->stream));
  }
  if (wide) HIP_TRY(ctx, LaunchTraceWide(MakeWide(ctx), query, dr, dh, n, (uint32_t*)ctx->dCounters, ctx->stream));
  else HIP_TRY(ctx, LaunchTraceRef(MakeRef(ctx), query, dr, dh, n, ctx->stream));
  if (!rays_on_device) {
    HIP_TRY(ctx, hipMemcpyAsync(hits, dh, n * sizeof(gpugi_hit), hipMemcpyDeviceToHost, ctx->stream));
    uint32_t flag = 0;
    HIP_TRY(ctx, hipMemcpyAsync(&flag, (uint32_t*)ctx->dCounters + 1, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (flag) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_trace: traversal stack overflow (bug in the stack bound)");
  }
  return GPUGI_OK;
}

int gpugi_trace_counts(gpugi_ctx* ctx, int query, const gpugi_ray* rays, uint32_t* box_tests, uint32_t* tri_tests,
                       size_t n, int rays_on_device) {
  if (!ctx || (n > 0 && (!rays || !box_tests || !tri_tests)))
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_trace_counts: null argument");
  if (query != GPUGI_CLOSEST_HIT && query != GPUGI_ANY_HIT) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_trace_counts: bad query");
  if (!ctx->hasScene) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_trace_counts: no scene loaded");
  if (n == 0) return GPUGI_OK;
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  const gpugi_ray* dr = rays;
  uint32_t* db = box_tests;
  uint32_t* dt = tri_tests;
  if (!rays_on_device) {
    if (int rc = EnsureScratch(ctx, n * (sizeof(gpugi_ray) + 8))) return rc;
    dr = (const gpugi_ray*)ctx->dScratch;
    db = (uint32_t*)((char*)ctx->dScratch + n * sizeof(gpugi_ray));
    dt = db + n;
    HIP_TRY(ctx, hipMemcpyAsync((void*)dr, rays, n * sizeof(gpugi_ray), hipMemcpyHostToDevice, ctx->stream));
  }
  HIP_TRY(ctx, LaunchTraceRefCounts(MakeRef(ctx), query, dr, db, dt, n, ctx->stream));
  if (!rays_on_device) {
    HIP_TRY(ctx, hipMemcpyAsync(box_tests, db, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(tri_tests, dt, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GPUGI_OK;
}

int gpugi_sync(gpugi_ctx* ctx) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_sync: null context");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return GPUGI_OK;
}

int gpugi_alloc(gpugi_ctx* ctx, size_t bytes, void** device_ptr) {
  if (!ctx || !device_ptr || bytes == 0) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_alloc: invalid argument");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) return Fail(ctx, GPUGI_ERR_NOMEM, std::string("gpugi_alloc: ") + hipGetErrorString(e));
  ctx->userAllocs.insert(p);
  *device_ptr = p;
  return GPUGI_OK;
}

int gpugi_free(gpugi_ctx* ctx, void* device_ptr) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_free: null context");
  auto it = ctx->userAllocs.find(device_ptr);
  if (it == ctx->userAllocs.end()) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_free: pointer not allocated by gpugi_alloc");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipFree(device_ptr));
  ctx->userAllocs.erase(it);
  return GPUGI_OK;
}

int gpugi_upload(gpugi_ctx* ctx, void* device_dst, const void* host_src, size_t bytes) {
  if (!ctx || !device_dst || !host_src) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_upload: null argument");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  HIP_TRY(ctx, hipMemcpyAsync(device_dst, host_src, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return GPUGI_OK;
}

int gpugi_download(gpugi_ctx* ctx, void* host_dst, const void* device_src, size_t bytes) {
  if (!ctx || !device_src || !host_dst) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_download: null argument");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  HIP_TRY(ctx, hipMemcpyAsync(host_dst, device_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return GPUGI_OK;
}

int gpugi_timer_begin(gpugi_ctx* ctx) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_timer_begin: null context");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  ctx->timerStarted = true;
  return GPUGI_OK;
}

int gpugi_timer_end(gpugi_ctx* ctx, float* milliseconds) {
  if (!ctx || !milliseconds) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_timer_end: null argument");
  if (!ctx->timerStarted) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_timer_end: timer not started");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
  HIP_TRY(ctx, hipEventElapsedTime(milliseconds, ctx->ev0, ctx->ev1));
  ctx->timerStarted = false;
  return GPUGI_OK;
}

int gpugi_raygen_primary(gpugi_ctx* ctx, const gpugi_camera* camera, int width, int height, uint32_t seed,
                         gpugi_ray* rays_dev, size_t n) {
  if (!ctx || !camera || !rays_dev || width <= 0 || height <= 0)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_raygen_primary: invalid argument");
  if (n > 0xffffffffull) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_raygen_primary: too many rays");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  HIP_TRY(ctx, LaunchRaygenPrimary(*camera, width, height, seed, rays_dev, n, ctx->stream));
  return GPUGI_OK;
}

static int HitList(gpugi_ctx* ctx, const gpugi_hit* hits, size_t n_primary, uint32_t** list, uint32_t** count) {
  size_t temp = 0;
  HIP_TRY(ctx, LaunchHitCompaction(hits, n_primary, nullptr, nullptr, nullptr, &temp, ctx->stream));
  const size_t listBytes = n_primary * 4;
  if (int rc = EnsureScratch(ctx, listBytes + temp + 256)) return rc;
  *list = (uint32_t*)ctx->dScratch;
  *count = (uint32_t*)ctx->dCounters + 2;
  void* tmp = (char*)ctx->dScratch + ((listBytes + 255) & ~(size_t)255);
  HIP_TRY(ctx, LaunchHitCompaction(hits, n_primary, *list, *count, tmp, &temp, ctx->stream));
  return GPUGI_OK;
}

int gpugi_raygen_diffuse(gpugi_ctx* ctx, const gpugi_ray* primary_dev, const gpugi_hit* hits_dev, size_t n_primary,
                         uint32_t seed, gpugi_ray* rays_dev, size_t n) {
  if (!ctx || !primary_dev || !hits_dev || !rays_dev || n_primary == 0 || n_primary > 0xffffffffull ||
      n > 0xffffffffull)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_raygen_diffuse: invalid argument");
  if (!ctx->hasScene) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_raygen_diffuse: no scene loaded");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  uint32_t *list = nullptr, *count = nullptr;
  if (int rc = HitList(ctx, hits_dev, n_primary, &list, &count)) return rc;
  HIP_TRY(ctx, LaunchRaygenDiffuse(MakeRef(ctx), primary_dev, hits_dev, list, count, seed, rays_dev, n, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // the list lives in scratch
  return GPUGI_OK;
}

int gpugi_raygen_shadow(gpugi_ctx* ctx, const gpugi_ray* primary_dev, const gpugi_hit* hits_dev, size_t n_primary,
                        const float* lights_xyz, uint32_t n_lights, gpugi_ray* rays_dev, size_t n) {
  if (!ctx || !primary_dev || !hits_dev || !rays_dev || !lights_xyz || n_lights == 0 || n_lights > 1024 ||
      n_primary == 0 || n_primary > 0xffffffffull || n > 0xffffffffull)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_raygen_shadow: invalid argument");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  uint32_t *list = nullptr, *count = nullptr;
  if (int rc = HitList(ctx, hits_dev, n_primary, &list, &count)) return rc;
  void* dl = nullptr;
  HIP_TRY(ctx, hipMalloc(&dl, (size_t)n_lights * 12));
  hipError_t e = hipMemcpyAsync(dl, lights_xyz, (size_t)n_lights * 12, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = LaunchRaygenShadow(primary_dev, hits_dev, list, count, (const float*)dl, n_lights, rays_dev, n, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(dl);
  if (e != hipSuccess) return Fail(ctx, GPUGI_ERR_HIP, std::string("gpugi_raygen_shadow: ") + hipGetErrorString(e));
  return GPUGI_OK;
}

}  // extern "C"
