// C ABI of the device context: scene upload and re-pack, traces, timers,
// device buffers and the microbenchmark ray generators.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <set>
#include <string>
#include <vector>

#include "../host/host_api.hpp"
#include "integrators.h"
#include "launch.h"
#include "mlp.h"
#include "scan.h"
#include "hierarchy.h"
#include "wavefront_pt.h"

using namespace gpugi;

struct gpugi_ctx {
  int device = 0;
  int numCUs = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // second stream (default priority) for work that overlaps the main stream's,
  // forked and joined with the two events; created on first use
  hipStream_t side = nullptr;
  // host-buffer traces: a copy stream and per-slot events of the chunk pipeline
  hipStream_t copy = nullptr;
  hipEvent_t evIn[2] = {nullptr, nullptr}, evTraced[2] = {nullptr, nullptr}, evStart = nullptr;
  std::set<void*> hostAllocs;  // pinned host memory from gpugi_host_alloc
  hipEvent_t evFork = nullptr, evJoin = nullptr, evInfer = nullptr;
  hipEvent_t evShade = nullptr, evResolve = nullptr;  // wavefront draw: shade -> shadow trace -> next shade
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
  void* dWideQueue = nullptr;  // the wide trace kernel's batch counters
  size_t spillBytes = 0;
  uint32_t wideMaxStack = 0;
  int wideLeaf = 2;     // triangles per merged wide leaf
  int wideStack = 10;   // LDS stack entries per lane of the trace kernel
  int wideRefill = 8;   // idle lanes that trigger a wave-level ray refill
  int wideWait = 16;    // lanes waiting with a pending leaf that trigger a leaf pass
  // host copy of the reference layout (re-packing after a configuration change)
  std::vector<float> hHier, hPos;
  std::vector<int32_t> hTris;
  uint32_t hNumLeaves = 0, hNumVertices = 0;
  // image state
  int width = 0, height = 0;
  void* dAccum = nullptr;
  uint32_t iterations = 0;
  gpugi_camera camera{};
  bool hasCamera = false;
  void* dLights = nullptr;
  uint32_t numLights = 0, lightCap = 0;
  int maxPathLength = 16;  // default pathLength (application.cpp:123)
  int rrStartDepth = 0;
  int specificPathLength = -1;  // SHOW_SPECIFIC_PATHLENGTH; -1: every path
  void* noAllocScratch = nullptr;  // device address of gpugi_trace_noalloc's scratch array
  hipGraphExec_t drawGraph = nullptr;  // the draw gpugi_capture_draw recorded (null: none)
  // HierarchyImportance passes: the tree levels, per node/slot importance {x, y},
  // per node subtree importance and packed material (12 words)
  bool hierTreeReady = false;
  HierTree hierTree;
  void* dHierParent = nullptr;
  void* dHierLevel = nullptr;
  void* dHierImp = nullptr;
  void* dHierSubtree = nullptr;
  void* dHierMat = nullptr;
  bool hasHierImportance = false, hasHierMaterials = false;
  int triangleTest = GPUGI_TRIANGLE_REFERENCE;
  int bvhType = GPUGI_BVH_AABOX;       // applies from the next scene load
  bool sceneObox = false;              // dNodes holds TreeNode<OBox> (3 texels per node)
  int integratorKernel = GPUGI_KERNEL_AUTO;
  int ptSchedule = GPUGI_PT_MEGAKERNEL;
  void* dWave = nullptr;         // wavefront path state and ray queues
  size_t waveCap = 0;
  // the wavefront shadow traces' own wide-kernel spill stacks and batch
  // counters (they run beside the closest-hit traces on the side stream)
  void* dWaveSpill = nullptr;
  size_t waveSpillBytes = 0;
  void* dWaveQueue = nullptr;
  int bptCapacity = -1;
  int bptAverageLength = 0;
  IntegratorState integ;
  // radiance-cache MLP
  bool hasMlp = false;
  MlpDev mlp{};
  uint32_t mlpStep = 0;
  float bboxMin[3] = {0, 0, 0}, bboxMax[3] = {0, 0, 0};
  // cached path tracer (GPUGI_PT_CACHE)
  CacheParams cacheCfg;
  uint32_t cacheMaxSamples = 65536;
  int cacheTrainSteps = 4;
  float cacheLr = 1e-3f;
  const uint32_t* dCacheCount = nullptr;  // training samples written by the last cached draw (device)
  float cacheLastLoss = 0.0f;
  void* dCache = nullptr;
  size_t cacheBytes = 0;
  float* dMlpLoss = nullptr;  // loss of the last gradient / training step
  void* dMlpIo = nullptr;     // host-pointer staging of features / targets / outputs
  size_t mlpIoCap = 0;
  // material textures: device texels per texture, their descriptors, and per
  // material the texture of each slot (-1: the record's constant)
  std::vector<void*> texBufs;
  std::vector<TextureDesc> texDescs;
  std::vector<int4> matTex;
  void* dTexDesc = nullptr;
  void* dMatTex = nullptr;
  bool anyTextureBound = false;
  // environment cube map (linear RGB float4 texels, faces x-, x+, y-, y+, z-, z+)
  void* dEnv = nullptr;
  int envSize = 0;
  // reference image of the MSE check (RGB32F) and the reduction's block sums
  void* dRefImage = nullptr;
  int refW = 0, refH = 0;
  void* dMseTemp = nullptr;
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

bool Watertight(const gpugi_ctx* c) { return c->triangleTest == GPUGI_TRIANGLE_WATERTIGHT; }

// Any call that changes what a recorded draw reads (buffers, their sizes, the
// values its kernels were given) drops the recording.
void DropGraph(gpugi_ctx* c) {
  if (c && c->drawGraph) {
    (void)hipGraphExecDestroy(c->drawGraph);
    c->drawGraph = nullptr;
  }
}

RefScene MakeRef(const gpugi_ctx* c) {
  RefScene s;
  s.nodes = (const float4*)c->dNodes;
  s.slots = (const int4*)c->dSlots;
  s.positions = (const float*)c->dPos;
  s.vinfo = (const float4*)c->dVinfo;
  s.obox = c->sceneObox ? 1 : 0;
  return s;
}

// Wide-scene view for a kernel whose LDS stack holds `stack` entries; the
// spill slice per thread covers the rest of the tree's worst case.
WideScene MakeWide(const gpugi_ctx* c, int stack) {
  WideScene s;
  s.nodes = (const uint4*)c->dWideNodes;
  s.tris = (const float4*)c->dWideTris;
  s.spill = (uint2*)c->dSpill;
  s.spillDepth = c->wideMaxStack > (uint32_t)stack ? c->wideMaxStack - (uint32_t)stack : 0;
  s.numThreads = 0;
  s.errorFlag = (uint32_t*)c->dCounters + 1;
  s.queue = (uint32_t*)c->dWideQueue;
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

// The side stream (default priority) and the events that fork work onto it
// and join it back, created on first use.
int EnsureSideStream(gpugi_ctx* ctx) {
  if (!ctx->side) HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
  for (hipEvent_t* e : {&ctx->evFork, &ctx->evJoin, &ctx->evInfer, &ctx->evShade, &ctx->evResolve})
    if (!*e) HIP_TRY(ctx, hipEventCreateWithFlags(e, hipEventDisableTiming));
  return GPUGI_OK;
}

// Persistent grid and spill buffer for the trace kernel variant `stack`.
int PrepareTraceWide(gpugi_ctx* ctx, int stack, WideScene* out) {
  int blocksPerCU = 0;
  HIP_TRY(ctx, WideBlocksPerCU(stack, &blocksPerCU));
  if (blocksPerCU < 1) blocksPerCU = 1;
  WideScene s = MakeWide(ctx, stack);
  // whole blocks per XCD for the static schedule of the trace kernel
  const uint32_t blocks = std::max<uint32_t>(8, (uint32_t)(ctx->numCUs * blocksPerCU) / 8 * 8);
  s.numThreads = blocks * (uint32_t)kWideBlock;
  const size_t bytes = (size_t)s.numThreads * s.spillDepth * sizeof(uint2);
  if (bytes > ctx->spillBytes) {
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    FreeDev(ctx->dSpill);
    ctx->spillBytes = 0;
    HIP_TRY(ctx, hipMalloc(&ctx->dSpill, bytes));
    ctx->spillBytes = bytes;
  }
  s.spill = (uint2*)ctx->dSpill;
  *out = s;
  return GPUGI_OK;
}

// Re-packs the host copy of the reference BVH into the wide layout with the
// context's leaf size and uploads it.
int RebuildWide(gpugi_ctx* ctx) {
  gpugi_scene_desc d{};
  d.hierarchy = ctx->hHier.data();
  d.num_nodes = ctx->numNodes;
  d.triangles = ctx->hTris.data();
  d.num_leaves = ctx->hNumLeaves;
  d.triangles_per_leaf = 2;
  d.positions = ctx->hPos.data();
  d.num_vertices = ctx->hNumVertices;
  WideBVHHost wide;
  std::string err = BuildWideBVH(d, (uint32_t)ctx->wideLeaf, &wide, ctx->triangleTest == GPUGI_TRIANGLE_WATERTIGHT);
  if (!err.empty()) return Fail(ctx, GPUGI_ERR_FORMAT, "wide BVH re-pack: " + err);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  FreeDev(ctx->dWideNodes);
  FreeDev(ctx->dWideTris);
  ctx->hasWide = false;
  HIP_TRY(ctx, hipMalloc(&ctx->dWideNodes, wide.nodes.size() * sizeof(WideNodeHost)));
  HIP_TRY(ctx, hipMalloc(&ctx->dWideTris, std::max<size_t>(1, wide.tris.size()) * sizeof(WideTriHost)));
  HIP_TRY(ctx, hipMemcpy(ctx->dWideNodes, wide.nodes.data(), wide.nodes.size() * sizeof(WideNodeHost),
                         hipMemcpyHostToDevice));
  if (!wide.tris.empty())
    HIP_TRY(ctx, hipMemcpy(ctx->dWideTris, wide.tris.data(), wide.tris.size() * sizeof(WideTriHost),
                           hipMemcpyHostToDevice));
  ctx->wideMaxStack = wide.maxStack;
  ctx->hasWide = true;
  return GPUGI_OK;
}

}  // namespace

const RefScene* gpugi_internal_refscene(const gpugi_ctx* c, RefScene* tmp) { *tmp = MakeRef(c); return tmp; }

static DrawParams MakeDrawParams(gpugi_ctx* ctx, uint32_t frame_seed);

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
      hipMalloc(&ctx->dCounters, 64) != hipSuccess || hipMemset(ctx->dCounters, 0, 64) != hipSuccess ||
      hipMalloc(&ctx->dWideQueue, kWideQueueWords * sizeof(uint32_t)) != hipSuccess) {
    gpugi_destroy(ctx);
    return Fail(nullptr, GPUGI_ERR_HIP, "gpugi_create: HIP stream/event/counter setup failed");
  }
  // the device address of gpugi_trace_noalloc's scratch array, once per device
  if (NoAllocScratchAddress(device, &ctx->noAllocScratch) != hipSuccess) {
    gpugi_destroy(ctx);
    return Fail(nullptr, GPUGI_ERR_HIP, "gpugi_create: scratch symbol lookup failed");
  }
  *out = ctx;
  return GPUGI_OK;
}

void gpugi_destroy(gpugi_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  DropGraph(ctx);
  if (ctx->side) hipStreamSynchronize(ctx->side);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  for (void** p : {&ctx->dNodes, &ctx->dSlots, &ctx->dPos, &ctx->dVinfo, &ctx->dMats, &ctx->dWideNodes,
                   &ctx->dWideTris, &ctx->dSpill, &ctx->dCounters, &ctx->dWideQueue, &ctx->dAccum, &ctx->dLights,
                   &ctx->dScratch, &ctx->dRefImage, &ctx->dMseTemp, &ctx->dEnv, &ctx->dWave, &ctx->dWaveSpill,
                   &ctx->dWaveQueue, &ctx->dHierParent, &ctx->dHierLevel, &ctx->dHierImp, &ctx->dHierSubtree,
                   &ctx->dHierMat})
    FreeDev(*p);
  FreeIntegratorState(&ctx->integ);
  for (void* p : ctx->texBufs) hipFree(p);
  FreeDev(ctx->dTexDesc);
  FreeDev(ctx->dMatTex);
  for (void** p : {(void**)&ctx->mlp.w, (void**)&ctx->mlp.m, (void**)&ctx->mlp.v, (void**)&ctx->mlp.wb,
                   (void**)&ctx->mlp.wtb, (void**)&ctx->mlp.wbi, (void**)&ctx->mlp.partials, (void**)&ctx->mlp.grad,
                   (void**)&ctx->mlp.loss, (void**)&ctx->dMlpLoss, &ctx->dMlpIo, &ctx->dCache})
    FreeDev(*p);
  for (void* p : ctx->userAllocs) hipFree(p);
  if (ctx->ev0) hipEventDestroy(ctx->ev0);
  if (ctx->ev1) hipEventDestroy(ctx->ev1);
  if (ctx->evFork) hipEventDestroy(ctx->evFork);
  if (ctx->evJoin) hipEventDestroy(ctx->evJoin);
  if (ctx->evInfer) hipEventDestroy(ctx->evInfer);
  if (ctx->evShade) hipEventDestroy(ctx->evShade);
  if (ctx->evResolve) hipEventDestroy(ctx->evResolve);
  if (ctx->copy) hipStreamSynchronize(ctx->copy);
  for (hipEvent_t e : {ctx->evIn[0], ctx->evIn[1], ctx->evTraced[0], ctx->evTraced[1], ctx->evStart})
    if (e) hipEventDestroy(e);
  if (ctx->copy) hipStreamDestroy(ctx->copy);
  for (void* p : ctx->hostAllocs) hipHostFree(p);
  if (ctx->side) hipStreamDestroy(ctx->side);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* gpugi_last_error(const gpugi_ctx* ctx) { return ctx ? ctx->error.c_str() : GlobalError(); }

int gpugi_load_scene(gpugi_ctx* ctx, const gpugi_scene_desc* d) {
  DropGraph(ctx);
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
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->side) HIP_TRY(ctx, hipStreamSynchronize(ctx->side));
  for (void** p : {&ctx->dNodes, &ctx->dSlots, &ctx->dPos, &ctx->dVinfo, &ctx->dMats, &ctx->dWideNodes,
                   &ctx->dWideTris, &ctx->dSpill, &ctx->dWaveSpill})
    FreeDev(*p);
  // textures belong to the scene's materials
  for (void* p : ctx->texBufs) hipFree(p);
  ctx->texBufs.clear();
  ctx->texDescs.clear();
  ctx->matTex.assign(d->num_materials, make_int4(-1, -1, -1, -1));
  ctx->anyTextureBound = false;
  FreeDev(ctx->dTexDesc);
  FreeDev(ctx->dMatTex);
  ctx->spillBytes = ctx->waveSpillBytes = 0;
  ctx->hasScene = ctx->hasWide = false;
  std::vector<float> oboxNodes;
  if (ctx->bvhType == GPUGI_BVH_OBOX) {
    try {
      oboxNodes = FitOBoxHierarchy(d->hierarchy, nn, d->triangles, nl, d->positions, nv);
    } catch (const std::bad_alloc&) {
      return Fail(ctx, GPUGI_ERR_NOMEM, "gpugi_load_scene: oriented-box fit");
    }
  }
  ctx->sceneObox = ctx->bvhType == GPUGI_BVH_OBOX;
  const size_t bNodes = (size_t)nn * (ctx->sceneObox ? 48 : 32), bSlots = (size_t)nl * 2 * 16, bPos = (size_t)nv * 12 + 4,
               bInfo = (size_t)nv * 16, bMats = (size_t)d->num_materials * sizeof(gpugi_material);
  HIP_TRY(ctx, hipMalloc(&ctx->dNodes, bNodes));
  HIP_TRY(ctx, hipMalloc(&ctx->dSlots, bSlots));
  HIP_TRY(ctx, hipMalloc(&ctx->dPos, bPos));
  HIP_TRY(ctx, hipMalloc(&ctx->dVinfo, bInfo));
  HIP_TRY(ctx, hipMalloc(&ctx->dMats, bMats));
  HIP_TRY(ctx, hipMemcpy(ctx->dNodes, ctx->sceneObox ? oboxNodes.data() : d->hierarchy, bNodes,
                         hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(ctx->dSlots, d->triangles, bSlots, hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(ctx->dPos, d->positions, (size_t)nv * 12, hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(ctx->dVinfo, d->vertex_info, bInfo, hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(ctx->dMats, d->materials, bMats, hipMemcpyHostToDevice));
  ctx->numNodes = nn;
  ctx->numSlots = nl * 2;
  ctx->numVertices = nv;
  ctx->numMaterials = d->num_materials;
  ctx->hasScene = true;
  // the hierarchy passes' buffers belong to the previous tree
  for (void** p : {&ctx->dHierParent, &ctx->dHierLevel, &ctx->dHierImp, &ctx->dHierSubtree, &ctx->dHierMat})
    FreeDev(*p);
  ctx->hierTreeReady = ctx->hasHierImportance = ctx->hasHierMaterials = false;
  try {
    ctx->hHier.assign(d->hierarchy, d->hierarchy + (size_t)nn * 8);
    ctx->hTris.assign(d->triangles, d->triangles + (size_t)nl * 8);
    ctx->hPos.assign(d->positions, d->positions + (size_t)nv * 3);
  } catch (const std::bad_alloc&) {
    return Fail(ctx, GPUGI_ERR_NOMEM, "gpugi_load_scene: host copy of the scene");
  }
  ctx->hNumLeaves = nl;
  ctx->hNumVertices = nv;
  for (int a = 0; a < 3; ++a) {
    ctx->bboxMin[a] = INFINITY;
    ctx->bboxMax[a] = -INFINITY;
  }
  for (uint32_t v = 0; v < nv; ++v)
    for (int a = 0; a < 3; ++a) {
      const float x = d->positions[(size_t)v * 3 + a];
      ctx->bboxMin[a] = std::min(ctx->bboxMin[a], x);
      ctx->bboxMax[a] = std::max(ctx->bboxMax[a], x);
    }
  if (int rc = RebuildWide(ctx)) return rc;
  HIP_TRY(ctx, hipMemset(ctx->dCounters, 0, 64));
  if (ctx->dAccum) {
    HIP_TRY(ctx, hipMemset(ctx->dAccum, 0, (size_t)ctx->width * ctx->height * 16));
    ctx->iterations = 0;
  }
  InvalidateIntegratorCaches(&ctx->integ);
  ctx->bptCapacity = -1;
  return GPUGI_OK;
}

int gpugi_set_camera(gpugi_ctx* ctx, const gpugi_camera* camera) {
  DropGraph(ctx);
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
  DropGraph(ctx);
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
  ctx->bptCapacity = -1;
  return GPUGI_OK;
}

int gpugi_set_light_samples(gpugi_ctx* ctx, const gpugi_light_sample* samples, uint32_t n) {
  DropGraph(ctx);
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
  DropGraph(ctx);
  if (!ctx || max_path_length < 1 || max_path_length > 64)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_max_path_length: must be in [1, 64]");
  ctx->maxPathLength = max_path_length;
  InvalidateIntegratorCaches(&ctx->integ);
  return GPUGI_OK;
}

int gpugi_set_integrator_kernel(gpugi_ctx* ctx, int kernel) {
  DropGraph(ctx);
  if (!ctx || kernel < GPUGI_KERNEL_AUTO || kernel > GPUGI_KERNEL_WIDE)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_integrator_kernel: bad kernel");
  ctx->integratorKernel = kernel;
  return GPUGI_OK;
}

int gpugi_set_pt_schedule(gpugi_ctx* ctx, int schedule) {
  DropGraph(ctx);
  if (!ctx || (schedule != GPUGI_PT_MEGAKERNEL && schedule != GPUGI_PT_WAVEFRONT))
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_pt_schedule: bad schedule");
  ctx->ptSchedule = schedule;
  return GPUGI_OK;
}

namespace {

size_t Align256(size_t b) { return (b + 255) & ~(size_t)255; }
int MlpEnsureTiles(gpugi_ctx* ctx, uint32_t tiles);

// Scene bounds as the cache's position mapping: min and 1 / extent per axis.
int SceneBounds(gpugi_ctx* ctx, float mn[3], float invExtent[3]) {
  if (!ctx->hasScene) return Fail(ctx, GPUGI_ERR_STATE, "no scene loaded");
  for (int a = 0; a < 3; ++a) {
    const float e = ctx->bboxMax[a] - ctx->bboxMin[a];
    mn[a] = ctx->bboxMin[a];
    invExtent[a] = e > 0.0f ? 1.0f / e : 0.0f;
  }
  return GPUGI_OK;
}

// One frame of the cached path tracer: short paths and training paths, cache
// inference at the termination and tail vertices, resolve into the
// accumulation buffer, training targets, then `cacheTrainSteps` Adam steps.
int DrawCached(gpugi_ctx* ctx, uint32_t frameSeed) {
  if (!ctx->hasMlp) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_draw(PT_CACHE): call gpugi_mlp_init first");
  DrawParams p = MakeDrawParams(ctx, frameSeed);
  CacheParams cp = ctx->cacheCfg;
  const size_t P = (size_t)ctx->width * ctx->height;
  const uint32_t NT = CacheTrainPaths(ctx->width, ctx->height, cp.trainStride);
  const size_t L = (size_t)p.maxPathLength;
  const uint32_t S = ctx->cacheMaxSamples;
  const bool atPrimary = cp.query == GPUGI_CACHE_PRIMARY_INDIRECT;
  const size_t PW = atPrimary ? P : 0;  // wavefront buffers: rays, hit records, random states, NEE
  // the training pass's own traversal spill (it runs beside the pixel pass)
  const size_t trainSpillBytes =
      p.useWide ? ((size_t)NT + 64) / 64 * 64 * p.wide.spillDepth * sizeof(uint2) : 8;  // 64-thread blocks
  size_t off[19], total = 0;
  const size_t sizes[19] = {P * 16, P * 64, P * 16, P * 16, (size_t)NT * L * 96, (size_t)NT * 4, (size_t)NT * 4,
                            (size_t)NT * 32, (size_t)NT * 64, (size_t)NT * 16, ExclusiveScanTempBytes(NT), 4,
                            (size_t)S * 64, (size_t)S * 16, PW * 32, PW * 16, PW * 4, PW * 16, trainSpillBytes};
  for (int i = 0; i < 19; ++i) {
    off[i] = total;
    total += Align256(sizes[i]);
  }
  if (ctx->cacheBytes < total) {
    ctx->dCacheCount = nullptr;
    FreeDev(ctx->dCache);
    ctx->cacheBytes = 0;
    HIP_TRY(ctx, hipMalloc(&ctx->dCache, total));
    ctx->cacheBytes = total;
  }
  char* b = (char*)ctx->dCache;
  cp.pixelColor = (float*)(b + off[0]);
  cp.queryFeat = (float*)(b + off[1]);
  cp.queryWeight = (float*)(b + off[2]);
  float* queryPred = (float*)(b + off[3]);
  cp.queryPred = queryPred;
  cp.trainVert = (float*)(b + off[4]);
  cp.trainCount = (uint32_t*)(b + off[5]);
  uint32_t* offsets = (uint32_t*)(b + off[6]);
  cp.trainPath = (float*)(b + off[7]);
  cp.tailFeat = (float*)(b + off[8]);
  float* tailPred = (float*)(b + off[9]);
  cp.tailPred = tailPred;
  void* scanTemp = b + off[10];
  uint32_t* dTotal = (uint32_t*)(b + off[11]);
  float* feat = (float*)(b + off[12]);
  float* targets = (float*)(b + off[13]);
  float4* wRays = (float4*)(b + off[14]);
  int4* wHits = (int4*)(b + off[15]);
  uint32_t* wSeeds = (uint32_t*)(b + off[16]);
  float4* wNee = (float4*)(b + off[17]);
  void* trainSpill = b + off[18];
  cp.trainOffset = frameSeed % cp.trainStride;
  if (int rc = SceneBounds(ctx, cp.bboxMin, cp.bboxInvExtent)) return rc;

  // The training paths (few, latency-bound threads) and the whole training
  // chain run on the side stream beside the pixels' passes; the Adam steps
  // wait until the frame's queries have been inferred (so every query sees the
  // weights from before the frame), and the main stream joins the side stream
  // at the end of the frame.
  if (int rc = EnsureSideStream(ctx)) return rc;
  HIP_TRY(ctx, hipEventRecord(ctx->evFork, ctx->stream));
  HIP_TRY(ctx, hipStreamWaitEvent(ctx->side, ctx->evFork, 0));
  const hipStream_t trainStream = ctx->side;
  HIP_TRY(ctx, LaunchFillU32(cp.trainCount, 0u, NT, trainStream));
  HIP_TRY(ctx, LaunchCachedPaths(p, cp, trainSpill, true, trainStream));
  if (!atPrimary) {
    void* spill = nullptr;
    if (p.useWide) {
      const size_t threads = (size_t)(ctx->width / 8) * (ctx->height / 8) * 64;
      HIP_TRY(ctx, EnsureBuffer(&ctx->integ.spill, &ctx->integ.spillBytes,
                                std::max<size_t>(8, threads * p.wide.spillDepth * sizeof(uint2))));
      spill = ctx->integ.spill;
    }
    HIP_TRY(ctx, LaunchCachedPaths(p, cp, spill, false, ctx->stream));
  }
  HIP_TRY(ctx, LaunchMlpInfer(ctx->mlp, cp.tailFeat, tailPred, NT, trainStream));
  HIP_TRY(ctx, LaunchExclusiveScanU32(cp.trainCount, offsets, NT, scanTemp, dTotal, trainStream));
  HIP_TRY(ctx, LaunchCacheTargets(p, cp, offsets, dTotal, S, (uint32_t)std::max(ctx->cacheTrainSteps, 1), feat,
                                  targets, trainStream));
  if (atPrimary) {
    // every pixel's primary hit and NEE as wavefront passes through the trace kernels
    WideScene ws;
    if (p.useWide)
      if (int rc = PrepareTraceWide(ctx, ctx->wideStack, &ws)) return rc;
    auto trace = [&](int query) -> int {
      if (p.useWide)
        HIP_TRY(ctx, LaunchTraceWide(ws, ctx->wideStack, query, (const gpugi_ray*)wRays, (gpugi_hit*)wHits, P,
                                     ctx->wideRefill, ctx->wideWait, Watertight(ctx), ctx->stream));
      else
        HIP_TRY(ctx, LaunchTraceRef(MakeRef(ctx), query, (const gpugi_ray*)wRays, (gpugi_hit*)wHits, P,
                                    Watertight(ctx), ctx->stream));
      return GPUGI_OK;
    };
    HIP_TRY(ctx, LaunchPrimaryRaygen(p, wRays, wSeeds, ctx->stream));
    if (int rc = trace(GPUGI_CLOSEST_HIT)) return rc;
    HIP_TRY(ctx, LaunchPrimaryShade(p, cp, wRays, wHits, wSeeds, wNee, ctx->stream));
    if (int rc = trace(GPUGI_ANY_HIT)) return rc;
    HIP_TRY(ctx, LaunchPrimaryFinish(p, cp, wHits, wNee, ctx->stream));
  }
  HIP_TRY(ctx, LaunchMlpInfer(ctx->mlp, cp.queryFeat, queryPred, P, ctx->stream));
  HIP_TRY(ctx, hipEventRecord(ctx->evInfer, ctx->stream));
  HIP_TRY(ctx, LaunchCacheResolve(p, cp, ctx->stream));
  // The sample count stays on the device: each step's kernels read it and
  // take their chunk of min(count, S) samples, so the host does not wait.
  ctx->dCacheCount = dTotal;
  const int steps = ctx->cacheTrainSteps;
  if (steps > 0) {
    if (int rc = MlpEnsureTiles(ctx, MlpTiles(S / (uint32_t)steps))) return rc;
    for (int s = 0; s < steps; ++s) {
      const MlpBatch batch{dTotal, S, (uint32_t)s, (uint32_t)steps};
      HIP_TRY(ctx, LaunchMlpGradientTiles(ctx->mlp, feat, targets, batch, trainStream));
      // the first weight update waits for the frame's query inference, which reads the weights
      if (s == 0) HIP_TRY(ctx, hipStreamWaitEvent(ctx->side, ctx->evInfer, 0));
      AdamParams ap{ctx->cacheLr, 0.9f, 0.99f, 1e-8f, ++ctx->mlpStep};
      HIP_TRY(ctx, LaunchMlpReduce(ctx->mlp, batch, ctx->dMlpLoss, &ap, trainStream));
    }
  }
  HIP_TRY(ctx, hipEventRecord(ctx->evJoin, ctx->side));
  HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream, ctx->evJoin, 0));
  return GPUGI_OK;
}

// Buffers and launch parameters of the wavefront path tracer (allocates what
// is missing, so a capture can follow without allocating).
int PrepareWavefront(gpugi_ctx* ctx, const DrawParams& p, WavefrontTrace* tr) {
  const size_t pixels = (size_t)ctx->width * ctx->height;
  const size_t bytes = WavefrontBytes(pixels);
  if (ctx->waveCap < bytes) {
    FreeDev(ctx->dWave);
    ctx->waveCap = 0;
    HIP_TRY(ctx, hipMalloc(&ctx->dWave, bytes));
    ctx->waveCap = bytes;
  }
  if (int rc = EnsureSideStream(ctx)) return rc;
  tr->wide = p.useWide;
  if (tr->wide) {
    if (int rc = PrepareTraceWide(ctx, ctx->wideStack, &tr->ws)) return rc;
    // the shadow traces' own spill stacks and batch counters (the grid's size)
    const size_t spill = std::max<size_t>(8, (size_t)tr->ws.numThreads * tr->ws.spillDepth * sizeof(uint2));
    if (ctx->waveSpillBytes < spill) {
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
      HIP_TRY(ctx, hipStreamSynchronize(ctx->side));
      HIP_TRY(ctx, EnsureBuffer(&ctx->dWaveSpill, &ctx->waveSpillBytes, spill));
    }
    if (!ctx->dWaveQueue) HIP_TRY(ctx, hipMalloc(&ctx->dWaveQueue, kWideQueueWords * sizeof(uint32_t)));
    tr->wsShadow = tr->ws;
    tr->wsShadow.spill = (uint2*)ctx->dWaveSpill;
    tr->wsShadow.queue = (uint32_t*)ctx->dWaveQueue;
  }
  tr->side = ctx->side;
  tr->evShade = ctx->evShade;
  tr->evResolve = ctx->evResolve;
  tr->stack = ctx->wideStack;
  tr->refill = ctx->wideRefill;
  tr->wait = ctx->wideWait;
  tr->passBlocks = (uint32_t)ctx->numCUs * 8u;  // 8 resident blocks of 256 lanes per CU
  return GPUGI_OK;
}

// GPUGI_PT, or the eye paths of GPUGI_BPT after its light pass, as per-bounce
// passes over device ray queues (wavefront_pt.hip).
int DrawWavefront(gpugi_ctx* ctx, int integrator, DrawParams p) {
  WavefrontTrace tr;
  if (int rc = PrepareWavefront(ctx, p, &tr)) return rc;
  const WavefrontBuffers b = CarveWavefront(ctx->dWave, (size_t)ctx->width * ctx->height);
  if (integrator == GPUGI_PT) {
    HIP_TRY(ctx, DrawPTWavefront(p, tr, b, ctx->stream));
    return GPUGI_OK;
  }
  std::string err;
  const LightCacheEntryDev* cache = nullptr;
  const uint32_t* total = nullptr;
  const hipError_t e = BptLightPass(&ctx->integ, &p, ctx->stream, &err, &cache, &total);
  if (e != hipSuccess) return Fail(ctx, GPUGI_ERR_HIP, std::string("gpugi_draw: ") + hipGetErrorString(e));
  if (!err.empty()) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_draw: " + err);
  HIP_TRY(ctx, DrawBPTWavefront(p, tr, b, cache, total, ctx->stream));
  return GPUGI_OK;
}

}  // namespace

int gpugi_draw(gpugi_ctx* ctx, int integrator, uint32_t frame_seed) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_draw: null context");
  if (!ctx->hasScene || !ctx->dAccum || !ctx->hasCamera || ctx->numLights == 0)
    return Fail(ctx, GPUGI_ERR_STATE, "gpugi_draw: scene, screen, camera and light samples must be set first");
  if (integrator < GPUGI_PT || integrator > GPUGI_PT_CACHE)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_draw: unknown integrator");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  if (integrator == GPUGI_PT_CACHE) {
    if (ctx->specificPathLength >= 0)
      return Fail(ctx, GPUGI_ERR_STATE, "gpugi_draw: the cached path tracer does not isolate path lengths");
    if (int rc = DrawCached(ctx, frame_seed)) return rc;
    ++ctx->iterations;
    return GPUGI_OK;
  }
  DrawParams p = MakeDrawParams(ctx, frame_seed);
  if ((integrator == GPUGI_PT || integrator == GPUGI_BPT) && ctx->ptSchedule == GPUGI_PT_WAVEFRONT) {
    if (int rc = DrawWavefront(ctx, integrator, p)) return rc;
    ++ctx->iterations;
    return GPUGI_OK;
  }
  std::string err;
  hipError_t e = DrawIntegrator(&ctx->integ, integrator, p, ctx->stream, &err);
  if (e != hipSuccess) return Fail(ctx, GPUGI_ERR_HIP, "gpugi_draw: " + err + hipGetErrorString(e));
  if (!err.empty()) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_draw: " + err);
  ++ctx->iterations;
  return GPUGI_OK;
}

static DrawParams MakeDrawParams(gpugi_ctx* ctx, uint32_t frame_seed) {
  DrawParams p;
  p.ref = MakeRef(ctx);
  p.wide = MakeWide(ctx, kWideLdsStack);
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
  p.rrStartDepth = ctx->rrStartDepth;
  p.specificPathLength = ctx->specificPathLength;
  p.watertight = Watertight(ctx);
  p.counters = (uint32_t*)ctx->dCounters;
  p.bptCapacity = ctx->bptCapacity;
  if (ctx->anyTextureBound) {
    p.tex.matTex = (const int4*)ctx->dMatTex;
    p.tex.desc = (const TextureDesc*)ctx->dTexDesc;
  }
  p.env = (const float4*)ctx->dEnv;
  p.envSize = ctx->envSize;
  return p;
}

int gpugi_capture_draw(gpugi_ctx* ctx, int integrator, uint32_t frame_seed) {
  DropGraph(ctx);
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_capture_draw: null context");
  if (!ctx->hasScene || !ctx->dAccum || !ctx->hasCamera || ctx->numLights == 0)
    return Fail(ctx, GPUGI_ERR_STATE, "gpugi_capture_draw: scene, screen, camera and light samples must be set first");
  if (integrator != GPUGI_PT || ctx->ptSchedule != GPUGI_PT_WAVEFRONT)
    return Fail(ctx, GPUGI_ERR_UNSUPPORTED, "gpugi_capture_draw: only GPUGI_PT with the wavefront schedule is recorded");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  const DrawParams p = MakeDrawParams(ctx, frame_seed);
  WavefrontTrace tr;
  if (int rc = PrepareWavefront(ctx, p, &tr)) return rc;  // every buffer exists before the capture
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const hipError_t e = DrawPTWavefront(p, tr, CarveWavefront(ctx->dWave, (size_t)ctx->width * ctx->height),
                                       ctx->stream);
  hipGraph_t graph = nullptr;
  const hipError_t ee = hipStreamEndCapture(ctx->stream, &graph);
  if (e != hipSuccess || ee != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return Fail(ctx, GPUGI_ERR_HIP, std::string("gpugi_capture_draw: ") + hipGetErrorString(e != hipSuccess ? e : ee));
  }
  const hipError_t ei = hipGraphInstantiate(&ctx->drawGraph, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    ctx->drawGraph = nullptr;
    return Fail(ctx, GPUGI_ERR_HIP, std::string("gpugi_capture_draw: ") + hipGetErrorString(ei));
  }
  return GPUGI_OK;
}

int gpugi_replay_draw(gpugi_ctx* ctx) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_replay_draw: null context");
  if (!ctx->drawGraph) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_replay_draw: no recorded draw (or a setter dropped it)");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  HIP_TRY(ctx, hipGraphLaunch(ctx->drawGraph, ctx->stream));
  ++ctx->iterations;
  return GPUGI_OK;
}

int gpugi_bpt_warmup(gpugi_ctx* ctx, uint32_t frame_seed, uint64_t* num_light_cache_entries,
                     uint64_t* light_path_length_sum) {
  if (!ctx || !num_light_cache_entries || !light_path_length_sum)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_bpt_warmup: null argument");
  if (!ctx->hasScene || !ctx->dAccum || !ctx->hasCamera || ctx->numLights == 0)
    return Fail(ctx, GPUGI_ERR_STATE, "gpugi_bpt_warmup: scene, screen, camera and light samples must be set first");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  BptWarmupCounts c{0, 0};
  HIP_TRY(ctx, BptWarmup(&ctx->integ, MakeDrawParams(ctx, frame_seed), ctx->stream, &c));
  *num_light_cache_entries = c.numLightCacheEntries;
  *light_path_length_sum = c.lightPathLengthSum;
  return GPUGI_OK;
}

int gpugi_bpt_set_light_cache(gpugi_ctx* ctx, int32_t capacity, int32_t average_light_path_length) {
  if (!ctx || capacity < 1) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_bpt_set_light_cache: capacity must be positive");
  ctx->bptCapacity = capacity;
  ctx->bptAverageLength = average_light_path_length;
  return GPUGI_OK;
}

int gpugi_bpt_read_light_cache(gpugi_ctx* ctx, void* entries, uint32_t max_entries, uint32_t* num_entries) {
  if (!ctx || !num_entries || (max_entries > 0 && !entries))
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_bpt_read_light_cache: null argument");
  if (!ctx->integ.lastTotal || !ctx->integ.lightCache)
    return Fail(ctx, GPUGI_ERR_STATE, "gpugi_bpt_read_light_cache: no BPT draw since the last scene/screen change");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  uint32_t total = 0;
  HIP_TRY(ctx, hipMemcpyAsync(&total, ctx->integ.lastTotal, 4, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  const uint32_t avail = (uint32_t)std::min<int64_t>((int64_t)total, (int64_t)ctx->bptCapacity);
  const uint32_t n = std::min(avail, max_entries);
  if (n > 0) {
    HIP_TRY(ctx, hipMemcpyAsync(entries, ctx->integ.lightCache, (size_t)n * 64, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  *num_entries = total;
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

// Re-uploads the texture descriptors and the material bindings.
static int UploadTextureTables(gpugi_ctx* ctx) {
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->side) HIP_TRY(ctx, hipStreamSynchronize(ctx->side));
  FreeDev(ctx->dTexDesc);
  FreeDev(ctx->dMatTex);
  ctx->anyTextureBound = false;
  for (const int4& t : ctx->matTex)
    if (t.x >= 0 || t.y >= 0 || t.z >= 0 || t.w >= 0) ctx->anyTextureBound = true;
  if (!ctx->anyTextureBound) return GPUGI_OK;
  HIP_TRY(ctx, hipMalloc(&ctx->dTexDesc, ctx->texDescs.size() * sizeof(TextureDesc)));
  HIP_TRY(ctx, hipMemcpy(ctx->dTexDesc, ctx->texDescs.data(), ctx->texDescs.size() * sizeof(TextureDesc),
                         hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMalloc(&ctx->dMatTex, ctx->matTex.size() * sizeof(int4)));
  HIP_TRY(ctx, hipMemcpy(ctx->dMatTex, ctx->matTex.data(), ctx->matTex.size() * sizeof(int4), hipMemcpyHostToDevice));
  return GPUGI_OK;
}

// The sRGB transfer function per byte value, evaluated in double.
static void SrgbTable(float lut[256]) {
  for (int i = 0; i < 256; ++i) {
    const double c = i / 255.0;
    lut[i] = (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
  }
}

int gpugi_add_texture(gpugi_ctx* ctx, int width, int height, int format, const void* texels, int* id) {
  DropGraph(ctx);
  if (!ctx || !texels || !id) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_add_texture: null argument");
  if (!ctx->hasScene) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_add_texture: load a scene first");
  if (width <= 0 || height <= 0 || width > 32768 || height > 32768)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_add_texture: bad size");
  if (format != GPUGI_TEXTURE_RGBA8_SRGB && format != GPUGI_TEXTURE_RGBA32F)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_add_texture: unknown format");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  const size_t n = (size_t)width * height;
  std::vector<float4> lin;
  try {
    lin.resize(n);
  } catch (const std::bad_alloc&) {
    return Fail(ctx, GPUGI_ERR_NOMEM, "gpugi_add_texture: host staging");
  }
  if (format == GPUGI_TEXTURE_RGBA8_SRGB) {
    // SRGB8_ALPHA8: colour channels decoded before filtering, alpha linear
    float lut[256];
    SrgbTable(lut);
    const uint8_t* b = static_cast<const uint8_t*>(texels);
    for (size_t i = 0; i < n; ++i)
      lin[i] = make_float4(lut[b[4 * i]], lut[b[4 * i + 1]], lut[b[4 * i + 2]], (float)b[4 * i + 3] / 255.0f);
  } else {
    std::memcpy(lin.data(), texels, n * sizeof(float4));
  }
  void* buf = nullptr;
  HIP_TRY(ctx, hipMalloc(&buf, n * sizeof(float4)));
  const hipError_t e = hipMemcpy(buf, lin.data(), n * sizeof(float4), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hipFree(buf);
    return Fail(ctx, GPUGI_ERR_HIP, std::string("gpugi_add_texture: ") + hipGetErrorString(e));
  }
  ctx->texBufs.push_back(buf);
  ctx->texDescs.push_back(TextureDesc{static_cast<const float4*>(buf), width, height});
  *id = (int)ctx->texDescs.size() - 1;
  return UploadTextureTables(ctx);
}

int gpugi_set_material_texture(gpugi_ctx* ctx, uint32_t material, int slot, int texture) {
  DropGraph(ctx);
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_material_texture: null context");
  if (!ctx->hasScene) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_set_material_texture: load a scene first");
  if (material >= ctx->matTex.size()) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_material_texture: bad material");
  if (slot < GPUGI_TEX_DIFFUSE || slot > GPUGI_TEX_EMISSIVITY)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_material_texture: bad slot");
  if (texture < -1 || texture >= (int)ctx->texDescs.size())
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_material_texture: unknown texture");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  int4& t = ctx->matTex[material];
  (slot == GPUGI_TEX_DIFFUSE ? t.x : slot == GPUGI_TEX_OPACITY ? t.y : slot == GPUGI_TEX_REFLECTIVENESS ? t.z : t.w) =
      texture;
  return UploadTextureTables(ctx);
}

int gpugi_set_environment_map(gpugi_ctx* ctx, int size, const uint8_t* faces) {
  DropGraph(ctx);
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_environment_map: null context");
  if (size < 0 || size > 16384 || (size > 0 && !faces))
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_environment_map: bad size or null faces");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // draws in flight may still read the old map
  if (ctx->side) HIP_TRY(ctx, hipStreamSynchronize(ctx->side));
  FreeDev(ctx->dEnv);
  ctx->envSize = 0;
  if (size == 0) return GPUGI_OK;  // the black default map (renderersystem.cpp:39-47)
  // SRGB8 texels are decoded to linear RGB before filtering, as GL does for
  // sRGB textures: the sRGB transfer function evaluated in double per byte.
  float lut[256];
  SrgbTable(lut);
  const size_t n = (size_t)6 * size * size;
  std::vector<float4> texels(n);
  for (size_t i = 0; i < n; ++i)
    texels[i] = make_float4(lut[faces[3 * i]], lut[faces[3 * i + 1]], lut[faces[3 * i + 2]], 1.0f);
  HIP_TRY(ctx, hipMalloc(&ctx->dEnv, n * sizeof(float4)));
  HIP_TRY(ctx, hipMemcpy(ctx->dEnv, texels.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  ctx->envSize = size;
  return GPUGI_OK;
}

int gpugi_set_reference_image(gpugi_ctx* ctx, const float* rgb, int width, int height) {
  if (!ctx || !rgb) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_reference_image: null argument");
  if (width <= 0 || height <= 0) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_reference_image: empty image");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  const size_t bytes = (size_t)width * height * 12;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // an MSE computation may still read the old image
  FreeDev(ctx->dRefImage);
  ctx->refW = ctx->refH = 0;
  HIP_TRY(ctx, hipMalloc(&ctx->dRefImage, bytes));
  if (!ctx->dMseTemp) HIP_TRY(ctx, hipMalloc(&ctx->dMseTemp, ImageMseTempBytes() + 256));
  HIP_TRY(ctx, hipMemcpy(ctx->dRefImage, rgb, bytes, hipMemcpyHostToDevice));
  ctx->refW = width;
  ctx->refH = height;
  return GPUGI_OK;
}

int gpugi_compute_mse(gpugi_ctx* ctx, uint32_t divisor, float* mse) {
  if (!ctx || !mse) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_compute_mse: null argument");
  *mse = INFINITY;  // no reference image: ComputeCurrentMSE returns +inf
  if (!ctx->dRefImage) return GPUGI_OK;
  if (!ctx->dAccum) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_compute_mse: no screen set");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  float* dOut = reinterpret_cast<float*>(static_cast<char*>(ctx->dMseTemp) + ImageMseTempBytes());
  HIP_TRY(ctx, LaunchImageMse(static_cast<const float4*>(ctx->dAccum), ctx->width, ctx->height, divisor,
                              static_cast<const float*>(ctx->dRefImage), ctx->refW, ctx->refH, ctx->dMseTemp, dOut,
                              ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(mse, dOut, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return GPUGI_OK;
}

// Pinned host-buffer trace in chunks through two device slots: the copy stream
// brings chunk k + 1 in and takes chunk k - 1 out while the main stream
// traces chunk k, so the copies overlap the traversal (the caller checked
// that both host buffers are pinned).
static int TraceHostChunked(gpugi_ctx* ctx, int query, bool wide, const gpugi_ray* rays, gpugi_hit* hits, size_t n) {
  constexpr size_t kChunk = (size_t)1 << 22;
  const size_t slotBytes = kChunk * (sizeof(gpugi_ray) + sizeof(gpugi_hit));
  if (int rc = EnsureScratch(ctx, 2 * slotBytes)) return rc;
  if (!ctx->copy) {
    HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->copy, hipStreamNonBlocking));
    for (int s = 0; s < 2; ++s) {
      HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->evIn[s], hipEventDisableTiming));
      HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->evTraced[s], hipEventDisableTiming));
    }
    HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->evStart, hipEventDisableTiming));
  }
  WideScene ws;
  if (wide)
    if (int rc = PrepareTraceWide(ctx, ctx->wideStack, &ws)) return rc;
  // the slots are free once the main stream's earlier work is done
  HIP_TRY(ctx, hipEventRecord(ctx->evStart, ctx->stream));
  HIP_TRY(ctx, hipStreamWaitEvent(ctx->copy, ctx->evStart, 0));
  const size_t numChunks = (n + kChunk - 1) / kChunk;
  auto slotRays = [&](int s) { return (gpugi_ray*)((char*)ctx->dScratch + s * slotBytes); };
  auto slotHits = [&](int s) { return (gpugi_hit*)((char*)ctx->dScratch + s * slotBytes + kChunk * sizeof(gpugi_ray)); };
  auto copyIn = [&](size_t k) -> int {
    const int s = (int)(k & 1);
    const size_t first = k * kChunk, m = std::min(kChunk, n - first);
    HIP_TRY(ctx, hipMemcpyAsync(slotRays(s), rays + first, m * sizeof(gpugi_ray), hipMemcpyHostToDevice, ctx->copy));
    HIP_TRY(ctx, hipEventRecord(ctx->evIn[s], ctx->copy));
    return GPUGI_OK;
  };
  auto copyOut = [&](size_t k) -> int {
    const int s = (int)(k & 1);
    const size_t first = k * kChunk, m = std::min(kChunk, n - first);
    HIP_TRY(ctx, hipStreamWaitEvent(ctx->copy, ctx->evTraced[s], 0));
    HIP_TRY(ctx, hipMemcpyAsync(hits + first, slotHits(s), m * sizeof(gpugi_hit), hipMemcpyDeviceToHost, ctx->copy));
    return GPUGI_OK;
  };
  // copy stream: in(0) in(1) out(0) in(2) out(1) in(3) ...; slot k & 1 is
  // reused by in(k) only after out(k - 2), which waited for trace(k - 2)
  if (int rc = copyIn(0)) return rc;
  for (size_t k = 0; k < numChunks; ++k) {
    if (k + 1 < numChunks)
      if (int rc = copyIn(k + 1)) return rc;
    const int s = (int)(k & 1);
    const size_t m = std::min(kChunk, n - k * kChunk);
    HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream, ctx->evIn[s], 0));
    if (wide)
      HIP_TRY(ctx, LaunchTraceWide(ws, ctx->wideStack, query, slotRays(s), slotHits(s), m, ctx->wideRefill,
                                   ctx->wideWait, Watertight(ctx), ctx->stream));
    else
      HIP_TRY(ctx, LaunchTraceRef(MakeRef(ctx), query, slotRays(s), slotHits(s), m, Watertight(ctx), ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->evTraced[s], ctx->stream));
    if (int rc = copyOut(k)) return rc;
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->copy));
  uint32_t flag = 0;
  HIP_TRY(ctx, hipMemcpyAsync(&flag, (uint32_t*)ctx->dCounters + 1, 4, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (flag) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_trace: traversal stack overflow (bug in the stack bound)");
  return GPUGI_OK;
}

// Whether [p, p + bytes) is pinned host memory (gpugi_host_alloc or any other
// page-locked allocation the runtime knows): both ends must be.
static bool IsPinned(const void* p, size_t bytes) {
  const char* ends[2] = {static_cast<const char*>(p), static_cast<const char*>(p) + bytes - 1};
  for (const char* q : ends) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, q) != hipSuccess) {
      (void)hipGetLastError();  // pageable memory: the query fails
      return false;
    }
    if (a.type != hipMemoryTypeHost) return false;
  }
  return true;
}

int gpugi_host_alloc(gpugi_ctx* ctx, size_t bytes, void** host_ptr) {
  if (!ctx || !host_ptr || bytes == 0) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_host_alloc: invalid argument");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  void* p = nullptr;
  hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
  if (e != hipSuccess) return Fail(ctx, GPUGI_ERR_NOMEM, std::string("gpugi_host_alloc: ") + hipGetErrorString(e));
  ctx->hostAllocs.insert(p);
  *host_ptr = p;
  return GPUGI_OK;
}

int gpugi_host_free(gpugi_ctx* ctx, void* host_ptr) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_host_free: null context");
  auto it = ctx->hostAllocs.find(host_ptr);
  if (it == ctx->hostAllocs.end())
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_host_free: pointer not allocated by gpugi_host_alloc");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->copy) HIP_TRY(ctx, hipStreamSynchronize(ctx->copy));
  HIP_TRY(ctx, hipHostFree(host_ptr));
  ctx->hostAllocs.erase(it);
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
  // the stack-overflow flag reports this call only (a fill kernel, so the call can be captured)
  HIP_TRY(ctx, LaunchFillU32((uint32_t*)ctx->dCounters + 1, 0u, 1, ctx->stream));
  const bool wide = kernel == GPUGI_KERNEL_WIDE || (kernel == GPUGI_KERNEL_AUTO && ctx->hasWide);
  // pinned host buffers: chunked copies that overlap the traversal; pageable
  // ones go through one staged copy each way, which the runtime does faster
  if (!rays_on_device && n > ((size_t)1 << 22) && IsPinned(rays, n * sizeof(gpugi_ray)) &&
      IsPinned(hits, n * sizeof(gpugi_hit)))
    return TraceHostChunked(ctx, query, wide, rays, hits, n);
  const gpugi_ray* dr = rays;
  gpugi_hit* dh = hits;
  if (!rays_on_device) {
    if (int rc = EnsureScratch(ctx, n * (sizeof(gpugi_ray) + sizeof(gpugi_hit)))) return rc;
    dr = (const gpugi_ray*)ctx->dScratch;
    dh = (gpugi_hit*)((char*)ctx->dScratch + n * sizeof(gpugi_ray));
    HIP_TRY(ctx, hipMemcpyAsync((void*)dr, rays, n * sizeof(gpugi_ray), hipMemcpyHostToDevice, ctx->stream));
  }
  if (wide) {
    WideScene ws;
    if (int rc = PrepareTraceWide(ctx, ctx->wideStack, &ws)) return rc;
    HIP_TRY(ctx, LaunchTraceWide(ws, ctx->wideStack, query, dr, dh, n, ctx->wideRefill, ctx->wideWait,
                                 Watertight(ctx), ctx->stream));
  } else {
    HIP_TRY(ctx, LaunchTraceRef(MakeRef(ctx), query, dr, dh, n, Watertight(ctx), ctx->stream));
  }
  if (!rays_on_device) {
    HIP_TRY(ctx, hipMemcpyAsync(hits, dh, n * sizeof(gpugi_hit), hipMemcpyDeviceToHost, ctx->stream));
    uint32_t flag = 0;
    HIP_TRY(ctx, hipMemcpyAsync(&flag, (uint32_t*)ctx->dCounters + 1, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (flag) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_trace: traversal stack overflow (bug in the stack bound)");
  }
  return GPUGI_OK;
}

int gpugi_trace_noalloc(gpugi_ctx* ctx, int query, int kernel, const gpugi_ray* rays, gpugi_hit* hits, size_t n) {
  if (!ctx || (n > 0 && (!rays || !hits))) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_trace_noalloc: null argument");
  if (query != GPUGI_CLOSEST_HIT && query != GPUGI_ANY_HIT)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_trace_noalloc: bad query");
  if (kernel < GPUGI_KERNEL_AUTO || kernel > GPUGI_KERNEL_WIDE)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_trace_noalloc: bad kernel");
  if (!ctx->hasScene) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_trace_noalloc: no scene loaded");
  if (n > 0xffffff00ull) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_trace_noalloc: at most 2^32 - 256 rays per call");
  if (n == 0) return GPUGI_OK;
  if (!ctx->noAllocScratch) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_trace_noalloc: scratch not set up");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  HIP_TRY(ctx, LaunchFillU32((uint32_t*)ctx->dCounters + 1, 0u, 1, ctx->stream));
  const bool wide = kernel == GPUGI_KERNEL_WIDE || (kernel == GPUGI_KERNEL_AUTO && ctx->hasWide);
  if (!wide) {  // the escape walk keeps no stack: nothing but the caller's buffers
    HIP_TRY(ctx, LaunchTraceRef(MakeRef(ctx), query, rays, hits, n, Watertight(ctx), ctx->stream));
    return GPUGI_OK;
  }
  if (!ctx->hasWide) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_trace_noalloc: no wide hierarchy");
  int blocksPerCU = 0;
  HIP_TRY(ctx, WideBlocksPerCU(ctx->wideStack, &blocksPerCU));
  if (blocksPerCU < 1) blocksPerCU = 1;
  WideScene s = MakeWide(ctx, ctx->wideStack);
  // the resident grid (whole blocks per XCD), shrunk until its spill stacks fit the array
  uint32_t blocks = std::max<uint32_t>(8, (uint32_t)(ctx->numCUs * blocksPerCU) / 8 * 8);
  const size_t perBlock = (size_t)s.spillDepth * kWideBlock * sizeof(uint2);
  const size_t avail = GPUGI_NOALLOC_SCRATCH_BYTES - kNoAllocQueueBytes;
  if (perBlock > 0) blocks = (uint32_t)std::min<size_t>(blocks, avail / perBlock / 8 * 8);
  if (blocks < 8)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_trace_noalloc: the scene's traversal stack needs " +
                                        std::to_string(8 * perBlock + kNoAllocQueueBytes) +
                                        " bytes of scratch for the smallest grid, more than the " +
                                        std::to_string(GPUGI_NOALLOC_SCRATCH_BYTES) + " the array holds");
  s.numThreads = blocks * (uint32_t)kWideBlock;
  s.queue = static_cast<uint32_t*>(ctx->noAllocScratch);
  s.spill = reinterpret_cast<uint2*>(static_cast<char*>(ctx->noAllocScratch) + kNoAllocQueueBytes);
  HIP_TRY(ctx, LaunchTraceWide(s, ctx->wideStack, query, rays, hits, n, ctx->wideRefill, ctx->wideWait,
                               Watertight(ctx), ctx->stream));
  return GPUGI_OK;
}

namespace {

// Parent pointers, the nodes grouped by tree level and the importance,
// subtree and material buffers of the current scene (made on first use).
int EnsureHierarchy(gpugi_ctx* ctx) {
  if (ctx->hierTreeReady) return GPUGI_OK;
  std::vector<int32_t> parent;
  std::vector<uint32_t> order, start;
  if (!BuildHierTree(ctx->hHier.data(), ctx->numNodes, &parent, &order, &start))
    return Fail(ctx, GPUGI_ERR_FORMAT, "hierarchy: the nodes are not in preorder");
  const size_t nn = ctx->numNodes;
  for (void** p : {&ctx->dHierParent, &ctx->dHierLevel, &ctx->dHierImp, &ctx->dHierSubtree, &ctx->dHierMat})
    FreeDev(*p);
  HIP_TRY(ctx, hipMalloc(&ctx->dHierParent, nn * 4));
  HIP_TRY(ctx, hipMalloc(&ctx->dHierLevel, nn * 4));
  HIP_TRY(ctx, hipMalloc(&ctx->dHierImp, (nn + ctx->numSlots) * 8));
  HIP_TRY(ctx, hipMalloc(&ctx->dHierSubtree, nn * 4));
  HIP_TRY(ctx, hipMalloc(&ctx->dHierMat, nn * 48));
  HIP_TRY(ctx, hipMemcpy(ctx->dHierParent, parent.data(), nn * 4, hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(ctx->dHierLevel, order.data(), nn * 4, hipMemcpyHostToDevice));
  ctx->hierTree.numNodes = ctx->numNodes;
  ctx->hierTree.parent = static_cast<const int32_t*>(ctx->dHierParent);
  ctx->hierTree.levelOrder = static_cast<const uint32_t*>(ctx->dHierLevel);
  ctx->hierTree.levelStart = std::move(start);
  ctx->hierTreeReady = true;
  return GPUGI_OK;
}

}  // namespace

int gpugi_hierarchy_update_importance(gpugi_ctx* ctx, int iterations) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_hierarchy_update_importance: null context");
  if (iterations < 1 || iterations > 4096)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_hierarchy_update_importance: iterations in [1, 4096]");
  if (!ctx->hasScene || !ctx->dAccum || !ctx->hasCamera || ctx->numLights == 0)
    return Fail(ctx, GPUGI_ERR_STATE,
                "gpugi_hierarchy_update_importance: scene, screen, camera and light samples must be set first");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  if (int rc = EnsureHierarchy(ctx)) return rc;
  ctx->hasHierImportance = false;
  const DrawParams p = MakeDrawParams(ctx, 0);
  const size_t entries = (size_t)ctx->numNodes + ctx->numSlots;
  unsigned long long* acc = static_cast<unsigned long long*>(ctx->dHierImp);
  HIP_TRY(ctx, hipMemsetAsync(acc, 0, entries * 8, ctx->stream));
  for (int i = 1; i <= iterations; ++i)  // NumImportanceIterations = i (hierarchyimportance.cpp:93-103)
    HIP_TRY(ctx, LaunchHierAcquire(p, i, acc, ctx->numNodes, ctx->stream));
  HIP_TRY(ctx, LaunchHierSubtreeSum(p.ref, ctx->hierTree, acc, ctx->stream));
  HIP_TRY(ctx, LaunchHierToFloat(acc, entries, ctx->stream));
  HIP_TRY(ctx, LaunchHierPropagate(p.ref, ctx->hierTree, iterations, static_cast<float2*>(ctx->dHierImp),
                                   static_cast<float*>(ctx->dHierSubtree), ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  ctx->hasHierImportance = true;
  return GPUGI_OK;
}

int gpugi_hierarchy_compute_materials(gpugi_ctx* ctx) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_hierarchy_compute_materials: null context");
  if (!ctx->hasScene) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_hierarchy_compute_materials: no scene loaded");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  if (int rc = EnsureHierarchy(ctx)) return rc;
  DrawParams p = MakeDrawParams(ctx, 0);
  HIP_TRY(ctx, LaunchHierMaterials(p, ctx->hierTree, static_cast<uint32_t*>(ctx->dHierMat), ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  ctx->hasHierMaterials = true;
  return GPUGI_OK;
}

int gpugi_hierarchy_read(gpugi_ctx* ctx, float* importance, size_t importance_floats, float* subtree,
                         size_t subtree_floats, uint32_t* materials, size_t material_words) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_hierarchy_read: null context");
  const size_t nn = ctx->numNodes;
  if ((importance && importance_floats < (nn + ctx->numSlots) * 2) || (subtree && subtree_floats < nn) ||
      (materials && material_words < nn * 12))
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_hierarchy_read: buffer too small");
  if ((importance || subtree) && !ctx->hasHierImportance)
    return Fail(ctx, GPUGI_ERR_STATE, "gpugi_hierarchy_read: no importance computed for this scene");
  if (materials && !ctx->hasHierMaterials)
    return Fail(ctx, GPUGI_ERR_STATE, "gpugi_hierarchy_read: no materials computed for this scene");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  if (importance)
    HIP_TRY(ctx, hipMemcpyAsync(importance, ctx->dHierImp, (nn + ctx->numSlots) * 8, hipMemcpyDeviceToHost,
                                ctx->stream));
  if (subtree) HIP_TRY(ctx, hipMemcpyAsync(subtree, ctx->dHierSubtree, nn * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (materials) HIP_TRY(ctx, hipMemcpyAsync(materials, ctx->dHierMat, nn * 48, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return GPUGI_OK;
}

int gpugi_trace_importance_break(gpugi_ctx* ctx, int query, const gpugi_ray* rays, gpugi_hit* hits,
                                 gpugi_node_hit* nodes, size_t n, const float* importance, uint32_t num_nodes,
                                 float threshold, float ray_threshold) {
  if (!ctx || (n > 0 && (!rays || !hits)))
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_trace_importance_break: null argument");
  if (!importance && !ctx->hasHierImportance)
    return Fail(ctx, GPUGI_ERR_STATE,
                "gpugi_trace_importance_break: no importance given and none computed (gpugi_hierarchy_update_importance)");
  if (query != GPUGI_CLOSEST_HIT && query != GPUGI_ANY_HIT)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_trace_importance_break: bad query");
  if (!ctx->hasScene) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_trace_importance_break: no scene loaded");
  if (importance && num_nodes != ctx->numNodes)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_trace_importance_break: need one importance pair per node");
  if (n > 0xffffff00ull) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_trace_importance_break: at most 2^32 - 256 rays");
  if (n == 0) return GPUGI_OK;
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  const size_t bRays = n * sizeof(gpugi_ray), bHits = n * sizeof(gpugi_hit), bNodes = n * sizeof(gpugi_node_hit);
  const size_t bImp = importance ? (size_t)num_nodes * 8 : 0;
  if (int rc = EnsureScratch(ctx, bRays + bHits + bNodes + bImp)) return rc;
  char* base = (char*)ctx->dScratch;
  gpugi_ray* dr = (gpugi_ray*)base;
  gpugi_hit* dh = (gpugi_hit*)(base + bRays);
  int4* dn = (int4*)(base + bRays + bHits);
  // the caller's importance, or the one gpugi_hierarchy_update_importance left on the device
  float2* di = importance ? (float2*)(base + bRays + bHits + bNodes) : static_cast<float2*>(ctx->dHierImp);
  HIP_TRY(ctx, hipMemcpyAsync(dr, rays, bRays, hipMemcpyHostToDevice, ctx->stream));
  if (importance) HIP_TRY(ctx, hipMemcpyAsync(di, importance, bImp, hipMemcpyHostToDevice, ctx->stream));
  ImportanceBreak ib;
  ib.importance = di;
  ib.threshold = threshold;
  ib.rayThreshold = ray_threshold;
  HIP_TRY(ctx, LaunchTraceRefImportance(MakeRef(ctx), query, dr, dh, nodes ? dn : nullptr, n, ib, Watertight(ctx),
                                        ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(hits, dh, bHits, hipMemcpyDeviceToHost, ctx->stream));
  if (nodes && query == GPUGI_CLOSEST_HIT)
    HIP_TRY(ctx, hipMemcpyAsync(nodes, dn, bNodes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return GPUGI_OK;
}

int gpugi_set_wide_config(gpugi_ctx* ctx, int max_leaf_triangles, int lds_stack_entries, int refill_lanes,
                          int wait_lanes) {
  DropGraph(ctx);
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_wide_config: null context");
  bool stackOk = false;
  for (int v : kWideStackVariants) stackOk |= v == lds_stack_entries;
  if (!stackOk || max_leaf_triangles < 1 || max_leaf_triangles > (int)kWideMaxLeafTris || refill_lanes < 1 ||
      refill_lanes > 64 || wait_lanes < 1 || wait_lanes > 64)
    return Fail(ctx, GPUGI_ERR_ARG,
                "gpugi_set_wide_config: leaf size in [1, 16], stack one of 4/8/10/12/16/24, refill and wait in [1, 64]");
  ctx->wideStack = lds_stack_entries;
  ctx->wideRefill = refill_lanes;
  ctx->wideWait = wait_lanes;
  if (max_leaf_triangles != ctx->wideLeaf) {
    ctx->wideLeaf = max_leaf_triangles;
    if (ctx->hasScene) {
      if (SetDev(ctx)) return GPUGI_ERR_HIP;
      if (int rc = RebuildWide(ctx)) return rc;
    }
  }
  return GPUGI_OK;
}

int gpugi_get_wide_config(gpugi_ctx* ctx, int* max_leaf_triangles, int* lds_stack_entries, int* refill_lanes,
                          int* wait_lanes, uint32_t* max_stack) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_get_wide_config: null context");
  if (max_leaf_triangles) *max_leaf_triangles = ctx->wideLeaf;
  if (lds_stack_entries) *lds_stack_entries = ctx->wideStack;
  if (refill_lanes) *refill_lanes = ctx->wideRefill;
  if (wait_lanes) *wait_lanes = ctx->wideWait;
  if (max_stack) *max_stack = ctx->wideMaxStack;
  return GPUGI_OK;
}

int gpugi_autotune_trace(gpugi_ctx* ctx, int query, const gpugi_ray* rays_dev, gpugi_hit* hits_dev, size_t n,
                         int repeats, float* best_ms) {
  DropGraph(ctx);
  if (!ctx || !rays_dev || !hits_dev || n == 0 || repeats < 1)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_autotune_trace: invalid argument");
  if (!ctx->hasScene) return Fail(ctx, GPUGI_ERR_STATE, "gpugi_autotune_trace: no scene loaded");
  if (query != GPUGI_CLOSEST_HIT && query != GPUGI_ANY_HIT)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_autotune_trace: bad query");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  const int leaves[] = {1, 2, 4, 8};
  const int refills[] = {8, 16, 32, 64};
  const int waits[] = {1, 8, 16, 24, 32, 48};
  float best = INFINITY;
  int bestLeaf = ctx->wideLeaf, bestStack = ctx->wideStack, bestRefill = ctx->wideRefill, bestWait = ctx->wideWait;
  // one untimed launch, then the average of `repeats`
  auto time = [&](const WideScene& ws, int stack, int refill, int wait, float* ms) -> int {
    HIP_TRY(ctx, LaunchTraceWide(ws, stack, query, rays_dev, hits_dev, n, refill, wait, Watertight(ctx), ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int r = 0; r < repeats; ++r)
      HIP_TRY(ctx, LaunchTraceWide(ws, stack, query, rays_dev, hits_dev, n, refill, wait, Watertight(ctx), ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    HIP_TRY(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    *ms /= (float)repeats;
    return GPUGI_OK;
  };
  // leaf size x stack depth x refill at the current wait, then the wait threshold
  for (int leaf : leaves) {
    if (int rc = gpugi_set_wide_config(ctx, leaf, ctx->wideStack, ctx->wideRefill, ctx->wideWait)) return rc;
    for (int stack : kWideStackVariants) {
      WideScene ws;
      if (int rc = PrepareTraceWide(ctx, stack, &ws)) return rc;
      for (int refill : refills) {
        float ms = 0.0f;
        if (int rc = time(ws, stack, refill, ctx->wideWait, &ms)) return rc;
        if (ms < best) { best = ms; bestLeaf = leaf; bestStack = stack; bestRefill = refill; }
      }
    }
  }
  if (int rc = gpugi_set_wide_config(ctx, bestLeaf, bestStack, bestRefill, bestWait)) return rc;
  {
    WideScene ws;
    if (int rc = PrepareTraceWide(ctx, bestStack, &ws)) return rc;
    for (int wait : waits) {
      float ms = 0.0f;
      if (int rc = time(ws, bestStack, bestRefill, wait, &ms)) return rc;
      if (ms < best) { best = ms; bestWait = wait; }
    }
  }
  if (int rc = gpugi_set_wide_config(ctx, bestLeaf, bestStack, bestRefill, bestWait)) return rc;
  if (best_ms) *best_ms = best;
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

int gpugi_probe_valu(gpugi_ctx* ctx, uint32_t blocks, uint32_t iters, float* out_device, float* milliseconds) {
  if (!ctx || !out_device || !milliseconds || blocks == 0)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_probe_valu: invalid argument");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  HIP_TRY(ctx, LaunchProbeValu(out_device, blocks, iters, ctx->stream));
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
  HIP_TRY(ctx, hipEventElapsedTime(milliseconds, ctx->ev0, ctx->ev1));
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

int gpugi_copy_device(gpugi_ctx* ctx, void* device_dst, const void* device_src, size_t bytes) {
  if (!ctx || !device_dst || !device_src) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_copy_device: null argument");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  HIP_TRY(ctx, hipMemcpyAsync(device_dst, device_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
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

// ---------------------------------------------------------------- radiance-cache MLP

namespace {

// He-uniform initial weights from a xorshift64* stream; rows of padded outputs are 0.
void MlpInitialWeights(uint64_t seed, std::vector<float>* w) {
  w->assign(kMlpParams, 0.0f);
  uint64_t x = seed ? seed : 0x9E3779B97F4A7C15ull;
  for (int l = 0; l < kMlpLayers; ++l) {
    const int in = kMlpLayerIn[l], out = kMlpLayerOut[l];
    const int used = l == kMlpLayers - 1 ? kMlpOutUsed : out;
    const float bound = (float)std::sqrt(6.0 / (l == 0 ? (double)kMlpInUsed : (double)in));
    for (int o = 0; o < used; ++o)
      for (int i = 0; i < in; ++i) {
        x ^= x >> 12;
        x ^= x << 25;
        x ^= x >> 27;
        const uint64_t r = x * 0x2545F4914F6CDD1Dull;
        const double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);
        (*w)[kMlpWOff[l] + o * in + i] = (float)((2.0 * u - 1.0) * bound);
      }
  }
}

int MlpEnsureTiles(gpugi_ctx* ctx, uint32_t tiles) {
  if (ctx->mlp.partialTiles >= tiles) return GPUGI_OK;
  FreeDev(*(void**)&ctx->mlp.partials);
  FreeDev(*(void**)&ctx->mlp.loss);
  ctx->mlp.partialTiles = 0;
  HIP_TRY(ctx, hipMalloc((void**)&ctx->mlp.partials, (size_t)tiles * kMlpParams * sizeof(float)));
  HIP_TRY(ctx, hipMalloc((void**)&ctx->mlp.loss, (size_t)tiles * sizeof(float)));
  ctx->mlp.partialTiles = tiles;
  return GPUGI_OK;
}

int MlpEnsureIo(gpugi_ctx* ctx, size_t bytes) {
  if (ctx->mlpIoCap >= bytes) return GPUGI_OK;
  FreeDev(ctx->dMlpIo);
  ctx->mlpIoCap = 0;
  HIP_TRY(ctx, hipMalloc(&ctx->dMlpIo, bytes));
  ctx->mlpIoCap = bytes;
  return GPUGI_OK;
}

int MlpUpload(gpugi_ctx* ctx, const float* w) {
  const size_t bytes = kMlpParams * sizeof(float);
  HIP_TRY(ctx, hipMemcpyAsync(ctx->mlp.w, w, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipMemsetAsync(ctx->mlp.m, 0, bytes, ctx->stream));
  HIP_TRY(ctx, hipMemsetAsync(ctx->mlp.v, 0, bytes, ctx->stream));
  HIP_TRY(ctx, LaunchMlpRefreshBf16(ctx->mlp, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  ctx->mlpStep = 0;
  return GPUGI_OK;
}

int MlpCheck(gpugi_ctx* ctx, const char* fn, size_t n) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, std::string(fn) + ": null context");
  if (!ctx->hasMlp) return Fail(ctx, GPUGI_ERR_STATE, std::string(fn) + ": call gpugi_mlp_init first");
  if (n > (size_t)1 << 28) return Fail(ctx, GPUGI_ERR_ARG, std::string(fn) + ": at most 2^28 samples per call");
  return SetDev(ctx);
}

}  // namespace

int gpugi_mlp_init(gpugi_ctx* ctx, uint64_t seed) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_mlp_init: null context");
  if (SetDev(ctx)) return GPUGI_ERR_HIP;
  if (!ctx->hasMlp) {
    const size_t fb = kMlpParams * sizeof(float);
    HIP_TRY(ctx, hipMalloc((void**)&ctx->mlp.w, fb));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->mlp.m, fb));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->mlp.v, fb));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->mlp.grad, fb));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->mlp.wb, kMlpParams * sizeof(uint16_t)));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->mlp.wtb, kMlpWTSize * sizeof(uint16_t)));
    HIP_TRY(ctx, hipMemset(ctx->mlp.wtb, 0, kMlpWTSize * sizeof(uint16_t)));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->mlp.wbi, kMlpInferSize * sizeof(uint16_t)));
    HIP_TRY(ctx, hipMemset(ctx->mlp.wbi, 0, kMlpInferSize * sizeof(uint16_t)));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->dMlpLoss, 256));
    ctx->hasMlp = true;
  }
  std::vector<float> w;
  MlpInitialWeights(seed, &w);
  return MlpUpload(ctx, w.data());
}

int gpugi_mlp_set_weights(gpugi_ctx* ctx, const float* weights, size_t n) {
  if (int rc = MlpCheck(ctx, "gpugi_mlp_set_weights", 0)) return rc;
  if (!weights || n != (size_t)kMlpParams)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_mlp_set_weights: need GPUGI_MLP_PARAMS floats");
  return MlpUpload(ctx, weights);
}

int gpugi_mlp_get_weights(gpugi_ctx* ctx, float* weights, size_t n) {
  if (int rc = MlpCheck(ctx, "gpugi_mlp_get_weights", 0)) return rc;
  if (!weights || n != (size_t)kMlpParams)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_mlp_get_weights: need GPUGI_MLP_PARAMS floats");
  HIP_TRY(ctx, hipMemcpyAsync(weights, ctx->mlp.w, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return GPUGI_OK;
}

int gpugi_mlp_encode(gpugi_ctx* ctx, const float* features, uint16_t* encoded, size_t n) {
  if (int rc = MlpCheck(ctx, "gpugi_mlp_encode", n)) return rc;
  if (n == 0) return GPUGI_OK;
  if (!features || !encoded) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_mlp_encode: null argument");
  const size_t fb = n * kMlpRaw * sizeof(float), eb = n * kMlpIn * sizeof(uint16_t);
  if (int rc = MlpEnsureIo(ctx, fb + eb)) return rc;
  float* df = (float*)ctx->dMlpIo;
  uint16_t* de = (uint16_t*)((char*)ctx->dMlpIo + fb);
  HIP_TRY(ctx, hipMemcpyAsync(df, features, fb, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, LaunchMlpEncode(df, de, n, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(encoded, de, eb, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return GPUGI_OK;
}

int gpugi_mlp_infer(gpugi_ctx* ctx, const float* features, float* radiance, size_t n, int on_device) {
  if (int rc = MlpCheck(ctx, "gpugi_mlp_infer", n)) return rc;
  if (n == 0) return GPUGI_OK;
  if (!features || !radiance) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_mlp_infer: null argument");
  const float* df = features;
  float* dr = radiance;
  const size_t fb = n * kMlpRaw * sizeof(float), rb = n * 4 * sizeof(float);
  if (!on_device) {
    if (int rc = MlpEnsureIo(ctx, fb + rb)) return rc;
    df = (const float*)ctx->dMlpIo;
    dr = (float*)((char*)ctx->dMlpIo + fb);
    HIP_TRY(ctx, hipMemcpyAsync((void*)df, features, fb, hipMemcpyHostToDevice, ctx->stream));
  }
  HIP_TRY(ctx, LaunchMlpInfer(ctx->mlp, df, dr, n, ctx->stream));
  if (!on_device) {
    HIP_TRY(ctx, hipMemcpyAsync(radiance, dr, rb, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GPUGI_OK;
}

namespace {

// Stages features/targets for host pointers and runs the gradient kernels.
// The gradient of n samples into ctx->mlp.grad and its loss; with `adam`
// the Adam step follows in the same reduction launch.
int MlpGradientCommon(gpugi_ctx* ctx, const char* fn, const float* features, const float* targets, size_t n,
                      int on_device, const AdamParams* adam = nullptr) {
  if (n == 0 || !features || !targets) return Fail(ctx, GPUGI_ERR_ARG, std::string(fn) + ": need samples");
  if (int rc = MlpEnsureTiles(ctx, MlpTiles(n))) return rc;
  const float* df = features;
  const float* dt = targets;
  const size_t fb = n * kMlpRaw * sizeof(float), tb = n * 4 * sizeof(float);
  if (!on_device) {
    if (int rc = MlpEnsureIo(ctx, fb + tb + 256)) return rc;
    df = (const float*)ctx->dMlpIo;
    dt = (const float*)((char*)ctx->dMlpIo + fb);
    HIP_TRY(ctx, hipMemcpyAsync((void*)df, features, fb, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync((void*)dt, targets, tb, hipMemcpyHostToDevice, ctx->stream));
  }
  const MlpBatch batch{nullptr, (uint32_t)n, 0, 1};
  HIP_TRY(ctx, LaunchMlpGradientTiles(ctx->mlp, df, dt, batch, ctx->stream));
  HIP_TRY(ctx, LaunchMlpReduce(ctx->mlp, batch, ctx->dMlpLoss, adam, ctx->stream));
  return GPUGI_OK;
}

}  // namespace

int gpugi_mlp_gradient(gpugi_ctx* ctx, const float* features, const float* targets, size_t n, float* grad,
                       float* loss, int on_device) {
  if (int rc = MlpCheck(ctx, "gpugi_mlp_gradient", n)) return rc;
  if (!grad) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_mlp_gradient: null gradient buffer");
  if (int rc = MlpGradientCommon(ctx, "gpugi_mlp_gradient", features, targets, n, on_device)) return rc;
  HIP_TRY(ctx, hipMemcpyAsync(grad, ctx->mlp.grad, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  float l = 0.0f;
  HIP_TRY(ctx, hipMemcpyAsync(&l, ctx->dMlpLoss, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (loss) *loss = l;
  return GPUGI_OK;
}

int gpugi_mlp_train_step(gpugi_ctx* ctx, const float* features, const float* targets, size_t n,
                         float learning_rate, float* loss, int on_device) {
  if (int rc = MlpCheck(ctx, "gpugi_mlp_train_step", n)) return rc;
  if (!(learning_rate > 0.0f)) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_mlp_train_step: learning rate must be > 0");
  AdamParams p{learning_rate, 0.9f, 0.99f, 1e-8f, ctx->mlpStep + 1};
  if (int rc = MlpGradientCommon(ctx, "gpugi_mlp_train_step", features, targets, n, on_device, &p)) return rc;
  ++ctx->mlpStep;
  if (loss) {
    HIP_TRY(ctx, hipMemcpyAsync(loss, ctx->dMlpLoss, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GPUGI_OK;
}

int gpugi_cache_config(gpugi_ctx* ctx, float spread_c, int extra_bounces, uint32_t train_stride,
                       uint32_t max_train_samples, int train_steps, float learning_rate) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_cache_config: null context");
  if (!(spread_c > 0.0f) || extra_bounces < 0 || extra_bounces > 16 || train_stride == 0 || max_train_samples == 0 ||
      max_train_samples > (1u << 24) || train_steps < 0 || train_steps > 64 || !(learning_rate > 0.0f))
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_cache_config: invalid argument");
  ctx->cacheCfg.spreadC = spread_c;
  ctx->cacheCfg.extraBounces = extra_bounces;
  ctx->cacheCfg.trainStride = train_stride;
  ctx->cacheMaxSamples = max_train_samples;
  ctx->cacheTrainSteps = train_steps;
  ctx->cacheLr = learning_rate;
  return GPUGI_OK;
}

int gpugi_cache_set_query(gpugi_ctx* ctx, int query) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_cache_set_query: null context");
  if (query != GPUGI_CACHE_SPREAD && query != GPUGI_CACHE_PRIMARY_INDIRECT)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_cache_set_query: unknown query rule");
  ctx->cacheCfg.query = query;
  return GPUGI_OK;
}

int gpugi_cache_stats(gpugi_ctx* ctx, uint32_t* train_samples, float* loss) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_cache_stats: null context");
  if (train_samples) {
    *train_samples = 0;
    if (ctx->dCacheCount) {
      uint32_t c = 0;
      HIP_TRY(ctx, hipMemcpyAsync(&c, ctx->dCacheCount, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
      *train_samples = c < ctx->cacheMaxSamples ? c : ctx->cacheMaxSamples;
    }
  }
  if (loss) {
    *loss = 0.0f;
    if (ctx->hasMlp) {
      HIP_TRY(ctx, hipMemcpyAsync(loss, ctx->dMlpLoss, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    }
  }
  return GPUGI_OK;
}

int gpugi_set_rr_start_depth(gpugi_ctx* ctx, int depth) {
  DropGraph(ctx);
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_rr_start_depth: null context");
  if (depth < 0 || depth > 64) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_rr_start_depth: depth in [0, 64]");
  ctx->rrStartDepth = depth;
  return GPUGI_OK;
}

int gpugi_set_specific_path_length(gpugi_ctx* ctx, int length) {
  DropGraph(ctx);
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_specific_path_length: null context");
  if (length < -1 || length > 64)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_specific_path_length: length in [-1, 64]");
  ctx->specificPathLength = length;
  return GPUGI_OK;
}

int gpugi_set_bvh_type(gpugi_ctx* ctx, int bvh_type) {
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_bvh_type: null context");
  if (bvh_type != GPUGI_BVH_AABOX && bvh_type != GPUGI_BVH_OBOX)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_bvh_type: unknown type");
  ctx->bvhType = bvh_type;
  return GPUGI_OK;
}

int gpugi_set_triangle_test(gpugi_ctx* ctx, int mode) {
  DropGraph(ctx);
  if (!ctx) return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_triangle_test: null context");
  if (mode != GPUGI_TRIANGLE_REFERENCE && mode != GPUGI_TRIANGLE_WATERTIGHT)
    return Fail(ctx, GPUGI_ERR_ARG, "gpugi_set_triangle_test: unknown mode");
  if (mode == ctx->triangleTest) return GPUGI_OK;
  ctx->triangleTest = mode;
  if (ctx->hasScene) {
    if (SetDev(ctx)) return GPUGI_ERR_HIP;
    if (int rc = RebuildWide(ctx)) return rc;
  }
  return GPUGI_OK;
}
