/*
 * gpugi_rt.h -- C ABI of the MI355X (gfx950) ray-tracing hot path.
 *
 * This is the drop-in boundary for the GLSL compute path of gpugi.  Every entry
 * point names the reference interface it replaces (paths relative to the
 * gpugi source tree).  Conventions:
 *   - every function returns an int status (GPUGI_OK == 0, negative on error);
 *     no C++ exception crosses this boundary, and gpugi_last_error() returns a
 *     message for the last failing call on a context (thread-local for calls
 *     without a context);
 *   - a context owns one HIP stream and all device memory it allocates; it is
 *     not thread-safe (one host thread per context), like the single GL context
 *     of the reference (application.cpp:287-301);
 *   - host inputs are copied during the call; calls are asynchronous on the
 *     context stream except readback, trace with host pointers, timers and
 *     gpugi_sync.
 */
#ifndef GPUGI_RT_H
#define GPUGI_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPUGI_ABI_VERSION 1

enum gpugi_status {
  GPUGI_OK = 0,
  GPUGI_ERR_ARG = -1,         /* invalid argument (null pointer, bad size, bad enum) */
  GPUGI_ERR_HIP = -2,         /* a HIP runtime call failed */
  GPUGI_ERR_STATE = -3,       /* call order violated (e.g. trace before a scene is loaded) */
  GPUGI_ERR_IO = -4,          /* file could not be opened/read/written */
  GPUGI_ERR_FORMAT = -5,      /* malformed .rawscene / json content */
  GPUGI_ERR_UNSUPPORTED = -6, /* valid request this build does not support */
  GPUGI_ERR_NOMEM = -7,       /* host or device allocation failed */
  GPUGI_ERR_NODEVICE = -8     /* no usable gfx950 device */
};

/* Which query a trace performs. */
enum gpugi_query {
  GPUGI_CLOSEST_HIT = 0,      /* TraceRay (traceray.glsl:25-38, HIT_INDEX_OUTPUT) */
  GPUGI_ANY_HIT = 1           /* TraceRayAnyHit (traceray.glsl:19-23, ANY_HIT) */
};

/* Which traversal kernel a trace uses. */
enum gpugi_kernel {
  GPUGI_KERNEL_AUTO = 0,      /* fastest kernel available for the loaded scene */
  GPUGI_KERNEL_REFERENCE = 1, /* escape-pointer walk on the reference BVH2 layout (bit-exact) */
  GPUGI_KERNEL_WIDE = 2       /* ordered walk on the re-packed quantized 4-wide BVH */
};

/* Path schedules of gpugi_set_pt_schedule. */
enum gpugi_pt_schedule { GPUGI_PT_MEGAKERNEL = 0, GPUGI_PT_WAVEFRONT = 1 };

/* Bounding volumes of the reference-layout hierarchy (the bvhType parameter,
 * application.cpp:135: Types3D BOX or OBOX). */
enum gpugi_bvh_type {
  GPUGI_BVH_AABOX = 0, /* TreeNode<Box>, FetchIntersectBoxNode (intersectiontests.glsl:100-111) */
  GPUGI_BVH_OBOX = 1   /* TreeNode<OBox>, FetchIntersectOBoxNode (intersectiontests.glsl:114-128) */
};

/* Which ray/triangle test traversal uses. */
enum gpugi_triangle_test {
  GPUGI_TRIANGLE_REFERENCE = 0,  /* IntersectTriangle (intersectiontests.glsl:63-79): not watertight */
  GPUGI_TRIANGLE_WATERTIGHT = 1  /* Woop, Benthin, Wald 2013 with exact-sign edge functions (no
                                    reference counterpart): no ray passes between triangles that
                                    share an edge or vertex */
};

/* Integrators a draw runs. */
enum gpugi_integrator {
  GPUGI_PT = 0,               /* Pathtracer::Draw (pathtracer.cpp:47-53) */
  GPUGI_LT = 1,               /* LightPathtracer::Draw (lightpathtracer.cpp:45-51) */
  GPUGI_BPT = 2,              /* BidirectionalPathtracer::Draw (bidirectionalpathtracer.cpp:182-203) */
  GPUGI_PT_CACHE = 3          /* path tracer terminating into the radiance-cache MLP, trained online
                                 each frame (no reference counterpart; needs gpugi_mlp_init) */
};

/* A ray query: 32 bytes.  tmax is the initial rayLength of TraceRay /
 * TraceRayAnyHit (traceray.glsl:19-25); RAY_MAX = 3.402823466e+38
 * (stdheader.glsl:6) means unbounded. */
typedef struct gpugi_ray {
  float origin[3];
  float tmax;
  float dir[3];
  uint32_t pad;
} gpugi_ray;

/* A hit record: 16 bytes.  slot is the triangle-slot index of the closest hit
 * (the _hitIndex.y of traceray.glsl:155, i.e. leaf * TRIANGLES_PER_LEAF +
 * position in the leaf), -1 on a miss; t is the final rayLength (tmax on a
 * miss); bary_y/bary_z are barycentricCoord.y/.z (intersectiontests.glsl:73-74).
 * For GPUGI_ANY_HIT, slot is 1 when an occluder was found and 0 otherwise;
 * t, bary_y and bary_z are left unspecified. */
typedef struct gpugi_hit {
  int32_t slot;
  float t;
  float bary_y;
  float bary_z;
} gpugi_hit;

/* Resolved material: the values SampleMaterialData returns for a constant
 * material (materials.glsl:19-30).  Diffuse and opacity are already quantized
 * to UNORM8 the way Scene::LoadMaterial uploads them (scene.cpp:164,168). */
typedef struct gpugi_material {
  float diffuse[3];
  float opacity[3];
  float reflectiveness[4];    /* .w = Phong exponent */
  float emissivity[3];
  float fresnel0[3];
  float refraction_index_avg;
  float fresnel1[3];
} gpugi_material;

/* LightSampler::LightSample (lightsampler.hpp:17-23): 32 bytes. */
typedef struct gpugi_light_sample {
  float position[3];
  float normal_phi;           /* atan2(n.y, n.x) */
  float intensity[3];
  float normal_theta_cos;     /* n.z; > 1.5 marks an omnidirectional point light */
} gpugi_light_sample;

/* Scene arrays in the reference GPU layout (renderersystem.cpp:138-147). */
typedef struct gpugi_scene_desc {
  const float* hierarchy;        /* num_nodes x 8 floats: TreeNode<Box> (scene.hpp:30-38):
                                    {min.xyz, bits(firstChild)}, {max.xyz, bits(escape)} */
  uint32_t num_nodes;
  const int32_t* triangles;      /* num_leaves x triangles_per_leaf x {v0,v1,v2,material}
                                    (scene.hpp:64-68); invalid slot has v0 == v1 */
  uint32_t num_leaves;
  uint32_t triangles_per_leaf;   /* must be 2 (scenedata.glsl:5) */
  const float* positions;        /* num_vertices x 3 floats (RGB32F, scenedata.glsl:34) */
  const float* vertex_info;      /* num_vertices x {atan2(ny,nx), nz, u, v} (scene.hpp:71-75) */
  uint32_t num_vertices;
  const gpugi_material* materials;
  uint32_t num_materials;        /* at most 256 (globaluniforms.glsl:23-27) */
} gpugi_scene_desc;

/* Camera UBO (globaluniforms.glsl:7-15) without the raster-only ViewProjection. */
typedef struct gpugi_camera {
  float position[3];
  float u[3];
  float v[3];
  float w[3];
  float pixel_area;
} gpugi_camera;

/* ------------------------------------------------------------------------ */
/* Device context                                                           */
/* ------------------------------------------------------------------------ */
typedef struct gpugi_ctx gpugi_ctx;

/* Creates a context on HIP device `device` (RendererSystem::RendererSystem,
 * renderersystem.cpp:23-48). */
int gpugi_create(int device, gpugi_ctx** out);
/* Releases all device memory of the context (RendererSystem::~RendererSystem). */
void gpugi_destroy(gpugi_ctx* ctx);
/* Message of the last failing call on ctx; ctx may be NULL for calls that
 * have no context (host utilities, gpugi_create). */
const char* gpugi_last_error(const gpugi_ctx* ctx);

/* Uploads a scene and re-packs it for the wide kernel (Scene::Scene,
 * UploadGeometry, UploadHierarchy, scene.cpp:15-130 + RendererSystem::SetScene,
 * renderersystem.cpp:132-171).  Resets the accumulation buffer. */
int gpugi_load_scene(gpugi_ctx* ctx, const gpugi_scene_desc* scene);
/* Camera UBO update (RendererSystem::SetCamera, renderersystem.cpp:188-220). */
int gpugi_set_camera(gpugi_ctx* ctx, const gpugi_camera* camera);
/* Backbuffer size (RendererSystem::SetScreenSize, renderersystem.cpp:222-233).
 * Allocates a zeroed RGBA32F accumulation buffer. */
int gpugi_set_screen(gpugi_ctx* ctx, int width, int height);
/* Initial light samples for the next draw (InitialLightSampleBuffer, TBO 5;
 * RendererSystem::PerIterationBufferUpdate, renderersystem.cpp:95-116). */
int gpugi_set_light_samples(gpugi_ctx* ctx, const gpugi_light_sample* samples, uint32_t n);
/* MAX_PATHLENGTH define (pathtracer.cpp:60). */
int gpugi_set_max_path_length(gpugi_ctx* ctx, int max_path_length);
/* Triangle test of every trace and integrator (default
 * GPUGI_TRIANGLE_REFERENCE).  Watertight mode re-packs the wide triangles
 * as raw vertices; barycentrics keep the reference meaning (bary_y, bary_z
 * weight the second and third vertex). */
int gpugi_set_triangle_test(gpugi_ctx* ctx, int mode);
/* bvhType (application.cpp:135): like the reference it must be set before the
 * scene is loaded and applies from the next gpugi_load_scene.  With
 * GPUGI_BVH_OBOX the load fits an oriented box to every node of the given
 * hierarchy (the reference reads them from its scene library, not in the
 * tree; see gpugi_hostscene_obox_hierarchy) and the reference walk
 * (GPUGI_KERNEL_REFERENCE, in gpugi_trace and the integrators) tests those;
 * the wide kernel keeps its own axis-aligned 4-wide tree. */
int gpugi_set_bvh_type(gpugi_ctx* ctx, int bvh_type);
/* Bounce from which Russian roulette runs (default 0: every bounce, as
 * pathtracer.comp:99-104 and lighttracer.comp:183-188 do; 3 gives the
 * "roulette after depth 3" variant, which has no reference counterpart). */
int gpugi_set_rr_start_depth(gpugi_ctx* ctx, int depth);
/* Keeps only the paths of `length` bounces (1 = direct lighting) in
 * GPUGI_PT, GPUGI_LT and GPUGI_BPT draws; -1 (default) keeps every path.
 * Replaces the SHOW_SPECIFIC_PATHLENGTH define (renderer.hpp:6-9) that the
 * reference compiles into pathtracer.comp:74-84, lighttracer.comp:133-157 and
 * pathtracer_bidir.comp:75-108; as there, the light tracer's splat of the
 * light samples themselves (showlightcaches.comp) is kept.  GPUGI_PT_CACHE
 * draws fail with GPUGI_ERR_STATE while it is set. */
int gpugi_set_specific_path_length(gpugi_ctx* ctx, int length);
/* Traversal the integrators use: GPUGI_KERNEL_REFERENCE (bit-exact escape
 * walk) or GPUGI_KERNEL_WIDE / GPUGI_KERNEL_AUTO (4-wide quantized BVH).
 * Replaces the AABOX_BVH define of the shader compile (pathtracer.cpp:55-64). */
int gpugi_set_integrator_kernel(gpugi_ctx* ctx, int kernel);
/* How gpugi_draw(GPUGI_PT) schedules the paths, and gpugi_draw(GPUGI_BPT) its
 * eye paths: GPUGI_PT_MEGAKERNEL (default; one lane carries a path through
 * all its bounces, as the reference's pathtracer.comp and
 * pathtracer_bidir.comp dispatches do) or GPUGI_PT_WAVEFRONT (one pass per
 * bounce over the live paths: closest-hit rays and shadow / light-cache
 * connection rays traced by the gpugi_trace kernels on compact device queues
 * whose lengths every pass reads on the device, so the host never waits; see
 * gpugi_capture_draw; each bounce's shadow or connection trace runs on a
 * second stream of the context beside the next closest-hit trace, joined back
 * into the context's stream).  Both give the same image with the reference
 * walk (BPT: up to the order of the light pass's float-atomic camera
 * splats); the wavefront schedule with the reference walk and
 * GPUGI_TRIANGLE_WATERTIGHT rebuilds b_x as 1 - (b_y + b_z) from the hit
 * record.  No counterpart in the reference. */
int gpugi_set_pt_schedule(gpugi_ctx* ctx, int schedule);
/* Measurement helper, no counterpart in the reference: launches the VALU
 * calibration kernel (blocks x 256 lanes, eight independent FP32 FMA chains
 * of 8 * iters each per lane, results into out_device, one float per lane) on
 * the context's stream and returns its duration in milliseconds.  Under
 * rocprofv3 its SQ_INSTS_VALU / GRBM_GUI_ACTIVE give the device's VALU issue
 * rate, the roofline of the traversal kernels. */
int gpugi_probe_valu(gpugi_ctx* ctx, uint32_t blocks, uint32_t iters, float* out_device, float* milliseconds);
/* Bytes of the per-device scratch array of gpugi_trace_noalloc (64 MiB). */
#define GPUGI_NOALLOC_SCRATCH_BYTES (64u << 20)
/* gpugi_trace for device-resident rays and hits that allocates no memory: no
 * hipMalloc, hipFree or allocating library call on any path, so it can run in
 * code that must not allocate (under HIP graph capture, in serving loops, in
 * harnesses that forbid allocation).  No counterpart in the reference; it is
 * for those callers.  The wide kernel's batch counters and spill stacks live
 * in a file-scope device array of GPUGI_NOALLOC_SCRATCH_BYTES that each GPU
 * holds until the process ends; its grid shrinks to the spill that fits, and
 * a scene whose traversal stack needs more than the array holds even for the
 * smallest grid fails with GPUGI_ERR_ARG.  The call is asynchronous on the
 * context's stream and reads nothing back (gpugi_sync waits for it).  Every
 * call on one GPU shares the array: calls that overlap on one GPU race, so
 * order them (one stream).  rays: n x 32 B and hits: n x 16 B device
 * pointers; query and kernel as gpugi_trace. */
int gpugi_trace_noalloc(gpugi_ctx* ctx, int query, int kernel, const gpugi_ray* rays, gpugi_hit* hits, size_t n);
/* One iteration of an integrator that adds into the accumulation buffer
 * (Renderer::Draw, renderer.hpp:42; FrameSeed UBO, globaluniforms.glsl:17-21). */
int gpugi_draw(gpugi_ctx* ctx, int integrator, uint32_t frame_seed);
/* Records one gpugi_draw into a HIP graph (stream capture) without running
 * it; gpugi_replay_draw launches the recording on the context's stream, each
 * replay adding that frame (same frame seed) to the accumulation buffer and
 * counting one iteration, with no host work per pass.  Only GPUGI_PT with the
 * wavefront schedule is recorded (GPUGI_ERR_UNSUPPORTED otherwise): its passes
 * read every queue length on the device.  Buffers are allocated before the
 * capture.  Any setter that changes the scene, screen, camera, light samples,
 * textures, environment or traversal settings drops the recording
 * (gpugi_reset_accum keeps it).  No counterpart in the reference, whose draws
 * are single dispatches. */
int gpugi_capture_draw(gpugi_ctx* ctx, int integrator, uint32_t frame_seed);
int gpugi_replay_draw(gpugi_ctx* ctx);
/* One warm-up light pass of the bidirectional path tracer with the current
 * light samples (lighttracer.comp with SAVE_LIGHT_CACHE_WARMUP): returns the
 * NumLightCacheEntries and LightPathLengthSum counters the reference reads
 * back in CreateLightCacheWithCapacityEstimate (bidirectionalpathtracer.cpp:65-104).
 * Leaves the accumulation buffer untouched.  Blocks until the counts are known. */
int gpugi_bpt_warmup(gpugi_ctx* ctx, uint32_t frame_seed, uint64_t* num_light_cache_entries,
                     uint64_t* light_path_length_sum);
/* LightCacheCapacity and AverageLightPathLength of the LightPathTrace UBO
 * (bidirectionalpathtracer.cpp:106-132); allocates capacity x 64-byte entries.
 * A BPT draw fails with GPUGI_ERR_STATE until this is set; changing the screen
 * or the scene clears it. */
int gpugi_bpt_set_light_cache(gpugi_ctx* ctx, int32_t capacity, int32_t average_light_path_length);
/* Copies the light cache of the last BPT draw (LightCacheEntry, 64 bytes each,
 * bidirpathtracer/lightcache.glsl:1-16): up to max_entries entries and the
 * NumLightCacheEntries counter (which may exceed the capacity). */
int gpugi_bpt_read_light_cache(gpugi_ctx* ctx, void* entries, uint32_t max_entries, uint32_t* num_entries);
/* Clears the accumulation buffer (Texture2D::ClearToZero in SetRenderer,
 * renderersystem.hpp:73-86). */
int gpugi_reset_accum(gpugi_ctx* ctx);
/* Reads the RGBA32F accumulation buffer (Texture2D::ReadImage in
 * Application::SaveImage, application.cpp:394-414).  bytes must be
 * >= width*height*16.  *iterations (optional) gets the number of draws since
 * the last reset. */
int gpugi_readback(gpugi_ctx* ctx, float* rgba, size_t bytes, uint32_t* iterations);

/* Material textures (Scene::GetBindlessHandle and the texture branches of
 * Scene::LoadMaterial, scene.cpp:161-178, 189-203).  A texture is width x
 * height texels, row 0 at t = 0.5 / height: GPUGI_TEXTURE_RGBA8_SRGB (4 bytes
 * per texel, colour sRGB-encoded, alpha linear, as an 8-bit texture the
 * reference loads) or GPUGI_TEXTURE_RGBA32F (16 bytes, linear).  Sampling is
 * textureLod(uv, 0) with the scene's sampler (scene.cpp:16-20): bilinear,
 * REPEAT.  Textures and bindings belong to the loaded scene: gpugi_load_scene
 * drops them.  gpugi_add_texture stores a copy and returns its index in *id;
 * gpugi_set_material_texture makes slot `slot` of material `material` sample
 * texture `texture` at the hit's texture coordinate (-1: back to the constant
 * of the material record). */
enum gpugi_texture_format { GPUGI_TEXTURE_RGBA8_SRGB = 0, GPUGI_TEXTURE_RGBA32F = 1 };
enum gpugi_material_slot {
  GPUGI_TEX_DIFFUSE = 0,        /* MaterialData.Diffuse = .rgb */
  GPUGI_TEX_OPACITY = 1,        /* MaterialData.Opacity = .rgb */
  GPUGI_TEX_REFLECTIVENESS = 2, /* MaterialData.Reflectiveness = .rgba (.a: Phong exponent) */
  GPUGI_TEX_EMISSIVITY = 3      /* MaterialData.Emissivity = .rgb */
};
int gpugi_add_texture(gpugi_ctx* ctx, int width, int height, int format, const void* texels, int* id);
int gpugi_set_material_texture(gpugi_ctx* ctx, uint32_t material, int slot, int texture);

/* RendererSystem::SetEnvironmentMap (renderersystem.cpp:173-186): six square
 * faces of size x size RGB8 texels, sRGB-encoded like the reference's SRGB8
 * cube map, in the order x-, x+, y-, y+, z-, z+ of the envMap script command
 * (application.cpp:151-155); texel (i, j) of a face is faces[((face * size +
 * j) * size + i) * 3 + c], the texel centred at (s, t) = ((i + 0.5) / size,
 * (j + 0.5) / size) of the OpenGL cube-map face coordinates.  The path tracers
 * add throughput * environment(direction) when a ray leaves the scene
 * (pathtracer.comp:56-58), with the face chosen by the major axis of the
 * direction and bilinear filtering clamped to the face.  size 0 (faces NULL)
 * restores the black default (renderersystem.cpp:39-47).  Does not reset the
 * accumulation, as the reference does not. */
int gpugi_set_environment_map(gpugi_ctx* ctx, int size, const uint8_t* faces);

/* Reference image of the MSE check (TextureMSE::LoadReferenceFromPFM,
 * texturemse.cpp:26-37): width*height RGB32F pixels in the row order LoadPfm
 * returns (hdrimage.cpp:35-63).  Replaces any earlier reference image; its
 * size may differ from the screen's. */
int gpugi_set_reference_image(gpugi_ctx* ctx, const float* rgb, int width, int height);
/* TextureMSE::ComputeCurrentMSE(backbuffer, divisor) (texturemse.cpp:39-66,
 * shader/utils/sqerror.frag): the mean over the reference image's pixels of
 * dot(d, d), d = accum.rgb / divisor - reference.rgb, with the accumulation
 * buffer sampled nearest at each reference pixel's centre.  The mean is taken
 * in double over all pixels (the reference averages a float mip chain).
 * *mse = +inf when no reference image is set; divisor 0 gives inf/NaN as the
 * shader's division does.  Blocks until the value is on the host. */
int gpugi_compute_mse(gpugi_ctx* ctx, uint32_t divisor, float* mse);

/* Per-ray extra output of the importance-break walk: _hitIndex.x,
 * _nodeImportance and _nodeSizeSq of TraceRay (traceray.glsl:25-31). */
typedef struct gpugi_node_hit {
  int32_t node;            /* last hit inner node (the node the walk stopped in after a break), -1: none */
  float node_importance;   /* importance.x of that node */
  float node_size_sq;      /* |box diagonal|^2 of that node */
  int32_t pad;
} gpugi_node_hit;

/* TraceRay / TraceRayAnyHit compiled with TRACERAY_IMPORTANCE_BREAK
 * (traceray.glsl:48-112), the escape walk of GPUGI_KERNEL_REFERENCE (either
 * bvh type).  importance: num_nodes pairs {x, y} (HierarchyImportanceBuffer);
 * threshold: the IMPORTANCE_THRESHOLD define; ray_threshold: the
 * _importanceThreshold argument.  When a node's box is hit at newHit (within
 * the ray) and x < threshold with newHit > 0, or y <= ray_threshold * newHit,
 * its subtree is skipped: a closest-hit query then ends there with slot -1 and
 * t = newHit + (exit - newHit) * (wanghash(bits(exit)) % 8388593) / 8388593
 * unless a closer triangle follows, an any-hit query reports a hit if the box
 * interval ends within the ray.  Host arrays; nodes (closest hit only) may be NULL; blocks until
 * the results are copied back.  importance == NULL uses the importance the last
 * gpugi_hierarchy_update_importance computed on the device (num_nodes is then
 * not read); GPUGI_ERR_STATE if there is none for the current scene. */
/* The importance update of HierarchyImportance::Draw on its first iteration
 * (hierarchyimportance.cpp:84-105, 120-152): `iterations` acquisition passes
 * (acquisition2.comp: one path of up to 8 vertices per pixel of the W/8*8 x
 * H/8*8 grid with the current camera and light samples, iteration i seeded
 * by wanghash(i); each ray adds the path's importance, halved at every
 * vertex, to every node box it enters and triangle it hits) then the
 * propagation (hierarchypropagation_init.comp, hierarchypropagation_nodes.comp):
 * per node x = rays through it per iteration, subtree importance = what its
 * triangles (or children) received, y = (x / subtree)^2 * a third of the box
 * surface.  The reference uses 10 iterations.  Runs the reference walk and
 * triangle test; blocks until done.  The result stays on the device for
 * gpugi_trace_importance_break and gpugi_hierarchy_read.  The sums are exact
 * and independent of the order of the additions (64-bit fixed point in units
 * of 2^-8, the smallest addend), so they equal the reference's FP32 sums
 * wherever those are exact, i.e. below 2^16 rays per node; beyond that an FP32
 * sum drops addends (1 + 2^24 rounds to 2^24), so the most-visited nodes of a
 * large scene get larger, exact values here. */
int gpugi_hierarchy_update_importance(gpugi_ctx* ctx, int iterations);
/* HierarchyImportance::ComputeHierarchyMaterials (hierarchyimportance.cpp:154-178):
 * per leaf node the area-weighted material of its triangles sampled at their
 * centroids (sampleleafmaterials.comp), pulled up the tree weighted by the
 * children's areas (pullmaterials.comp); each node keeps six pairs of packed
 * halves (hierarchymaterialbuffer.glsl: reflectiveness xy|zw, opacity xy|z
 * area, diffuse, emissivity, fresnel0 xy|z refraction index, fresnel1).  The
 * oriented-box layout reads each child's escape from its own record (the
 * reference's pull pass assumes two texels per node there). */
int gpugi_hierarchy_compute_materials(gpugi_ctx* ctx);
/* Copies the hierarchy passes' results to host arrays (any may be NULL):
 * importance (num_nodes + 2 * num_leaves) pairs {x, y} (nodes, then triangle
 * slots), subtree num_nodes floats, materials num_nodes x 12 words. */
int gpugi_hierarchy_read(gpugi_ctx* ctx, float* importance, size_t importance_floats, float* subtree,
                         size_t subtree_floats, uint32_t* materials, size_t material_words);
int gpugi_trace_importance_break(gpugi_ctx* ctx, int query, const gpugi_ray* rays, gpugi_hit* hits,
                                 gpugi_node_hit* nodes, size_t n, const float* importance, uint32_t num_nodes,
                                 float threshold, float ray_threshold);

/* Pinned host memory (hipHostMalloc) owned by the context: host ray and hit
 * buffers in it let gpugi_trace overlap its copies with the traversal.
 * gpugi_destroy frees what is left. */
int gpugi_host_alloc(gpugi_ctx* ctx, size_t bytes, void** host_ptr);
int gpugi_host_free(gpugi_ctx* ctx, void* host_ptr);

/* Ray queries for parity and the traversal microbenchmark (the GLSL
 * TraceRay/TraceRayAnyHit functions, traceray.glsl:18-39).  rays and hits
 * are host pointers when rays_on_device == 0 (the call then blocks until the
 * hits are copied back; above 2^22 rays in pinned memory they move in chunks
 * on a second stream while earlier chunks are traced), device pointers
 * otherwise (asynchronous). */
int gpugi_trace(gpugi_ctx* ctx, int query, int kernel, const gpugi_ray* rays, gpugi_hit* hits,
                size_t n, int rays_on_device);
/* Configuration of the wide traversal: triangles per merged leaf (1..16; the
 * BVH is re-packed when it changes), LDS stack entries per lane of the trace
 * kernel (4, 8, 10, 12, 16 or 24; deeper entries spill to global memory), the
 * number of idle lanes (1..64) at which a wave refills them with new rays and
 * the number of lanes (1..64) waiting with a pending leaf at which a wave runs
 * a triangle pass instead of a box pass.  Default (2, 10, 8, 16): the fastest
 * incoherent closest-hit configuration on the 2,097,152-triangle heightfield.  No
 * reference counterpart: the GLSL path has one fixed layout
 * (TRIANGLES_PER_LEAF, scenedata.glsl:5). */
int gpugi_set_wide_config(gpugi_ctx* ctx, int max_leaf_triangles, int lds_stack_entries, int refill_lanes,
                          int wait_lanes);
int gpugi_get_wide_config(gpugi_ctx* ctx, int* max_leaf_triangles, int* lds_stack_entries, int* refill_lanes,
                          int* wait_lanes, uint32_t* max_stack);
/* Autotuner: times every (leaf size, stack depth, refill) triple on the given device
 * ray set (one warm-up launch + `repeats` timed launches each), then the wait
 * threshold at the best triple, keeps the fastest configuration active and
 * returns its average launch time. */
int gpugi_autotune_trace(gpugi_ctx* ctx, int query, const gpugi_ray* rays_dev, gpugi_hit* hits_dev, size_t n,
                         int repeats, float* best_ms);
/* Reference escape-walk visit counters (TRACERAY_DEBUG_VARS,
 * traceray.glsl:59-61,131-133): per ray, box tests and valid-slot triangle
 * tests.  box_tests / tri_tests are device arrays of n uint32 each when
 * rays_on_device != 0, host arrays otherwise. */
int gpugi_trace_counts(gpugi_ctx* ctx, int query, const gpugi_ray* rays, uint32_t* box_tests,
                       uint32_t* tri_tests, size_t n, int rays_on_device);
/* Waits for all work queued on the context stream (glFinish). */
int gpugi_sync(gpugi_ctx* ctx);

/* Device memory owned by the context (freed by gpugi_destroy at the latest). */
int gpugi_alloc(gpugi_ctx* ctx, size_t bytes, void** device_ptr);
int gpugi_free(gpugi_ctx* ctx, void* device_ptr);
int gpugi_upload(gpugi_ctx* ctx, void* device_dst, const void* host_src, size_t bytes);
int gpugi_download(gpugi_ctx* ctx, void* host_dst, const void* device_src, size_t bytes);
/* Device-to-device copy on the context stream (asynchronous). */
int gpugi_copy_device(gpugi_ctx* ctx, void* device_dst, const void* device_src, size_t bytes);

/* HIP-event timer on the context stream (replaces the glFinish + clock()
 * bracket of RendererSystem::Draw, renderersystem.cpp:235-258). */
int gpugi_timer_begin(gpugi_ctx* ctx);
int gpugi_timer_end(gpugi_ctx* ctx, float* milliseconds);

/* Ray generators of the traversal microbenchmark, all on the device.
 * Primary: ray i -> pixel i mod (width*height), jitter Random2 of
 * InitRandomSeed(seed, i), direction as pathtracer.comp:39-46.
 * Diffuse: ray i starts at the (i mod n_primary)-th primary hit, skipping
 * misses cyclically; origin = hit + 1e-4 * d, d cosine-distributed about the
 * geometric normal facing the incoming ray (random.glsl:62-71), seed `seed`.
 * Shadow: from the same hit points toward lights[i mod n_lights],
 * origin + 1e-4 * d, tmax = dist - 2e-4 (lightsource.glsl:38-40). */
int gpugi_raygen_primary(gpugi_ctx* ctx, const gpugi_camera* camera, int width, int height,
                         uint32_t seed, gpugi_ray* rays_dev, size_t n);
int gpugi_raygen_diffuse(gpugi_ctx* ctx, const gpugi_ray* primary_dev, const gpugi_hit* hits_dev,
                         size_t n_primary, uint32_t seed, gpugi_ray* rays_dev, size_t n);
int gpugi_raygen_shadow(gpugi_ctx* ctx, const gpugi_ray* primary_dev, const gpugi_hit* hits_dev,
                        size_t n_primary, const float* lights_xyz, uint32_t n_lights,
                        gpugi_ray* rays_dev, size_t n);

/* ------------------------------------------------------------------------ */
/* Host-side scene: .rawscene files, the SAH sweep builder, generators       */
/* ------------------------------------------------------------------------ */
typedef struct gpugi_hostscene gpugi_hostscene;

/* Material parameters of the material json (Scene::LoadMaterial key names,
 * scene.cpp:138-180): albedo, specularColor, roughness, opacity, emissivity,
 * refractionIndex. */
typedef struct gpugi_material_params {
  char name[32];
  float albedo[3];
  float specular_color[3];
  float roughness;
  float opacity;
  float emissivity[3];
  float refraction_index;
  /* "reflectiveness" constant (Vec4, scene.cpp:170-172) replacing the default
   * (1, 1, 1, 1 / (roughness^2 + 1e-20)) when has_reflectiveness != 0 */
  float reflectiveness[4];
  int32_t has_reflectiveness;
} gpugi_material_params;

/* Textures a loaded scene's materials reference: the material json's string
 * values for albedo, opacity, reflectiveness and emissivity name texture files
 * relative to the scene file's directory (Scene::LoadMaterial and
 * GetBindlessHandle, scene.cpp:155-203), read here as TGA, binary PPM or PFM
 * (rows bottom first, t = 0.5 / height for row 0).  A caller uploads them with
 * gpugi_add_texture (texels, format and size from gpugi_hostscene_texture,
 * valid while the scene lives) and binds them with gpugi_set_material_texture
 * (slots from gpugi_hostscene_material_textures, -1 for a constant slot). */
/* The oriented-box hierarchy gpugi_load_scene builds under GPUGI_BVH_OBOX for
 * this scene: num_nodes x 12 floats {centre.xyz, firstChild}{halfSides.xyz,
 * escape}{rotationInv quaternion xyzw} (TreeNode<OBox>, scene.hpp:39-49).  Each
 * node's box holds its subtree's triangles: principal or world axes, whichever
 * gives the smaller surface, exact under the float quaternion, widened by
 * 4e-6 x the scene's largest coordinate for the walk's float arithmetic. */
int gpugi_hostscene_obox_hierarchy(const gpugi_hostscene* scene, float* out, size_t num_floats);
int gpugi_hostscene_num_textures(const gpugi_hostscene* scene, uint32_t* n);
int gpugi_hostscene_texture(const gpugi_hostscene* scene, uint32_t i, int* width, int* height, int* format,
                            const void** texels);
int gpugi_hostscene_material_textures(const gpugi_hostscene* scene, uint32_t material, int32_t slots[4]);

/* Builds a scene from a triangle soup with the SAH sweep builder (restating
 * bvhmake/buildmethods/sweep.cpp:14-274 with leaf size triangles_per_leaf).
 * positions/normals: num_vertices x 3 floats; texcoords may be NULL;
 * indices: num_triangles x 3; material_ids: num_triangles. */
int gpugi_hostscene_build(const float* positions, const float* normals, const float* texcoords,
                          uint32_t num_vertices, const uint32_t* indices,
                          const uint32_t* material_ids, uint32_t num_triangles,
                          const gpugi_material_params* materials, uint32_t num_materials,
                          uint32_t triangles_per_leaf, gpugi_hostscene** out);
/* Synthetic Cornell room: 36 triangles, 10 walls (left red, right green),
 * 2-triangle emissive ceiling quad, two rotated boxes. */
int gpugi_hostscene_cornell(gpugi_hostscene** out);
/* Heightfield of cells x cells quads (2 triangles each) over [0,1]^2, heights
 * 0.2 * fBm (6 octaves of value noise on a wanghash lattice seeded `seed`). */
int gpugi_hostscene_heightfield(uint32_t cells, uint32_t seed, gpugi_hostscene** out);
/* Clutter room: a 20 x 20 x 8 room holding `spheres` seeded icospheres
 * (subdivision 3, 1280 triangles each, radii log-uniform in [0.05, 0.8];
 * 4000 spheres give 5,120,016 triangles), 60% Lambertian (albedo uniform in
 * [0.2, 0.9]) and 40% Phong-glossy (exponent log-uniform in [10, 1000])
 * materials binned into 240, one 4-triangle ceiling area light and 3 point
 * lights.  No reference counterpart: a synthetic scene of the integrator
 * benchmark. */
int gpugi_hostscene_clutter(uint32_t spheres, uint32_t seed, gpugi_hostscene** out);
/* Loads a .rawscene (documentation/fileformat.tex:37-168) and its material json. */
int gpugi_hostscene_load(const char* rawscene_path, const char* json_path, gpugi_hostscene** out);
/* Writes the .rawscene chunks the way bvhmake exports them
 * (bvhbuilder.cpp:569-632) plus the material json. */
int gpugi_hostscene_save(const gpugi_hostscene* scene, const char* rawscene_path,
                         const char* json_path);
/* Reference-layout arrays of the scene; pointers stay valid while the scene
 * lives and until the next call that modifies it. */
int gpugi_hostscene_desc(const gpugi_hostscene* scene, gpugi_scene_desc* out);
/* Counts and bounds: num_triangles (valid slots), max tree depth. */
int gpugi_hostscene_info(const gpugi_hostscene* scene, uint32_t* num_triangles,
                         uint32_t* max_depth, float bbox_min[3], float bbox_max[3]);
/* Scene::AddPointLight (scene.hpp:144). */
int gpugi_hostscene_add_point_light(gpugi_hostscene* scene, const float position[3],
                                    const float intensity[3]);
void gpugi_hostscene_destroy(gpugi_hostscene* scene);

/* Camera::ComputeCameraParams (camera.cpp:76-85) plus the PixelArea of
 * RendererSystem::SetCamera (renderersystem.cpp:208). */
int gpugi_camera_compute(const float position[3], const float lookat[3], const float up[3],
                         float vfov_degrees, int width, int height, gpugi_camera* out);

/* LightSampler (scene/lightsampler.cpp:18-95): area lights by summed-area CDF,
 * point lights by summed log-flux CDF, xorshift64* seeded 123 by default. */
typedef struct gpugi_lightsampler gpugi_lightsampler;
int gpugi_lightsampler_create(const gpugi_hostscene* scene, uint64_t seed, gpugi_lightsampler** out);
int gpugi_lightsampler_generate(gpugi_lightsampler* sampler, gpugi_light_sample* out, uint32_t n,
                                float position_bias);
/* LightSampler::SetScene (scene/lightsampler.cpp:13-16): samples another
 * scene from now on; the random state carries over. */
int gpugi_lightsampler_set_scene(gpugi_lightsampler* sampler, const gpugi_hostscene* scene);
void gpugi_lightsampler_destroy(gpugi_lightsampler* sampler);

/* ------------------------------------------------------------------------ */
/* Radiance-cache MLP (no reference counterpart: the learned cache that      */
/* lets paths terminate early; the plain path tracer stays the judge)       */
/* ------------------------------------------------------------------------ */
/* 96 encoded inputs -> 4 hidden layers of 64 (ReLU) -> 3 outputs, no biases;
 * BF16 weights on the matrix cores, FP32 accumulation, FP32 master weights
 * and Adam (beta1 0.9, beta2 0.99, eps 1e-8).  Weights are the five matrices
 * W_l[out][in] row-major, concatenated: 64x96, 64x64, 64x64, 64x64, 16x64
 * (rows 3..15 of the last are padding and stay 0). */
#define GPUGI_MLP_PARAMS 19456
/* Per-sample raw features (16 floats): position in [0,1]^3 (scene bounds),
 * direction, shading normal, diffuse albedo, log2(1 + Phong exponent) / 16,
 * 3 unused.  The network encodes them into 96 inputs: sin and cos of
 * 2^k * pi * position for k < 12, then the other 10 values, then zeros. */
#define GPUGI_MLP_FEATURES 16
#define GPUGI_MLP_INPUTS 96
/* Allocates the cache on the context and sets He-uniform initial weights from
 * a xorshift64* stream seeded `seed`; resets Adam. */
int gpugi_mlp_init(gpugi_ctx* ctx, uint64_t seed);
/* FP32 master weights in and out (n must be GPUGI_MLP_PARAMS); set resets Adam. */
int gpugi_mlp_set_weights(gpugi_ctx* ctx, const float* weights, size_t n);
int gpugi_mlp_get_weights(gpugi_ctx* ctx, float* weights, size_t n);
/* The BF16 input encoding of n samples (host buffers; n x 96 BF16 out). */
int gpugi_mlp_encode(gpugi_ctx* ctx, const float* features, uint16_t* encoded, size_t n);
/* Predicted radiance, n x 4 floats (rgb, 0). */
int gpugi_mlp_infer(gpugi_ctx* ctx, const float* features, float* radiance, size_t n, int on_device);
/* Gradient (GPUGI_MLP_PARAMS floats, host) of the relative L2 loss
 * mean over samples and rgb of (y - t)^2 / (y^2 + 0.01), the denominator held
 * constant, and the loss itself; targets are n x 4 floats.  No update. */
int gpugi_mlp_gradient(gpugi_ctx* ctx, const float* features, const float* targets, size_t n, float* grad,
                       float* loss, int on_device);
/* One Adam step on that gradient with the given learning rate; *loss
 * (optional) receives the loss before the step. */
int gpugi_mlp_train_step(gpugi_ctx* ctx, const float* features, const float* targets, size_t n,
                         float learning_rate, float* loss, int on_device);

/* Where cached paths query the cache (gpugi_cache_set_query). */
enum gpugi_cache_query {
  GPUGI_CACHE_SPREAD = 0,           /* footprint-spread termination, the cache predicts reflected radiance */
  GPUGI_CACHE_PRIMARY_INDIRECT = 1  /* at the primary hit after its NEE, the cache predicts the bounce part */
};

/* Cached path tracer (GPUGI_PT_CACHE): a path terminates into the cache at
 * the first vertex from the second hit on whose footprint spread exceeds
 * spread_c times the primary one (default 0.01); one pixel in train_stride
 * (default 32) also traces a training path extra_bounces (default 2) vertices
 * further; up to max_train_samples (default 65536) vertices of those paths
 * train the cache with train_steps (default 4) Adam steps of
 * learning_rate (default 1e-3) after every frame. */
int gpugi_cache_config(gpugi_ctx* ctx, float spread_c, int extra_bounces, uint32_t train_stride,
                       uint32_t max_train_samples, int train_steps, float learning_rate);
/* Query rule of the cached path tracer (default GPUGI_CACHE_SPREAD).  With
 * GPUGI_CACHE_PRIMARY_INDIRECT every path stops at its primary hit after the
 * hit's emission and NEE, and the cache supplies the radiance the hit
 * reflects of light arriving by its sampled bounce (training paths do NEE at
 * their tail vertex for the same reason).  Changing the rule changes what
 * the network predicts: re-initialize or retrain it. */
int gpugi_cache_set_query(gpugi_ctx* ctx, int query);
/* Training samples and loss of the last cached frame. */
int gpugi_cache_stats(gpugi_ctx* ctx, uint32_t* train_samples, float* loss);

/* ------------------------------------------------------------------------ */
/* Batched rendering (no reference counterpart: the reference renders one   */
/* scene per GL context)                                                    */
/* ------------------------------------------------------------------------ */
/* Many scenes, each with its own BVH, camera, resolution, light samples and
 * accumulation buffer, rendered by the path tracer (pathtracer.comp) in one
 * persistent launch whose workgroups take 8x8 pixel tiles of all scenes from
 * one shared queue.  A frame k draws with FrameSeed = WangHash(k) for every
 * scene, derived on the device, so one captured frame replays as a hipGraph.
 * A batch has its own HIP stream; it is not thread-safe. */
typedef struct gpugi_batch gpugi_batch;
int gpugi_batch_create(int device, gpugi_batch** out);
void gpugi_batch_destroy(gpugi_batch* batch);
const char* gpugi_batch_last_error(const gpugi_batch* batch);
/* Adds a scene (arrays are copied) before gpugi_batch_finalize. */
int gpugi_batch_add(gpugi_batch* batch, const gpugi_scene_desc* scene, const gpugi_camera* camera, int width,
                    int height, const gpugi_light_sample* lights, uint32_t num_lights, int max_path_length);
/* Packs and uploads all scenes; then gpugi_batch_reset(batch, 0). */
int gpugi_batch_finalize(gpugi_batch* batch);
/* Traversal (GPUGI_KERNEL_REFERENCE: the BVH2 escape walk, GPUGI_KERNEL_WIDE:
 * the 4-wide BVH), tiles taken per queue access, the persistent grid in
 * quarters of full residency (1..16) and the wide walk's LDS stack entries
 * per lane (16, 24 or 32; deeper entries spill to global memory; the escape
 * walk ignores it).  Defaults: WIDE, 2, 2, 16. */
int gpugi_batch_config(gpugi_batch* batch, int kernel, int tiles_per_grab, int grid_multiplier, int stack_depth);
/* Clears every accumulation buffer and the ray counter; the next frame is frame_index. */
int gpugi_batch_reset(gpugi_batch* batch, uint32_t frame_index);
/* One frame: persistent != 0 as one persistent launch, else one launch per scene. */
int gpugi_batch_draw(gpugi_batch* batch, int persistent);
/* Captures one persistent frame as a hipGraph; gpugi_batch_replay launches it `times` times. */
int gpugi_batch_capture(gpugi_batch* batch);
int gpugi_batch_replay(gpugi_batch* batch, int times);
/* Runs `frames` frames (mode 0: per-scene launches, 1: persistent, 2: graph
 * replay of a persistent frame) and returns their GPU time in *ms. */
int gpugi_batch_run(gpugi_batch* batch, int mode, int frames, float* ms);
/* Accumulation buffer (RGBA32F, width x height) of scene `scene`. */
int gpugi_batch_readback(gpugi_batch* batch, int scene, float* rgba, size_t bytes);
/* Ray queries (closest + any hit) traced since the last reset. */
int gpugi_batch_rays(gpugi_batch* batch, uint64_t* rays);
/* Times every (kernel, tiles per grab in {1,2,4,8}, grid multiplier in
 * {1,2,4}, and for the wide kernel stack depth in {16,24,32}) on `frames`
 * persistent frames and keeps the fastest. */
int gpugi_batch_autotune(gpugi_batch* batch, int frames, int* kernel, int* tiles_per_grab, int* grid_multiplier,
                         int* stack_depth, float* best_ms);

/* WangHash (utilities/random.cpp:5-13): the per-iteration FrameSeed. */
uint32_t gpugi_wanghash(uint32_t x);

#ifdef __cplusplus
}
#endif

#endif /* GPUGI_RT_H */
