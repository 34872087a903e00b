/*
 * CPU oracle of the gpugi ray-tracing hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithms, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * The product (gpugi_amd/) never links, loads or calls it.
 *
 * Compiled with gcc -O2 -ffp-contract=off (SSE scalar FP32, no FMA, no excess
 * precision), so every FP32 expression rounds exactly as written.
 *
 * Pinning (see tests/): RNG against golden vectors from the reference's own
 * utilities/random.cpp (built into oracle/_ref) and closed forms; math
 * functions against libm / FP64 within stated ulp bounds; box and triangle
 * tests against FP64 geometry and special cases; traversal against brute
 * force over all triangles on small scenes; integrators against analytic
 * cases (white furnace, single-bounce Lambertian, path-length slices).
 */
#ifndef GPUGI_ORACLE_H
#define GPUGI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

typedef struct {
  const float* hierarchy;   /* num_nodes x 8 */
  uint32_t num_nodes;
  const int32_t* triangles; /* num_leaves x 2 x 4 */
  uint32_t num_leaves;
  const float* positions;   /* num_vertices x 3 */
  const float* vertex_info; /* num_vertices x 4 */
  uint32_t num_vertices;
  const float* materials;   /* num_materials x 20 (gpugi_material) */
  uint32_t num_materials;
  const float* obox;        /* NULL, or num_nodes x 12 (TreeNode<OBox>): the walk tests these boxes */
} or_scene;

typedef struct {
  float position[3], u[3], v[3], w[3], pixel_area;
} or_camera;

/* RNG (random.glsl:1-49, utilities/random.cpp:5-43) */
uint32_t or_wanghash(uint32_t x);
uint32_t or_xorshift32(uint32_t* s);
float or_random01(uint32_t* s);
uint64_t or_xorshift64star(uint64_t s);
float or_xorshiftf(uint64_t* s);

/* FP32 math (Cephes-style kernels; the semantics the renderer defines) */
float or_sinf(float x);
float or_cosf(float x);
float or_log2f(float x);
float or_exp2f(float x);
float or_powf(float x, float y);
float or_atanf(float x);
float or_atan2f(float y, float x);

/* Traversal (traceray.glsl:18-180).  rays: n x 8 floats {o, tmax, d, pad};
 * hits: n x 4 int32 {slot, bits(t), bits(by), bits(bz)} (closest) or {flag,0,0,0}
 * (any).  nbox/ntri may be NULL. */
void or_trace(const or_scene* s, int any, const float* rays, int32_t* hits, uint32_t* nbox, uint32_t* ntri,
              size_t n);
/* Brute force: every valid slot with the reference triangle test, closest
 * t > 1e-4 below tmax, ties to the smaller slot; any-hit: first accepted. */
void or_bruteforce(const or_scene* s, int any, const float* rays, int32_t* hits, size_t n);
/* The same with the watertight triangle test when wt != 0 (Woop et al. 2013;
 * no reference counterpart). */
void or_trace_mode(const or_scene* s, int any, int wt, const float* rays, int32_t* hits, uint32_t* nbox,
                   uint32_t* ntri, size_t n);
void or_bruteforce_mode(const or_scene* s, int any, int wt, const float* rays, int32_t* hits, size_t n);
/* FP64 classification of a ray against slots a and b (-1 = none).
 * Returns flags: 1 = a within rel_eps of an edge, 2 = b within rel_eps of an
 * edge, 4 = tie (|ta - tb| <= rel_eps * max), 8 = a or b within rel_eps of
 * the t bounds (1e-4 or tmax), 16 = ray/triangle plane nearly parallel. */
int or_classify(const or_scene* s, const float* ray, int32_t slot_a, int32_t slot_b, double rel_eps);

/* Ray generators of the microbenchmark (restating the device generators). */
void or_raygen_primary(const or_camera* c, int width, int height, uint32_t seed, float* rays, size_t first,
                       size_t n);
void or_raygen_diffuse(const or_scene* s, const float* prim, const int32_t* prim_hits, size_t n_prim,
                       uint32_t seed, float* rays, size_t first, size_t n);
void or_raygen_shadow(const float* prim, const int32_t* prim_hits, size_t n_prim, const float* lights,
                      uint32_t n_lights, float* rays, size_t first, size_t n);

/* BSDF probes (__SampleBSDF / __BSDF, materials.glsl:118-391) for a 20-float
 * material record: n samples as {dir3, throughput3, pdf, diffuse} and n
 * evaluations at the given excident directions as {bsdf3, pdf}. */
void or_bsdf_samples(const float* mat20, const float* inc, const float* N, uint32_t seed, int adjoint, size_t n,
                     float* out);
void or_bsdf_evals(const float* mat20, const float* inc, const float* N, int adjoint, const float* exc, size_t n,
                   float* out);

/* Environment cube map used by or_pathtracer on a miss (NULL / 0: black):
 * texels as 6 faces (x-, x+, y-, y+, z-, z+) of size x size linear RGB, 4
 * floats each; the array must outlive the draws. */
void or_set_environment(const float* texels, int size);
/* TraceRay / TraceRayAnyHit with TRACERAY_IMPORTANCE_BREAK (traceray.glsl:48-112):
 * importance holds {x, y} per node; nodes (closest hit, optional) get
 * {node, importance.x bits, size^2 bits, 0} per ray. */
void or_trace_importance(const or_scene* s, int any, const float* rays, int32_t* hits, int32_t* nodes, size_t n,
                         const float* importance, float threshold, float rayThreshold);
/* Material textures for every integrator (NULL: none): matTex holds 4 texture
 * indices per material (diffuse, opacity, reflectiveness, emissivity; -1 for
 * the record's constant), texels[i] the linear RGBA floats of texture i and
 * sizes[2i], sizes[2i+1] its width and height; the arrays must outlive the draws. */
void or_set_textures(const int32_t* matTex, uint32_t numMaterials, const float* const* texels, const int32_t* sizes);
/* The texture lookup of texture id (set by or_set_textures) at n coordinates (u, v) into out (4 floats each). */
void or_texture_samples(int id, const float* uvs, size_t n, float* out);
/* The cube-map lookup for n directions (3 floats each) into out (3 floats each). */
void or_env_lookups(const float* tex, int size, const float* dirs, size_t n, float* out);

/* Options of every integrator below (global state; defaults -1 and 0):
 * specificPathLength >= 0 keeps only paths of that many bounces (the reference's
 * SHOW_SPECIFIC_PATHLENGTH define, renderer.hpp:6-9, 1 = direct lighting);
 * rrStartDepth is the first bounce at which Russian roulette runs (0, every
 * bounce, is the reference; a later start has no reference counterpart). */
void or_set_path_options(int specificPathLength, int rrStartDepth);
extern int or_spl, or_rr_start;
/* Pixel that ProjectToScreen (screenreproject.glsl:55-96) splats world point p
 * into for the camera: returns 0 and (*ix, *iy) when it lies in front of the
 * camera inside the screen, -1 otherwise. */
int or_screen_pixel(const or_camera* c, int width, int height, const float* p, int* ix, int* iy);

/* Integrators: one draw adds into accum (width x height x 4 floats). */
void or_pathtracer(const or_scene* s, const float* lights, uint32_t num_lights, const or_camera* c, int width,
                   int height, uint32_t frame_seed, int max_path_length, float* accum, int row_begin, int row_end);

/* Light tracer draw: light paths then the light-sample splat (LightPathtracer::Draw). */
void or_lighttracer(const or_scene* s, const float* lights, uint32_t num_lights, const or_camera* c, int width,
                    int height, uint32_t frame_seed, int max_path_length, float* accum);
/* BPT warm-up counters (bidirectionalpathtracer.cpp:65-104). */
void or_bpt_warmup(const or_scene* s, const float* lights, uint32_t num_lights, const or_camera* c, int width,
                   int height, uint32_t frame_seed, int max_path_length, uint64_t* num_entries,
                   uint64_t* path_length_sum);
/* BPT light pass: splats into accum, fills cache (capacity x 16 floats); returns NumLightCacheEntries. */
uint32_t or_bpt_light_pass(const or_scene* s, const float* lights, uint32_t num_lights, const or_camera* c,
                           int width, int height, uint32_t frame_seed, int max_path_length, int capacity,
                           float* accum, float* cache);
/* Parts of the LT draw and the BPT light pass for splitting light paths over
 * threads: paths [first, last) into accum; the light-sample splat; BPT paths
 * [first, last) with initial entries into cache and bounce vertices staged per
 * path (staged: (last - first) x max_path_length x 16 floats, counts per path). */
void or_lighttracer_paths(const or_scene* s, const float* lights, uint32_t num_lights, const or_camera* c, int width,
                          int height, uint32_t frame_seed, int max_path_length, uint32_t first, uint32_t last,
                          float* accum);
void or_showlightcaches(const or_scene* s, const float* lights, uint32_t num_lights, const or_camera* c, int width,
                        int height, float* accum);
void or_bpt_light_paths(const or_scene* s, const float* lights, uint32_t num_lights, const or_camera* c, int width,
                        int height, uint32_t frame_seed, int max_path_length, int capacity, uint32_t first,
                        uint32_t last, float* accum, float* cache, float* staged, uint32_t* counts);
/* BPT eye pass over rows [row_begin, row_end). */
void or_bpt_eye(const or_scene* s, const or_camera* c, int width, int height, uint32_t frame_seed,
                int max_path_length, int capacity, const float* cache, uint32_t num_entries, float* accum,
                int row_begin, int row_end);

/* HierarchyImportance passes (hierarchyimportance.cpp:84-178, the shader/hierarchy passes).
 * importance: (num_nodes + 2 * num_leaves) x {x, y}; or_count_hits_rays adds
 * w[i] for ray i (rays: n x 8 floats {o, length, d, pad}); or_slot_hits: per ray
 * and valid slot 1 if IntersectTriangle accepts it at t <= length; parents: num_nodes. */
void or_count_hits_rays(const or_scene* s, const float* rays, const float* w, size_t n, float* importance);
void or_slot_hits(const or_scene* s, const float* rays, size_t n, uint8_t* out);
void or_hier_parents(const or_scene* s, int32_t* parent);
void or_hier_acquire(const or_scene* s, const float* lights, uint32_t num_lights, const or_camera* c, int width,
                     int height, int iteration, float* importance);
void or_hier_propagate(const or_scene* s, int iterations, float* importance, float* subtree);
void or_hier_materials(const or_scene* s, uint32_t* out);
uint16_t or_half_bits(float f);
float or_half_value(uint16_t h);

#endif
