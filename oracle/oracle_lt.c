/*
 * CPU oracle: light tracer (lighttracer.comp, screenreproject.glsl,
 * showlightcaches.comp) and the bidirectional path tracer's light pass and eye
 * pass (lighttracer.comp with SAVE_LIGHT_CACHE, bidirpathtracer/pathtracer_bidir.comp).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The light cache is appended in (path, bounce) order, one of the orders the
 * reference's atomicAdd append (lighttracer.comp:120) can produce.
 */
#include <string.h>

#include "oracle.h"
#include "oracle_vec.h"

static int rays_per_sample(int W, int H, uint32_t n) {
  int r = (int)(((uint32_t)(W * H) / n / 64u) * 64u);  /* lightpathtracer.cpp:29-31 */
  return r < 64 ? 64 : r;
}

static float mis_h(float p) { return fminf(1e18f, p); }            /* lightcache.glsl:61-73 */
static float dist_mis(float d, float c) { (void)c; return mis_h(d * d); }  /* lightcache.glsl:84-89 */

/* AdjointBSDFShadingNormalCorrectedOutDotN (materials.glsl:415-426) */
static float corrected_out_dot_n(vec3 in, vec3 cam, vec3 gN, vec3 sN) {
  if (vdot(in, gN) * vdot(in, sN) < 0.0f || vdot(cam, gN) * vdot(cam, sN) < 0.0f) return 0.0f;
  return fsat(fabsf(vdot(in, sN) * vdot(cam, gN)) / fmaxf(fabsf(vdot(in, gN)), 1e-5f));
}
/* AdjointBSDFShadingNormalCorrection (materials.glsl:427-436) */
static float adjoint_correction(vec3 in, vec3 out, vec3 gN, vec3 sN) {
  if (vdot(in, gN) * vdot(in, sN) < 0.0f || vdot(out, gN) * vdot(out, sN) < 0.0f) return 0.0f;
  return fsat(fabsf(vdot(in, sN) * vdot(out, gN)) / fmaxf(fabsf(vdot(in, gN) * vdot(out, sN)), 1e-5f));
}
/* AdjointBSDF (materials.glsl:405-408) */
static vec3 adjoint_bsdf(vec3 inc, vec3 exc, const or_material* m, vec3 N, float* pdf) {
  return or_bsdf(vneg(exc), vneg(inc), m, N, pdf, 1, 0);
}

static void cam_vectors(const or_camera* c, vec3* U, vec3* V, vec3* W, vec3* P) {
  *U = v3(c->u[0], c->u[1], c->u[2]);
  *V = v3(c->v[0], c->v[1], c->v[2]);
  *W = v3(c->w[0], c->w[1], c->w[2]);
  *P = v3(c->position[0], c->position[1], c->position[2]);
}

/* The screen position of ProjectToScreen (screenreproject.glsl:71-95):
 * crd = camera - p, the projection onto U/V over the view distance, the
 * half-open screen window, and the pixel by round-to-nearest.  Returns 0 and
 * the pixel when p is in front of the camera and inside the window.
 *
 * The OR_PLANT_* macros plant a known mistake into a separate build of the
 * oracle (oracle.planted() in oracle.py); tests use those builds to show that
 * their checks reject such a mistake.  The normal build defines none. */
static int screen_pixel(const or_camera* c, int Wd, int Ht, vec3 p, vec3* crdOut, float* camViewDotRayOut,
                        int* ix, int* iy) {
  vec3 U, V, W, P;
  cam_vectors(c, &U, &V, &W, &P);
  vec3 crd = vsub(P, p);
  float camViewDotRay = -vdot(W, crd);
  if (!(camViewDotRay > 0.0f)) return -1;
  float sx = -vdot(U, crd) / camViewDotRay;
  float sy = -vdot(V, crd) / camViewDotRay;
  sx = sx / vdot(U, U);
  sy = sy / vdot(V, V);
#ifdef OR_PLANT_YFLIP
  sy = -sy;
#endif
  if (!(sx >= -1.0f && sx < 1.0f && sy > -1.0f && sy <= 1.0f)) return -1;
  *ix = (int)roundf((sx * 0.5f + 0.5f) * (float)Wd);
  *iy = (int)roundf((sy * 0.5f + 0.5f) * (float)Ht);
#ifdef OR_PLANT_XSHIFT
  *ix += 1;
#endif
  if (crdOut) *crdOut = crd;
  if (camViewDotRayOut) *camViewDotRayOut = camViewDotRay;
  return 0;
}

int or_screen_pixel(const or_camera* c, int width, int height, const float* p, int* ix, int* iy) {
  return screen_pixel(c, width, height, v3(p[0], p[1], p[2]), NULL, NULL, ix, iy);
}

/* ProjectToScreen, box filter, no depth-buffer test, no chained MIS (screenreproject.glsl:55-227) */
static void project_to_screen(const or_scene* s, const or_camera* c, int Wd, int Ht, float* accum, vec3 rayO,
                              vec3 rayD, vec3 gN, vec3 sN, const or_material* m, vec3 tp, int computeMIS) {
  vec3 crd;
  float camViewDotRay;
  int ix, iy;
  if (screen_pixel(c, Wd, Ht, rayO, &crd, &camViewDotRay, &ix, &iy) != 0) return;
  float distSq = vdot(crd, crd);
  float dist = sqrtf(distSq);
  crd = vdiv(crd, dist);
  vec3 co = vadd(vscalev(crd, 0.0001f), rayO);
  if (or_any_hit(s, co, crd, dist)) return;
  float cosAtCamera = fsat(camViewDotRay / dist);
  float f2i = corrected_out_dot_n(rayD, crd, gN, sN) / (cosAtCamera * cosAtCamera * cosAtCamera * c->pixel_area * distSq);
  float pdf;
  vec3 bsdf = adjoint_bsdf(rayD, crd, m, sN, &pdf);
  if (computeMIS) {
    float eyeToLight = 1.0f, lightToEye = mis_h(pdf);
    float mis = eyeToLight / fmaxf(lightToEye + eyeToLight, 1e-5f);
    bsdf = vscalev(bsdf, mis);
  }
  vec3 color = vmul(vscalev(tp, f2i), bsdf);
  if (ix < 0 || ix >= Wd || iy < 0 || iy >= Ht) return;
  float* px = accum + ((size_t)iy * Wd + ix) * 4;
  px[0] += color.x;
  px[1] += color.y;
  px[2] += color.z;
}

/* one 64-byte LightCacheEntry (bidirpathtracer/lightcache.glsl:1-16) */
static void put_entry(float* e, vec3 flux, int mat, vec3 pos, float n0, float n1, float u, float v, vec3 inc, int len) {
  int32_t mi = mat, li = len;
  e[0] = flux.x; e[1] = flux.y; e[2] = flux.z; memcpy(&e[3], &mi, 4);
  e[4] = pos.x; e[5] = pos.y; e[6] = pos.z; e[7] = n0;
  e[8] = u; e[9] = v; e[10] = n1; memcpy(&e[11], &li, 4);
  or_pack_normal(inc, &e[12], &e[13]);
  e[14] = 0.0f; e[15] = 0.0f;
}

/* mode 0: LT; 1: BPT light pass (stores vertices into `staged`); 2: warm-up (counts only).
 * Returns the number of stored vertices of this path. */
static int light_path(const or_scene* s, const float* lights, const or_camera* c, int Wd, int Ht, uint32_t frameSeed,
                      int maxLen, int R, uint32_t gid, int mode, float* accum, float* staged) {
  uint32_t seed = or_wanghash(frameSeed + gid);
  const float* ls = lights + (size_t)(gid / (uint32_t)R) * 8;
  vec3 lpos = v3(ls[0], ls[1], ls[2]), lint = v3(ls[4], ls[5], ls[6]);
  vec3 tp = vscalev(lint, OR_PI / (float)R);
  vec3 o = lpos, d;
  if (ls[7] > 1.5f) {
    float r0 = or_random01(&seed), r1 = or_random01(&seed);
    d = or_sample_direction(r0, r1);
    tp = vscalev(tp, 4.0f);
  } else {
    vec3 n = or_unpack_normal(ls[3], ls[7]), U, V;
    or_create_onb(n, &U, &V);
    float r0 = or_random01(&seed), r1 = or_random01(&seed);
    d = or_hemisphere_cosine(r0, r1, U, V, n);
  }
  int stored = 0;
  for (int i = 0; i < maxLen; ++i) {
    float t = 3.402823466e+38f;
    int32_t slot, tri[4];
    vec3 b, gN;
    or_trace_one(s, 0, o, d, &t, &slot, &b, &gN, tri, NULL, NULL);
    if (t == 3.402823466e+38f) break;
    o = vadd(o, vscale(t, d));
    vec3 sN = or_shading_normal(s, tri, b);
    const float* va = s->vertex_info + (size_t)tri[0] * 4;
    const float* vb = s->vertex_info + (size_t)tri[1] * 4;
    const float* vc = s->vertex_info + (size_t)tri[2] * 4;
    float u = va[2] * b.x + vb[2] * b.y + vc[2] * b.z;
    float v = va[3] * b.x + vb[3] * b.y + vc[3] * b.z;
    or_material m = or_mat_at(s, tri[3], u, v);
    if (mode != 0 && m.refl[3] < 100.0f) {
      if (mode == 1) {
        float n0, n1;
        or_pack_normal(sN, &n0, &n1);
        put_entry(staged + (size_t)stored * 16, tp, tri[3], o, n0, n1, u, v, d, i + 1);
      }
      ++stored;
    }
    /* the camera connection closes a path of i+1 bounces (lighttracer.comp:133-157) */
    if (mode != 2 && (or_spl < 0 || i + 1 == or_spl))
      project_to_screen(s, c, Wd, Ht, accum, o, d, gN, sN, &m, tp, mode == 1);
    vec3 t1 = v3(1, 1, 1);
    float pdf;
    int dif;
    vec3 nd = or_sample_bsdf(d, &m, &seed, sN, &t1, 1, &pdf, &dif);
    t1 = vscalev(t1, adjoint_correction(d, nd, gN, sN));
    tp = vmul(tp, t1);
    if (i >= or_rr_start) { /* Russian roulette (lighttracer.comp:183-188) from or_rr_start on */
      float q = fsat(or_luminance(t1));
      if (or_random01(&seed) >= q) break;
      tp = vdivv(tp, v3(q, q, q));
    }
    d = nd;
    o = vadd(o, vscalev(d, 0.0001f));
  }
  return stored;
}

/* showlightcaches.comp:9-65 for light sample i */
static void show_light_sample(const or_scene* s, const float* lights, uint32_t i, const or_camera* c, int Wd, int Ht,
                              float* accum) {
  vec3 U, V, W, P;
  cam_vectors(c, &U, &V, &W, &P);
  const float* ls = lights + (size_t)i * 8;
  vec3 lpos = v3(ls[0], ls[1], ls[2]);
  vec3 crd = vsub(P, lpos);
  float camViewDotRay = -vdot(W, crd);
  if (!(camViewDotRay > 0.0f)) return;
  float sx = -vdot(U, crd) / camViewDotRay;
  float sy = -vdot(V, crd) / camViewDotRay;
  sx = sx / vdot(U, U);
  sy = sy / vdot(V, V);
  if (!(sx >= -1.0f && sx < 1.0f && sy > -1.0f && sy <= 1.0f)) return;
  float distSq = vdot(crd, crd);
  float dist = sqrtf(distSq);
  crd = vdiv(crd, dist);
  vec3 co = vadd(vscalev(crd, 0.0001f), lpos);
  if (or_any_hit(s, co, crd, dist)) return;
  vec3 n = or_unpack_normal(ls[3], ls[7]);
  float cosTheta = vdot(n, crd);
  float cosAtCamera = fsat(camViewDotRay / dist);
  float i2r = cosTheta / (OR_PI_2 * cosAtCamera * cosAtCamera * cosAtCamera * c->pixel_area * distSq);
  vec3 color = vscalev(v3(ls[4], ls[5], ls[6]), i2r);
  int ix = (int)((sx * 0.5f + 0.5f) * (float)Wd);
  int iy = (int)((sy * 0.5f + 0.5f) * (float)Ht);
  if (ix < 0 || ix >= Wd || iy < 0 || iy >= Ht) return;
  float* px = accum + ((size_t)iy * Wd + ix) * 4;
  px[0] += color.x;
  px[1] += color.y;
  px[2] += color.z;
}

void or_lighttracer(const or_scene* s, const float* lights, uint32_t numLights, const or_camera* c, int W, int H,
                    uint32_t frameSeed, int maxLen, float* accum) {
  int R = rays_per_sample(W, H, numLights);
  uint32_t n = numLights * (uint32_t)R;
  for (uint32_t gid = 0; gid < n; ++gid) light_path(s, lights, c, W, H, frameSeed, maxLen, R, gid, 0, accum, NULL);
  for (uint32_t i = 0; i < numLights; ++i) show_light_sample(s, lights, i, c, W, H, accum);
}

void or_bpt_warmup(const or_scene* s, const float* lights, uint32_t numLights, const or_camera* c, int W, int H,
                   uint32_t frameSeed, int maxLen, uint64_t* numEntries, uint64_t* pathLengthSum) {
  int R = rays_per_sample(W, H, numLights);
  uint32_t n = numLights * (uint32_t)R;
  uint64_t stored = 0;
  for (uint32_t gid = 0; gid < n; ++gid)
    stored += (uint64_t)light_path(s, lights, c, W, H, frameSeed, maxLen, R, gid, 2, NULL, NULL);
  *pathLengthSum = stored;
  *numEntries = n + stored;
}

/* Light pass: initial vertices at [0, R*N), bounce vertices appended in
 * (path, bounce) order; entries at or beyond `capacity` are not written. */
uint32_t or_bpt_light_pass(const or_scene* s, const float* lights, uint32_t numLights, const or_camera* c, int W,
                           int H, uint32_t frameSeed, int maxLen, int capacity, float* accum, float* cache) {
  int R = rays_per_sample(W, H, numLights);
  uint32_t n = numLights * (uint32_t)R;
  float* staged = (float*)malloc((size_t)maxLen * 64);
  uint32_t next = n;
  for (uint32_t gid = 0; gid < n; ++gid) {
    const float* ls = lights + (size_t)(gid / (uint32_t)R) * 8;
    if ((int)gid < capacity) {
      vec3 flux = vdiv(vdiv(v3(ls[4], ls[5], ls[6]), OR_PI_2), (float)R);
      put_entry(cache + (size_t)gid * 16, flux, -1, v3(ls[0], ls[1], ls[2]), ls[3], ls[7], 0.0f, 0.0f,
                v3(0, 0, 0), 0);
    }
    int k = light_path(s, lights, c, W, H, frameSeed, maxLen, R, gid, 1, accum, staged);
    for (int j = 0; j < k; ++j, ++next)
      if ((int64_t)next < (int64_t)capacity) memcpy(cache + (size_t)next * 16, staged + (size_t)j * 16, 64);
  }
  free(staged);
  return next;
}

/* Parts of the draws above, so a caller can split the light paths over
 * threads (each into its own accumulation buffer): the light paths
 * [first, last) of a light tracer draw, the light-sample splat of a light
 * tracer draw, and the light paths [first, last) of a BPT light pass with
 * their initial cache entries (at [first, last), below capacity) and their
 * stored bounce vertices staged per path (staged: (last - first) x maxLen
 * entries of 16 floats, counts: stored vertices per path). */
void or_lighttracer_paths(const or_scene* s, const float* lights, uint32_t numLights, const or_camera* c, int W,
                          int H, uint32_t frameSeed, int maxLen, uint32_t first, uint32_t last, float* accum) {
  const int R = rays_per_sample(W, H, numLights);
  const uint32_t n = numLights * (uint32_t)R;
  if (last > n) last = n;
  for (uint32_t gid = first; gid < last; ++gid) light_path(s, lights, c, W, H, frameSeed, maxLen, R, gid, 0, accum, NULL);
}

void or_showlightcaches(const or_scene* s, const float* lights, uint32_t numLights, const or_camera* c, int W,
                        int H, float* accum) {
  for (uint32_t i = 0; i < numLights; ++i) show_light_sample(s, lights, i, c, W, H, accum);
}

void or_bpt_light_paths(const or_scene* s, const float* lights, uint32_t numLights, const or_camera* c, int W,
                        int H, uint32_t frameSeed, int maxLen, int capacity, uint32_t first, uint32_t last,
                        float* accum, float* cache, float* staged, uint32_t* counts) {
  const int R = rays_per_sample(W, H, numLights);
  const uint32_t n = numLights * (uint32_t)R;
  if (last > n) last = n;
  for (uint32_t gid = first; gid < last; ++gid) {
    const float* ls = lights + (size_t)(gid / (uint32_t)R) * 8;
    if ((int)gid < capacity) {
      vec3 flux = vdiv(vdiv(v3(ls[4], ls[5], ls[6]), OR_PI_2), (float)R);
      put_entry(cache + (size_t)gid * 16, flux, -1, v3(ls[0], ls[1], ls[2]), ls[3], ls[7], 0.0f, 0.0f,
                v3(0, 0, 0), 0);
    }
    counts[gid - first] = (uint32_t)light_path(s, lights, c, W, H, frameSeed, maxLen, R, gid, 1, accum,
                                               staged + (size_t)(gid - first) * (size_t)maxLen * 16);
  }
}

/* pathtracer_bidir.comp:16-251, no chained MIS; rows [r0, r1) */
void or_bpt_eye(const or_scene* s, const or_camera* c, int W, int H, uint32_t frameSeed, int maxLen, int capacity,
                const float* cache, uint32_t numEntries, float* accum, int row_begin, int row_end) {
  const int gw = (W / 8) * 8, gh = (H / 8) * 8;
  vec3 U, V, Wv, P;
  cam_vectors(c, &U, &V, &Wv, &P);
  const uint32_t actualN = (uint32_t)((int)numEntries < capacity ? (int)numEntries : capacity);
  const float lsw = (float)actualN;
  if (row_end > gh) row_end = gh;
  for (int y = row_begin; y < row_end; ++y) {
    for (int x = 0; x < gw; ++x) {
      uint32_t seed = or_wanghash(frameSeed + (uint32_t)(x + y * W));
      float j0 = or_random01(&seed), j1 = or_random01(&seed);
      vec3 o = P, d = or_primary_dir(j0, j1, x, y, W, H, U, V, Wv);
      vec3 color = v3(0, 0, 0), tp = v3(1, 1, 1);
      float lastP = 1.0f, lastCos = 1.0f, mis = 0.0f;
      for (int i = 0; i < maxLen; ++i) {
        float t = 3.402823466e+38f;
        int32_t slot, tri[4];
        vec3 b, gN;
        or_trace_one(s, 0, o, d, &t, &slot, &b, &gN, tri, NULL, NULL);
        if (t == 3.402823466e+38f) break;
        float rayHit = t;
        o = vadd(o, vscale(rayHit, d));
        vec3 n = or_shading_normal(s, tri, b);
        float tu, tv;
        or_hit_texcoord(s, tri, b, &tu, &tv);
        or_material m = or_mat_at(s, tri[3], tu, tv);
        float inCos = fsat(-vdot(n, d));
        float l2e = dist_mis(rayHit, lastCos) * mis_h(inCos / OR_PI);
#ifdef OR_PLANT_NO_HIT_DIST_MIS
        l2e = mis_h(inCos / OR_PI);
#endif
        /* SHOW_SPECIFIC_PATHLENGTH (pathtracer_bidir.comp:75-92): emission at vertex i */
        if (or_spl < 0 || i == or_spl) {
          mis = (i == 0) ? 1.0f : (lastP / fmaxf(lastP + l2e, 1e-5f));
          color = vadd(color, vmul(vscale(mis, tp), m.emissivity));
        }
        uint32_t idx = or_xorshift32(&seed) % actualN;
        const float* e = cache + (size_t)idx * 16;
        int32_t emat, elen;
        memcpy(&emat, &e[3], 4);
        memcpy(&elen, &e[11], 4);
        int pathLength = elen + i + 1;
        /* a connection closes a path of elen + i + 1 bounces (pathtracer_bidir.comp:101-104) */
        if (pathLength <= maxLen && (or_spl < 0 || pathLength == or_spl)) {
          vec3 ld = vsub(v3(e[4], e[5], e[6]), o);
          float distSq = vdot(ld, ld);
          float dist = sqrtf(distSq);
          ld = vdiv(ld, fmaxf(dist, 1e-5f));
          vec3 lo = vadd(vscale(0.0001f, ld), o);
          if (!or_any_hit(s, lo, ld, dist - 0.0001f * 2.0f)) {
            vec3 flux = v3(e[0], e[1], e[2]);
            vec3 lnormal = v3(0.0f, 1.0f, 0.0f);
            float cosLight = 0.0f;
            vec3 lint;
            float cosCamera = fabsf(vdot(ld, n));
            if (e[10] < 1.5f) {
              lnormal = or_unpack_normal_raw(e[7], e[10]);
              cosLight = fabsf(vdot(ld, lnormal));
            }
            float pdf;
            vec3 bsdf = or_bsdf(d, ld, &m, n, &pdf, 0, 0);
            float e2l = mis_h(pdf);
            if (i > 0) {
              float dm = dist_mis(rayHit, lastCos);
              e2l = e2l * dm;
              mis = lastP / fmaxf(lastP + e2l, 1e-5f);
              if (m.refl[3] >= 100.0f) mis = 1.0f;
            } else {
              mis = e2l / fmaxf(lastP + e2l, 1e-5f);
            }
            float misLight = 1.0f;
            if (e[10] < 1.5f) {
              float lightToEye;
              if (emat != -1) {
                or_material lm = or_mat_at(s, emat, e[8], e[9]);
                vec3 inc = or_unpack_normal(e[12], e[13]);
                vec3 lb = adjoint_bsdf(ld, vneg(inc), &lm, lnormal, &pdf);
                lint = vscalev(vmul(flux, lb), cosLight);
                lightToEye = mis_h(pdf) * dist_mis(dist, cosCamera);
              } else {
                lint = vscalev(flux, cosLight * OR_PI_2);
                lightToEye = mis_h(cosLight / OR_PI) * dist_mis(dist, cosCamera);
              }
#ifdef OR_PLANT_NO_LIGHT_DIST_MIS
              lightToEye = lightToEye / dist_mis(dist, cosCamera);
#endif
              misLight = lightToEye / fmaxf(lightToEye + e2l, 1e-5f);
            } else {
              lint = vscalev(flux, OR_PI_2);
              misLight = 1.0f;
            }
            mis = mis * misLight;
            vec3 irr = vscale(mis * lsw * cosCamera / distSq, lint);
            color = vadd(color, vmul(vmul(irr, tp), bsdf));
          }
        }
        float spdf;
        vec3 t1 = v3(1, 1, 1);
        int dif;
        d = or_sample_bsdf(d, &m, &seed, n, &t1, 0, &spdf, &dif);
        o = vadd(o, vscalev(d, 0.0001f));
        tp = vmul(tp, t1);
        if (i >= or_rr_start) { /* Russian roulette (pathtracer_bidir.comp:213-218) from or_rr_start on */
          float q = fsat(or_luminance(t1));
          if (or_random01(&seed) >= q) break;
          tp = vdivv(tp, v3(q, q, q));
        }
        float rev = mis_h(spdf);
        if (i == 0) {
          mis = rev / fmaxf(rev + lastP, 1e-5f);
        } else {
          float dm = dist_mis(rayHit, lastCos);
          rev = rev * dm;
          mis = lastP / fmaxf(lastP + dm * rev, 1e-5f);
          if (m.refl[3] >= 100.0f) mis = 1.0f;
        }
        tp = vscalev(tp, mis);
        lastP = rev;
        lastCos = fabsf(vdot(d, n));
      }
      float* px = accum + ((size_t)y * W + x) * 4;
      px[0] = color.x + px[0];
      px[1] = color.y + px[1];
      px[2] = color.z + px[2];
      px[3] = 1.0f + px[3];
    }
  }
}
