"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
leg as the checker; the product package never imports it.
"""
import contextlib
import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "librandom_ref.so")


def build(reference=True):
    """Builds liboracle.so (and the reference RNG library when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if reference and os.path.exists("/root/reference/gpugi/utilities/random.cpp"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


class _Scene(ctypes.Structure):
    _fields_ = [("hierarchy", ctypes.c_void_p), ("num_nodes", ctypes.c_uint32),
                ("triangles", ctypes.c_void_p), ("num_leaves", ctypes.c_uint32),
                ("positions", ctypes.c_void_p), ("vertex_info", ctypes.c_void_p),
                ("num_vertices", ctypes.c_uint32), ("materials", ctypes.c_void_p),
                ("num_materials", ctypes.c_uint32), ("obox", ctypes.c_void_p)]


class _Camera(ctypes.Structure):
    _fields_ = [("position", ctypes.c_float * 3), ("u", ctypes.c_float * 3), ("v", ctypes.c_float * 3),
                ("w", ctypes.c_float * 3), ("pixel_area", ctypes.c_float)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(reference=False)
        _lib = _bind(ctypes.CDLL(LIB))
    return _lib


@contextlib.contextmanager
def planted(mistake):
    """Runs the body against a separate oracle build with a planted mistake
    (YFLIP, XSHIFT, NO_LIGHT_DIST_MIS or NO_HIT_DIST_MIS, the OR_PLANT_* macros
    of oracle_lt.c), so a test can show that its check rejects that mistake."""
    global _lib
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "liboracle_%s.so" % mistake.lower())
        subprocess.run(["make", "-s", "-C", HERE, "planted", "PLANT=" + mistake, "OUT=" + out], check=True)
        saved = lib()
        _lib = _bind(ctypes.CDLL(out))
        try:
            yield _lib
        finally:
            _lib = saved


def _bind(l):
    """Declares the argument and result types of every oracle entry point on library l."""
    P = ctypes.c_void_p
    l.or_wanghash.argtypes, l.or_wanghash.restype = [ctypes.c_uint32], ctypes.c_uint32
    l.or_xorshift32.argtypes, l.or_xorshift32.restype = [ctypes.POINTER(ctypes.c_uint32)], ctypes.c_uint32
    l.or_random01.argtypes, l.or_random01.restype = [ctypes.POINTER(ctypes.c_uint32)], ctypes.c_float
    l.or_xorshift64star.argtypes, l.or_xorshift64star.restype = [ctypes.c_uint64], ctypes.c_uint64
    l.or_xorshiftf.argtypes, l.or_xorshiftf.restype = [ctypes.POINTER(ctypes.c_uint64)], ctypes.c_float
    for f in ("or_sinf", "or_cosf", "or_log2f", "or_exp2f", "or_atanf"):
        getattr(l, f).argtypes, getattr(l, f).restype = [ctypes.c_float], ctypes.c_float
    for f in ("or_powf", "or_atan2f"):
        getattr(l, f).argtypes, getattr(l, f).restype = [ctypes.c_float, ctypes.c_float], ctypes.c_float
    l.or_trace.argtypes = [P, ctypes.c_int, P, P, P, P, ctypes.c_size_t]
    l.or_trace.restype = None
    l.or_bruteforce.argtypes = [P, ctypes.c_int, P, P, ctypes.c_size_t]
    l.or_bruteforce.restype = None
    l.or_trace_mode.argtypes = [P, ctypes.c_int, ctypes.c_int, P, P, P, P, ctypes.c_size_t]
    l.or_trace_mode.restype = None
    l.or_bruteforce_mode.argtypes = [P, ctypes.c_int, ctypes.c_int, P, P, ctypes.c_size_t]
    l.or_bruteforce_mode.restype = None
    l.or_classify.argtypes = [P, P, ctypes.c_int32, ctypes.c_int32, ctypes.c_double]
    l.or_classify.restype = ctypes.c_int
    l.or_raygen_primary.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, P, ctypes.c_size_t,
                                    ctypes.c_size_t]
    l.or_raygen_primary.restype = None
    l.or_raygen_diffuse.argtypes = [P, P, P, ctypes.c_size_t, ctypes.c_uint32, P, ctypes.c_size_t, ctypes.c_size_t]
    l.or_raygen_diffuse.restype = None
    l.or_raygen_shadow.argtypes = [P, P, ctypes.c_size_t, P, ctypes.c_uint32, P, ctypes.c_size_t, ctypes.c_size_t]
    l.or_raygen_shadow.restype = None
    l.or_bsdf_samples.argtypes = [P, P, P, ctypes.c_uint32, ctypes.c_int, ctypes.c_size_t, P]
    l.or_bsdf_samples.restype = None
    l.or_bsdf_evals.argtypes = [P, P, P, ctypes.c_int, P, ctypes.c_size_t, P]
    l.or_bsdf_evals.restype = None
    l.or_set_environment.argtypes, l.or_set_environment.restype = [P, ctypes.c_int], None
    l.or_trace_importance.argtypes = [P, ctypes.c_int, P, P, P, ctypes.c_size_t, P, ctypes.c_float,
                                      ctypes.c_float]
    l.or_trace_importance.restype = None
    l.or_set_textures.argtypes, l.or_set_textures.restype = [P, ctypes.c_uint32, P, P], None
    l.or_texture_samples.argtypes, l.or_texture_samples.restype = [ctypes.c_int, P, ctypes.c_size_t, P], None
    l.or_env_lookups.argtypes, l.or_env_lookups.restype = [P, ctypes.c_int, P, ctypes.c_size_t, P], None
    l.or_pathtracer.argtypes = [P, P, ctypes.c_uint32, P, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                ctypes.c_int, P, ctypes.c_int, ctypes.c_int]
    l.or_pathtracer.restype = None
    l.or_lighttracer.argtypes = [P, P, ctypes.c_uint32, P, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                 ctypes.c_int, P]
    l.or_lighttracer.restype = None
    l.or_bpt_warmup.argtypes = [P, P, ctypes.c_uint32, P, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                ctypes.c_int, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
    l.or_bpt_warmup.restype = None
    l.or_bpt_light_pass.argtypes = [P, P, ctypes.c_uint32, P, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                    ctypes.c_int, ctypes.c_int, P, P]
    l.or_bpt_light_pass.restype = ctypes.c_uint32
    l.or_bpt_eye.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, P,
                             ctypes.c_uint32, P, ctypes.c_int, ctypes.c_int]
    l.or_bpt_eye.restype = None
    l.or_set_path_options.argtypes, l.or_set_path_options.restype = [ctypes.c_int, ctypes.c_int], None
    l.or_count_hits_rays.argtypes, l.or_count_hits_rays.restype = [P, P, P, ctypes.c_size_t, P], None
    l.or_lighttracer_paths.argtypes = [P, P, ctypes.c_uint32, P, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                       ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, P]
    l.or_lighttracer_paths.restype = None
    l.or_showlightcaches.argtypes = [P, P, ctypes.c_uint32, P, ctypes.c_int, ctypes.c_int, P]
    l.or_showlightcaches.restype = None
    l.or_bpt_light_paths.argtypes = [P, P, ctypes.c_uint32, P, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, P, P, P, P]
    l.or_bpt_light_paths.restype = None
    l.or_slot_hits.argtypes, l.or_slot_hits.restype = [P, P, ctypes.c_size_t, P], None
    l.or_hier_parents.argtypes, l.or_hier_parents.restype = [P, P], None
    l.or_hier_acquire.argtypes = [P, P, ctypes.c_uint32, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
    l.or_hier_acquire.restype = None
    l.or_hier_propagate.argtypes, l.or_hier_propagate.restype = [P, ctypes.c_int, P, P], None
    l.or_hier_materials.argtypes, l.or_hier_materials.restype = [P, P], None
    l.or_half_bits.argtypes, l.or_half_bits.restype = [ctypes.c_float], ctypes.c_uint16
    l.or_half_value.argtypes, l.or_half_value.restype = [ctypes.c_uint16], ctypes.c_float
    l.or_screen_pixel.argtypes = [P, ctypes.c_int, ctypes.c_int, P, ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_int)]
    l.or_screen_pixel.restype = ctypes.c_int
    return l


@contextlib.contextmanager
def path_options(specific_path_length=-1, rr_start_depth=0):
    """Every integrator in the body keeps only paths of specific_path_length
    bounces (>= 0; the reference's SHOW_SPECIFIC_PATHLENGTH) and runs Russian
    roulette from bounce rr_start_depth on (0: every bounce, the reference)."""
    lib().or_set_path_options(int(specific_path_length), int(rr_start_depth))
    try:
        yield
    finally:
        lib().or_set_path_options(-1, 0)


def screen_pixel(cam, w, h, point):
    """The pixel ProjectToScreen splats a world point into, or None outside the screen."""
    c = camera(cam)
    p = np.ascontiguousarray(point, dtype=np.float32).reshape(3)
    ix, iy = ctypes.c_int(), ctypes.c_int()
    r = lib().or_screen_pixel(ctypes.byref(c), w, h, _p(p), ctypes.byref(ix), ctypes.byref(iy))
    return None if r != 0 else (ix.value, iy.value)


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


class Scene:
    """Holds numpy copies of the reference-layout arrays and the C view of them."""

    def __init__(self, arrays):
        self.a = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in arrays.items()}
        mats = self.a["materials"]
        self.mat_floats = np.ascontiguousarray(mats.view(np.float32).reshape(len(mats), -1))
        assert self.mat_floats.shape[1] == 20
        obox = self.a.get("obox")  # optional TreeNode<OBox> records (num_nodes x 12 float32)
        if obox is not None:
            assert obox.dtype == np.float32 and obox.shape == (len(self.a["hierarchy"]), 12)
        self.c = _Scene(_p(self.a["hierarchy"]), len(self.a["hierarchy"]), _p(self.a["triangles"]),
                        len(self.a["triangles"]) // 2, _p(self.a["positions"]), _p(self.a["vertex_info"]),
                        len(self.a["positions"]), _p(self.mat_floats), len(mats),
                        _p(obox) if obox is not None else None)

    @property
    def ptr(self):
        return ctypes.byref(self.c)


def camera(cam):
    c = _Camera()
    for k in ("position", "u", "v", "w"):
        getattr(c, k)[:] = list(getattr(cam, k))
    c.pixel_area = cam.pixel_area
    return c


def _rays_f32(rays):
    return np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)


def trace(scene, rays, any_hit=False, counts=False, watertight=False):
    r = _rays_f32(rays)
    n = len(r)
    hits = np.zeros((n, 4), dtype=np.int32)
    nb = np.zeros(n, dtype=np.uint32) if counts else None
    nt = np.zeros(n, dtype=np.uint32) if counts else None
    lib().or_trace_mode(scene.ptr, int(any_hit), int(watertight), _p(r), _p(hits), _p(nb) if counts else None,
                        _p(nt) if counts else None, n)
    return (hits, nb, nt) if counts else hits


def trace_importance(scene, rays, importance, threshold, ray_threshold, any_hit=False):
    """The TRACERAY_IMPORTANCE_BREAK walk: returns (hits n x 4, nodes n x 4 {node,
    importance bits, size^2 bits, 0}) for closest hit, hits for any hit."""
    r = _rays_f32(rays)
    n = len(r)
    hits = np.zeros((n, 4), dtype=np.int32)
    nodes = np.zeros((n, 4), dtype=np.int32)
    imp = np.ascontiguousarray(importance, dtype=np.float32).reshape(-1, 2)
    assert len(imp) == scene.c.num_nodes
    lib().or_trace_importance(scene.ptr, int(any_hit), _p(r), _p(hits), _p(nodes), n, _p(imp), float(threshold),
                              float(ray_threshold))
    return hits if any_hit else (hits, nodes)


def bruteforce(scene, rays, any_hit=False, watertight=False):
    r = _rays_f32(rays)
    hits = np.zeros((len(r), 4), dtype=np.int32)
    lib().or_bruteforce_mode(scene.ptr, int(any_hit), int(watertight), _p(r), _p(hits), len(r))
    return hits


def classify(scene, rays, slots_a, slots_b, rel_eps=1e-6):
    r = _rays_f32(rays)
    out = np.zeros(len(r), dtype=np.int32)
    for i in range(len(r)):
        row = np.ascontiguousarray(r[i])
        out[i] = lib().or_classify(scene.ptr, _p(row), int(slots_a[i]), int(slots_b[i]), rel_eps)
    return out


def raygen_primary(cam, w, h, seed, n, first=0):
    out = np.zeros((n, 8), dtype=np.float32)
    c = camera(cam)
    lib().or_raygen_primary(ctypes.byref(c), w, h, seed, _p(out), first, n)
    return out


def raygen_diffuse(scene, prim, prim_hits, seed, n, first=0):
    out = np.zeros((n, 8), dtype=np.float32)
    p = _rays_f32(prim)
    h = np.ascontiguousarray(prim_hits, dtype=np.int32).reshape(-1, 4)
    lib().or_raygen_diffuse(scene.ptr, _p(p), _p(h), len(p), seed, _p(out), first, n)
    return out


def raygen_shadow(prim, prim_hits, lights, n, first=0):
    out = np.zeros((n, 8), dtype=np.float32)
    p = _rays_f32(prim)
    h = np.ascontiguousarray(prim_hits, dtype=np.int32).reshape(-1, 4)
    l = np.ascontiguousarray(lights, dtype=np.float32).reshape(-1, 3)
    lib().or_raygen_shadow(_p(p), _p(h), len(p), _p(l), len(l), _p(out), first, n)
    return out


def _mat20(material):
    m = np.ascontiguousarray(material, dtype=np.float32).reshape(-1)
    assert m.size == 20
    return m


def bsdf_samples(material, inc, normal, seed, n, adjoint=False):
    """n draws of __SampleBSDF (materials.glsl:118-277): returns (dirs n x 3,
    throughput n x 3, pdf n, diffuse-branch flag n)."""
    out = np.zeros((n, 8), np.float32)
    m = _mat20(material)
    i = np.ascontiguousarray(inc, np.float32)
    nn = np.ascontiguousarray(normal, np.float32)
    lib().or_bsdf_samples(_p(m), _p(i), _p(nn), seed, int(adjoint), n, _p(out))
    return out[:, 0:3], out[:, 3:6], out[:, 6], out[:, 7] > 0.5


def bsdf_evals(material, inc, normal, exc, adjoint=False):
    """__BSDF (materials.glsl:306-391) at the excident directions exc (n x 3):
    returns (bsdf n x 3, pdf n)."""
    e = np.ascontiguousarray(exc, np.float32).reshape(-1, 3)
    out = np.zeros((len(e), 4), np.float32)
    m = _mat20(material)
    i = np.ascontiguousarray(inc, np.float32)
    nn = np.ascontiguousarray(normal, np.float32)
    lib().or_bsdf_evals(_p(m), _p(i), _p(nn), int(adjoint), _p(e), len(e), _p(out))
    return out[:, 0:3], out[:, 3]


_env_keep = None


def set_environment(texels):
    """The path tracer's environment map: texels (6, size, size, 4) float32
    linear RGB (faces x-, x+, y-, y+, z-, z+), or None for the black default."""
    global _env_keep
    if texels is None:
        lib().or_set_environment(None, 0)
        _env_keep = None
        return
    t = np.ascontiguousarray(texels, dtype=np.float32)
    assert t.ndim == 4 and t.shape[0] == 6 and t.shape[1] == t.shape[2] and t.shape[3] == 4
    _env_keep = t
    lib().or_set_environment(_p(t), t.shape[1])


_tex_keep = None


def set_textures(mat_tex, textures):
    """Material textures of every integrator: mat_tex (materials, 4) texture
    indices (diffuse, opacity, reflectiveness, emissivity; -1: constant) and
    textures, a list of (h, w, 4) float32 linear RGBA arrays; None clears."""
    global _tex_keep
    if mat_tex is None:
        lib().or_set_textures(None, 0, None, None)
        _tex_keep = None
        return
    mt = np.ascontiguousarray(mat_tex, dtype=np.int32).reshape(-1, 4)
    texs = [np.ascontiguousarray(t, dtype=np.float32) for t in textures]
    ptrs = (ctypes.c_void_p * max(1, len(texs)))(*[t.ctypes.data for t in texs])
    sizes = np.array([[t.shape[1], t.shape[0]] for t in texs] or [[0, 0]], np.int32)
    _tex_keep = (mt, texs, ptrs, sizes)
    lib().or_set_textures(_p(mt), len(mt), ctypes.cast(ptrs, ctypes.c_void_p), _p(sizes))


def texture_samples(tex_id, uvs):
    """The material-texture lookup of texture tex_id (of the last set_textures) at uvs (n x 2)."""
    uv = np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 2)
    out = np.zeros((len(uv), 4), np.float32)
    lib().or_texture_samples(int(tex_id), _p(uv), len(uv), _p(out))
    return out


def env_lookups(texels, dirs):
    """The cube-map lookup of the path tracer's miss for directions dirs (n x 3)."""
    t = np.ascontiguousarray(texels, dtype=np.float32)
    d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
    out = np.zeros((len(d), 3), np.float32)
    lib().or_env_lookups(_p(t), t.shape[1], _p(d), len(d), _p(out))
    return out


def srgb_decode(faces):
    """RGB8 sRGB faces (6, size, size, 3) -> linear float texels (6, size, size, 4),
    the sRGB transfer function in double rounded to float per byte value."""
    c = np.arange(256, dtype=np.float64) / 255.0
    lut = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4).astype(np.float32)
    f = np.asarray(faces, dtype=np.uint8)
    out = np.ones(f.shape[:3] + (4,), np.float32)
    out[..., :3] = lut[f]
    return out


def pathtrace(scene, lights, cam, w, h, frame_seed, max_path_length, accum=None, rows=None):
    accum = np.zeros((h, w, 4), dtype=np.float32) if accum is None else accum
    lt = np.ascontiguousarray(lights).view(np.float32).reshape(-1, 8)
    c = camera(cam)
    r0, r1 = rows if rows else (0, h)
    lib().or_pathtracer(scene.ptr, _p(lt), len(lt), ctypes.byref(c), w, h, frame_seed, max_path_length,
                        _p(accum), r0, r1)
    return accum


def _lights(lights):
    return np.ascontiguousarray(lights).view(np.float32).reshape(-1, 8)


def lighttrace(scene, lights, cam, w, h, frame_seed, max_path_length, accum=None):
    accum = np.zeros((h, w, 4), dtype=np.float32) if accum is None else accum
    lt = _lights(lights)
    c = camera(cam)
    lib().or_lighttracer(scene.ptr, _p(lt), len(lt), ctypes.byref(c), w, h, frame_seed, max_path_length, _p(accum))
    return accum


def _split(n, workers):
    step = max(1, (n + workers - 1) // workers)
    return [(a, min(n, a + step)) for a in range(0, n, step)]


def _rays_per_sample(w, h, num_lights):
    r = (w * h // num_lights // 64) * 64  # NumRaysPerLightSample (lightpathtracer.cpp:29-31)
    return max(64, r)


def lighttrace_parallel(scene, lights, cam, w, h, frame_seed, max_path_length, workers):
    """The light tracer draw with its light paths split over `workers` threads, each
    into its own buffer (float64 sum of the parts; the splat order differs from a
    sequential draw only in float rounding)."""
    import concurrent.futures
    lt = _lights(lights)
    c = camera(cam)
    n = len(lt) * _rays_per_sample(w, h, len(lt))

    def part(rng):
        acc = np.zeros((h, w, 4), np.float32)
        lib().or_lighttracer_paths(scene.ptr, _p(lt), len(lt), ctypes.byref(c), w, h, frame_seed, max_path_length,
                                   rng[0], rng[1], _p(acc))
        return acc.astype(np.float64)

    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
        total = sum(ex.map(part, _split(n, workers)))
    splat = np.zeros((h, w, 4), np.float32)
    lib().or_showlightcaches(scene.ptr, _p(lt), len(lt), ctypes.byref(c), w, h, _p(splat))
    return total + splat


def bpt_draw_parallel(scene, lights, cam, w, h, frame_seed, max_path_length, capacity, workers):
    """One BPT draw (light pass, then eye pass) with the light paths and the eye
    rows split over `workers` threads.  The cache equals the sequential draw's
    bit for bit (bounce vertices appended in path order); the image equals it up
    to the order of the light splats.  Returns (accum float64, cache, NumLightCacheEntries)."""
    import concurrent.futures
    lt = _lights(lights)
    c = camera(cam)
    R = _rays_per_sample(w, h, len(lt))
    n = len(lt) * R
    cache = np.zeros((capacity, 16), np.float32)

    def part(rng):
        acc = np.zeros((h, w, 4), np.float32)
        staged = np.zeros((rng[1] - rng[0], max_path_length, 16), np.float32)
        counts = np.zeros(rng[1] - rng[0], np.uint32)
        lib().or_bpt_light_paths(scene.ptr, _p(lt), len(lt), ctypes.byref(c), w, h, frame_seed, max_path_length,
                                 capacity, rng[0], rng[1], _p(acc), _p(cache), _p(staged), _p(counts))
        return acc.astype(np.float64), staged, counts

    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
        parts = list(ex.map(part, _split(n, workers)))
    accum = sum(p[0] for p in parts)
    nxt = n
    for _, staged, counts in parts:
        for j, k in enumerate(counts):
            for b in range(int(k)):
                if nxt < capacity:
                    cache[nxt] = staged[j, b]
                nxt += 1
    eye = np.zeros((h, w, 4), np.float32)

    def rows(rng):
        lib().or_bpt_eye(scene.ptr, ctypes.byref(c), w, h, frame_seed, max_path_length, capacity, _p(cache), nxt,
                         _p(eye), rng[0], rng[1])

    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(rows, _split(h, workers)))
    return accum + eye, cache, nxt


def bpt_warmup(scene, lights, cam, w, h, frame_seed, max_path_length):
    lt = _lights(lights)
    c = camera(cam)
    a, b = ctypes.c_uint64(), ctypes.c_uint64()
    lib().or_bpt_warmup(scene.ptr, _p(lt), len(lt), ctypes.byref(c), w, h, frame_seed, max_path_length,
                        ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def bpt_draw(scene, lights, cam, w, h, frame_seed, max_path_length, capacity, accum=None):
    """One BPT draw: light pass then eye pass.  Returns (accum, cache, NumLightCacheEntries)."""
    accum = np.zeros((h, w, 4), dtype=np.float32) if accum is None else accum
    lt = _lights(lights)
    c = camera(cam)
    cache = np.zeros((capacity, 16), dtype=np.float32)
    n = lib().or_bpt_light_pass(scene.ptr, _p(lt), len(lt), ctypes.byref(c), w, h, frame_seed, max_path_length,
                                capacity, _p(accum), _p(cache))
    lib().or_bpt_eye(scene.ptr, ctypes.byref(c), w, h, frame_seed, max_path_length, capacity, _p(cache), n,
                     _p(accum), 0, h)
    return accum, cache, int(n)


def count_hits(scene, rays, w):
    """TraceRayCountHits of every ray (n x 8 {o, length, d, pad}) with weight w[i]:
    returns the importance array ((nodes + slots) x 2, only .x written)."""
    r = _rays_f32(rays)
    w = np.ascontiguousarray(w, dtype=np.float32).reshape(-1)
    assert len(w) == len(r)
    out = np.zeros((scene.c.num_nodes + 2 * scene.c.num_leaves, 2), np.float32)
    lib().or_count_hits_rays(scene.ptr, _p(r), _p(w), len(r), _p(out))
    return out


def slot_hits(scene, rays):
    """(n, slots) bool: IntersectTriangle accepts the valid slot at t <= the ray's length."""
    r = _rays_f32(rays)
    out = np.zeros((len(r), 2 * scene.c.num_leaves), np.uint8)
    lib().or_slot_hits(scene.ptr, _p(r), len(r), _p(out))
    return out.astype(bool)


def hier_parents(scene):
    out = np.zeros(scene.c.num_nodes, np.int32)
    lib().or_hier_parents(scene.ptr, _p(out))
    return out


def hier_acquire(scene, lights, cam, w, h, iteration, importance):
    """One acquisition2.comp pass (iteration is NumImportanceIterations, 1-based)
    adding into importance ((nodes + slots) x 2 float32, in place)."""
    assert importance.dtype == np.float32 and importance.flags.c_contiguous
    assert importance.shape == (scene.c.num_nodes + 2 * scene.c.num_leaves, 2)
    lt = _lights(lights)
    c = camera(cam)
    lib().or_hier_acquire(scene.ptr, _p(lt), len(lt), ctypes.byref(c), w, h, iteration, _p(importance))
    return importance


def hier_propagate(scene, iterations, importance):
    """The propagation passes: returns (importance {x / iterations, y}, subtree importance)."""
    imp = np.ascontiguousarray(importance, dtype=np.float32).copy()
    sub = np.zeros(scene.c.num_nodes, np.float32)
    lib().or_hier_propagate(scene.ptr, int(iterations), _p(imp), _p(sub))
    return imp, sub


def hier_materials(scene):
    """The node materials: (nodes, 12) uint32, six pairs of packed halves per node."""
    out = np.zeros((scene.c.num_nodes, 12), np.uint32)
    lib().or_hier_materials(scene.ptr, _p(out))
    return out


def half_bits(x):
    return int(lib().or_half_bits(float(x)))


def half_value(bits):
    return float(lib().or_half_value(int(bits)))


def image_mse(accum, divisor, ref):
    """TextureMSE::ComputeCurrentMSE (imageut/texturemse.cpp:39-66) with the
    per-pixel error of shader/utils/sqerror.frag:13-17.

    accum: (h, w, 4) float32 accumulation buffer; ref: (rh, rw, 3) float32.
    Every reference pixel samples the render nearest at its centre (texcoord
    (x + 0.5) / rw, the sampler the reference documents, texturemse.hpp:14-16),
    diff = render.rgb / divisor - ref.rgb in float32, error = dot(diff, diff)
    summed in the GLSL order x, y, z; the result is the mean over the reference's
    pixels (taken in float64; the reference averages a float mip chain).
    """
    accum = np.asarray(accum, dtype=np.float32)
    ref = np.asarray(ref, dtype=np.float32)
    h, w = accum.shape[:2]
    rh, rw = ref.shape[:2]
    xs = ((2 * np.arange(rw, dtype=np.int64) + 1) * w) // (2 * rw)
    ys = ((2 * np.arange(rh, dtype=np.int64) + 1) * h) // (2 * rh)
    c = accum[ys[:, None], xs[None, :], :3]
    d = c / np.float32(divisor) - ref
    e = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    return np.float32(e.astype(np.float64).sum() / e.size)


# ---- reference RNG built from the reference's own utilities/random.cpp ----
_ref = None


def ref_lib():
    """The reference RNG library (oracle/_ref), or None where it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        l = ctypes.CDLL(REF_LIB)
        l._Z8WangHashj.argtypes, l._Z8WangHashj.restype = [ctypes.c_uint32], ctypes.c_uint32
        l._Z8Xorshiftj.argtypes, l._Z8Xorshiftj.restype = [ctypes.c_uint32], ctypes.c_uint32
        l._Z8Xorshiftm.argtypes, l._Z8Xorshiftm.restype = [ctypes.c_uint64], ctypes.c_uint64
        l._Z9XorshiftFRm.argtypes, l._Z9XorshiftFRm.restype = [ctypes.POINTER(ctypes.c_uint64)], ctypes.c_float
        _ref = l
    return _ref
