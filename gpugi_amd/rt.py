"""ctypes binding of the C ABI in include/gpugi_rt.h.

This is glue for tests and the benchmark driver; the renderer itself is the
native library.  Loading fails loudly when libgpugi_rt.so has not been built:
there is no Python or CPU fallback for any entry point.
"""
import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgpugi_rt.so")

GPUGI_OK = 0
GPUGI_ERR_ARG = -1  # gpugi_status in gpugi_rt.h
GPUGI_ERR_STATE = -3
CLOSEST_HIT, ANY_HIT = 0, 1
KERNEL_AUTO, KERNEL_REFERENCE, KERNEL_WIDE = 0, 1, 2
PT_MEGAKERNEL, PT_WAVEFRONT = 0, 1
PT, LT, BPT, PT_CACHE = 0, 1, 2, 3
TRIANGLE_REFERENCE, TRIANGLE_WATERTIGHT = 0, 1
CACHE_SPREAD, CACHE_PRIMARY_INDIRECT = 0, 1
TEXTURE_RGBA8_SRGB, TEXTURE_RGBA32F = 0, 1
BVH_AABOX, BVH_OBOX = 0, 1
TEX_DIFFUSE, TEX_OPACITY, TEX_REFLECTIVENESS, TEX_EMISSIVITY = 0, 1, 2, 3
RAY_MAX = np.float32(3.402823466e38)

RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("tmax", "<f4"), ("dir", "<f4", 3), ("pad", "<u4")])
HIT_DTYPE = np.dtype([("slot", "<i4"), ("t", "<f4"), ("bary_y", "<f4"), ("bary_z", "<f4")])
NODE_HIT_DTYPE = np.dtype([("node", "<i4"), ("node_importance", "<f4"), ("node_size_sq", "<f4"), ("pad", "<i4")])
LIGHT_DTYPE = np.dtype([("position", "<f4", 3), ("normal_phi", "<f4"), ("intensity", "<f4", 3),
                        ("normal_theta_cos", "<f4")])
LIGHT_CACHE_DTYPE = np.dtype([("flux", "<f4", 3), ("material", "<i4"), ("position", "<f4", 3), ("normal0", "<f4"),
                              ("texcoord", "<f4", 2), ("normal1", "<f4"), ("path_length", "<i4"),
                              ("incident", "<f4", 2), ("mis_n", "<f4"), ("mis_t", "<f4")])
MATERIAL_DTYPE = np.dtype([("diffuse", "<f4", 3), ("opacity", "<f4", 3), ("reflectiveness", "<f4", 4),
                           ("emissivity", "<f4", 3), ("fresnel0", "<f4", 3), ("refraction_index_avg", "<f4"),
                           ("fresnel1", "<f4", 3)])


class SceneDesc(ctypes.Structure):
    _fields_ = [("hierarchy", ctypes.c_void_p), ("num_nodes", ctypes.c_uint32),
                ("triangles", ctypes.c_void_p), ("num_leaves", ctypes.c_uint32),
                ("triangles_per_leaf", ctypes.c_uint32), ("positions", ctypes.c_void_p),
                ("vertex_info", ctypes.c_void_p), ("num_vertices", ctypes.c_uint32),
                ("materials", ctypes.c_void_p), ("num_materials", ctypes.c_uint32)]


class Camera(ctypes.Structure):
    _fields_ = [("position", ctypes.c_float * 3), ("u", ctypes.c_float * 3), ("v", ctypes.c_float * 3),
                ("w", ctypes.c_float * 3), ("pixel_area", ctypes.c_float)]

    def as_dict(self):
        return {k: list(getattr(self, k)) if k != "pixel_area" else self.pixel_area
                for k, _ in self._fields_}


class MaterialParams(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("albedo", ctypes.c_float * 3), ("specular_color", ctypes.c_float * 3),
                ("roughness", ctypes.c_float), ("opacity", ctypes.c_float), ("emissivity", ctypes.c_float * 3),
                ("refraction_index", ctypes.c_float), ("reflectiveness", ctypes.c_float * 4),
                ("has_reflectiveness", ctypes.c_int32)]


class GpugiError(RuntimeError):
    def __init__(self, func, code, msg):
        super().__init__("%s failed with status %d: %s" % (func, code, msg))
        self.code = code


# Exported symbols with their argument types (restype int unless stated).
_P = ctypes.c_void_p
_SIGNATURES = {
    "gpugi_create": [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)],
    "gpugi_destroy": ([_P], None),
    "gpugi_last_error": ([_P], ctypes.c_char_p),
    "gpugi_load_scene": [_P, ctypes.POINTER(SceneDesc)],
    "gpugi_set_camera": [_P, ctypes.POINTER(Camera)],
    "gpugi_set_screen": [_P, ctypes.c_int, ctypes.c_int],
    "gpugi_set_light_samples": [_P, _P, ctypes.c_uint32],
    "gpugi_set_max_path_length": [_P, ctypes.c_int],
    "gpugi_set_integrator_kernel": [_P, ctypes.c_int],
    "gpugi_set_pt_schedule": [_P, ctypes.c_int],
    "gpugi_set_rr_start_depth": [_P, ctypes.c_int],
    "gpugi_capture_draw": [_P, ctypes.c_int, ctypes.c_uint32],
    "gpugi_hierarchy_update_importance": [_P, ctypes.c_int],
    "gpugi_hierarchy_compute_materials": [_P],
    "gpugi_hierarchy_read": [_P, _P, ctypes.c_size_t, _P, ctypes.c_size_t, _P, ctypes.c_size_t],
    "gpugi_replay_draw": [_P],
    "gpugi_set_specific_path_length": [_P, ctypes.c_int],
    "gpugi_set_triangle_test": [_P, ctypes.c_int],
    "gpugi_draw": [_P, ctypes.c_int, ctypes.c_uint32],
    "gpugi_bpt_warmup": [_P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)],
    "gpugi_bpt_set_light_cache": [_P, ctypes.c_int32, ctypes.c_int32],
    "gpugi_bpt_read_light_cache": [_P, _P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)],
    "gpugi_reset_accum": [_P],
    "gpugi_readback": [_P, _P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint32)],
    "gpugi_set_reference_image": [_P, _P, ctypes.c_int, ctypes.c_int],
    "gpugi_set_environment_map": [_P, ctypes.c_int, _P],
    "gpugi_add_texture": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, ctypes.POINTER(ctypes.c_int)],
    "gpugi_set_material_texture": [_P, ctypes.c_uint32, ctypes.c_int, ctypes.c_int],
    "gpugi_compute_mse": [_P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float)],
    "gpugi_trace": [_P, ctypes.c_int, ctypes.c_int, _P, _P, ctypes.c_size_t, ctypes.c_int],
    "gpugi_probe_valu": [_P, ctypes.c_uint32, ctypes.c_uint32, _P, ctypes.POINTER(ctypes.c_float)],
    "gpugi_trace_noalloc": [_P, ctypes.c_int, ctypes.c_int, _P, _P, ctypes.c_size_t],
    "gpugi_trace_counts": [_P, ctypes.c_int, _P, _P, _P, ctypes.c_size_t, ctypes.c_int],
    "gpugi_set_wide_config": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "gpugi_get_wide_config": [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(ctypes.c_uint32)],
    "gpugi_autotune_trace": [_P, ctypes.c_int, _P, _P, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_float)],
    "gpugi_sync": [_P],
    "gpugi_alloc": [_P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)],
    "gpugi_free": [_P, _P],
    "gpugi_upload": [_P, _P, _P, ctypes.c_size_t],
    "gpugi_download": [_P, _P, _P, ctypes.c_size_t],
    "gpugi_copy_device": [_P, _P, _P, ctypes.c_size_t],
    "gpugi_timer_begin": [_P],
    "gpugi_timer_end": [_P, ctypes.POINTER(ctypes.c_float)],
    "gpugi_raygen_primary": [_P, ctypes.POINTER(Camera), ctypes.c_int, ctypes.c_int, ctypes.c_uint32, _P,
                             ctypes.c_size_t],
    "gpugi_raygen_diffuse": [_P, _P, _P, ctypes.c_size_t, ctypes.c_uint32, _P, ctypes.c_size_t],
    "gpugi_raygen_shadow": [_P, _P, _P, ctypes.c_size_t, _P, ctypes.c_uint32, _P, ctypes.c_size_t],
    "gpugi_hostscene_build": [_P, _P, _P, ctypes.c_uint32, _P, _P, ctypes.c_uint32, _P, ctypes.c_uint32,
                              ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)],
    "gpugi_hostscene_cornell": [ctypes.POINTER(ctypes.c_void_p)],
    "gpugi_hostscene_heightfield": [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)],
    "gpugi_hostscene_clutter": [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)],
    "gpugi_hostscene_load": [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)],
    "gpugi_hostscene_save": [_P, ctypes.c_char_p, ctypes.c_char_p],
    "gpugi_hostscene_desc": [_P, ctypes.POINTER(SceneDesc)],
    "gpugi_hostscene_num_textures": [_P, ctypes.POINTER(ctypes.c_uint32)],
    "gpugi_hostscene_obox_hierarchy": [_P, _P, ctypes.c_size_t],
    "gpugi_set_bvh_type": [_P, ctypes.c_int],
    "gpugi_host_alloc": [_P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)],
    "gpugi_host_free": [_P, _P],
    "gpugi_trace_importance_break": [_P, ctypes.c_int, _P, _P, _P, ctypes.c_size_t, _P, ctypes.c_uint32,
                                     ctypes.c_float, ctypes.c_float],
    "gpugi_hostscene_texture": [_P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_void_p)],
    "gpugi_hostscene_material_textures": [_P, ctypes.c_uint32, _P],
    "gpugi_hostscene_info": [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                             ctypes.POINTER(ctypes.c_float * 3), ctypes.POINTER(ctypes.c_float * 3)],
    "gpugi_hostscene_add_point_light": [_P, ctypes.POINTER(ctypes.c_float * 3), ctypes.POINTER(ctypes.c_float * 3)],
    "gpugi_hostscene_destroy": ([_P], None),
    "gpugi_camera_compute": [ctypes.POINTER(ctypes.c_float * 3), ctypes.POINTER(ctypes.c_float * 3),
                             ctypes.POINTER(ctypes.c_float * 3), ctypes.c_float, ctypes.c_int, ctypes.c_int,
                             ctypes.POINTER(Camera)],
    "gpugi_lightsampler_create": [_P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)],
    "gpugi_lightsampler_generate": [_P, _P, ctypes.c_uint32, ctypes.c_float],
    "gpugi_lightsampler_set_scene": [_P, _P],
    "gpugi_lightsampler_destroy": ([_P], None),
    "gpugi_wanghash": ([ctypes.c_uint32], ctypes.c_uint32),
    "gpugi_batch_create": [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)],
    "gpugi_batch_destroy": ([_P], None),
    "gpugi_batch_last_error": ([_P], ctypes.c_char_p),
    "gpugi_batch_add": [_P, ctypes.POINTER(SceneDesc), ctypes.POINTER(Camera), ctypes.c_int, ctypes.c_int, _P,
                        ctypes.c_uint32, ctypes.c_int],
    "gpugi_batch_finalize": [_P],
    "gpugi_batch_config": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "gpugi_batch_reset": [_P, ctypes.c_uint32],
    "gpugi_batch_draw": [_P, ctypes.c_int],
    "gpugi_batch_capture": [_P],
    "gpugi_batch_replay": [_P, ctypes.c_int],
    "gpugi_batch_run": [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)],
    "gpugi_batch_readback": [_P, ctypes.c_int, _P, ctypes.c_size_t],
    "gpugi_batch_rays": [_P, ctypes.POINTER(ctypes.c_uint64)],
    "gpugi_batch_autotune": [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                             ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                             ctypes.POINTER(ctypes.c_float)],
    "gpugi_cache_config": [_P, ctypes.c_float, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                           ctypes.c_float],
    "gpugi_cache_stats": [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_float)],
    "gpugi_cache_set_query": [_P, ctypes.c_int],
    "gpugi_mlp_init": [_P, ctypes.c_uint64],
    "gpugi_mlp_set_weights": [_P, _P, ctypes.c_size_t],
    "gpugi_mlp_get_weights": [_P, _P, ctypes.c_size_t],
    "gpugi_mlp_encode": [_P, _P, _P, ctypes.c_size_t],
    "gpugi_mlp_infer": [_P, _P, _P, ctypes.c_size_t, ctypes.c_int],
    "gpugi_mlp_gradient": [_P, _P, _P, ctypes.c_size_t, _P, ctypes.POINTER(ctypes.c_float), ctypes.c_int],
    "gpugi_mlp_train_step": [_P, _P, _P, ctypes.c_size_t, ctypes.c_float, ctypes.POINTER(ctypes.c_float),
                             ctypes.c_int],
}

MLP_PARAMS = 19456
MLP_FEATURES = 16
MLP_INPUTS = 96

_lib = None


def lib():
    """Loads libgpugi_rt.so (built by __graft_entry__.build()); raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgpugi_rt.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        l = ctypes.CDLL(LIB_PATH)
        for name, sig in _SIGNATURES.items():
            f = getattr(l, name)
            if isinstance(sig, tuple):
                f.argtypes, f.restype = sig
            else:
                f.argtypes, f.restype = sig, ctypes.c_int
        _lib = l
    return _lib


def _check(func, code, ctx=None):
    if code != GPUGI_OK:
        msg = lib().gpugi_last_error(ctx)
        raise GpugiError(func, code, msg.decode() if msg else "")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _v3(x):
    return (ctypes.c_float * 3)(*[float(v) for v in x])


def wanghash(x):
    return int(lib().gpugi_wanghash(ctypes.c_uint32(x & 0xffffffff)))


def camera(position, lookat, up, vfov_deg, width, height):
    cam = Camera()
    _check("gpugi_camera_compute", lib().gpugi_camera_compute(_v3(position), _v3(lookat), _v3(up),
                                                              float(vfov_deg), int(width), int(height),
                                                              ctypes.byref(cam)))
    return cam


class HostScene:
    """A host-side scene (.rawscene content + materials + derived GPU layout)."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)

    @classmethod
    def cornell(cls):
        h = ctypes.c_void_p()
        _check("gpugi_hostscene_cornell", lib().gpugi_hostscene_cornell(ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def heightfield(cls, cells, seed):
        h = ctypes.c_void_p()
        _check("gpugi_hostscene_heightfield", lib().gpugi_hostscene_heightfield(cells, seed, ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def clutter(cls, spheres, seed):
        h = ctypes.c_void_p()
        _check("gpugi_hostscene_clutter", lib().gpugi_hostscene_clutter(spheres, seed, ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def load(cls, path, json_path=None):
        h = ctypes.c_void_p()
        _check("gpugi_hostscene_load", lib().gpugi_hostscene_load(
            path.encode(), json_path.encode() if json_path else None, ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def build(cls, positions, indices, normals=None, texcoords=None, material_ids=None, materials=None,
              triangles_per_leaf=2):
        positions = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        indices = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
        normals = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        texcoords = None if texcoords is None else np.ascontiguousarray(texcoords, dtype=np.float32).reshape(-1, 2)
        mids = None if material_ids is None else np.ascontiguousarray(material_ids, dtype=np.uint32)
        mats = materials or [default_material("Default")]
        marr = (MaterialParams * len(mats))(*mats)
        h = ctypes.c_void_p()
        _check("gpugi_hostscene_build", lib().gpugi_hostscene_build(
            _ptr(positions), _ptr(normals), _ptr(texcoords), len(positions), _ptr(indices), _ptr(mids),
            len(indices), ctypes.cast(marr, ctypes.c_void_p), len(mats), triangles_per_leaf, ctypes.byref(h)))
        return cls(h.value)

    def save(self, path, json_path=None):
        _check("gpugi_hostscene_save", lib().gpugi_hostscene_save(
            self._h, path.encode(), json_path.encode() if json_path else None))

    def add_point_light(self, position, intensity):
        _check("gpugi_hostscene_add_point_light",
               lib().gpugi_hostscene_add_point_light(self._h, _v3(position), _v3(intensity)))

    def desc(self):
        d = SceneDesc()
        _check("gpugi_hostscene_desc", lib().gpugi_hostscene_desc(self._h, ctypes.byref(d)))
        return d

    def arrays(self):
        """Copies of the reference-layout arrays as numpy arrays."""
        d = self.desc()

        def arr(ptr, count, dtype):
            if count == 0:
                return np.zeros(0, dtype=dtype)
            buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr)
            return np.frombuffer(buf, dtype=dtype).copy()

        return {
            "hierarchy": arr(d.hierarchy, d.num_nodes * 8, np.float32).reshape(-1, 8),
            "triangles": arr(d.triangles, d.num_leaves * d.triangles_per_leaf * 4, np.int32).reshape(-1, 4),
            "positions": arr(d.positions, d.num_vertices * 3, np.float32).reshape(-1, 3),
            "vertex_info": arr(d.vertex_info, d.num_vertices * 4, np.float32).reshape(-1, 4),
            "materials": arr(d.materials, d.num_materials, MATERIAL_DTYPE),
            "triangles_per_leaf": d.triangles_per_leaf,
        }

    def obox_hierarchy(self):
        """The TreeNode<OBox> records (num_nodes x 12 float32) a GPUGI_BVH_OBOX load builds."""
        n = self.desc().num_nodes
        out = np.zeros((n, 12), np.float32)
        _check("gpugi_hostscene_obox_hierarchy", lib().gpugi_hostscene_obox_hierarchy(self._h, _ptr(out), out.size))
        return out

    def textures(self):
        """The material textures the scene's json references: a list of (h, w, 4)
        arrays (uint8 sRGB or float32 linear), row 0 at t = 0.5 / h."""
        n = ctypes.c_uint32()
        _check("gpugi_hostscene_num_textures", lib().gpugi_hostscene_num_textures(self._h, ctypes.byref(n)))
        out = []
        for i in range(n.value):
            w, h, f, p = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_void_p()
            _check("gpugi_hostscene_texture", lib().gpugi_hostscene_texture(
                self._h, i, ctypes.byref(w), ctypes.byref(h), ctypes.byref(f), ctypes.byref(p)))
            dtype, size = (np.uint8, 4) if f.value == TEXTURE_RGBA8_SRGB else (np.float32, 16)
            buf = (ctypes.c_char * (w.value * h.value * size)).from_address(p.value)
            out.append(np.frombuffer(buf, dtype=dtype).copy().reshape(h.value, w.value, 4))
        return out

    def material_textures(self):
        """(materials, 4) texture indices of the diffuse, opacity, reflectiveness and
        emissivity slots (-1: the material's constant)."""
        nm = self.desc().num_materials
        out = np.full((nm, 4), -1, np.int32)
        for m in range(nm):
            _check("gpugi_hostscene_material_textures",
                   lib().gpugi_hostscene_material_textures(self._h, m, _ptr(out[m])))
        return out

    def info(self):
        nt, dep = ctypes.c_uint32(), ctypes.c_uint32()
        lo, hi = (ctypes.c_float * 3)(), (ctypes.c_float * 3)()
        _check("gpugi_hostscene_info", lib().gpugi_hostscene_info(self._h, ctypes.byref(nt), ctypes.byref(dep),
                                                                  ctypes.byref(lo), ctypes.byref(hi)))
        return {"num_triangles": nt.value, "max_depth": dep.value, "bbox_min": list(lo), "bbox_max": list(hi)}

    def close(self):
        if self._h:
            lib().gpugi_hostscene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_material(name, albedo=(0.5, 0.5, 0.5), specular=(0.03, 0.03, 0.03), roughness=1.0, opacity=1.0,
                     emissivity=(0.0, 0.0, 0.0), refraction_index=1.5):
    m = MaterialParams()
    m.name = name.encode()[:31]
    m.albedo[:] = albedo
    m.specular_color[:] = specular
    m.roughness = roughness
    m.opacity = opacity
    m.emissivity[:] = emissivity
    m.refraction_index = refraction_index
    return m


class LightSampler:
    def __init__(self, scene, seed=123):
        self._scene = scene  # keep alive
        h = ctypes.c_void_p()
        _check("gpugi_lightsampler_create", lib().gpugi_lightsampler_create(scene._h, seed, ctypes.byref(h)))
        self._h = h

    def generate(self, n, position_bias=1.0e-4):
        out = np.zeros(n, dtype=LIGHT_DTYPE)
        _check("gpugi_lightsampler_generate", lib().gpugi_lightsampler_generate(self._h, _ptr(out), n,
                                                                                 float(position_bias)))
        return out

    def set_scene(self, scene):
        _check("gpugi_lightsampler_set_scene", lib().gpugi_lightsampler_set_scene(self._h, scene._h))
        self._scene = scene

    def __del__(self):
        try:
            if self._h:
                lib().gpugi_lightsampler_destroy(self._h)
        except Exception:
            pass


class _PinnedBlock:
    """One gpugi_host_alloc allocation, freed once (by host_free or when its last array goes)."""

    def __init__(self, ctx, nbytes):
        p = ctypes.c_void_p()
        ctx._call("gpugi_host_alloc", ctypes.c_size_t(nbytes), ctypes.byref(p))
        self.ctx = ctx
        self.ptr = p.value

    def free(self):
        if self.ptr and self.ctx._h:
            self.ctx._call("gpugi_host_free", ctypes.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceBuffer:
    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = nbytes
        p = ctypes.c_void_p()
        ctx._call("gpugi_alloc", ctypes.c_size_t(nbytes), ctypes.byref(p))
        self.ptr = p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.ctx._call("gpugi_upload", ctypes.c_void_p(self.ptr), _ptr(arr), ctypes.c_size_t(arr.nbytes))

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        self.ctx._call("gpugi_download", _ptr(out), ctypes.c_void_p(self.ptr), ctypes.c_size_t(out.nbytes))
        return out

    def copy_from(self, other, nbytes):
        assert nbytes <= self.nbytes and nbytes <= other.nbytes
        self.ctx._call("gpugi_copy_device", ctypes.c_void_p(self.ptr), ctypes.c_void_p(other.ptr),
                       ctypes.c_size_t(nbytes))

    def free(self):
        if self.ptr:
            self.ctx._call("gpugi_free", ctypes.c_void_p(self.ptr))
            self.ptr = None


class Context:
    """One device context (gpugi_ctx) on a gfx950 device."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check("gpugi_create", lib().gpugi_create(device, ctypes.byref(h)))
        self._h = h
        self._scene = None
        self._pinned = weakref.WeakSet()  # live host_empty allocations

    def _call(self, name, *args):
        _check(name, getattr(lib(), name)(self._h, *args), self._h)

    def load_scene(self, scene):
        """gpugi_load_scene, then the scene's material textures (if any) uploaded and bound."""
        d = scene.desc()
        self._call("gpugi_load_scene", ctypes.byref(d))
        self._scene = scene
        texs = scene.textures()
        if texs:
            ids = [self.add_texture(t) for t in texs]
            for m, slots in enumerate(scene.material_textures()):
                for k, t in enumerate(slots):
                    if t >= 0:
                        self.set_material_texture(m, k, ids[t])

    def set_camera(self, cam):
        self._call("gpugi_set_camera", ctypes.byref(cam))

    def set_screen(self, w, h):
        self._call("gpugi_set_screen", int(w), int(h))

    def set_light_samples(self, samples):
        samples = np.ascontiguousarray(samples, dtype=LIGHT_DTYPE)
        self._call("gpugi_set_light_samples", _ptr(samples), len(samples))

    def set_max_path_length(self, n):
        self._call("gpugi_set_max_path_length", int(n))

    def set_integrator_kernel(self, k):
        self._call("gpugi_set_integrator_kernel", int(k))

    def set_pt_schedule(self, schedule):
        """PT_MEGAKERNEL (default) or PT_WAVEFRONT: how draw(PT) schedules its paths."""
        self._call("gpugi_set_pt_schedule", int(schedule))

    def set_rr_start_depth(self, depth):
        self._call("gpugi_set_rr_start_depth", int(depth))

    def capture_draw(self, integrator, frame_seed):
        """Records one draw into a HIP graph (wavefront PT only); replay_draw runs it."""
        self._call("gpugi_capture_draw", int(integrator), ctypes.c_uint32(frame_seed))

    def replay_draw(self):
        self._call("gpugi_replay_draw")

    def set_specific_path_length(self, length):
        """Only paths of `length` bounces in PT/LT/BPT draws; -1 keeps every path."""
        self._call("gpugi_set_specific_path_length", int(length))

    def set_triangle_test(self, mode):
        self._call("gpugi_set_triangle_test", int(mode))

    def draw(self, integrator, frame_seed):
        self._call("gpugi_draw", int(integrator), ctypes.c_uint32(frame_seed & 0xffffffff))

    def bpt_warmup(self, frame_seed):
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        self._call("gpugi_bpt_warmup", ctypes.c_uint32(frame_seed & 0xffffffff), ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def bpt_set_light_cache(self, capacity, average_path_length=0):
        self._call("gpugi_bpt_set_light_cache", int(capacity), int(average_path_length))

    def bpt_read_light_cache(self, max_entries):
        out = np.zeros(max_entries, dtype=LIGHT_CACHE_DTYPE)
        n = ctypes.c_uint32()
        self._call("gpugi_bpt_read_light_cache", _ptr(out), int(max_entries), ctypes.byref(n))
        return out, n.value

    def reset_accum(self):
        self._call("gpugi_reset_accum")

    def readback(self, w, h):
        out = np.zeros((h, w, 4), dtype=np.float32)
        it = ctypes.c_uint32()
        self._call("gpugi_readback", _ptr(out), ctypes.c_size_t(out.nbytes), ctypes.byref(it))
        return out, it.value

    def add_texture(self, texels):
        """Scene::GetBindlessHandle: texels (h, w, 4) uint8 (sRGB colour, linear alpha)
        or float32 (linear).  Returns the texture index."""
        t = np.ascontiguousarray(texels)
        if t.ndim != 3 or t.shape[2] != 4:
            raise ValueError("texture must be (h, w, 4)")
        fmt = TEXTURE_RGBA8_SRGB if t.dtype == np.uint8 else TEXTURE_RGBA32F
        if fmt == TEXTURE_RGBA32F:
            t = np.ascontiguousarray(t, dtype=np.float32)
        out = ctypes.c_int()
        self._call("gpugi_add_texture", int(t.shape[1]), int(t.shape[0]), fmt, _ptr(t), ctypes.byref(out))
        return out.value

    def set_material_texture(self, material, slot, texture):
        self._call("gpugi_set_material_texture", int(material), int(slot), int(texture))

    def set_bvh_type(self, bvh_type):
        """bvhType (BVH_AABOX / BVH_OBOX); applies from the next load_scene."""
        self._call("gpugi_set_bvh_type", int(bvh_type))

    def set_environment_map(self, faces):
        """RendererSystem::SetEnvironmentMap: faces (6, size, size, 3) uint8 sRGB in the
        order x-, x+, y-, y+, z-, z+, or None for the black default."""
        if faces is None:
            self._call("gpugi_set_environment_map", 0, None)
            return
        f = np.ascontiguousarray(faces, dtype=np.uint8)
        if f.ndim != 4 or f.shape[0] != 6 or f.shape[1] != f.shape[2] or f.shape[3] != 3:
            raise ValueError("environment faces must be (6, size, size, 3)")
        self._call("gpugi_set_environment_map", int(f.shape[1]), _ptr(f))

    def set_reference_image(self, rgb):
        """TextureMSE::LoadReferenceFromPFM's upload: rgb is (h, w, 3) float32 in PFM row order."""
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        if rgb.ndim != 3 or rgb.shape[2] != 3:
            raise ValueError("reference image must be (h, w, 3)")
        self._call("gpugi_set_reference_image", _ptr(rgb), int(rgb.shape[1]), int(rgb.shape[0]))

    def compute_mse(self, divisor):
        """TextureMSE::ComputeCurrentMSE(backbuffer, divisor); +inf without a reference image."""
        v = ctypes.c_float()
        self._call("gpugi_compute_mse", int(divisor), ctypes.byref(v))
        return v.value

    def trace(self, rays, query=CLOSEST_HIT, kernel=KERNEL_AUTO, out=None):
        """Host-buffer trace; out (optional, HIT_DTYPE, e.g. from host_empty) receives the hits."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(len(rays), dtype=HIT_DTYPE) if out is None else out
        assert hits.dtype == HIT_DTYPE and len(hits) >= len(rays) and hits.flags.c_contiguous
        self._call("gpugi_trace", int(query), int(kernel), _ptr(rays), _ptr(hits), ctypes.c_size_t(len(rays)), 0)
        return hits

    def host_empty(self, count, dtype):
        """An uninitialised numpy array of `count` elements in pinned host memory
        (gpugi_host_alloc).  The array owns the allocation: it is released when
        the last view of it goes (or earlier by host_free), and the context
        refuses to close while such an array is alive."""
        dtype = np.dtype(dtype)
        block = _PinnedBlock(self, max(1, count * dtype.itemsize))
        buf = (ctypes.c_char * (count * dtype.itemsize)).from_address(block.ptr)
        buf._owner = block  # numpy keeps `buf` as the array's base, `buf` keeps the block
        self._pinned.add(block)
        return np.frombuffer(buf, dtype=dtype, count=count)

    def host_free(self, arr):
        """Releases a host_empty array's pinned memory now; the array must not be used afterwards."""
        base = arr
        while base is not None and not hasattr(base, "_owner"):
            base = getattr(base, "base", None) if not isinstance(base, memoryview) else base.obj
        if base is None:
            raise GpugiError("host_free", GPUGI_ERR_ARG, "not an array from host_empty")
        base._owner.free()

    def trace_importance_break(self, rays, importance, threshold, ray_threshold, query=CLOSEST_HIT):
        """TraceRay with TRACERAY_IMPORTANCE_BREAK: importance (num_nodes, 2) float32, or None
        for the importance hierarchy_update_importance left on the device.
        Returns (hits, nodes) for closest hit (nodes: NODE_HIT_DTYPE), hits for any hit."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        imp = None if importance is None else np.ascontiguousarray(importance, dtype=np.float32).reshape(-1, 2)
        hits = np.zeros(len(rays), dtype=HIT_DTYPE)
        nodes = np.zeros(len(rays), dtype=NODE_HIT_DTYPE)
        self._call("gpugi_trace_importance_break", int(query), _ptr(rays), _ptr(hits), _ptr(nodes),
                   ctypes.c_size_t(len(rays)), _ptr(imp), 0 if imp is None else len(imp),
                   ctypes.c_float(threshold), ctypes.c_float(ray_threshold))
        return hits if query == ANY_HIT else (hits, nodes)

    def hierarchy_update_importance(self, iterations=10):
        """HierarchyImportance's acquisition passes and propagation (result stays on the device)."""
        self._call("gpugi_hierarchy_update_importance", int(iterations))

    def hierarchy_compute_materials(self):
        self._call("gpugi_hierarchy_compute_materials")

    def _hier_sizes(self):
        d = self._scene.desc()
        return d.num_nodes, d.num_leaves * d.triangles_per_leaf

    def hierarchy_read(self):
        """(importance ((nodes + slots), 2) float32, subtree importance (nodes,) float32)."""
        nn, ns = self._hier_sizes()
        imp = np.zeros((nn + ns, 2), np.float32)
        sub = np.zeros(nn, np.float32)
        self._call("gpugi_hierarchy_read", _ptr(imp), ctypes.c_size_t(imp.size), _ptr(sub), ctypes.c_size_t(nn),
                   None, ctypes.c_size_t(0))
        return imp, sub

    def hierarchy_read_materials(self):
        """(nodes, 12) uint32: six pairs of packed halves per node."""
        nn, _ = self._hier_sizes()
        mats = np.zeros((nn, 12), np.uint32)
        self._call("gpugi_hierarchy_read", None, ctypes.c_size_t(0), None, ctypes.c_size_t(0), _ptr(mats),
                   ctypes.c_size_t(mats.size))
        return mats

    def trace_device(self, rays_buf, hits_buf, n, query=CLOSEST_HIT, kernel=KERNEL_AUTO):
        self._call("gpugi_trace", int(query), int(kernel), ctypes.c_void_p(rays_buf.ptr),
                   ctypes.c_void_p(hits_buf.ptr), ctypes.c_size_t(n), 1)

    def probe_valu(self, out_buf, blocks, iters):
        """Runs the VALU calibration kernel (blocks x 256 lanes, 64 * iters FMAs each); returns ms."""
        ms = ctypes.c_float()
        self._call("gpugi_probe_valu", ctypes.c_uint32(blocks), ctypes.c_uint32(iters), ctypes.c_void_p(out_buf.ptr),
                   ctypes.byref(ms))
        return ms.value

    def trace_noalloc(self, rays_buf, hits_buf, n, query=CLOSEST_HIT, kernel=KERNEL_AUTO):
        """gpugi_trace_noalloc: device-resident trace that allocates nothing (asynchronous)."""
        self._call("gpugi_trace_noalloc", int(query), int(kernel), ctypes.c_void_p(rays_buf.ptr),
                   ctypes.c_void_p(hits_buf.ptr), ctypes.c_size_t(n))

    def set_wide_config(self, max_leaf_triangles, lds_stack_entries, refill_lanes=8, wait_lanes=16):
        self._call("gpugi_set_wide_config", int(max_leaf_triangles), int(lds_stack_entries), int(refill_lanes),
                   int(wait_lanes))

    def wide_config(self):
        a, b, r, w, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_uint32()
        self._call("gpugi_get_wide_config", ctypes.byref(a), ctypes.byref(b), ctypes.byref(r), ctypes.byref(w),
                   ctypes.byref(c))
        return {"max_leaf_triangles": a.value, "lds_stack_entries": b.value, "refill_lanes": r.value,
                "wait_lanes": w.value, "max_stack": c.value}

    def autotune_trace(self, rays_buf, hits_buf, n, query=CLOSEST_HIT, repeats=3):
        ms = ctypes.c_float()
        self._call("gpugi_autotune_trace", int(query), ctypes.c_void_p(rays_buf.ptr), ctypes.c_void_p(hits_buf.ptr),
                   ctypes.c_size_t(n), int(repeats), ctypes.byref(ms))
        return ms.value

    def trace_counts(self, rays, query=CLOSEST_HIT):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        nb = np.zeros(len(rays), dtype=np.uint32)
        nt = np.zeros(len(rays), dtype=np.uint32)
        self._call("gpugi_trace_counts", int(query), _ptr(rays), _ptr(nb), _ptr(nt), ctypes.c_size_t(len(rays)), 0)
        return nb, nt

    def trace_counts_device(self, rays_buf, nbox_buf, ntri_buf, n, query=CLOSEST_HIT):
        self._call("gpugi_trace_counts", int(query), ctypes.c_void_p(rays_buf.ptr), ctypes.c_void_p(nbox_buf.ptr),
                   ctypes.c_void_p(ntri_buf.ptr), ctypes.c_size_t(n), 1)

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def sync(self):
        self._call("gpugi_sync")

    def timer_begin(self):
        self._call("gpugi_timer_begin")

    def timer_end(self):
        ms = ctypes.c_float()
        self._call("gpugi_timer_end", ctypes.byref(ms))
        return ms.value

    def raygen_primary(self, cam, w, h, seed, rays_buf, n):
        self._call("gpugi_raygen_primary", ctypes.byref(cam), int(w), int(h), ctypes.c_uint32(seed),
                   ctypes.c_void_p(rays_buf.ptr), ctypes.c_size_t(n))

    def raygen_diffuse(self, prim_buf, hits_buf, n_prim, seed, rays_buf, n):
        self._call("gpugi_raygen_diffuse", ctypes.c_void_p(prim_buf.ptr), ctypes.c_void_p(hits_buf.ptr),
                   ctypes.c_size_t(n_prim), ctypes.c_uint32(seed), ctypes.c_void_p(rays_buf.ptr), ctypes.c_size_t(n))

    def raygen_shadow(self, prim_buf, hits_buf, n_prim, lights_xyz, rays_buf, n):
        lights = np.ascontiguousarray(lights_xyz, dtype=np.float32).reshape(-1, 3)
        self._call("gpugi_raygen_shadow", ctypes.c_void_p(prim_buf.ptr), ctypes.c_void_p(hits_buf.ptr),
                   ctypes.c_size_t(n_prim), _ptr(lights), len(lights), ctypes.c_void_p(rays_buf.ptr),
                   ctypes.c_size_t(n))

    # -- radiance-cache MLP
    def cache_config(self, spread_c=0.01, extra_bounces=2, train_stride=32, max_train_samples=65536,
                     train_steps=4, learning_rate=1e-3):
        self._call("gpugi_cache_config", ctypes.c_float(spread_c), int(extra_bounces), int(train_stride),
                   int(max_train_samples), int(train_steps), ctypes.c_float(learning_rate))

    def cache_set_query(self, query):
        self._call("gpugi_cache_set_query", int(query))

    def cache_stats(self):
        n = ctypes.c_uint32()
        loss = ctypes.c_float()
        self._call("gpugi_cache_stats", ctypes.byref(n), ctypes.byref(loss))
        return n.value, loss.value

    def mlp_init(self, seed=1):
        self._call("gpugi_mlp_init", ctypes.c_uint64(seed))

    def mlp_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == MLP_PARAMS
        self._call("gpugi_mlp_set_weights", _ptr(w), ctypes.c_size_t(w.size))

    def mlp_get_weights(self):
        w = np.zeros(MLP_PARAMS, np.float32)
        self._call("gpugi_mlp_get_weights", _ptr(w), ctypes.c_size_t(w.size))
        return w

    def mlp_encode(self, features):
        f = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, MLP_FEATURES)
        out = np.zeros((len(f), MLP_INPUTS), np.uint16)
        self._call("gpugi_mlp_encode", _ptr(f), _ptr(out), ctypes.c_size_t(len(f)))
        return out

    def mlp_infer(self, features):
        f = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, MLP_FEATURES)
        out = np.zeros((len(f), 4), np.float32)
        self._call("gpugi_mlp_infer", _ptr(f), _ptr(out), ctypes.c_size_t(len(f)), 0)
        return out

    def mlp_gradient(self, features, targets):
        f = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, MLP_FEATURES)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, 4)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = ctypes.c_float()
        self._call("gpugi_mlp_gradient", _ptr(f), _ptr(t), ctypes.c_size_t(len(f)), _ptr(g), ctypes.byref(loss), 0)
        return g, loss.value

    def mlp_train_step(self, features, targets, lr=1e-3):
        f = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, MLP_FEATURES)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, 4)
        loss = ctypes.c_float()
        self._call("gpugi_mlp_train_step", _ptr(f), _ptr(t), ctypes.c_size_t(len(f)), ctypes.c_float(lr),
                   ctypes.byref(loss), 0)
        return loss.value

    def close(self):
        if self._h:
            live = [blk for blk in self._pinned if blk.ptr]
            if live:
                raise GpugiError("close", GPUGI_ERR_STATE, "%d pinned host arrays from host_empty are still "
                                 "alive; drop them or release them with host_free first" % len(live))
            lib().gpugi_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many scenes rendered by one persistent launch (gpugi_batch_*)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        code = lib().gpugi_batch_create(device, ctypes.byref(h))
        if code != GPUGI_OK:
            raise GpugiError("gpugi_batch_create", code, "no batch")
        self._h = h
        self._scenes = []

    def _call(self, name, *args):
        code = getattr(lib(), name)(self._h, *args)
        if code != GPUGI_OK:
            raise GpugiError(name, code, lib().gpugi_batch_last_error(self._h).decode())

    def add(self, scene, cam, width, height, lights, max_path_length=5):
        lights = np.ascontiguousarray(lights, dtype=LIGHT_DTYPE)
        self._call("gpugi_batch_add", ctypes.byref(scene.desc()), ctypes.byref(cam), int(width), int(height),
                   _ptr(lights), len(lights), int(max_path_length))
        self._scenes.append((width, height))
        return len(self._scenes) - 1

    def finalize(self):
        self._call("gpugi_batch_finalize")

    def config(self, kernel=KERNEL_WIDE, tiles_per_grab=2, grid_multiplier=2, stack_depth=16):
        self._call("gpugi_batch_config", int(kernel), int(tiles_per_grab), int(grid_multiplier), int(stack_depth))

    def reset(self, frame_index=0):
        self._call("gpugi_batch_reset", ctypes.c_uint32(frame_index))

    def draw(self, persistent=True):
        self._call("gpugi_batch_draw", int(bool(persistent)))

    def capture(self):
        self._call("gpugi_batch_capture")

    def replay(self, times=1):
        self._call("gpugi_batch_replay", int(times))

    def run(self, mode, frames):
        ms = ctypes.c_float()
        self._call("gpugi_batch_run", int(mode), int(frames), ctypes.byref(ms))
        return ms.value

    def readback(self, scene):
        w, h = self._scenes[scene]
        out = np.zeros((h, w, 4), np.float32)
        self._call("gpugi_batch_readback", int(scene), _ptr(out), ctypes.c_size_t(out.nbytes))
        return out

    def rays(self):
        r = ctypes.c_uint64()
        self._call("gpugi_batch_rays", ctypes.byref(r))
        return r.value

    def autotune(self, frames=2):
        k, g, m, st, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
        self._call("gpugi_batch_autotune", int(frames), ctypes.byref(k), ctypes.byref(g), ctypes.byref(m),
                   ctypes.byref(st), ctypes.byref(ms))
        return {"kernel": k.value, "tiles_per_grab": g.value, "grid_multiplier": m.value, "stack_depth": st.value,
                "ms_per_frame": ms.value}

    def close(self):
        if self._h:
            lib().gpugi_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def exported_symbols():
    return sorted(_SIGNATURES)
