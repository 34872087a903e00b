"""Synthetic workloads of the benchmark and the parity tests.

heightfield: 1024 x 1024 cells (2,097,152 triangles) over [0,1]^2 with
  heights 0.2 * fBm (6 octaves of wanghash value noise, seed MASTER_SEED);
  1920 x 1080 pinhole camera, vertical FOV 40 degrees, 30 degrees elevation,
  distance 0.6 from (0.5, 0.1, 0.5); three ray sets of 2^24 rays:
  coherent primary rays (seed S), incoherent cosine-distributed rays from the
  primary hits (seed S + 1), shadow rays toward 4 seeded point lights.
cornell: the 36-triangle room, 256 x 256, vertical FOV 40 degrees, path length 5.
clutter: the 20 x 20 x 8 room with 4000 icospheres (5,120,016 triangles), 1920 x 1080,
  vertical FOV 60 degrees, path length 8, Russian roulette from bounce 3.
"""
import math

import numpy as np

from . import rt

MASTER_SEED = 0x9E3779B9

HF_CELLS = 1024
HF_WIDTH, HF_HEIGHT = 1920, 1080
HF_RAYS = 1 << 24
HF_ELEVATION_DEG = 30.0
HF_AZIMUTH_DEG = 30.0
HF_DISTANCE = 0.6
HF_TARGET = (0.5, 0.1, 0.5)

CORNELL_SIZE = 256
CORNELL_PATH_LENGTH = 5


def heightfield_camera(width=HF_WIDTH, height=HF_HEIGHT):
    e, a = math.radians(HF_ELEVATION_DEG), math.radians(HF_AZIMUTH_DEG)
    pos = [HF_TARGET[0] + HF_DISTANCE * math.cos(e) * math.sin(a),
           HF_TARGET[1] + HF_DISTANCE * math.sin(e),
           HF_TARGET[2] + HF_DISTANCE * math.cos(e) * math.cos(a)]
    return rt.camera(pos, HF_TARGET, (0.0, 1.0, 0.0), 40.0, width, height)


def heightfield_lights(seed=MASTER_SEED, count=4):
    """Point lights above the terrain: x, z in [0.1, 0.9], y in [0.5, 1.0]."""
    out = []
    h = seed
    for _ in range(count):
        vals = []
        for _ in range(3):
            h = rt.wanghash(h + 0x6D2B79F5)
            vals.append(h / 4294967295.0)
        out.append([0.1 + 0.8 * vals[0], 0.5 + 0.5 * vals[1], 0.1 + 0.8 * vals[2]])
    return np.array(out, dtype=np.float32)


def cornell_camera(size=CORNELL_SIZE):
    return rt.camera((0.0, 1.0, 3.7), (0.0, 1.0, 0.0), (0.0, 1.0, 0.0), 40.0, size, size)


class SceneRays:
    """The three ray sets of a camera and point lights, generated on the device;
    buffers stay resident in HBM.  Coherent: primary rays of pixel i mod
    (width * height) with seeded jitter (seed); incoherent: cosine-distributed
    from the i-th primary hit (seed + 1); shadow: toward light i mod count."""

    def __init__(self, ctx, cam, lights, n, seed, width, height):
        self.ctx, self.n = ctx, n
        self.cam = cam
        self.lights = np.ascontiguousarray(lights, dtype=np.float32)
        self.primary = ctx.alloc(n * 32)
        self.primary_hits = ctx.alloc(n * 16)
        self.incoherent = ctx.alloc(n * 32)
        self.shadow = ctx.alloc(n * 32)
        self.hits = ctx.alloc(n * 16)
        ctx.raygen_primary(self.cam, width, height, seed, self.primary, n)
        ctx.trace_device(self.primary, self.primary_hits, n, rt.CLOSEST_HIT, rt.KERNEL_REFERENCE)
        ctx.raygen_diffuse(self.primary, self.primary_hits, n, (seed + 1) & 0xffffffff, self.incoherent, n)
        ctx.raygen_shadow(self.primary, self.primary_hits, n, self.lights, self.shadow, n)
        ctx.sync()

    def free(self):
        for b in (self.primary, self.primary_hits, self.incoherent, self.shadow, self.hits):
            b.free()


class HeightfieldRays(SceneRays):
    """The heightfield benchmark's ray sets (camera and lights above)."""

    def __init__(self, ctx, n=HF_RAYS, seed=MASTER_SEED, width=HF_WIDTH, height=HF_HEIGHT):
        super().__init__(ctx, heightfield_camera(width, height), heightfield_lights(seed), n, seed, width, height)


CLUTTER_SPHERES = 4000
CLUTTER_PATH_LENGTH = 8
CLUTTER_RR_START = 3
CLUTTER_LIGHTS = np.array([[-6.0, 6.5, -5.0], [5.0, 5.5, 6.0], [6.0, 7.0, -6.0]], np.float32)


def clutter_camera(width=HF_WIDTH, height=HF_HEIGHT):
    """Inside the 20 x 20 x 8 clutter room, near the +z wall, looking at its middle."""
    return rt.camera((0.0, 4.0, 9.5), (0.0, 3.0, 0.0), (0.0, 1.0, 0.0), 60.0, width, height)


BATCH_SCENES = 16
BATCH_PATH_LENGTH = 5


def batch_scenes(count=BATCH_SCENES, seed=MASTER_SEED, max_triangles=2_000_000, min_triangles=10_000):
    """Scenes of the batched-rendering benchmark: heightfields whose triangle
    counts are log-uniform in [min_triangles, max_triangles] (cells =
    sqrt(triangles / 2)), square resolutions of 256..1024 pixels in steps of
    128, a camera at a seeded azimuth and 4 seeded point lights each (128
    light samples).  Returns [(HostScene, camera, width, height, lights)]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        tris = math.exp(math.log(min_triangles) + (math.log(max_triangles) - math.log(min_triangles)) * rng.random())
        cells = max(2, int(round(math.sqrt(tris / 2.0))))
        size = int(256 + 128 * rng.integers(0, 7))
        hs = rt.HostScene.heightfield(cells, (seed + 977 * i) & 0xffffffff)
        for pos in heightfield_lights((seed + 31 * i) & 0xffffffff):
            hs.add_point_light(pos, (0.5, 0.5, 0.5))
        az = math.radians(360.0 * rng.random())
        e = math.radians(HF_ELEVATION_DEG)
        eye = [HF_TARGET[0] + HF_DISTANCE * math.cos(e) * math.sin(az), HF_TARGET[1] + HF_DISTANCE * math.sin(e),
               HF_TARGET[2] + HF_DISTANCE * math.cos(e) * math.cos(az)]
        cam = rt.camera(eye, HF_TARGET, (0.0, 1.0, 0.0), 40.0, size, size)
        lights = rt.LightSampler(hs, 123).generate(128)
        out.append((hs, cam, size, size, lights))
    return out
