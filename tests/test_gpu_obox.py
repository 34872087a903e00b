"""Oriented-box hierarchy on the GPU (gpugi_set_bvh_type(GPUGI_BVH_OBOX)):
the reference walk over TreeNode<OBox> equals the oracle's walk over the same
records bit for bit (closest hit and any hit), the path tracer with the
reference traversal equals the oracle's, and the axis-aligned hierarchy comes
back with the next load under GPUGI_BVH_AABOX."""
import numpy as np
import pytest

from gpugi_amd import workloads
from test_oracle_geometry import make_rays
from test_oracle_obox import SCENES

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def octx(rt):
    c = rt.Context(0)  # its own context: the bvh type stays out of other modules' tests
    yield c
    c.close()


def hits_int(h):
    return np.ascontiguousarray(h).view(np.int32).reshape(-1, 4)


@pytest.mark.parametrize("name", list(SCENES))
def test_obox_trace_matches_oracle(rt, oracle, octx, name):
    hs = SCENES[name](rt)
    a = hs.arrays()
    sco = oracle.Scene(dict(a, obox=hs.obox_hierarchy()))
    octx.set_bvh_type(rt.BVH_OBOX)
    octx.load_scene(hs)
    lo, hi = a["positions"].min(0), a["positions"].max(0)
    rng = np.random.default_rng(31)
    o = rng.uniform(lo, hi, (8192, 3))
    target = a["positions"][rng.integers(0, len(a["positions"]), 8192)]
    rays = make_rays(o, np.where(np.arange(8192)[:, None] % 2 == 0, target - o, rng.normal(size=(8192, 3))))
    print("case: obox trace %s" % name, flush=True)
    got = hits_int(octx.trace(np.ascontiguousarray(rays).view(rt.RAY_DTYPE).reshape(-1), rt.CLOSEST_HIT,
                              rt.KERNEL_REFERENCE))
    want = oracle.trace(sco, rays)
    assert (want[:, 0] >= 0).mean() > 0.2
    assert np.array_equal(got, want)
    gany = hits_int(octx.trace(np.ascontiguousarray(rays).view(rt.RAY_DTYPE).reshape(-1), rt.ANY_HIT,
                               rt.KERNEL_REFERENCE))
    wany = oracle.trace(sco, rays, any_hit=True)
    assert np.array_equal(gany[:, 0], wany[:, 0])


def test_obox_pathtracer_matches_oracle(rt, oracle, octx):
    hs = rt.HostScene.cornell()
    sco = oracle.Scene(dict(hs.arrays(), obox=hs.obox_hierarchy()))
    octx.set_bvh_type(rt.BVH_OBOX)
    octx.load_scene(hs)
    w = h = 48
    cam = workloads.cornell_camera(w)
    octx.set_screen(w, h)
    octx.set_camera(cam)
    octx.set_max_path_length(5)
    octx.set_integrator_kernel(rt.KERNEL_REFERENCE)
    sampler = rt.LightSampler(hs, 123)
    acc = np.zeros((h, w, 4), np.float32)
    print("case: obox PT", flush=True)
    for it in range(2):
        ls = sampler.generate(128)
        octx.set_light_samples(ls)
        octx.draw(rt.PT, rt.wanghash(it))
        oracle.pathtrace(sco, ls, cam, w, h, rt.wanghash(it), 5, acc)
    img, _ = octx.readback(w, h)
    assert img[..., :3].mean() > 0.05
    assert np.array_equal(img.view(np.uint32), acc.view(np.uint32))


def test_aabox_returns_with_the_next_load(rt, oracle, octx):
    hs = rt.HostScene.heightfield(16, workloads.MASTER_SEED)
    sc = oracle.Scene(hs.arrays())
    rng = np.random.default_rng(2)
    rays = make_rays(rng.uniform(0, 1, (2048, 3)) + [0, 0.3, 0], rng.normal(size=(2048, 3)))
    octx.set_bvh_type(rt.BVH_OBOX)
    octx.load_scene(hs)
    octx.set_bvh_type(rt.BVH_AABOX)  # applies from the next load only
    r = np.ascontiguousarray(rays).view(rt.RAY_DTYPE).reshape(-1)
    ob = hits_int(octx.trace(r, rt.CLOSEST_HIT, rt.KERNEL_REFERENCE))
    assert np.array_equal(ob, oracle.trace(oracle.Scene(dict(hs.arrays(), obox=hs.obox_hierarchy())), rays))
    octx.load_scene(hs)
    aa = hits_int(octx.trace(r, rt.CLOSEST_HIT, rt.KERNEL_REFERENCE))
    assert np.array_equal(aa, oracle.trace(sc, rays))


def test_bvh_type_script_command(rt, oracle, tmp_path):
    """gpugi_render's bvhType = 1 before sceneFilename renders with the oriented
    boxes: the PT image equals the oracle's over the same records."""
    import os
    import subprocess
    app = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpugi_amd", "bin", "gpugi_render")
    out = str(tmp_path / "obox.pfm")
    w = 32
    lines = ["bvhType = 1", "sceneFilename = cornell", "cameraPos = 0, 1, 3.7", "cameraLookAt = 0, 1, 0",
             "cameraFOV = 40", "resolution = %d, %d" % (w, w), "pathLength = 5", "traversal = 1", "renderer = 0",
             "waitIterations = 1", "namedscreenshot = " + out, "exit"]
    args = [app]
    for l in lines:
        args += ["-e", l]
    print("gpugi_render bvhType 1", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stdout + r.stderr
    with open(out, "rb") as f:
        f.readline(), f.readline(), f.readline()
        img = np.frombuffer(f.read(), dtype="<f4").reshape(w, w, 3)
    hs = rt.HostScene.cornell()
    sco = oracle.Scene(dict(hs.arrays(), obox=hs.obox_hierarchy()))
    ls = rt.LightSampler(hs, 123)
    # RendererSystem: the Pathtracer's constructor draws light-sample batch 1 with FrameSeed WangHash(0)
    acc = oracle.pathtrace(sco, ls.generate(128), workloads.cornell_camera(w), w, w, rt.wanghash(0), 5)
    assert img.mean() > 0.05
    assert np.array_equal(img, np.ascontiguousarray(acc[..., :3]))
