"""Environment map on the GPU (gpugi_set_environment_map; pathtracer.comp:56-58)
against the oracle: the path tracer bit for bit with the reference traversal,
nearly everywhere with the wide one; the black default comes back with size 0;
the cached path tracer's exact limits still hold with a map set."""
import numpy as np
import pytest

from gpugi_amd import workloads

pytestmark = pytest.mark.gpu

W = H = 48


@pytest.fixture(scope="module")
def ectx(rt):
    c = rt.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def scene(rt):
    hs = rt.HostScene.heightfield(16, 77)
    for p in workloads.heightfield_lights():
        hs.add_point_light(p, (0.5, 0.5, 0.5))
    faces = np.random.default_rng(3).integers(0, 256, size=(6, 8, 8, 3), dtype=np.uint8)
    return hs, faces


def setup(rt, c, hs, kernel, path_length=4):
    c.load_scene(hs)
    c.set_screen(W, H)
    c.set_camera(workloads.heightfield_camera(W, H))
    c.set_max_path_length(path_length)
    c.set_integrator_kernel(kernel)


def render(rt, c, hs, integrator, frames):
    sampler = rt.LightSampler(hs, 11)
    c.reset_accum()
    for k in range(frames):
        c.set_light_samples(sampler.generate(128))
        c.draw(integrator, rt.wanghash(k))
    return c.readback(W, H)[0]


@pytest.mark.parametrize("kernel", ["reference", "wide"])
def test_pt_environment_matches_oracle(rt, oracle, ectx, scene, kernel):
    hs, faces = scene
    setup(rt, ectx, hs, rt.KERNEL_REFERENCE if kernel == "reference" else rt.KERNEL_WIDE)
    ectx.set_environment_map(faces)
    print("case: environment PT %s" % kernel, flush=True)
    img = render(rt, ectx, hs, rt.PT, 2)
    sc = oracle.Scene(hs.arrays())
    cam = workloads.heightfield_camera(W, H)
    sampler = rt.LightSampler(hs, 11)
    acc = np.zeros((H, W, 4), np.float32)
    oracle.set_environment(oracle.srgb_decode(faces))
    try:
        for k in range(2):
            oracle.pathtrace(sc, sampler.generate(128), cam, W, H, rt.wanghash(k), 4, acc)
    finally:
        oracle.set_environment(None)
    ectx.set_environment_map(None)
    plain = render(rt, ectx, hs, rt.PT, 2)
    assert (img[..., :3] - plain[..., :3]).mean() > 0.01  # the map adds light
    if kernel == "reference":
        assert np.array_equal(img.view(np.uint32), acc.view(np.uint32))
    else:
        same = (img.view(np.uint32) == acc.view(np.uint32)).all(-1)
        assert same.mean() > 0.98
        assert abs(img[..., :3].mean() / acc[..., :3].mean() - 1) < 2e-3


def test_black_default_is_restored(rt, ectx, scene):
    hs, faces = scene
    setup(rt, ectx, hs, rt.KERNEL_REFERENCE)
    a = render(rt, ectx, hs, rt.PT, 1)
    ectx.set_environment_map(faces)
    b = render(rt, ectx, hs, rt.PT, 1)
    ectx.set_environment_map(None)
    c = render(rt, ectx, hs, rt.PT, 1)
    assert not np.array_equal(a, b)
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32))


def test_cache_limits_hold_with_environment(rt, ectx, scene):
    hs, faces = scene
    setup(rt, ectx, hs, rt.KERNEL_REFERENCE)
    ectx.set_environment_map(faces)
    try:
        ectx.mlp_init(5)
        ectx.cache_set_query(rt.CACHE_SPREAD)
        ectx.cache_config(spread_c=1e30, train_steps=1)
        print("case: environment cache spread limit", flush=True)
        a = render(rt, ectx, hs, rt.PT_CACHE, 2)
        b = render(rt, ectx, hs, rt.PT, 2)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        # primary-hit query with an all-zero network: the path tracer with path length 1
        ectx.mlp_set_weights(np.zeros_like(ectx.mlp_get_weights()))
        ectx.cache_set_query(rt.CACHE_PRIMARY_INDIRECT)
        ectx.cache_config(train_steps=0)
        print("case: environment cache primary-indirect limit", flush=True)
        c = render(rt, ectx, hs, rt.PT_CACHE, 2)
        ectx.set_max_path_length(1)
        d = render(rt, ectx, hs, rt.PT, 2)
        assert np.array_equal(c.view(np.uint32), d.view(np.uint32))
    finally:
        ectx.set_environment_map(None)
        ectx.cache_set_query(rt.CACHE_SPREAD)


def test_envmap_script_command(rt, oracle, tmp_path):
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    app = os.path.join(root, "gpugi_amd", "bin", "gpugi_render")
    names = []
    colours = [(255, 0, 0), (0, 255, 0), (0, 0, 255), (200, 100, 50), (10, 20, 30), (90, 180, 240)]
    for k, col in enumerate(colours):
        path = str(tmp_path / ("face%d.ppm" % k))
        with open(path, "wb") as f:
            f.write(b"P6\n# face\n4 4\n255\n" + bytes(col) * 16)
        names.append(path)
    out = str(tmp_path / "sky.pfm")
    lines = ["sceneFilename = heightfield:8", "cameraPos = 0.5, 0.3, 0.5", "cameraLookAt = 0.5, 5, 0.9",
             "cameraFOV = 30", "resolution = 16, 16", "pathLength = 2", "envMap = 4, " + ", ".join(names),
             "renderer = 0", "waitIterations = 1", "namedscreenshot = " + out, "exit"]
    args = [app]
    for l in lines:
        args += ["-e", l]
    print("gpugi_render envMap", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stdout + r.stderr
    with open(out, "rb") as f:
        f.readline(), f.readline(), f.readline()
        img = np.frombuffer(f.read(), dtype="<f4").reshape(16, 16, 3)
    # looking straight up (+y): every pixel sees the y+ face (index 3)
    want = oracle.srgb_decode(np.array(colours, np.uint8).reshape(6, 1, 1, 3))[3, 0, 0, :3]
    assert np.array_equal(img.reshape(-1, 3), np.tile(want, (256, 1)))
