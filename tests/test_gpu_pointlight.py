"""GPU path tracer and light tracer under a point light, against the oracle
whose point-light branches the closed form of test_oracle_pointlight pins:
the path tracer bit for bit with the reference traversal, the light tracer
to the order of its float-atomic splats."""
import numpy as np
import pytest

from test_oracle_pointlight import H, W, plane_scene

pytestmark = pytest.mark.gpu


def setup(rt, ctx, hs, cam, kernel):
    ctx.load_scene(hs)
    ctx.set_screen(W, H)
    ctx.set_camera(cam)
    ctx.set_max_path_length(3)
    ctx.set_integrator_kernel(kernel)


def camera(rt):
    return rt.camera([0.0, 3.0, 0.5], [0.0, 0.0, 0.0], [0, 1, 0], 40, W, H)


@pytest.mark.parametrize("kernel", ["reference", "wide"])
def test_pt_point_light_matches_oracle(rt, oracle, ctx, kernel):
    hs = plane_scene(rt)
    cam = camera(rt)
    setup(rt, ctx, hs, cam, rt.KERNEL_REFERENCE if kernel == "reference" else rt.KERNEL_WIDE)
    sc = oracle.Scene(hs.arrays())
    sampler = rt.LightSampler(hs, 123)
    acc = np.zeros((H, W, 4), np.float32)
    print("case: point-light PT %s" % kernel, flush=True)
    for it in range(3):
        ls = sampler.generate(128)
        ctx.set_light_samples(ls)
        ctx.draw(rt.PT, rt.wanghash(it))
        oracle.pathtrace(sc, ls, cam, W, H, rt.wanghash(it), 3, acc)
    img, n = ctx.readback(W, H)
    assert n == 3 and acc[..., :3].mean() > 0.05
    if kernel == "reference":
        assert np.array_equal(img.view(np.uint32), acc.view(np.uint32))
    else:  # two triangles: only a ray through their shared edge could pick the other slot
        same = (img.view(np.uint32) == acc.view(np.uint32)).all(-1)
        assert same.mean() > 0.99
        assert abs(img[..., :3].mean() / acc[..., :3].mean() - 1) < 1e-3


def test_lt_point_light_matches_oracle(rt, oracle, ctx):
    hs = plane_scene(rt)
    cam = camera(rt)
    setup(rt, ctx, hs, cam, rt.KERNEL_REFERENCE)
    sc = oracle.Scene(hs.arrays())
    sampler = rt.LightSampler(hs, 321)
    acc = np.zeros((H, W, 4), np.float32)
    print("case: point-light LT", flush=True)
    for it in range(2):
        ls = sampler.generate(256)
        ctx.set_light_samples(ls)
        ctx.draw(rt.LT, rt.wanghash(it))
        oracle.lighttrace(sc, ls, cam, W, H, rt.wanghash(it), 3, acc)
    img, n = ctx.readback(W, H)
    assert n == 2 and acc[..., :3].mean() > 0
    ok = np.abs(img - acc) <= 1e-4 * np.abs(acc) + 1e-6
    assert ok.all(), np.argwhere(~ok)[:5]
