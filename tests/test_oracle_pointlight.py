"""Point-light branches of the integrator oracles pinned by a closed form.

A Lambertian plane (albedo a, no specular part) under one point light of
intensity I: the radiance leaving a point x toward the camera is
    L(x) = a / pi * I * cos(theta) / r^2
(lightsource.glsl:25-47: irradiance (cos / r^2) * I times the diffuse BSDF;
r^2 carries the shader's DIVISOR_EPSILON 1e-5).  Every other path escapes to
the black environment, so the path tracer's pixel is the integral of L over
the pixel's footprint (pathtracer.comp:39-46 jitters the primary ray over
[x - 0.5, x + 0.5) around pixel x), which the test evaluates by 8x8
supersampling.  The light tracer
must converge to the same image (its point-light emission, lighttracer.comp:
75-93, is what it pins).
"""
import numpy as np
import pytest

A = 0.6  # 153 / 255: exact through the UNORM8 quantization of the albedo
LIGHT = np.array([0.3, 1.5, -0.2])
INTENSITY = np.array([2.0, 3.0, 4.0])
W = H = 32


def plane_scene(rt):
    pos = np.array([[-20, 0, -20], [20, 0, -20], [20, 0, 20], [-20, 0, 20]], np.float32)
    nrm = np.tile(np.array([[0, 1, 0]], np.float32), (4, 1))
    m = rt.default_material("Plane", albedo=(A, A, A), specular=(0, 0, 0))
    hs = rt.HostScene.build(pos, np.array([[0, 2, 1], [0, 3, 2]]), normals=nrm, materials=[m])
    hs.add_point_light(LIGHT, INTENSITY)
    return hs


def closed_form(cam, sub=8):
    U, V, Wv = (np.array(getattr(cam, k), np.float64) for k in ("u", "v", "w"))
    o = np.array(cam.position[:], np.float64)
    out = np.zeros((H, W, 3))
    ys, xs = np.mgrid[0:H, 0:W]
    for j in range(sub):
        for i in range(sub):
            sx = (xs + (i + 0.5) / sub - 0.5) / W * 2 - 1
            sy = (ys + (j + 0.5) / sub - 0.5) / H * 2 - 1
            d = sx[..., None] * U + sy[..., None] * V + Wv
            t = -o[1] / d[..., 1]
            x = o + t[..., None] * d
            l = LIGHT - x
            r2 = (l * l).sum(-1) + 1e-5
            cos = np.clip(l[..., 1] / np.sqrt(r2), 0, None)
            out += (A / np.pi) * INTENSITY * (cos / r2)[..., None]
    return out / (sub * sub)


@pytest.fixture(scope="module")
def setup(rt, oracle):
    hs = plane_scene(rt)
    sc = oracle.Scene(hs.arrays())
    cam = rt.camera([0.0, 3.0, 0.5], [0.0, 0.0, 0.0], [0, 1, 0], 40, W, H)
    return hs, sc, cam, closed_form(cam)


def test_pathtracer_point_light_closed_form(rt, oracle, setup):
    hs, sc, cam, want = setup
    sampler = rt.LightSampler(hs, 123)
    acc = np.zeros((H, W, 4), np.float32)
    n = 64
    for it in range(n):
        oracle.pathtrace(sc, sampler.generate(128), cam, W, H, rt.wanghash(it), 3, acc)
    got = acc[..., :3] / n
    assert want.min() > 0.03
    assert np.allclose(got.mean((0, 1)), want.mean((0, 1)), rtol=2e-3), (got.mean((0, 1)), want.mean((0, 1)))
    assert np.allclose(got, want, rtol=0.02)


def test_lighttracer_point_light_closed_form(rt, oracle, setup):
    hs, sc, cam, want = setup
    sampler = rt.LightSampler(hs, 321)
    acc = np.zeros((H, W, 4), np.float32)
    n = 64
    for it in range(n):
        oracle.lighttrace(sc, sampler.generate(256), cam, W, H, rt.wanghash(it), 3, acc)
    # showlightcaches.comp:9-65 also splats every light sample the camera sees
    # into its pixel, point lights included (with the normal decoded from the
    # marker value), so the pixel under the light is left out; so is the border
    # row, which takes half footprints (screenreproject.glsl)
    U, V, Wv = (np.array(getattr(cam, k), np.float64) for k in ("u", "v", "w"))
    c = np.array(cam.position[:], np.float64) - LIGHT
    sx = -U.dot(c) / -Wv.dot(c) / U.dot(U)
    sy = -V.dot(c) / -Wv.dot(c) / V.dot(V)
    lx, ly = int((sx * 0.5 + 0.5) * W), int((sy * 0.5 + 0.5) * H)
    keep = np.zeros((H, W), bool)
    keep[1:-1, 1:-1] = True
    keep[ly, lx] = False
    got = acc[..., :3][keep] / n
    ref = want[keep]
    assert acc[ly, lx, :3].min() / n > 100 * want[ly, lx].max()  # the splat is there
    assert np.allclose(got.mean(0), ref.mean(0), rtol=0.01), (got.mean(0), ref.mean(0))
