"""The oracle's material textures (SampleMaterialData, materials.glsl:19-30,
with the scene's bilinear REPEAT sampler, scene.cpp:16-20) pinned by closed
forms:

- a texture whose texels are a linear function of their centres reproduces
  that function exactly at interior coordinates, and REPEAT wraps both the
  coordinate (u and u + 1 agree) and the filter (u = 0 lies halfway between
  the last and the first column);
- a Lambertian plane under a point light with an albedo texture that is a
  linear ramp in u renders (alpha + beta u) / pi * I * cos / r^2 per pixel,
  where u is the plane's interpolated texture coordinate.
"""
import numpy as np
import pytest

from test_oracle_pointlight import INTENSITY, LIGHT


@pytest.fixture
def ramp(oracle):
    w, h = 16, 8
    xs = (np.arange(w) + 0.5) / w
    ys = (np.arange(h) + 0.5) / h
    tex = np.zeros((h, w, 4), np.float32)
    tex[..., 0] = 0.25 + 0.5 * xs[None, :]
    tex[..., 1] = 0.75 - 0.5 * ys[:, None]
    tex[..., 2] = 0.5
    tex[..., 3] = 2.0 * xs[None, :] + ys[:, None]
    oracle.set_textures(np.array([[0, -1, -1, -1]]), [tex])
    yield tex
    oracle.set_textures(None, None)


def test_linear_texture_is_reproduced(oracle, ramp):
    rng = np.random.default_rng(4)
    uv = np.stack([rng.uniform(0.5 / 16, 1 - 0.5 / 16, 4000), rng.uniform(0.5 / 8, 1 - 0.5 / 8, 4000)], 1)
    got = oracle.texture_samples(0, uv).astype(np.float64)
    u, v = uv[:, 0].astype(np.float32).astype(np.float64), uv[:, 1].astype(np.float32).astype(np.float64)
    assert np.allclose(got[:, 0], 0.25 + 0.5 * u, atol=2e-6)
    assert np.allclose(got[:, 1], 0.75 - 0.5 * v, atol=2e-6)
    assert np.allclose(got[:, 2], 0.5, atol=1e-7)
    assert np.allclose(got[:, 3], 2.0 * u + v, atol=4e-6)


def test_repeat_wraps_coordinates_and_filter(oracle, ramp):
    rng = np.random.default_rng(5)
    uv = rng.uniform(0, 1, (1000, 2)).astype(np.float32)
    a = oracle.texture_samples(0, uv)
    b = oracle.texture_samples(0, uv + np.float32(1))
    c = oracle.texture_samples(0, uv - np.float32(3))
    assert np.allclose(a, b, atol=1e-5) and np.allclose(a, c, atol=1e-5)
    # u = 0 is halfway between column 15 and column 0 (texel centres at 15.5/16 and 0.5/16)
    mid = oracle.texture_samples(0, np.array([[0.0, 0.5]], np.float32))[0, 0]
    assert abs(mid - 0.5 * (ramp[3, 15, 0] + ramp[3, 0, 0])) < 1e-6
    # NaN coordinates sample nothing
    assert (oracle.texture_samples(0, np.array([[np.nan, 0.5]], np.float32)) == 0).all()


def test_plane_with_albedo_ramp_under_point_light(rt, oracle):
    w = h = 32
    pos = np.array([[-2, 0, -2], [2, 0, -2], [2, 0, 2], [-2, 0, 2]], np.float32)
    nrm = np.tile(np.array([[0, 1, 0]], np.float32), (4, 1))
    # u runs from 0.1 at x = -2 to 0.9 at x = +2; v is constant 0.5
    tc = np.array([[0.1, 0.5], [0.9, 0.5], [0.9, 0.5], [0.1, 0.5]], np.float32)
    m = rt.default_material("Plane", albedo=(0.5, 0.5, 0.5), specular=(0, 0, 0))
    hs = rt.HostScene.build(pos, np.array([[0, 2, 1], [0, 3, 2]]), normals=nrm, texcoords=tc, materials=[m])
    hs.add_point_light(LIGHT, INTENSITY)
    sc = oracle.Scene(hs.arrays())
    tw = 64
    xs = (np.arange(tw) + 0.5) / tw
    tex = np.ones((1, tw, 4), np.float32)
    tex[0, :, 0:3] = (0.2 + 0.6 * xs)[:, None]  # albedo(u) = 0.2 + 0.6 u on all channels
    cam = rt.camera([0.0, 3.0, 0.5], [0.0, 0.0, 0.0], [0, 1, 0], 40, w, h)
    sampler = rt.LightSampler(hs, 123)
    acc = np.zeros((h, w, 4), np.float32)
    n = 64
    oracle.set_textures(np.array([[0, -1, -1, -1]]), [tex])
    try:
        for it in range(n):
            oracle.pathtrace(sc, sampler.generate(128), cam, w, h, rt.wanghash(it), 3, acc)
    finally:
        oracle.set_textures(None, None)
    got = acc[..., :3] / n
    # closed form, 8x8 supersampled over the footprint [x - 0.5, x + 0.5) of each pixel
    U, V, W = (np.array(getattr(cam, k), np.float64) for k in ("u", "v", "w"))
    o = np.array(cam.position[:], np.float64)
    want = np.zeros((h, w, 3))
    ys, xs_ = np.mgrid[0:h, 0:w]
    for j in range(8):
        for i in range(8):
            sx = (xs_ + (i + 0.5) / 8 - 0.5) / w * 2 - 1
            sy = (ys + (j + 0.5) / 8 - 0.5) / h * 2 - 1
            d = sx[..., None] * U + sy[..., None] * V + W
            x = o + (-o[1] / d[..., 1])[..., None] * d
            u = 0.1 + 0.8 * (x[..., 0] + 2) / 4
            l = LIGHT - x
            r2 = (l * l).sum(-1) + 1e-5
            cos = np.clip(l[..., 1] / np.sqrt(r2), 0, None)
            want += ((0.2 + 0.6 * u) / np.pi * cos / r2)[..., None] * INTENSITY
    want /= 64
    inside = np.zeros((h, w), bool)
    inside[2:-2, 2:-2] = True  # the plane's edges lie outside the view at this distance; keep away anyway
    assert np.allclose(got[inside], want[inside], rtol=0.02)
    assert np.allclose(got[inside].mean(0), want[inside].mean(0), rtol=2e-3)
