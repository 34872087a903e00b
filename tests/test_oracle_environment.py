"""The oracle's environment map (pathtracer.comp:56-58, the SRGB8 cube map of
renderersystem.cpp:173-186) pinned by closed forms:

- the sRGB decode at its end points, on its linear segment and at the knee;
- constant faces give that constant everywhere; one colour per face gives the
  face of the major axis;
- a map filled with a smooth function of direction, sampled at texel centres
  through the inverse of the OpenGL face table (written here from the
  specification, independently of the forward lookup), is reproduced by the
  bilinear lookup to within interpolation error: a wrong face, a flipped or
  swapped (s, t) axis or a half-texel shift fails it;
- a Lambertian plane under a constant sky of radiance Le renders Le where the
  camera sees the sky and a * Le on the plane (cosine sampling of the upper
  hemisphere, every bounce escapes).
"""
import numpy as np
import pytest


def test_srgb_decode(oracle):
    faces = np.arange(256, dtype=np.uint8).reshape(1, 16, 16, 1).repeat(6, 0).repeat(3, 3)
    lin = oracle.srgb_decode(faces)[0, ..., 0].reshape(-1).astype(np.float64)
    assert lin[0] == 0.0 and lin[255] == 1.0
    assert abs(lin[10] - 10 / 255 / 12.92) < 1e-8  # linear segment below 0.04045
    assert abs(lin[11] - ((11 / 255 + 0.055) / 1.055) ** 2.4) < 1e-7
    assert abs(lin[188] - 0.5029) < 2e-4  # 188 is the sRGB code of linear 0.5
    assert (np.diff(lin) > 0).all()


def face_dirs(size, shift=0.5):
    """Direction of the centre of texel (i, j) of each face, faces x-, x+, y-, y+, z-, z+:
    the inverse of the OpenGL table (+x: sc = -z, tc = -y; -x: sc = z, tc = -y;
    +y: sc = x, tc = z; -y: sc = x, tc = -z; +z: sc = x, tc = -y; -z: sc = -x, tc = -y).
    shift 0 gives the texel corners instead of the centres."""
    c = (np.arange(size) + shift) / size * 2 - 1
    tc, sc = np.meshgrid(c, c, indexing="ij")  # row j is t, column i is s
    one = np.ones_like(sc)
    d = np.stack([
        np.stack([-one, -tc, sc], -1),   # x-
        np.stack([one, -tc, -sc], -1),   # x+
        np.stack([sc, -one, -tc], -1),   # y-
        np.stack([sc, one, tc], -1),     # y+
        np.stack([-sc, -tc, -one], -1),  # z-
        np.stack([sc, -tc, one], -1),    # z+
    ])
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def random_dirs(n, seed):
    d = np.random.default_rng(seed).normal(size=(n, 3))
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


def test_constant_and_per_face_maps(oracle):
    tex = np.ones((6, 8, 8, 4), np.float32)
    tex[..., :3] = [0.25, 0.5, 0.75]
    d = random_dirs(5000, 1)
    assert (oracle.env_lookups(tex, d) == np.array([0.25, 0.5, 0.75], np.float32)).all()
    for f in range(6):
        tex[f, ..., :3] = f + 1
    got = oracle.env_lookups(tex, d)[:, 0]
    axis = np.abs(d).argmax(1)
    want = 2 * axis + (d[np.arange(len(d)), axis] >= 0) + 1
    assert (got == want).all()
    for f, v in enumerate([(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]):
        assert oracle.env_lookups(tex, np.array([v], np.float32))[0, 0] == f + 1


def test_smooth_map_is_reproduced(oracle):
    size = 64
    q = random_dirs(20000, 2)
    errs = {}
    for name, shift in (("centres", 0.5), ("shifted", 0.0)):
        tex = np.ones((6, size, size, 4), np.float32)
        tex[..., :3] = 0.5 + 0.5 * face_dirs(size, shift)
        errs[name] = np.abs(oracle.env_lookups(tex, q) - (0.5 + 0.5 * q)).max()
    # bilinear over texel centres: second-order inside a face, clamped to the
    # face within half a texel of its edges
    assert errs["centres"] < 4e-3, errs
    assert errs["shifted"] > 2 * errs["centres"], errs


def test_lambertian_plane_under_constant_sky(rt, oracle):
    faces = np.full((6, 4, 4, 3), [140, 203, 243], np.uint8)
    tex = oracle.srgb_decode(faces)
    le = tex[0, 0, 0, :3].astype(np.float64)
    pos = np.array([[-50, 0, -50], [50, 0, -50], [50, 0, 50], [-50, 0, 50]], np.float32)
    nrm = np.tile(np.array([[0, 1, 0]], np.float32), (4, 1))
    m = rt.default_material("Plane", albedo=(0.6, 0.6, 0.6), specular=(0, 0, 0))
    hs = rt.HostScene.build(pos, np.array([[0, 2, 1], [0, 3, 2]]), normals=nrm, materials=[m])
    sc = oracle.Scene(hs.arrays())
    w = h = 32
    cam = rt.camera([0.0, 1.0, 0.0], [0.0, 0.6, -1.0], [0, 1, 0], 70, w, h)
    lights = rt.LightSampler(hs, 1).generate(128)  # no emitters: black samples
    acc = np.zeros((h, w, 4), np.float32)
    n = 96
    oracle.set_environment(tex)
    try:
        for it in range(n):
            oracle.pathtrace(sc, lights, cam, w, h, rt.wanghash(it), 3, acc)
    finally:
        oracle.set_environment(None)
    img = acc[..., :3] / n
    # the pixel centre's ray decides sky or plane except on the horizon rows
    ys, xs = np.mgrid[0:h, 0:w]
    U, V, W = (np.array(getattr(cam, k), np.float64) for k in ("u", "v", "w"))
    dy = ((ys + 0.5 - 0.5) / h * 2 - 1) * V[1] + W[1]
    sky = dy > 0.05
    plane = dy < -0.05
    assert sky.sum() > 50 and plane.sum() > 50
    assert np.allclose(img[sky], le.astype(np.float32), rtol=1e-6)
    assert np.allclose(img[plane].mean(0), 0.6 * le, rtol=0.015), (img[plane].mean(0), 0.6 * le)
