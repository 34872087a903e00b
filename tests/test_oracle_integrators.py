"""Integrator oracles pinned by closed forms and by mutual consistency.

- PT white furnace: a closed cube whose walls all emit E with albedo a renders
  E * (1 + a + ... + a^L) for path length L (NEE at every vertex, emission at
  the first hit only, Russian roulette unbiased).
- LT and BPT agree with PT (image means within 2.5 %) on the Cornell room away
  from directly visible emitters and the half-footprint border row (the reference's light-sample
  splat divides by 2*pi and rounds pixel positions, showlightcaches.comp:61,
  screenreproject.glsl:288).
- BPT warm-up counters: NumLightCacheEntries - R*N equals LightPathLengthSum.
"""
import numpy as np
import pytest

from gpugi_amd import workloads


def furnace_scene(rt, **material):
    P = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float32)
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    pos, nrm, idx = [], [], []
    for f in faces:
        q = P[list(f)]
        c = q.mean(0)
        n = -c / np.linalg.norm(c)
        b = len(pos)
        for v in q:
            pos.append(v)
            nrm.append(n)
        # wound so that cross(v0 - v1, v0 - v2) points into the cube: the emission side
        idx += [[b, b + 2, b + 1], [b, b + 3, b + 2]]
    args = dict(albedo=(0.5, 0.5, 0.5), specular=(0, 0, 0), emissivity=(1.0, 2.0, 3.0))
    args.update(material)
    m = rt.default_material("Furnace", **args)
    return rt.HostScene.build(np.array(pos), np.array(idx), normals=np.array(nrm), materials=[m])


@pytest.mark.parametrize("path_length", [1, 3])
def test_pt_white_furnace(rt, oracle, path_length):
    hs = furnace_scene(rt)
    sc = oracle.Scene(hs.arrays())
    a = float(np.float32(128) / np.float32(255))  # albedo 0.5 through UNORM8
    cam = rt.camera([0.1, 0.2, 0.3], [0.5, -0.3, -1.0], [0, 1, 0], 60, 16, 16)
    sampler = rt.LightSampler(hs, 123)
    acc = np.zeros((16, 16, 4), np.float32)
    n = 160
    for it in range(n):
        oracle.pathtrace(sc, sampler.generate(128), cam, 16, 16, rt.wanghash(it), path_length, acc)
    got = acc[..., :3].reshape(-1, 3).mean(0) / n
    want = np.array([1.0, 2.0, 3.0]) * sum(a ** k for k in range(path_length + 1))
    assert np.allclose(got, want, rtol=0.03), (got, want)


@pytest.mark.parametrize("path_length", [1, 3])
def test_pt_mirror_furnace(rt, oracle, path_length):
    # Roughness 0 gives the exponent 1e20 >= 500: no next-event estimation
    # (pathtracer.comp:81), so emission is counted at every vertex the path
    # reaches; each bounce keeps rho = s + (1 - s) * a on average (mirror lobe
    # with probability s, diffuse otherwise), and the last of the L vertices
    # cannot see the emission of a next one: E * (1 + rho + ... + rho^(L-1)).
    hs = furnace_scene(rt, specular=(0.5, 0.5, 0.5), roughness=0.0)
    sc = oracle.Scene(hs.arrays())
    a = float(np.float32(128) / np.float32(255))
    rho = 0.5 + 0.5 * a
    cam = rt.camera([0.1, 0.2, 0.3], [0.5, -0.3, -1.0], [0, 1, 0], 60, 16, 16)
    sampler = rt.LightSampler(hs, 123)
    acc = np.zeros((16, 16, 4), np.float32)
    n = 200
    for it in range(n):
        oracle.pathtrace(sc, sampler.generate(128), cam, 16, 16, rt.wanghash(it), path_length, acc)
    got = acc[..., :3].reshape(-1, 3).mean(0) / n
    want = np.array([1.0, 2.0, 3.0]) * sum(rho ** k for k in range(path_length))
    assert np.allclose(got, want, rtol=0.02), (got, want)


def emitter_mask(rt, oracle, sc, cam, w, h):
    """Pixels whose centre ray sees an emitter, dilated by one pixel in every direction, plus the border."""
    ys, xs = np.mgrid[0:h, 0:w]
    U, V, W = (np.array(getattr(cam, k), np.float64) for k in ("u", "v", "w"))
    sx = (xs / w) * 2 - 1
    sy = (ys / h) * 2 - 1
    d = sx[..., None] * U + sy[..., None] * V + W
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    rays = np.zeros((h * w, 8), np.float32)
    rays[:, 0:3] = cam.position[:]
    rays[:, 3] = 3.4e38
    rays[:, 4:7] = d.reshape(-1, 3)
    hits = oracle.trace(sc, rays)
    tri = sc.a["triangles"]
    em = sc.a["materials"]["emissivity"].sum(-1) > 0
    lit = np.zeros(h * w, bool)
    ok = hits[:, 0] >= 0
    lit[ok] = em[tri[hits[ok, 0], 3]]
    lit = lit.reshape(h, w)
    # dilated over the 8-neighbourhood: a pixel diagonal to an emitter pixel can
    # still see part of the emitter, which only the path tracer renders
    dil = lit.copy()
    dil[1:] |= lit[:-1]
    dil[:-1] |= lit[1:]
    row = dil.copy()
    dil[:, 1:] |= row[:, :-1]
    dil[:, :-1] |= row[:, 1:]
    dil[0, :] = dil[-1, :] = True
    dil[:, 0] = dil[:, -1] = True
    return dil


def test_lt_and_bpt_agree_with_pt(rt, oracle):
    hs = rt.HostScene.cornell()
    sc = oracle.Scene(hs.arrays())
    w = h = 32
    L = 5
    cam = workloads.cornell_camera(w)
    mask = ~emitter_mask(rt, oracle, sc, cam, w, h)
    n = 48
    imgs = {}
    for name, nls in (("pt", 128), ("lt", 256), ("bpt", 128)):
        sampler = rt.LightSampler(hs, 123)
        acc = np.zeros((h, w, 4), np.float32)
        cap = None
        if name == "bpt":
            counts = [oracle.bpt_warmup(sc, sampler.generate(nls), cam, w, h, rt.wanghash(k + 1), L) for k in range(8)]
            mean = sum(c[0] for c in counts) // 8
            cap = mean + mean // 10
        for it in range(n):
            ls = sampler.generate(nls)
            if name == "pt":
                oracle.pathtrace(sc, ls, cam, w, h, rt.wanghash(it), L, acc)
            elif name == "lt":
                oracle.lighttrace(sc, ls, cam, w, h, rt.wanghash(it), L, acc)
            else:
                acc, _, _ = oracle.bpt_draw(sc, ls, cam, w, h, rt.wanghash(it), L, cap, acc)
        imgs[name] = acc[..., :3].sum(-1)[mask].mean() / n
    assert abs(imgs["lt"] / imgs["pt"] - 1) < 0.025, imgs
    assert abs(imgs["bpt"] / imgs["pt"] - 1) < 0.025, imgs


def test_bpt_warmup_and_cache_structure(rt, oracle):
    hs = rt.HostScene.cornell()
    sc = oracle.Scene(hs.arrays())
    w = h = 32
    cam = workloads.cornell_camera(w)
    ls = rt.LightSampler(hs, 123).generate(128)
    R = max(64, (w * h // 128 // 64) * 64)
    entries, stored = oracle.bpt_warmup(sc, ls, cam, w, h, 7, 5)
    assert entries - stored == 128 * R and stored > 0
    acc, cache, n = oracle.bpt_draw(sc, ls, cam, w, h, 7, 5, entries + 100)
    assert n == entries  # the draw with the same seed stores the same vertices
    c = cache[:n]
    mat = c[:, 3].view(np.int32)
    plen = c[:, 11].view(np.int32)
    # initial vertices: one per light path, material -1, length 0, flux I / (2 pi) / R
    assert (mat[:128 * R] == -1).all() and (plen[:128 * R] == 0).all()
    ls_idx = np.arange(128 * R) // R
    want = (ls["intensity"][ls_idx] / np.float32(6.28318530717958)) / np.float32(R)
    assert np.array_equal(c[:128 * R, 0:3], want.astype(np.float32))
    # appended vertices: path lengths 1..5, materials of the scene, only Phong exponent < 100
    assert ((plen[128 * R:] >= 1) & (plen[128 * R:] <= 5)).all()
    assert (mat[128 * R:] >= 0).all()
    # capacity clips the appended part but keeps the counter
    acc2, cache2, n2 = oracle.bpt_draw(sc, ls, cam, w, h, 7, 5, 128 * R + 10)
    assert n2 == entries
    assert np.array_equal(cache2[:128 * R + 10].view(np.uint32), cache[:128 * R + 10].view(np.uint32))


def test_parallel_lt_and_bpt_equal_the_sequential_draws(rt, oracle):
    """The thread-split light tracer and BPT draws (used on the large scenes)
    give the sequential draws' cache bit for bit and their images up to the
    order of the light splats."""
    hs = rt.HostScene.cornell()
    sc = oracle.Scene(hs.arrays())
    w = h = 32
    cam = workloads.cornell_camera(w)
    ls = rt.LightSampler(hs, 5).generate(128)
    seq = oracle.lighttrace(sc, ls, cam, w, h, 17, 5)
    par = oracle.lighttrace_parallel(sc, ls, cam, w, h, 17, 5, workers=4)
    assert np.allclose(par, seq, rtol=1e-5, atol=1e-7) and seq[..., :3].max() > 0
    entries, _ = oracle.bpt_warmup(sc, ls, cam, w, h, 17, 5)
    cap = entries - 50  # some vertices past the capacity
    acc, cache, n = oracle.bpt_draw(sc, ls, cam, w, h, 17, 5, cap)
    pacc, pcache, pn = oracle.bpt_draw_parallel(sc, ls, cam, w, h, 17, 5, cap, workers=4)
    assert pn == n and np.array_equal(pcache.view(np.uint32), cache.view(np.uint32))
    assert np.allclose(pacc, acc, rtol=1e-5, atol=1e-7)
