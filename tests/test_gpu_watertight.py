"""Watertight triangle mode on the GPU against the oracle's watertight test.

The escape-walk kernel with the watertight test matches the oracle walk bit
for bit; the wide kernel finds the same triangles (its boxes are
conservative, so it never loses a hit at a box boundary) and leaves no gap
for rays aimed exactly at shared vertices and edges of the heightfield."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def to_rays(rt, o, d):
    r = np.zeros(len(d), dtype=rt.RAY_DTYPE)
    r["origin"] = o
    r["dir"] = d
    r["tmax"] = rt.RAY_MAX
    return r


def as_hits(h):
    return h.view(np.int32).reshape(-1, 4)


def vertex_edge_rays(rt, a, cells, seed):
    pos = a["positions"]
    n1 = cells + 1
    targets = []
    for j in range(1, cells):
        for i in range(1, cells):
            k = j * n1 + i
            targets += [pos[k], (pos[k] + pos[k + 1]) * np.float32(0.5), (pos[k] + pos[k + n1]) * np.float32(0.5),
                        (pos[k] + pos[k + n1 + 1]) * np.float32(0.5)]
    targets = np.array(targets, np.float32)
    rng = np.random.default_rng(seed)
    o = targets + np.float32([0, 1, 0]) + (rng.random((len(targets), 3)) - 0.5).astype(np.float32)
    d = targets - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return to_rays(rt, o, d.astype(np.float32))


@pytest.fixture
def wt_ctx(rt):
    c = rt.Context(0)
    yield c
    c.close()


def test_escape_walk_watertight_bitexact(rt, oracle, wt_ctx):
    cells = 32
    hs = rt.HostScene.heightfield(cells, 9)
    a = hs.arrays()
    sc = oracle.Scene(a)
    wt_ctx.load_scene(hs)
    wt_ctx.set_triangle_test(rt.TRIANGLE_WATERTIGHT)
    rays = vertex_edge_rays(rt, a, cells, 1)
    rng = np.random.default_rng(3)
    extra = to_rays(rt, rng.random((4000, 3)).astype(np.float32) * np.float32([1, 0.5, 1]) + np.float32([0, 0.4, 0]),
                    (lambda v: (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32))(
                        rng.normal(size=(4000, 3)) - np.float32([0, 1.5, 0])))
    rays = np.concatenate([rays, extra])
    print("case: escape walk, watertight, %d rays" % len(rays), flush=True)
    g = as_hits(wt_ctx.trace(rays, rt.CLOSEST_HIT, rt.KERNEL_REFERENCE))
    ref = oracle.trace(sc, rays.view(np.float32).reshape(-1, 8), watertight=True)
    assert np.array_equal(g, ref)
    ga = as_hits(wt_ctx.trace(rays, rt.ANY_HIT, rt.KERNEL_REFERENCE))[:, 0]
    assert np.array_equal(ga, oracle.trace(sc, rays.view(np.float32).reshape(-1, 8), any_hit=True,
                                           watertight=True)[:, 0])


def test_wide_watertight_has_no_gaps(rt, oracle, wt_ctx):
    cells = 64
    hs = rt.HostScene.heightfield(cells, 12)
    a = hs.arrays()
    sc = oracle.Scene(a)
    wt_ctx.load_scene(hs)
    rays = vertex_edge_rays(rt, a, cells, 2)
    wt_ctx.set_triangle_test(rt.TRIANGLE_WATERTIGHT)
    print("case: wide, watertight, %d vertex/edge rays" % len(rays), flush=True)
    g = as_hits(wt_ctx.trace(rays, rt.CLOSEST_HIT, rt.KERNEL_WIDE))
    bf = oracle.bruteforce(sc, rays.view(np.float32).reshape(-1, 8), watertight=True)
    assert (g[:, 0] >= 0).all() and (bf[:, 0] >= 0).all()
    assert np.array_equal(g, bf)  # ties resolve to the smaller slot in both
    # the reference test leaves gaps on the same rays; switching back re-packs the wide BVH
    wt_ctx.set_triangle_test(rt.TRIANGLE_REFERENCE)
    gr = as_hits(wt_ctx.trace(rays, rt.CLOSEST_HIT, rt.KERNEL_WIDE))
    print("reference-test misses on the same rays: %d" % (gr[:, 0] < 0).sum())
    assert np.array_equal(gr[:, 0], oracle.bruteforce(sc, rays.view(np.float32).reshape(-1, 8))[:, 0])


def test_path_tracer_in_watertight_mode(rt, oracle, wt_ctx):
    from gpugi_amd import workloads
    hs = rt.HostScene.cornell()
    wt_ctx.load_scene(hs)
    wt_ctx.set_screen(32, 32)
    wt_ctx.set_camera(workloads.cornell_camera(32))
    wt_ctx.set_max_path_length(5)
    imgs = []
    for mode in (rt.TRIANGLE_REFERENCE, rt.TRIANGLE_WATERTIGHT):
        for kernel in (rt.KERNEL_REFERENCE, rt.KERNEL_WIDE):
            wt_ctx.set_triangle_test(mode)
            wt_ctx.set_integrator_kernel(kernel)
            wt_ctx.reset_accum()
            sampler = rt.LightSampler(hs, 123)
            for k in range(16):
                wt_ctx.set_light_samples(sampler.generate(128))
                wt_ctx.draw(rt.PT, rt.wanghash(k))
            imgs.append(wt_ctx.readback(32, 32)[0][..., :3])
    base = imgs[0].mean()
    for im in imgs[1:]:
        assert abs(im.mean() - base) < 0.02 * base
    # both traversals agree pixel for pixel unless one of the pixel's 16 paths met a
    # shared edge or corner (the room's walls meet at axis-aligned edges), where
    # equal distances resolve to different but equally valid triangles
    same = (imgs[2] == imgs[3]).all(-1).mean()
    assert same > 0.85
