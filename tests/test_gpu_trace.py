"""GPU traversal parity against the CPU oracle (through the C ABI).

Reference kernel: bit-exact (slot, t, barycentrics) on every ray.
Wide kernel: same slot, t and barycentrics except on rays the FP64 classifier
places within 1e-6 (relative) of an edge, a tie or the t bounds; those must be
rare.  Full-size (2^24 rays) checks use size-independent properties.
"""
import numpy as np
import pytest

from gpugi_amd import workloads

pytestmark = pytest.mark.gpu


def rays_struct(rt, r):
    return np.ascontiguousarray(r, dtype=np.float32).view(rt.RAY_DTYPE).reshape(-1)


def hits_int(h):
    return np.ascontiguousarray(h).view(np.int32).reshape(-1, 4)


def scene_rays(rt, oracle, hs, cam, w, h, n):
    sc = oracle.Scene(hs.arrays())
    prim = oracle.raygen_primary(cam, w, h, workloads.MASTER_SEED, n)
    ph = oracle.trace(sc, prim)
    diff = oracle.raygen_diffuse(sc, prim, ph, workloads.MASTER_SEED + 1, n)
    return sc, prim, ph, diff


def check_wide(oracle, sc, rays, ref, wide, max_frac):
    diff = np.nonzero((ref != wide).any(1))[0]
    if len(diff):
        flags = oracle.classify(sc, rays[diff], ref[diff, 0], wide[diff, 0])
        bad = diff[flags == 0]
        assert len(bad) == 0, [(int(i), ref[i].tolist(), wide[i].tolist()) for i in bad[:5]]
    assert len(diff) <= max(2, max_frac * len(rays)), len(diff)


@pytest.mark.parametrize("name", ["cornell", "hf64", "hf1024"])
def test_closest_hit_parity(rt, oracle, ctx, name):
    print("case", name, flush=True)
    if name == "cornell":
        hs, cam, w, h, n = rt.HostScene.cornell(), workloads.cornell_camera(256), 256, 256, 65536
    elif name == "hf64":
        hs, cam, w, h, n = rt.HostScene.heightfield(64, workloads.MASTER_SEED), workloads.heightfield_camera(), 1920, 1080, 65536
    else:
        hs, cam, w, h, n = rt.HostScene.heightfield(1024, workloads.MASTER_SEED), workloads.heightfield_camera(), 1920, 1080, 32768
    ctx.load_scene(hs)
    sc, prim, ph, diff = scene_rays(rt, oracle, hs, cam, w, h, n)
    for rays in (prim, diff):
        want = oracle.trace(sc, rays)
        got_ref = hits_int(ctx.trace(rays_struct(rt, rays), rt.CLOSEST_HIT, rt.KERNEL_REFERENCE))
        assert np.array_equal(got_ref, want), np.nonzero((got_ref != want).any(1))[0][:10]
        got_wide = hits_int(ctx.trace(rays_struct(rt, rays), rt.CLOSEST_HIT, rt.KERNEL_WIDE))
        check_wide(oracle, sc, rays, want, got_wide, 1e-3)


def test_any_hit_parity_shadow_rays(rt, oracle, ctx):
    hs = rt.HostScene.heightfield(64, workloads.MASTER_SEED)
    ctx.load_scene(hs)
    sc, prim, ph, diff = scene_rays(rt, oracle, hs, workloads.heightfield_camera(), 1920, 1080, 32768)
    lights = workloads.heightfield_lights()
    shadow = oracle.raygen_shadow(prim, ph, lights, 32768)
    want = oracle.trace(sc, shadow, any_hit=True)
    assert 0.01 < want[:, 0].mean() < 0.99, "the shadow set must contain both outcomes"
    ref = hits_int(ctx.trace(rays_struct(rt, shadow), rt.ANY_HIT, rt.KERNEL_REFERENCE))
    assert np.array_equal(ref[:, 0], want[:, 0])
    wide = hits_int(ctx.trace(rays_struct(rt, shadow), rt.ANY_HIT, rt.KERNEL_WIDE))
    diff_idx = np.nonzero(wide[:, 0] != want[:, 0])[0]
    assert len(diff_idx) <= max(2, 1e-3 * len(shadow))
    if len(diff_idx):
        # the occluder the wide walk found (or missed) must sit on an edge / the t bounds
        full = oracle.bruteforce(sc, shadow[diff_idx], any_hit=False)
        flags = oracle.classify(sc, shadow[diff_idx], full[:, 0], full[:, 0])
        assert (flags != 0).all()


def test_visit_counts_match_oracle(rt, oracle, ctx):
    hs = rt.HostScene.heightfield(64, workloads.MASTER_SEED)
    ctx.load_scene(hs)
    sc, prim, ph, diff = scene_rays(rt, oracle, hs, workloads.heightfield_camera(), 1920, 1080, 16384)
    for rays, anyh in ((prim, False), (diff, False), (diff, True)):
        _, nb, nt = oracle.trace(sc, rays, any_hit=anyh, counts=True)
        gb, gt = ctx.trace_counts(rays_struct(rt, rays), rt.ANY_HIT if anyh else rt.CLOSEST_HIT)
        assert np.array_equal(gb, nb) and np.array_equal(gt, nt)


def test_raygen_parity(rt, oracle, ctx):
    hs = rt.HostScene.heightfield(64, workloads.MASTER_SEED)
    ctx.load_scene(hs)
    n = 65536
    cam = workloads.heightfield_camera()
    sc = oracle.Scene(hs.arrays())
    prim_o = oracle.raygen_primary(cam, 1920, 1080, workloads.MASTER_SEED, n)
    ph_o = oracle.trace(sc, prim_o)
    lights = workloads.heightfield_lights()
    bufs = [ctx.alloc(n * 32) for _ in range(4)] + [ctx.alloc(n * 16)]
    prim_b, diff_b, shad_b, _, hits_b = bufs
    ctx.raygen_primary(cam, 1920, 1080, workloads.MASTER_SEED, prim_b, n)
    ctx.trace_device(prim_b, hits_b, n, rt.CLOSEST_HIT, rt.KERNEL_REFERENCE)
    ctx.raygen_diffuse(prim_b, hits_b, n, workloads.MASTER_SEED + 1, diff_b, n)
    ctx.raygen_shadow(prim_b, hits_b, n, lights, shad_b, n)
    got_p = prim_b.download(np.float32, n * 8).reshape(n, 8)
    assert np.array_equal(got_p.view(np.uint32), prim_o.view(np.uint32))
    assert np.array_equal(hits_b.download(np.int32, n * 4).reshape(n, 4), ph_o)
    diff_o = oracle.raygen_diffuse(sc, prim_o, ph_o, workloads.MASTER_SEED + 1, n)
    assert np.array_equal(diff_b.download(np.float32, n * 8).reshape(n, 8).view(np.uint32), diff_o.view(np.uint32))
    shad_o = oracle.raygen_shadow(prim_o, ph_o, lights, n)
    assert np.array_equal(shad_b.download(np.float32, n * 8).reshape(n, 8).view(np.uint32), shad_o.view(np.uint32))
    for b in bufs:
        b.free()


def test_edge_cases(rt, oracle, ctx):
    hs = rt.HostScene.cornell()
    ctx.load_scene(hs)
    sc = oracle.Scene(hs.arrays())
    # zero rays
    assert len(ctx.trace(np.zeros(0, rt.RAY_DTYPE))) == 0
    r = np.zeros((7, 8), np.float32)
    r[:, 3] = 3.4e38
    r[0, 0:3], r[0, 4:7] = [0, 1, 0], [0, -1, 0]       # axis-parallel, hits the floor
    r[1, 0:3], r[1, 4:7] = [0, 1, 0], [1, 0, 0]        # axis-parallel, hits the red/green wall
    r[2, 0:3], r[2, 4:7] = [-1, 1, 0.5], [0, 0, -1]    # origin exactly on the wall plane
    r[3, 0:3], r[3, 4:7] = [0, 1, 10], [0, 0, 1]       # points away from everything
    r[4, 0:3], r[4, 4:7] = [0, 1, 0], [0, -1, 0]
    r[4, 3] = 0.0                                      # tmax 0: nothing can be accepted
    r[5, 0:3], r[5, 4:7] = [0.1, 1.99, 0.1], [0, 1, 0]  # just below the light, up through it
    r[6, 0:3], r[6, 4:7] = [0.2, 0.5, 0.2], [1e-30, -1, 1e-30]  # near-axis direction
    want = oracle.trace(sc, r)
    for k in (rt.KERNEL_REFERENCE, rt.KERNEL_WIDE):
        got = hits_int(ctx.trace(rays_struct(rt, r), rt.CLOSEST_HIT, k))
        if k == rt.KERNEL_REFERENCE:
            assert np.array_equal(got, want)
        else:
            check_wide(oracle, sc, r, want, got, 0.3)
    assert want[3, 0] == -1 and want[4, 0] == -1 and want[0, 0] >= 0


def test_full_size_incoherent_set_properties(rt, oracle, ctx):
    """2^24 incoherent rays on the 2M-triangle heightfield: every hit is a valid
    slot with 1e-4 < t, wide and reference kernels agree on a strided sample,
    and two launches give identical bits."""
    hs = rt.HostScene.heightfield(1024, workloads.MASTER_SEED)
    ctx.load_scene(hs)
    a = hs.arrays()
    valid = a["triangles"][:, 0] != a["triangles"][:, 1]
    R = workloads.HeightfieldRays(ctx)
    n = R.n
    ctx.trace_device(R.incoherent, R.hits, n, rt.CLOSEST_HIT, rt.KERNEL_WIDE)
    h1 = R.hits.download(np.int32, n * 4).reshape(n, 4)
    ctx.trace_device(R.incoherent, R.hits, n, rt.CLOSEST_HIT, rt.KERNEL_WIDE)
    h2 = R.hits.download(np.int32, n * 4).reshape(n, 4)
    assert np.array_equal(h1, h2)
    hit = h1[:, 0] >= 0
    assert hit.mean() > 0.005
    assert valid[h1[hit, 0]].all() and (h1[:, 0] >= -1).all()
    t = h1[hit, 1].view(np.float32)
    assert (t > 1e-4).all() and np.isfinite(t).all()
    rays = R.incoherent.download(np.float32, n * 8).reshape(n, 8)
    sample = rays[:: 4099]
    ref = hits_int(ctx.trace(rays_struct(rt, sample), rt.CLOSEST_HIT, rt.KERNEL_REFERENCE))
    sc = oracle.Scene(a)
    check_wide(oracle, sc, sample, ref, h1[:: 4099], 1e-3)
    R.free()


@pytest.mark.parametrize("leaf,stack,refill,wait", [(1, 4, 64, 1), (8, 4, 1, 64), (2, 24, 16, 32), (4, 8, 8, 16)])
def test_wide_configurations(rt, oracle, ctx, leaf, stack, refill, wait):
    """Every wide configuration gives the same answers: a 4-entry LDS stack
    spills to global memory on most rays, wait 1 runs a leaf pass as soon as a
    lane waits, wait 64 only when no lane can traverse."""
    print("case: wide config", leaf, stack, refill, wait, flush=True)
    hs = rt.HostScene.heightfield(64, workloads.MASTER_SEED)
    ctx.load_scene(hs)
    sc, prim, ph, diff = scene_rays(rt, oracle, hs, workloads.heightfield_camera(), 1920, 1080, 32768)
    ctx.set_wide_config(leaf, stack, refill, wait)
    cfg = ctx.wide_config()
    assert (cfg["max_leaf_triangles"], cfg["lds_stack_entries"], cfg["refill_lanes"], cfg["wait_lanes"]) == \
        (leaf, stack, refill, wait)
    if stack == 4:
        assert cfg["max_stack"] > stack  # the spill path is exercised
    for rays in (prim, diff):
        want = oracle.trace(sc, rays)
        got = hits_int(ctx.trace(rays_struct(rt, rays), rt.CLOSEST_HIT, rt.KERNEL_WIDE))
        check_wide(oracle, sc, rays, want, got, 1e-3)
    shadow = oracle.raygen_shadow(prim, ph, workloads.heightfield_lights(), 32768)
    want_any = oracle.trace(sc, shadow, any_hit=True)[:, 0]
    got_any = hits_int(ctx.trace(rays_struct(rt, shadow), rt.ANY_HIT, rt.KERNEL_WIDE))[:, 0]
    assert (got_any != want_any).sum() <= max(2, 1e-3 * len(shadow))
    ctx.set_wide_config(2, 10, 8, 16)


@pytest.mark.parametrize("query", ["closest", "any"])
def test_autotune_trace_keeps_answers(rt, oracle, ctx, query):
    """gpugi_autotune_trace searches leaf size, LDS stack depth, refill and
    wait thresholds on the given rays and keeps the fastest; the default
    configuration (leaf 2, stack 10, refill 8, wait 16) is one of the timed
    points, so the result is never slower than it on those rays.  After the
    search the scene's wide BVH is rebuilt for the chosen leaf size and the
    wide kernel must still give the oracle's answers."""
    print("case: autotune", query, flush=True)
    hs = rt.HostScene.heightfield(64, workloads.MASTER_SEED)
    ctx.load_scene(hs)
    sc, prim, ph, diff = scene_rays(rt, oracle, hs, workloads.heightfield_camera(), 1920, 1080, 32768)
    rays = diff if query == "closest" else oracle.raygen_shadow(prim, ph, workloads.heightfield_lights(), 32768)
    q = rt.CLOSEST_HIT if query == "closest" else rt.ANY_HIT
    n = len(rays)
    rb, hb = ctx.alloc(n * 32), ctx.alloc(n * 16)
    try:
        rb.upload(rays_struct(rt, rays))
        best = ctx.autotune_trace(rb, hb, n, q, repeats=1)
        assert 0 < best < 1e3
        cfg = ctx.wide_config()
        assert cfg["max_leaf_triangles"] in (1, 2, 4, 8)
        assert cfg["lds_stack_entries"] in (4, 8, 10, 12, 16, 24)
        assert cfg["refill_lanes"] in (8, 16, 32, 64)
        assert cfg["wait_lanes"] in (1, 8, 16, 24, 32, 48)
        ctx.trace_device(rb, hb, n, q, rt.KERNEL_WIDE)
        got = hb.download(np.int32, n * 4).reshape(n, 4)
        if q == rt.CLOSEST_HIT:
            check_wide(oracle, sc, rays, oracle.trace(sc, rays), got, 1e-3)
        else:
            want = oracle.trace(sc, rays, any_hit=True)[:, 0]
            assert (got[:, 0] != want).sum() <= max(2, 1e-3 * n)
    finally:
        ctx.set_wide_config(2, 10, 8, 16)


@pytest.mark.parametrize("scene", ["heightfield", "clutter"])
def test_full_size_sets_wide_against_reference_walk(rt, oracle, ctx, scene):
    """Every ray of the three 2^24-ray sets on the 2,097,152-triangle heightfield
    and on the 5,120,016-triangle clutter room (its camera and lights at
    1920x1080): the wide kernel's record equals the reference walk's (itself
    bit-exact with the oracle) except on rays the FP64 classifier flags, at most
    1e-3 of a set.  For any hit, a differing ray's closest hit (oracle walk)
    must be flagged."""
    if scene == "heightfield":
        hs = rt.HostScene.heightfield(1024, workloads.MASTER_SEED)
    else:
        hs = rt.HostScene.clutter(workloads.CLUTTER_SPHERES, workloads.MASTER_SEED)
    ctx.load_scene(hs)
    sc = oracle.Scene(hs.arrays())
    if scene == "heightfield":
        R = workloads.HeightfieldRays(ctx)
    else:
        w, h = workloads.HF_WIDTH, workloads.HF_HEIGHT
        R = workloads.SceneRays(ctx, workloads.clutter_camera(w, h), workloads.CLUTTER_LIGHTS, workloads.HF_RAYS,
                                workloads.MASTER_SEED, w, h)
    n = R.n
    try:
        for name, buf, query in (("incoherent", R.incoherent, rt.CLOSEST_HIT), ("coherent", R.primary, rt.CLOSEST_HIT),
                                 ("shadow", R.shadow, rt.ANY_HIT)):
            print("case: full-size %s %s set" % (scene, name), flush=True)
            ctx.trace_device(buf, R.hits, n, query, rt.KERNEL_REFERENCE)
            ref = R.hits.download(np.int32, n * 4).reshape(n, 4)
            ctx.trace_device(buf, R.hits, n, query, rt.KERNEL_WIDE)
            wide = R.hits.download(np.int32, n * 4).reshape(n, 4)
            if query == rt.CLOSEST_HIT:
                diff = np.nonzero((ref != wide).any(1))[0]
            else:
                diff = np.nonzero(ref[:, 0] != wide[:, 0])[0]
            print("  %d of %d rays differ" % (len(diff), n), flush=True)
            assert len(diff) <= 1e-3 * n
            if len(diff) == 0:
                continue
            rays = buf.download(np.float32, n * 8).reshape(n, 8)[diff]
            if query == rt.CLOSEST_HIT:
                flags = oracle.classify(sc, rays, ref[diff, 0], wide[diff, 0])
            else:
                closest = oracle.trace(sc, rays)[:, 0]
                flags = oracle.classify(sc, rays, closest, closest)
            assert (flags != 0).all(), [(int(diff[i]), ref[diff[i]].tolist(), wide[diff[i]].tolist())
                                        for i in np.nonzero(flags == 0)[0][:5]]
    finally:
        R.free()
