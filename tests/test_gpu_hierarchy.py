"""HierarchyImportance passes on the GPU against the CPU oracle.

gpugi_hierarchy_update_importance (acquisition2.comp + propagation) and
gpugi_hierarchy_compute_materials (sampleleafmaterials.comp + pullmaterials.comp)
must equal the oracle bit for bit: the acquisition sums powers of two in exact
fixed point, which equals the oracle's FP32 sums while a node sees fewer than
2^16 rays (those are exact too), and the propagation and material passes are
the same FP32 and half operations in the same order.  Beyond that range the
sums stay exact, so two updates give the same bits.  The importance then drives gpugi_trace_importance_break from
the device and must give the oracle's importance-break walk.
"""
import numpy as np
import pytest

from gpugi_amd import workloads

pytestmark = pytest.mark.gpu


def same(a, b):
    """Bitwise equal, any NaN equal to any NaN (0/0 has a different payload on the GPU)."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    nan = np.isnan(a) & np.isnan(b)
    return np.array_equal(np.where(nan, 0, a).view(np.uint32), np.where(nan, 0, b).view(np.uint32))


@pytest.mark.parametrize("name,bvh", [("cornell", "aabox"), ("cornell", "obox"), ("hf64", "aabox")])
def test_importance_and_materials_match_oracle(rt, oracle, ctx, name, bvh):
    print("case: hierarchy %s %s" % (name, bvh), flush=True)
    if name == "cornell":
        hs, cam, w, h = rt.HostScene.cornell(), workloads.cornell_camera(32), 32, 32
    else:
        hs, cam, w, h = rt.HostScene.heightfield(64, workloads.MASTER_SEED), workloads.heightfield_camera(64, 40), 64, 40
    ctx.set_bvh_type(rt.BVH_OBOX if bvh == "obox" else rt.BVH_AABOX)
    try:
        ctx.load_scene(hs)
        arrays = hs.arrays()
        if bvh == "obox":
            arrays["obox"] = hs.obox_hierarchy()
        sc = oracle.Scene(arrays)
        ctx.set_screen(w, h)
        ctx.set_camera(cam)
        ls = rt.LightSampler(hs, 11).generate(128)
        ctx.set_light_samples(ls)
        N = 10
        ctx.hierarchy_update_importance(N)
        imp, sub = ctx.hierarchy_read()
        nn = len(sc.a["hierarchy"])
        want = np.zeros((nn + len(sc.a["triangles"]), 2), np.float32)
        for it in range(1, N + 1):
            oracle.hier_acquire(sc, ls, cam, w, h, it, want)
        assert want[0, 0] < 2 ** 16 * N  # exact sums in any order below this
        want, want_sub = oracle.hier_propagate(sc, N, want)
        assert same(imp, want), np.argwhere(imp.view(np.uint32) != want.view(np.uint32))[:5]
        assert same(sub, want_sub)
        assert imp[0, 0] > 0 and (imp[:nn, 1] > 0).any()
        ctx.hierarchy_compute_materials()
        mats = ctx.hierarchy_read_materials()
        assert np.array_equal(mats, oracle.hier_materials(sc))
        # the importance-break walk from the device importance
        rays = oracle.raygen_primary(cam, w, h, workloads.MASTER_SEED, 4096)
        r = np.ascontiguousarray(rays).view(rt.RAY_DTYPE).reshape(-1)
        for thr, rthr in ((0.5, 0.0), (0.0, 1e-3)):
            gh, gn = ctx.trace_importance_break(r, None, thr, rthr)
            wh, wn = oracle.trace_importance(sc, rays, want[:nn], thr, rthr)
            assert np.array_equal(np.ascontiguousarray(gh).view(np.int32).reshape(-1, 4), wh)
            assert np.array_equal(np.ascontiguousarray(gn).view(np.int32).reshape(-1, 4)[:, :3], wn[:, :3])
    finally:
        ctx.set_bvh_type(rt.BVH_AABOX)


def test_hierarchy_state_errors(rt, ctx):
    ctx.load_scene(rt.HostScene.cornell())
    with pytest.raises(rt.GpugiError):
        ctx.hierarchy_read()  # nothing computed for this scene
    with pytest.raises(rt.GpugiError):
        ctx.hierarchy_update_importance(0)


def test_importance_sums_are_exact_at_scale(rt, ctx):
    """On a 256 x 256-cell heightfield at 512 x 512 pixels the root sees about
    2^22 rays per update, far past the range where FP32 sums are exact: the
    fixed-point sums still give the same bits twice, and every camera ray of
    every iteration enters the root box with importance 1."""
    hs = rt.HostScene.heightfield(256, workloads.MASTER_SEED)
    for p in workloads.heightfield_lights():
        hs.add_point_light(p, (0.5, 0.5, 0.5))
    w = h = 512
    ctx.set_bvh_type(rt.BVH_AABOX)
    ctx.load_scene(hs)
    ctx.set_screen(w, h)
    ctx.set_camera(workloads.heightfield_camera(w, h))
    ctx.set_light_samples(rt.LightSampler(hs, 5).generate(128))
    print("case: hierarchy importance at scale", flush=True)
    runs = []
    for _ in range(2):
        ctx.hierarchy_update_importance(10)
        runs.append(ctx.hierarchy_read())
    (imp1, sub1), (imp2, sub2) = runs
    assert np.array_equal(imp1.view(np.uint32), imp2.view(np.uint32))
    assert np.array_equal(sub1.view(np.uint32), sub2.view(np.uint32))
    assert imp1[0, 0] >= w * h  # per iteration: every camera ray, importance 1
