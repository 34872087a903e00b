"""Every benchmark configuration exercised on the GPU against the CPU oracle.

- C1: the Cornell room at 256 x 256, path length 5, 64 iterations, reference
  walk: the accumulated image equals the oracle's bit for bit.
- C2: the wide any-hit walk on the committed golden shadow rays equals the
  fixture's flags (the reference walk's), except rays the FP64 classifier
  puts within 1e-6 of an edge, a tie or the ray's end.
- C3: the 5.12 M-triangle clutter room at 1920 x 1080, path length 8,
  roulette from bounce 3: one PT iteration with the reference walk equals the
  oracle on three bands of rows; the wide walk's image mean agrees to 1e-3; a
  second draw with the same frame seed repeats the image bit for bit; one LT
  and one BPT iteration with the reference walk equal the oracle's draws to
  the order of their splats (1e-4 relative), the BPT light cache bit for bit.
  (Their image means need not equal PT's here: PT makes no light-sample
  connection at Phong exponents of 500 and more, so the point lights light
  those spheres only in LT and BPT, and the light tracer's host pass also
  splats the point lights themselves, showlightcaches.comp.)
- C4: the 16 batched heightfields (10 k - 2 M triangles, 256^2 - 1024^2
  pixels), one persistent launch with the reference walk: every scene equals
  the oracle's path tracer on a band of rows.
- C5: the radiance cache's exact limits on the 2 M-triangle heightfield at
  1920 x 1080: an unreachable spread threshold is the plain path tracer, an
  all-zero network at the primary hit is the plain path tracer cut after the
  first vertex, both bit for bit.
The oracle runs on a thread pool where the work splits by rows.
"""
import concurrent.futures
import os

import numpy as np
import pytest

from gpugi_amd import workloads

pytestmark = pytest.mark.gpu

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
WORKERS = min(16, os.cpu_count() or 1)


def oracle_pt_rows(oracle, sc, ls, cam, w, h, seed, path_length, rows, acc):
    """or_pathtracer over `rows` (begin, end) split across threads (disjoint rows)."""
    r0, r1 = rows
    step = max(8, (r1 - r0 + WORKERS - 1) // WORKERS)
    parts = [(a, min(r1, a + step)) for a in range(r0, r1, step)]
    with concurrent.futures.ThreadPoolExecutor(max_workers=WORKERS) as ex:
        list(ex.map(lambda p: oracle.pathtrace(sc, ls, cam, w, h, seed, path_length, acc, rows=p), parts))
    return acc


def bitexact_rows(img, acc, r0, r1):
    a, b = img[r0:r1], acc[r0:r1]
    diff = np.argwhere(a.view(np.uint32) != b.view(np.uint32))
    assert len(diff) == 0, diff[:5]


def test_c1_cornell_full_size(rt, oracle, ctx):
    print("case: C1 cornell 256x256 x 64", flush=True)
    hs = rt.HostScene.cornell()
    w = h = workloads.CORNELL_SIZE
    cam = workloads.cornell_camera(w)
    ctx.load_scene(hs)
    ctx.set_screen(w, h)
    ctx.set_camera(cam)
    ctx.set_max_path_length(workloads.CORNELL_PATH_LENGTH)
    ctx.set_integrator_kernel(rt.KERNEL_REFERENCE)
    sc = oracle.Scene(hs.arrays())
    sampler = rt.LightSampler(hs, 123)
    acc = np.zeros((h, w, 4), np.float32)
    its = 64
    for it in range(its):
        ls = sampler.generate(128)
        ctx.set_light_samples(ls)
        ctx.draw(rt.PT, rt.wanghash(it))
        oracle_pt_rows(oracle, sc, ls, cam, w, h, rt.wanghash(it), workloads.CORNELL_PATH_LENGTH, (0, h), acc)
    img, n = ctx.readback(w, h)
    assert n == its and img[..., :3].mean() > 0.05
    bitexact_rows(img, acc, 0, h)


def test_c2_wide_any_hit_on_golden_shadow_rays(rt, oracle, ctx):
    print("case: C2 wide any hit", flush=True)
    with np.load(os.path.join(GOLDEN_DIR, "c2_golden.npz"), allow_pickle=False) as z:
        shadow, want = z["shadow"], z["shadow_any"]
    hs = rt.HostScene.heightfield(workloads.HF_CELLS, workloads.MASTER_SEED)
    ctx.load_scene(hs)
    rays = np.ascontiguousarray(shadow).view(rt.RAY_DTYPE).reshape(-1)
    got = ctx.trace(rays, rt.ANY_HIT, rt.KERNEL_WIDE).view(np.int32).reshape(-1, 4)[:, 0]
    assert 10 < want.sum() < len(want) - 10  # both outcomes occur (1.3 % of these rays are blocked)
    diff = np.nonzero(got != want)[0]
    if len(diff):
        sc = oracle.Scene(hs.arrays())
        closest = oracle.bruteforce(sc, shadow[diff])[:, 0]
        flags = oracle.classify(sc, shadow[diff], closest, -np.ones(len(diff), np.int32))
        assert (flags != 0).all(), diff[flags == 0][:5]
    assert len(diff) <= 2


@pytest.fixture(scope="module")
def clutter(rt, oracle):
    hs = rt.HostScene.clutter(workloads.CLUTTER_SPHERES, workloads.MASTER_SEED)
    return hs, oracle.Scene(hs.arrays())


def test_c3_clutter_room(rt, oracle, ctx, clutter):
    hs, sc = clutter
    w, h = workloads.HF_WIDTH, workloads.HF_HEIGHT
    cam = workloads.clutter_camera(w, h)
    ctx.load_scene(hs)
    ctx.set_screen(w, h)
    ctx.set_camera(cam)
    ctx.set_max_path_length(workloads.CLUTTER_PATH_LENGTH)
    ctx.set_rr_start_depth(workloads.CLUTTER_RR_START)
    try:
        sampler = rt.LightSampler(hs, 123)
        ls = sampler.generate(128)
        ctx.set_light_samples(ls)
        seed = rt.wanghash(1)
        print("case: C3 PT reference walk", flush=True)
        ctx.set_integrator_kernel(rt.KERNEL_REFERENCE)
        ctx.reset_accum()
        ctx.draw(rt.PT, seed)
        ref, _ = ctx.readback(w, h)
        oracle.lib().or_set_path_options(-1, workloads.CLUTTER_RR_START)
        try:
            acc = np.zeros((h, w, 4), np.float32)
            for band in ((0, 8), (536, 544), (1064, 1072)):
                oracle_pt_rows(oracle, sc, ls, cam, w, h, seed, workloads.CLUTTER_PATH_LENGTH, band, acc)
                bitexact_rows(ref, acc, *band)
        finally:
            oracle.lib().or_set_path_options(-1, 0)
        assert ref[..., :3].mean() > 0
        print("case: C3 PT wide walk", flush=True)
        ctx.set_integrator_kernel(rt.KERNEL_WIDE)
        ctx.reset_accum()
        ctx.draw(rt.PT, seed)
        wide, _ = ctx.readback(w, h)
        assert abs(wide[..., :3].mean() / ref[..., :3].mean() - 1) < 1e-3
        ctx.reset_accum()
        ctx.draw(rt.PT, seed)
        again, _ = ctx.readback(w, h)
        assert np.array_equal(again.view(np.uint32), wide.view(np.uint32))
        # LT and BPT with the reference walk against the oracle's draws (light paths and
        # eye rows split over threads): splats agree to their order of addition, the BPT
        # light cache bit for bit
        ctx.set_integrator_kernel(rt.KERNEL_REFERENCE)
        oracle.lib().or_set_path_options(-1, workloads.CLUTTER_RR_START)
        try:
            print("case: C3 LT reference walk", flush=True)
            lls = rt.LightSampler(hs, 321).generate(256)
            ctx.set_light_samples(lls)
            ctx.reset_accum()
            ctx.draw(rt.LT, seed)
            img, _ = ctx.readback(w, h)
            want = oracle.lighttrace_parallel(sc, lls, cam, w, h, seed, workloads.CLUTTER_PATH_LENGTH, WORKERS)
            assert want[..., :3].max() > 0
            ok = np.abs(img - want) <= 1e-4 * np.abs(want) + 1e-6
            assert ok.all(), np.argwhere(~ok)[:5]
            print("case: C3 BPT reference walk", flush=True)
            sampler = rt.LightSampler(hs, 321)
            entries = 0
            for k in range(8):
                ctx.set_light_samples(sampler.generate(128))
                entries += ctx.bpt_warmup(rt.wanghash(100 + k))[0]
            mean = entries // 8
            cap = mean + mean // 10
            ctx.bpt_set_light_cache(cap)
            bls = sampler.generate(128)
            ctx.set_light_samples(bls)
            ctx.reset_accum()
            ctx.draw(rt.BPT, seed)
            img, _ = ctx.readback(w, h)
            gcache, ngpu = ctx.bpt_read_light_cache(cap)
            want, cache, n = oracle.bpt_draw_parallel(sc, bls, cam, w, h, seed, workloads.CLUTTER_PATH_LENGTH, cap,
                                                      WORKERS)
            assert ngpu == n
            m = min(n, cap)
            assert np.array_equal(gcache[:m].view(np.uint32).reshape(m, 16), cache[:m].view(np.uint32))
            ok = np.abs(img - want) <= 1e-4 * np.abs(want) + 1e-6
            assert ok.all(), np.argwhere(~ok)[:5]
        finally:
            oracle.lib().or_set_path_options(-1, 0)
    finally:
        ctx.set_rr_start_depth(0)
        ctx.set_integrator_kernel(rt.KERNEL_AUTO)


def test_c4_batched_scenes_against_oracle(rt, oracle):
    scenes = workloads.batch_scenes()
    b = rt.Batch(0)
    try:
        for hs, cam, w, h, lights in scenes:
            b.add(hs, cam, w, h, lights, workloads.BATCH_PATH_LENGTH)
        b.finalize()
        b.config(rt.KERNEL_REFERENCE, 1, 1)
        print("case: C4 batch of %d scenes" % len(scenes), flush=True)
        b.reset(0)
        b.draw(True)
        for i, (hs, cam, w, h, lights) in enumerate(scenes):
            img = b.readback(i)
            sc = oracle.Scene(hs.arrays())
            acc = np.zeros((h, w, 4), np.float32)
            band = (h // 2 - 4, h // 2 + 4)
            oracle_pt_rows(oracle, sc, lights, cam, w, h, rt.wanghash(0), workloads.BATCH_PATH_LENGTH, band, acc)
            assert acc[band[0]:band[1], :, :3].max() > 0, i
            bitexact_rows(img, acc, *band)
    finally:
        b.close()


def test_c5_cache_limits_on_the_heightfield(rt, ctx):
    hs = rt.HostScene.heightfield(workloads.HF_CELLS, workloads.MASTER_SEED)
    for p in workloads.heightfield_lights():
        hs.add_point_light(p, (0.5, 0.5, 0.5))
    w, h = workloads.HF_WIDTH, workloads.HF_HEIGHT
    ctx.load_scene(hs)
    ctx.set_screen(w, h)
    ctx.set_camera(workloads.heightfield_camera(w, h))
    ctx.set_integrator_kernel(rt.KERNEL_WIDE)
    ls = rt.LightSampler(hs, 123).generate(128)
    ctx.set_light_samples(ls)

    def frame(integrator, path_length):
        ctx.set_max_path_length(path_length)
        ctx.reset_accum()
        ctx.draw(integrator, rt.wanghash(3))
        return ctx.readback(w, h)[0]

    try:
        print("case: C5 unreachable threshold", flush=True)
        ctx.mlp_init(5)
        ctx.cache_config(spread_c=1e30, train_steps=1)
        a = frame(rt.PT_CACHE, 8)
        b = frame(rt.PT, 8)
        assert b[..., :3].mean() > 0
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        print("case: C5 zero network at the primary hit", flush=True)
        ctx.mlp_set_weights(np.zeros(rt.MLP_PARAMS, np.float32))
        ctx.cache_config(train_steps=0)
        ctx.cache_set_query(rt.CACHE_PRIMARY_INDIRECT)
        a = frame(rt.PT_CACHE, 8)
        b = frame(rt.PT, 1)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    finally:
        ctx.cache_set_query(rt.CACHE_SPREAD)
        ctx.cache_config()
        ctx.set_integrator_kernel(rt.KERNEL_AUTO)
