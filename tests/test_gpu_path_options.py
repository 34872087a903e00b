"""GPU integrators with path-length isolation and a late roulette start,
against the CPU oracle with the same options (reference traversal).

gpugi_set_specific_path_length replaces the reference's SHOW_SPECIFIC_PATHLENGTH
define (pathtracer.comp:74-84, lighttracer.comp:133-157,
pathtracer_bidir.comp:75-108); gpugi_set_rr_start_depth starts Russian
roulette at a later bounce (the clutter room's roulette from bounce 3).  PT is
bit-exact in both schedules; LT and BPT (both eye-path schedules) splat with
float atomics and agree to 1e-4 relative, their light caches bit for bit.
"""
import numpy as np
import pytest

from gpugi_amd import workloads

pytestmark = pytest.mark.gpu

W = H = 32
L = 5


@pytest.fixture
def options(rt, ctx, oracle):
    """Sets both options on the GPU and the oracle; restores the defaults after."""
    def set_(spl, rr):
        ctx.set_specific_path_length(spl)
        ctx.set_rr_start_depth(rr)
        oracle.lib().or_set_path_options(spl, rr)
    yield set_
    ctx.set_specific_path_length(-1)
    ctx.set_rr_start_depth(0)
    ctx.set_pt_schedule(rt.PT_MEGAKERNEL)
    oracle.lib().or_set_path_options(-1, 0)


def setup(ctx, hs, rt):
    ctx.load_scene(hs)
    ctx.set_screen(W, H)
    ctx.set_camera(workloads.cornell_camera(W))
    ctx.set_max_path_length(L)
    ctx.set_integrator_kernel(rt.KERNEL_REFERENCE)
    ctx.reset_accum()


def close(img, ref):
    return np.abs(img - ref) <= 1e-4 * np.abs(ref) + 1e-6


@pytest.mark.parametrize("schedule", ["megakernel", "wavefront"])
@pytest.mark.parametrize("spl,rr", [(-1, 3), (0, 0), (1, 0), (2, 0), (3, 3)])
def test_pt_path_options_bitexact(rt, oracle, ctx, options, schedule, spl, rr):
    print("case: PT %s spl=%d rr=%d" % (schedule, spl, rr), flush=True)
    hs = rt.HostScene.cornell()
    setup(ctx, hs, rt)
    options(spl, rr)
    ctx.set_pt_schedule(rt.PT_WAVEFRONT if schedule == "wavefront" else rt.PT_MEGAKERNEL)
    sc = oracle.Scene(hs.arrays())
    cam = workloads.cornell_camera(W)
    sampler = rt.LightSampler(hs, 123)
    acc = np.zeros((H, W, 4), np.float32)
    for it in range(2):
        ls = sampler.generate(128)
        ctx.set_light_samples(ls)
        ctx.draw(rt.PT, rt.wanghash(it))
        oracle.pathtrace(sc, ls, cam, W, H, rt.wanghash(it), L, acc)
    img, n = ctx.readback(W, H)
    assert n == 2 and acc[..., :3].max() > 0
    diff = np.argwhere(img.view(np.uint32) != acc.view(np.uint32))
    assert len(diff) == 0, diff[:5]


@pytest.mark.parametrize("spl,rr", [(2, 0), (-1, 3)])
def test_lt_path_options_match_oracle(rt, oracle, ctx, options, spl, rr):
    print("case: LT spl=%d rr=%d" % (spl, rr), flush=True)
    hs = rt.HostScene.cornell()
    setup(ctx, hs, rt)
    options(spl, rr)
    sc = oracle.Scene(hs.arrays())
    cam = workloads.cornell_camera(W)
    sampler = rt.LightSampler(hs, 123)
    acc = np.zeros((H, W, 4), np.float32)
    for it in range(2):
        ls = sampler.generate(256)
        ctx.set_light_samples(ls)
        ctx.draw(rt.LT, rt.wanghash(it))
        oracle.lighttrace(sc, ls, cam, W, H, rt.wanghash(it), L, acc)
    img, _ = ctx.readback(W, H)
    assert acc[..., :3].max() > 0
    ok = close(img, acc)
    assert ok.all(), np.argwhere(~ok)[:5]


@pytest.mark.parametrize("schedule", ["megakernel", "wavefront"])
@pytest.mark.parametrize("spl,rr", [(2, 0), (3, 3)])
def test_bpt_path_options_match_oracle(rt, oracle, ctx, options, schedule, spl, rr):
    print("case: BPT %s spl=%d rr=%d" % (schedule, spl, rr), flush=True)
    hs = rt.HostScene.cornell()
    setup(ctx, hs, rt)
    options(spl, rr)
    ctx.set_pt_schedule(rt.PT_WAVEFRONT if schedule == "wavefront" else rt.PT_MEGAKERNEL)
    sc = oracle.Scene(hs.arrays())
    cam = workloads.cornell_camera(W)
    sampler = rt.LightSampler(hs, 123)
    cap = 60000
    ctx.bpt_set_light_cache(cap)
    acc = np.zeros((H, W, 4), np.float32)
    for it in range(2):
        ls = sampler.generate(128)
        ctx.set_light_samples(ls)
        ctx.draw(rt.BPT, rt.wanghash(it))
        acc, cache, ncpu = oracle.bpt_draw(sc, ls, cam, W, H, rt.wanghash(it), L, cap, acc)
        gcache, ngpu = ctx.bpt_read_light_cache(cap)
        assert ngpu == ncpu
        m = min(ngpu, cap)
        assert np.array_equal(gcache[:m].view(np.uint32).reshape(m, 16), cache[:m].view(np.uint32))
    img, _ = ctx.readback(W, H)
    assert acc[..., :3].max() > 0
    ok = close(img, acc)
    assert ok.all(), np.argwhere(~ok)[:5]


def test_cached_pt_rejects_path_length_isolation(rt, ctx, options):
    hs = rt.HostScene.cornell()
    setup(ctx, hs, rt)
    ctx.set_light_samples(rt.LightSampler(hs, 1).generate(128))
    options(2, 0)
    with pytest.raises(rt.GpugiError):
        ctx.draw(rt.PT_CACHE, 1)
    with pytest.raises(rt.GpugiError):
        ctx.set_specific_path_length(65)
