"""GPU light tracer and bidirectional path tracer against the CPU oracle.

Camera splats are float atomics, so per-pixel sums agree up to the order of
the additions (relative 1e-4 plus a tiny absolute floor).  The BPT warm-up
counters and the light cache are deterministic and compared exactly.
"""
import numpy as np
import pytest

from gpugi_amd import workloads

pytestmark = pytest.mark.gpu

W = H = 32
L = 5


def setup(rt, ctx, hs, kernel):
    ctx.load_scene(hs)
    ctx.set_screen(W, H)
    ctx.set_camera(workloads.cornell_camera(W))
    ctx.set_max_path_length(L)
    ctx.set_integrator_kernel(kernel)


def close(img, ref):
    return np.abs(img - ref) <= 1e-4 * np.abs(ref) + 1e-6


def test_lt_matches_oracle(rt, oracle, ctx):
    hs = rt.HostScene.cornell()
    setup(rt, ctx, hs, rt.KERNEL_REFERENCE)
    sc = oracle.Scene(hs.arrays())
    cam = workloads.cornell_camera(W)
    sampler = rt.LightSampler(hs, 123)
    acc = np.zeros((H, W, 4), np.float32)
    for it in range(2):
        ls = sampler.generate(256)  # LightPathtracer uses 256 initial light samples (lightpathtracer.cpp:16)
        ctx.set_light_samples(ls)
        ctx.draw(rt.LT, rt.wanghash(it))
        oracle.lighttrace(sc, ls, cam, W, H, rt.wanghash(it), L, acc)
    img, n = ctx.readback(W, H)
    assert n == 2 and acc[..., :3].mean() > 0
    ok = close(img, acc)
    assert ok.all(), np.argwhere(~ok)[:5]


def bpt_capacity(rt, oracle, ctx, sc, sampler, cam):
    counts_gpu, counts_cpu = [], []
    for k in range(8):
        ls = sampler.generate(128)
        ctx.set_light_samples(ls)
        counts_gpu.append(ctx.bpt_warmup(rt.wanghash(k + 1)))
        counts_cpu.append(oracle.bpt_warmup(sc, ls, cam, W, H, rt.wanghash(k + 1), L))
    assert counts_gpu == counts_cpu
    mean = sum(c[0] for c in counts_gpu) // 8
    return mean + mean // 10


@pytest.mark.parametrize("schedule", ["megakernel", "wavefront"])
def test_bpt_matches_oracle(rt, oracle, ctx, schedule):
    """The eye paths run as one kernel or as the wavefront passes (their connection
    rays traced by the any-hit kernel); either way the image matches the oracle."""
    hs = rt.HostScene.cornell()
    setup(rt, ctx, hs, rt.KERNEL_REFERENCE)
    sc = oracle.Scene(hs.arrays())
    cam = workloads.cornell_camera(W)
    sampler = rt.LightSampler(hs, 123)
    cap = bpt_capacity(rt, oracle, ctx, sc, sampler, cam)
    ctx.bpt_set_light_cache(cap)
    ctx.reset_accum()
    ctx.set_pt_schedule(rt.PT_WAVEFRONT if schedule == "wavefront" else rt.PT_MEGAKERNEL)
    print("case: BPT %s" % schedule, flush=True)
    acc = np.zeros((H, W, 4), np.float32)
    try:
        for it in range(2):
            ls = sampler.generate(128)
            ctx.set_light_samples(ls)
            ctx.draw(rt.BPT, rt.wanghash(it))
            acc, cache, ncpu = oracle.bpt_draw(sc, ls, cam, W, H, rt.wanghash(it), L, cap, acc)
            gcache, ngpu = ctx.bpt_read_light_cache(cap)
            assert ngpu == ncpu
            m = min(ngpu, cap)
            assert np.array_equal(gcache[:m].view(np.uint32).reshape(m, 16), cache[:m].view(np.uint32))
    finally:
        ctx.set_pt_schedule(rt.PT_MEGAKERNEL)
    img, n = ctx.readback(W, H)
    assert n == 2 and (img[..., 3] == 2.0).all()
    ok = close(img, acc)
    assert ok.all(), np.argwhere(~ok)[:5]


def test_bpt_capacity_overflow_is_clipped(rt, oracle, ctx):
    hs = rt.HostScene.cornell()
    setup(rt, ctx, hs, rt.KERNEL_REFERENCE)
    sc = oracle.Scene(hs.arrays())
    cam = workloads.cornell_camera(W)
    ls = rt.LightSampler(hs, 5).generate(128)
    ctx.set_light_samples(ls)
    R = max(64, (W * H // 128 // 64) * 64)
    cap = 128 * R + 17  # far below what the paths store
    with pytest.raises(rt.GpugiError):
        ctx.draw(rt.BPT, 3)  # no capacity yet -> GPUGI_ERR_STATE
    ctx.bpt_set_light_cache(cap)
    ctx.reset_accum()
    ctx.draw(rt.BPT, 3)
    acc, cache, ncpu = oracle.bpt_draw(sc, ls, cam, W, H, 3, L, cap)
    gcache, ngpu = ctx.bpt_read_light_cache(cap)
    assert ngpu == ncpu and ngpu > cap
    assert np.array_equal(gcache.view(np.uint32).reshape(cap, 16), cache.view(np.uint32))
    img, _ = ctx.readback(W, H)
    assert close(img, acc).all()


@pytest.mark.parametrize("integrator", ["lt", "bpt"])
def test_wide_traversal_integrators_match_oracle_statistically(rt, oracle, ctx, integrator):
    hs = rt.HostScene.cornell()
    setup(rt, ctx, hs, rt.KERNEL_WIDE)
    sc = oracle.Scene(hs.arrays())
    cam = workloads.cornell_camera(W)
    sampler = rt.LightSampler(hs, 123)
    acc = np.zeros((H, W, 4), np.float32)
    if integrator == "bpt":
        cap = 60000
        ctx.bpt_set_light_cache(cap)
    for it in range(2):
        ls = sampler.generate(256 if integrator == "lt" else 128)
        ctx.set_light_samples(ls)
        if integrator == "lt":
            ctx.draw(rt.LT, rt.wanghash(it))
            oracle.lighttrace(sc, ls, cam, W, H, rt.wanghash(it), L, acc)
        else:
            ctx.draw(rt.BPT, rt.wanghash(it))
            acc, _, _ = oracle.bpt_draw(sc, ls, cam, W, H, rt.wanghash(it), L, cap, acc)
    img, _ = ctx.readback(W, H)
    assert close(img, acc).mean() > 0.98
    assert abs(img[..., :3].sum() / acc[..., :3].sum() - 1) < 2e-3


@pytest.mark.parametrize("variant", ["glossy", "refractive"])
def test_lt_specular_branches_match_oracle(rt, oracle, ctx, variant):
    # adjoint sampling and evaluation of the reflect / refract lobes
    from test_oracle_integrators import furnace_scene
    params = {"glossy": dict(specular=(0.3, 0.3, 0.3), roughness=0.4),
              "refractive": dict(specular=(0.1, 0.1, 0.1), roughness=0.2, opacity=0.4)}[variant]
    hs = furnace_scene(rt, **params)
    cam = rt.camera([0.1, 0.2, 0.3], [0.5, -0.3, -1.0], [0, 1, 0], 60, W, H)
    ctx.load_scene(hs)
    ctx.set_screen(W, H)
    ctx.set_camera(cam)
    ctx.set_max_path_length(L)
    ctx.set_integrator_kernel(rt.KERNEL_REFERENCE)
    sc = oracle.Scene(hs.arrays())
    sampler = rt.LightSampler(hs, 123)
    acc = np.zeros((H, W, 4), np.float32)
    print("case: LT specular branch %s" % variant, flush=True)
    ls = sampler.generate(256)
    ctx.set_light_samples(ls)
    ctx.draw(rt.LT, rt.wanghash(0))
    oracle.lighttrace(sc, ls, cam, W, H, rt.wanghash(0), L, acc)
    img, n = ctx.readback(W, H)
    assert n == 1 and acc[..., :3].mean() > 0
    ok = close(img, acc)
    assert ok.all(), np.argwhere(~ok)[:5]
