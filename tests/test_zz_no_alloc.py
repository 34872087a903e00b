"""gpugi_trace_noalloc against gpugi_trace (device-resident rays) through the C ABI.

The allocation-free variant runs the same traversal kernels with their batch
counters and spill stacks in the library's static scratch array, so every hit
record must equal the existing entry point's bit for bit (each ray's walk
does not depend on which wave runs it).  Two calls in a row, the second with
fewer rays (a ragged count) into a hits buffer still holding the first call's
records, check that the shared array carries nothing over.
"""
import numpy as np
import pytest

from gpugi_amd import workloads

pytestmark = pytest.mark.gpu


def scene_rays(rt, oracle, hs, n):
    sc = oracle.Scene(hs.arrays())
    cam = workloads.heightfield_camera()
    prim = oracle.raygen_primary(cam, 1920, 1080, workloads.MASTER_SEED, n)
    ph = oracle.trace(sc, prim)
    diff = oracle.raygen_diffuse(sc, prim, ph, workloads.MASTER_SEED + 1, n)
    shadow = oracle.raygen_shadow(prim, ph, workloads.heightfield_lights(), n)
    return {"primary": prim, "diffuse": diff, "shadow": shadow}


@pytest.mark.parametrize("kernel", ["wide", "reference"])
def test_noalloc_trace_matches_trace(rt, oracle, ctx, kernel):
    hs = rt.HostScene.heightfield(256, workloads.MASTER_SEED)
    ctx.load_scene(hs)
    k = rt.KERNEL_WIDE if kernel == "wide" else rt.KERNEL_REFERENCE
    n = 1 << 17
    sets = scene_rays(rt, oracle, hs, n)
    rays_buf = rt.DeviceBuffer(ctx, n * 32)
    want_buf = rt.DeviceBuffer(ctx, n * 16)
    got_buf = rt.DeviceBuffer(ctx, n * 16)
    try:
        for name, rays in sets.items():
            query = rt.ANY_HIT if name == "shadow" else rt.CLOSEST_HIT
            for m in (n, n // 3 + 7):  # the second call is smaller and ragged
                print("case: noalloc %s %s n=%d" % (kernel, name, m), flush=True)
                rays_buf.upload(np.ascontiguousarray(rays[:m], dtype=np.float32))
                ctx.trace_device(rays_buf, want_buf, m, query, k)
                ctx.trace_noalloc(rays_buf, got_buf, m, query, k)
                ctx.sync()
                want = want_buf.download(np.int32, 4 * m).reshape(m, 4)
                got = got_buf.download(np.int32, 4 * m).reshape(m, 4)
                assert np.array_equal(got, want), np.nonzero((got != want).any(1))[0][:10]
                if query == rt.CLOSEST_HIT:
                    assert (want[:, 0] >= 0).mean() > 0.05  # the rays do hit the terrain
    finally:
        for b in (rays_buf, want_buf, got_buf):
            b.free()


def test_noalloc_trace_rejects_bad_arguments(rt, ctx):
    hs = rt.HostScene.cornell()
    ctx.load_scene(hs)
    buf = rt.DeviceBuffer(ctx, 64)
    try:
        with pytest.raises(rt.GpugiError):
            ctx.trace_noalloc(buf, buf, 1, 7, rt.KERNEL_WIDE)  # unknown query
        ctx.trace_noalloc(buf, buf, 0)  # nothing to do
    finally:
        buf.free()
