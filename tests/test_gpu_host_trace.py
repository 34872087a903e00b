"""gpugi_trace with host buffers: above 2^22 rays in pinned memory
(gpugi_host_alloc) it runs chunked through two device slots with the copies
on a second stream, from pageable memory in one staged copy each way.  Both
must give exactly the hits of the same rays traced from device memory, with
2 and 3 chunks (a slot reused)."""
import numpy as np
import pytest

from gpugi_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [(1 << 22) + 1000, 2 * (1 << 22) + 5])
def test_chunked_host_trace_equals_device_trace(rt, ctx, n):
    hs = rt.HostScene.heightfield(64, workloads.MASTER_SEED)
    ctx.load_scene(hs)
    rng = np.random.default_rng(n)
    rays = np.zeros(n, rt.RAY_DTYPE)
    rays["origin"] = rng.uniform(0, 1, (n, 3)) + [0, 0.3, 0]
    rays["dir"] = rng.normal(size=(n, 3))
    rays["tmax"] = rt.RAY_MAX
    print("case: chunked host trace n=%d" % n, flush=True)
    dr, dh = ctx.alloc(n * 32), ctx.alloc(n * 16)
    try:
        dr.upload(rays)
        for kernel in (rt.KERNEL_WIDE, rt.KERNEL_REFERENCE):
            ctx.trace_device(dr, dh, n, rt.CLOSEST_HIT, kernel)
            want = dh.download(np.int32, n * 4).reshape(-1, 4)
            got = ctx.trace(rays, rt.CLOSEST_HIT, kernel).view(np.int32).reshape(-1, 4)
            assert np.array_equal(got, want)
            pr, ph = ctx.host_empty(n, rt.RAY_DTYPE), ctx.host_empty(n, rt.HIT_DTYPE)
            try:
                pr[:] = rays
                ctx.trace(pr, rt.CLOSEST_HIT, kernel, out=ph)
                assert np.array_equal(ph.view(np.int32).reshape(-1, 4), want)
            finally:
                ctx.host_free(pr)
                ctx.host_free(ph)
        assert (want[:, 0] >= 0).mean() > 0.1
    finally:
        dr.free()
        dh.free()


def test_pinned_arrays_own_their_memory(rt):
    """A host_empty array keeps its pinned allocation alive through its views
    and releases it when the last one goes; the context refuses to close
    while one is alive, and host_free releases it at once (also from a view)."""
    c = rt.Context(0)
    a = c.host_empty(1000, np.float32)
    view = a[10:20]
    del a
    view[:] = 3.0  # the allocation outlives the array it was cut from
    assert (view == 3.0).all()
    with pytest.raises(rt.GpugiError):
        c.close()
    del view
    b = c.host_empty(64, np.uint32)
    c.host_free(b[4:8])
    c.host_free(b)  # a second release is a no-op
    with pytest.raises(rt.GpugiError):
        c.host_free(np.zeros(4, np.float32))
    del b
    c.close()
