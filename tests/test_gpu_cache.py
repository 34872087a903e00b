"""Cached path tracer (GPUGI_PT_CACHE).

- With an unreachable spread threshold no path terminates into the cache and
  the image equals the plain path tracer bit for bit (same estimator, same
  random-number order).
- With the smallest threshold every path stops at its second hit; with an
  all-zero network that is the plain path tracer cut after the primary
  vertex (path length 1), again bit for bit.
- Trained online, the cache cuts the per-pixel error of an equal number of
  frames against a converged plain render while its mean stays close.
- The same two checks for the query at the primary hit (after its NEE): an
  all-zero network leaves the plain path tracer with path length 1, a trained
  one cuts the error.
"""
import numpy as np
import pytest

from gpugi_amd import workloads

pytestmark = pytest.mark.gpu

W = H = 64


def setup(rt, ctx, hs, path_length=5):
    ctx.load_scene(hs)
    ctx.set_screen(W, H)
    ctx.set_camera(workloads.cornell_camera(W))
    ctx.set_max_path_length(path_length)
    ctx.set_integrator_kernel(rt.KERNEL_REFERENCE)


def render(rt, ctx, integrator, hs, frames, first=0):
    sampler = rt.LightSampler(hs, 123)
    ctx.reset_accum()
    for k in range(first + frames):
        ls = sampler.generate(128)
        if k < first:
            continue
        ctx.set_light_samples(ls)
        ctx.draw(integrator, rt.wanghash(k))
    img, n = ctx.readback(W, H)
    assert n == frames
    return img


def test_no_termination_equals_plain_pt(rt, ctx):
    hs = rt.HostScene.cornell()
    setup(rt, ctx, hs)
    ctx.mlp_init(5)
    ctx.cache_config(spread_c=1e30, train_steps=1)
    a = render(rt, ctx, rt.PT_CACHE, hs, 2)
    b = render(rt, ctx, rt.PT, hs, 2)
    assert b[..., :3].mean() > 0.05
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_zero_cache_equals_primary_vertex_only(rt, ctx):
    hs = rt.HostScene.cornell()
    setup(rt, ctx, hs)
    ctx.mlp_init(5)
    ctx.mlp_set_weights(np.zeros(rt.MLP_PARAMS, np.float32))
    ctx.cache_config(spread_c=1e-30, train_steps=0)
    a = render(rt, ctx, rt.PT_CACHE, hs, 1)
    ctx.set_max_path_length(1)
    b = render(rt, ctx, rt.PT, hs, 1)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_trained_cache_reduces_error(rt, ctx):
    hs = rt.HostScene.cornell()
    setup(rt, ctx, hs)
    ref = render(rt, ctx, rt.PT, hs, 1024, first=5000)[..., :3] / 1024.0
    frames = 64
    plain = render(rt, ctx, rt.PT, hs, frames)[..., :3] / frames
    ctx.mlp_init(9)
    ctx.cache_config(train_stride=4, max_train_samples=16384, train_steps=4, learning_rate=1e-2)
    # warm the cache up, then accumulate a fresh image with it
    render(rt, ctx, rt.PT_CACHE, hs, 200, first=10000)
    n, loss = ctx.cache_stats()
    assert n > 1000 and np.isfinite(loss)
    cached = render(rt, ctx, rt.PT_CACHE, hs, frames)[..., :3] / frames
    err_plain = np.sqrt(np.mean((plain - ref) ** 2))
    err_cached = np.sqrt(np.mean((cached - ref) ** 2))
    print("rmse plain %.5f cached %.5f mean ref %.4f cached %.4f" % (err_plain, err_cached, ref.mean(),
                                                                     cached.mean()))
    assert abs(cached.mean() - ref.mean()) < 0.05 * ref.mean()
    assert err_cached < err_plain


def test_primary_indirect_zero_cache_equals_path_length_one(rt, ctx):
    """Querying at the primary hit after its NEE with an all-zero network leaves
    exactly the primary vertex's emission and NEE: the plain path tracer with
    path length 1."""
    hs = rt.HostScene.cornell()
    setup(rt, ctx, hs)
    ctx.mlp_init(5)
    ctx.mlp_set_weights(np.zeros(rt.MLP_PARAMS, np.float32))
    ctx.cache_config(train_steps=0)
    ctx.cache_set_query(rt.CACHE_PRIMARY_INDIRECT)
    a = render(rt, ctx, rt.PT_CACHE, hs, 2)
    ctx.cache_set_query(rt.CACHE_SPREAD)
    ctx.set_max_path_length(1)
    b = render(rt, ctx, rt.PT, hs, 2)
    assert b[..., :3].mean() > 0.05
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_primary_indirect_cache_reduces_error(rt, ctx):
    hs = rt.HostScene.cornell()
    setup(rt, ctx, hs)
    ref = render(rt, ctx, rt.PT, hs, 1024, first=5000)[..., :3] / 1024.0
    frames = 64
    plain = render(rt, ctx, rt.PT, hs, frames)[..., :3] / frames
    ctx.mlp_init(9)
    ctx.cache_config(train_stride=4, max_train_samples=16384, train_steps=4, learning_rate=1e-2)
    ctx.cache_set_query(rt.CACHE_PRIMARY_INDIRECT)
    try:
        render(rt, ctx, rt.PT_CACHE, hs, 200, first=10000)
        cached = render(rt, ctx, rt.PT_CACHE, hs, frames)[..., :3] / frames
    finally:
        ctx.cache_set_query(rt.CACHE_SPREAD)
    err_plain = np.sqrt(np.mean((plain - ref) ** 2))
    err_cached = np.sqrt(np.mean((cached - ref) ** 2))
    print("rmse plain %.5f cached %.5f mean ref %.4f cached %.4f" % (err_plain, err_cached, ref.mean(),
                                                                     cached.mean()))
    assert abs(cached.mean() - ref.mean()) < 0.05 * ref.mean()
    assert err_cached < err_plain
