"""Batched rendering: one persistent launch over the tiles of several scenes
gives, scene by scene, exactly the single-scene path tracer's image; so do
the per-scene launches and the hipGraph replay of a captured frame."""
import numpy as np
import pytest

from gpugi_amd import workloads

pytestmark = pytest.mark.gpu


def small_scenes(rt):
    out = []
    hs = rt.HostScene.cornell()
    out.append((hs, workloads.cornell_camera(32), 32, 32, rt.LightSampler(hs, 123).generate(128)))
    for cells, size in ((8, 48), (24, 64)):
        hs = rt.HostScene.heightfield(cells, 3 + cells)
        for p in workloads.heightfield_lights():
            hs.add_point_light(p, (0.5, 0.5, 0.5))
        out.append((hs, workloads.heightfield_camera(size, size), size, size, rt.LightSampler(hs, 9).generate(128)))
    return out


def single_scene(rt, ctx, hs, cam, w, h, lights, frames, kernel):
    ctx.load_scene(hs)
    ctx.set_screen(w, h)
    ctx.set_camera(cam)
    ctx.set_max_path_length(5)
    ctx.set_integrator_kernel(kernel)
    ctx.set_light_samples(lights)
    for k in range(frames):
        ctx.draw(rt.PT, rt.wanghash(k))
    return ctx.readback(w, h)[0]


@pytest.mark.parametrize("kernel", ["reference", "wide"])
def test_batch_matches_single_scene_renders(rt, ctx, kernel):
    k = rt.KERNEL_REFERENCE if kernel == "reference" else rt.KERNEL_WIDE
    scenes = small_scenes(rt)
    b = rt.Batch(0)
    try:
        for hs, cam, w, h, lights in scenes:
            b.add(hs, cam, w, h, lights, 5)
        b.finalize()
        b.config(k, 3, 1)
        print("case: batch %s persistent" % kernel, flush=True)
        for _ in range(2):
            b.draw(True)
        persistent = [b.readback(i) for i in range(len(scenes))]
        rays = b.rays()
        b.reset(0)
        print("case: batch %s separate" % kernel, flush=True)
        for _ in range(2):
            b.draw(False)
        separate = [b.readback(i) for i in range(len(scenes))]
        assert b.rays() == rays > 0
        for i, (hs, cam, w, h, lights) in enumerate(scenes):
            single = single_scene(rt, ctx, hs, cam, w, h, lights, 2, k)
            assert single[..., :3].mean() > 0
            assert np.array_equal(persistent[i].view(np.uint32), separate[i].view(np.uint32))
            assert np.array_equal(persistent[i].view(np.uint32), single.view(np.uint32)), i
    finally:
        b.close()


@pytest.mark.parametrize("stack", [16, 24, 32])
def test_batch_stack_depths_match_single_scene(rt, ctx, stack):
    """The LDS stack depth of the wide walk (16, 24 or 32 entries per lane;
    deeper entries spill to global memory) changes where the traversal stack
    lives, not what it holds: every depth gives the single-scene image bit for
    bit.  The 24-cell heightfield needs 19 entries (in LDS from depth 24 on),
    the 30-sphere clutter room 36 (it spills at every depth)."""
    scenes = small_scenes(rt)
    hs = rt.HostScene.clutter(30, 5)
    scenes.append((hs, workloads.clutter_camera(32, 32), 32, 32, rt.LightSampler(hs, 7).generate(128)))
    ctx.load_scene(hs)
    ctx.set_wide_config(2, 16, 8, 16)
    assert ctx.wide_config()["max_stack"] > 32  # the spill path is exercised at every depth
    b = rt.Batch(0)
    try:
        for hs, cam, w, h, lights in scenes:
            b.add(hs, cam, w, h, lights, 5)
        b.finalize()
        b.config(rt.KERNEL_WIDE, 2, 2, stack)
        print("case: batch stack depth %d" % stack, flush=True)
        for _ in range(2):
            b.draw(True)
        got = [b.readback(i) for i in range(len(scenes))]
        for i, (hs, cam, w, h, lights) in enumerate(scenes):
            single = single_scene(rt, ctx, hs, cam, w, h, lights, 2, rt.KERNEL_WIDE)
            assert single[..., :3].mean() > 0, i
            assert np.array_equal(got[i].view(np.uint32), single.view(np.uint32)), i
        with pytest.raises(rt.GpugiError):
            b.config(rt.KERNEL_WIDE, 2, 2, 20)
    finally:
        ctx.set_wide_config(2, 10, 8, 16)
        b.close()


def test_graph_replay_matches_eager_frames(rt):
    scenes = small_scenes(rt)
    b = rt.Batch(0)
    try:
        for hs, cam, w, h, lights in scenes:
            b.add(hs, cam, w, h, lights, 5)
        b.finalize()
        eager = []
        for f in range(3):
            b.draw(True)
            eager.append([b.readback(i) for i in range(len(scenes))])
        b.reset(0)
        b.capture()
        for f in range(3):
            print("case: graph replay %d" % f, flush=True)
            b.replay(1)
            got = [b.readback(i) for i in range(len(scenes))]  # eager work between replays
            for i in range(len(scenes)):
                assert np.array_equal(got[i].view(np.uint32), eager[f][i].view(np.uint32)), (f, i)
        tuned = b.autotune(1)
        assert tuned["kernel"] in (rt.KERNEL_REFERENCE, rt.KERNEL_WIDE) and tuned["ms_per_frame"] > 0
        assert tuned["tiles_per_grab"] in (1, 2, 4, 8) and tuned["grid_multiplier"] in (1, 2, 4)
        assert tuned["stack_depth"] in (16, 24, 32)
        # the tuned grab, grid and stack depth render what the defaults render with the tuned traversal
        frames = []
        for cfg in (None, (tuned["kernel"], 2, 2, 16)):
            if cfg:
                b.config(*cfg)
            b.reset(0)
            for f in range(3):
                b.draw(True)
            frames.append([b.readback(i) for i in range(len(scenes))])
        for i in range(len(scenes)):
            assert np.array_equal(frames[0][i].view(np.uint32), frames[1][i].view(np.uint32)), i
    finally:
        b.close()
