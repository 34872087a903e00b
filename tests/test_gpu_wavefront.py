"""Wavefront schedule of the path tracer (gpugi_set_pt_schedule(GPUGI_PT_WAVEFRONT)):
per-bounce passes over device ray queues, the rays traced by the gpugi_trace
kernels.  With the reference walk the image equals the oracle's
pathtracer.comp bit for bit (Cornell box, a truncated 8x8 dispatch grid, the
mirror / glossy / refractive branches, a heightfield with many misses);
with the wide walk it equals the one-kernel schedule's wide image on all but
tie pixels (at least 99 % of the pixels bit for bit, the means within 1e-3).
With the reference walk, roulette from depth 3, path length 1, the clutter
room, the environment map on misses, point lights and textured materials
equal the one-kernel schedule bit for bit.  A frame recorded into a HIP graph
replays to the eager frame bit for bit (Cornell box, both walks).  BPT's eye
paths under this schedule match the one-kernel eye pass on the clutter room
(per pixel to the float-atomic order of the light pass's camera splats)."""
import numpy as np
import pytest

from gpugi_amd import workloads
from test_gpu_pt import run_pt

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wctx(rt):
    c = rt.Context(0)  # its own context: the schedule stays out of other modules' tests
    c.set_pt_schedule(rt.PT_WAVEFRONT)
    yield c
    c.close()


def bitexact(img, acc):
    diff = np.nonzero(img.view(np.uint32) != acc.view(np.uint32))
    assert len(diff[0]) == 0, (diff[0][:5], diff[1][:5])


def test_wavefront_cornell_bitexact(rt, oracle, wctx):
    hs = rt.HostScene.cornell()
    print("case: wavefront cornell 64x64", flush=True)
    img, acc = run_pt(rt, oracle, wctx, hs, workloads.cornell_camera(64), 64, 64, 3, rt.KERNEL_REFERENCE)
    assert img.mean() > 0.05
    bitexact(img, acc)


def test_wavefront_truncated_grid(rt, oracle, wctx):
    hs = rt.HostScene.cornell()
    print("case: wavefront cornell 60x44", flush=True)
    img, acc = run_pt(rt, oracle, wctx, hs, workloads.cornell_camera(60), 60, 44, 1, rt.KERNEL_REFERENCE)
    assert (img[40:, :, :3] == 0).all() and (img[:, 56:, :3] == 0).all()
    bitexact(img, acc)


@pytest.mark.parametrize("variant", ["mirror", "glossy", "refractive"])
def test_wavefront_specular_branches(rt, oracle, wctx, variant):
    from test_oracle_integrators import furnace_scene
    params = {"mirror": dict(specular=(0.5, 0.5, 0.5), roughness=0.0),
              "glossy": dict(specular=(0.3, 0.3, 0.3), roughness=0.4),
              "refractive": dict(specular=(0.1, 0.1, 0.1), roughness=0.2, opacity=0.4)}[variant]
    hs = furnace_scene(rt, **params)
    cam = rt.camera([0.1, 0.2, 0.3], [0.5, -0.3, -1.0], [0, 1, 0], 60, 32, 32)
    print("case: wavefront specular %s" % variant, flush=True)
    img, acc = run_pt(rt, oracle, wctx, hs, cam, 32, 32, 2, rt.KERNEL_REFERENCE)
    assert img[..., :3].mean() > 0.5
    bitexact(img, acc)


def test_wavefront_heightfield_misses(rt, oracle, wctx):
    hs = rt.HostScene.heightfield(24, workloads.MASTER_SEED)
    cam = rt.camera([0.5, 1.2, -0.6], [0.5, 0.1, 0.5], [0, 1, 0], 70, 48, 40)
    print("case: wavefront heightfield 48x40", flush=True)
    img, acc = run_pt(rt, oracle, wctx, hs, cam, 48, 40, 2, rt.KERNEL_REFERENCE, path_length=6)
    bitexact(img, acc)


def draw_both(rt, ctx_a, ctx_b, hs, cam, w, h, kernel, path_length, rr_depth=0, iters=2):
    imgs = []
    for c in (ctx_a, ctx_b):
        c.load_scene(hs)
        c.set_screen(w, h)
        c.set_camera(cam)
        c.set_max_path_length(path_length)
        c.set_rr_start_depth(rr_depth)
        c.set_integrator_kernel(kernel)
        sampler = rt.LightSampler(hs, 7)
        for it in range(iters):
            c.set_light_samples(sampler.generate(128))
            c.draw(rt.PT, rt.wanghash(it + 11))
        imgs.append(c.readback(w, h)[0])
        c.set_rr_start_depth(0)
    return imgs


def test_wavefront_wide_matches_megakernel(rt, ctx, wctx):
    hs = rt.HostScene.cornell()
    print("case: wavefront wide cornell", flush=True)
    a, b = draw_both(rt, ctx, wctx, hs, workloads.cornell_camera(64), 64, 64, rt.KERNEL_WIDE, 5)
    same = (a.view(np.uint32) == b.view(np.uint32)).all(-1)
    assert same.mean() > 0.99
    assert abs(a[..., :3].mean() - b[..., :3].mean()) <= 1e-3 * a[..., :3].mean()
    ctx.set_integrator_kernel(rt.KERNEL_AUTO)


@pytest.mark.parametrize("path_length,rr_depth", [(1, 0), (12, 3)])
def test_wavefront_depth_settings_match_megakernel(rt, ctx, wctx, path_length, rr_depth):
    hs = rt.HostScene.cornell()
    print("case: wavefront path length %d rr depth %d" % (path_length, rr_depth), flush=True)
    a, b = draw_both(rt, ctx, wctx, hs, workloads.cornell_camera(32), 32, 32, rt.KERNEL_REFERENCE, path_length,
                     rr_depth)
    assert a[..., :3].mean() > 0.01
    bitexact(b, a)
    ctx.set_integrator_kernel(rt.KERNEL_AUTO)


def test_schedule_rejects_unknown_value(rt, wctx):
    with pytest.raises(Exception):
        wctx.set_pt_schedule(7)


@pytest.mark.parametrize("kernel", ["reference", "wide"])
def test_wavefront_bpt_clutter_matches_megakernel(rt, ctx, wctx, kernel):
    """BPT on the clutter room with its eye paths under both schedules.  The light
    pass is the same code; its camera splats are float atomics, so the images
    agree per pixel to 1e-4 relative (reference walk), and on all but tie pixels
    (wide walk)."""
    hs = rt.HostScene.clutter(300, workloads.MASTER_SEED)
    w, h = 192, 112
    k = rt.KERNEL_REFERENCE if kernel == "reference" else rt.KERNEL_WIDE
    print("case: wavefront BPT clutter %s" % kernel, flush=True)
    imgs = []
    for c in (ctx, wctx):
        c.load_scene(hs)
        c.set_screen(w, h)
        c.set_camera(workloads.clutter_camera(w, h))
        c.set_max_path_length(workloads.CLUTTER_PATH_LENGTH)
        c.set_rr_start_depth(workloads.CLUTTER_RR_START)
        c.set_integrator_kernel(k)
        sampler = rt.LightSampler(hs, 7)
        c.set_light_samples(sampler.generate(128))
        entries = c.bpt_warmup(rt.wanghash(1))[0] & 0xffffffff  # deterministic: the same on both contexts
        c.bpt_set_light_cache(entries + entries // 10)
        c.reset_accum()
        for it in range(2):
            c.set_light_samples(sampler.generate(128))
            c.draw(rt.BPT, rt.wanghash(it + 5))
        imgs.append(c.readback(w, h)[0])
        c.set_rr_start_depth(0)
    ctx.set_integrator_kernel(rt.KERNEL_AUTO)
    a, b = imgs
    print("means", a[..., :3].mean(), b[..., :3].mean(), flush=True)
    assert a[..., :3].mean() > 1e-3 and (a[..., 3] == 2.0).all() and (b[..., 3] == 2.0).all()
    ok = (np.abs(b - a) <= 1e-4 * np.abs(a) + 1e-6).all(-1)
    if kernel == "reference":
        assert ok.all(), np.argwhere(~ok)[:5]
    else:
        assert ok.mean() > 0.99
        assert abs(a[..., :3].mean() - b[..., :3].mean()) <= 1e-3 * a[..., :3].mean()


@pytest.mark.parametrize("kernel", ["reference", "wide"])
def test_wavefront_clutter_matches_megakernel(rt, ctx, wctx, kernel):
    """The clutter room (its emissive panels, path length 8, roulette from bounce 3)."""
    hs = rt.HostScene.clutter(300, workloads.MASTER_SEED)
    w, h = 192, 112
    k = rt.KERNEL_REFERENCE if kernel == "reference" else rt.KERNEL_WIDE
    print("case: wavefront clutter %s" % kernel, flush=True)
    a, b = draw_both(rt, ctx, wctx, hs, workloads.clutter_camera(w, h), w, h, k, workloads.CLUTTER_PATH_LENGTH,
                     workloads.CLUTTER_RR_START)
    print("means", a[..., :3].mean(), b[..., :3].mean(), flush=True)
    assert a[..., :3].mean() > 1e-3
    if kernel == "reference":
        bitexact(b, a)
    else:
        same = (a.view(np.uint32) == b.view(np.uint32)).all(-1)
        assert same.mean() > 0.99
        assert abs(a[..., :3].mean() - b[..., :3].mean()) <= 1e-3 * a[..., :3].mean()
    ctx.set_integrator_kernel(rt.KERNEL_AUTO)


def both_schedules(rt, c, hs, cam, w, h, path_length):
    """The same frames under the one-kernel and the wavefront schedule on context c."""
    imgs = []
    for sched in (rt.PT_MEGAKERNEL, rt.PT_WAVEFRONT):
        c.set_pt_schedule(sched)
        c.set_screen(w, h)
        c.set_camera(cam)
        c.set_max_path_length(path_length)
        c.reset_accum()
        sampler = rt.LightSampler(hs, 5)
        for k in range(2):
            c.set_light_samples(sampler.generate(128))
            c.draw(rt.PT, rt.wanghash(k + 3))
        imgs.append(c.readback(w, h)[0])
    c.set_pt_schedule(rt.PT_MEGAKERNEL)
    return imgs


@pytest.mark.parametrize("kernel", ["reference", "wide"])
def test_wavefront_environment_and_point_lights(rt, kernel):
    """Misses that add the environment map, NEE toward point lights."""
    hs = rt.HostScene.heightfield(16, 77)
    for p in workloads.heightfield_lights():
        hs.add_point_light(p, (0.5, 0.5, 0.5))
    faces = np.random.default_rng(3).integers(0, 256, size=(6, 8, 8, 3), dtype=np.uint8)
    c = rt.Context(0)
    try:
        c.load_scene(hs)
        c.set_integrator_kernel(rt.KERNEL_REFERENCE if kernel == "reference" else rt.KERNEL_WIDE)
        c.set_environment_map(faces)
        print("case: wavefront environment %s" % kernel, flush=True)
        a, b = both_schedules(rt, c, hs, workloads.heightfield_camera(96, 64), 96, 64, 4)
    finally:
        c.close()
    assert a[..., :3].mean() > 0.01
    if kernel == "reference":
        bitexact(b, a)
    else:
        assert (a.view(np.uint32) == b.view(np.uint32)).all(-1).mean() > 0.99


def test_wavefront_textures(rt):
    """Textured diffuse / reflectiveness / emissivity slots read at the hit's texcoord."""
    from test_gpu_textures import textured_cube, textures
    hs = textured_cube(rt)
    d8, refl, emis = textures()
    c = rt.Context(0)
    try:
        c.load_scene(hs)
        c.set_integrator_kernel(rt.KERNEL_REFERENCE)
        for slot, tex in ((rt.TEX_DIFFUSE, d8), (rt.TEX_REFLECTIVENESS, refl), (rt.TEX_EMISSIVITY, emis)):
            c.set_material_texture(0, slot, c.add_texture(tex))
        print("case: wavefront textures", flush=True)
        cam = rt.camera([0.1, 0.2, 0.3], [0.5, -0.3, -1.0], [0, 1, 0], 60, 32, 32)
        a, b = both_schedules(rt, c, hs, cam, 32, 32, 4)
    finally:
        c.close()
    assert a[..., :3].mean() > 0.1
    bitexact(b, a)


@pytest.mark.parametrize("kernel", ["reference", "wide"])
def test_wavefront_frame_graph_replay(rt, oracle, wctx, kernel):
    """One wavefront frame recorded into a HIP graph (every pass reads its queue
    length on the device) replays to the eager frame bit for bit, replay after
    replay with eager draws of other frames in between; a setter drops it."""
    print("case: wavefront graph replay %s" % kernel, flush=True)
    hs = rt.HostScene.cornell()
    w = h = 64
    cam = workloads.cornell_camera(w)
    wctx.load_scene(hs)
    wctx.set_screen(w, h)
    wctx.set_camera(cam)
    wctx.set_max_path_length(5)
    wctx.set_integrator_kernel(rt.KERNEL_REFERENCE if kernel == "reference" else rt.KERNEL_WIDE)
    ls = rt.LightSampler(hs, 123).generate(128)
    wctx.set_light_samples(ls)
    seed = rt.wanghash(7)
    wctx.reset_accum()
    wctx.draw(rt.PT, seed)
    eager, _ = wctx.readback(w, h)
    if kernel == "reference":
        sc = oracle.Scene(hs.arrays())
        bitexact(eager, oracle.pathtrace(sc, ls, cam, w, h, seed, 5))
    wctx.capture_draw(rt.PT, seed)
    for k in range(3):
        wctx.reset_accum()
        wctx.draw(rt.PT, rt.wanghash(100 + k))  # eager work between the replays
        wctx.sync()
        wctx.reset_accum()
        wctx.replay_draw()
        img, n = wctx.readback(w, h)
        assert n == 1
        bitexact(img, eager)
    # two replays add the frame twice
    wctx.reset_accum()
    wctx.replay_draw()
    wctx.replay_draw()
    img, n = wctx.readback(w, h)
    assert n == 2
    bitexact(img, eager + eager)
    wctx.set_camera(cam)
    with pytest.raises(rt.GpugiError):
        wctx.replay_draw()
    wctx.set_pt_schedule(rt.PT_MEGAKERNEL)
    with pytest.raises(rt.GpugiError):
        wctx.capture_draw(rt.PT, seed)
    wctx.set_pt_schedule(rt.PT_WAVEFRONT)
    with pytest.raises(rt.GpugiError):  # only PT frames are recorded, also under this schedule
        wctx.capture_draw(rt.BPT, seed)
