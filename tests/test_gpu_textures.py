"""Material textures on the GPU (gpugi_add_texture / gpugi_set_material_texture)
against the oracle: a closed emitting cube whose faces carry a sRGB diffuse
texture, a float reflectiveness texture (Phong exponent in alpha) and a float
emissivity texture, texture coordinates spanning two periods per face
(REPEAT).  PT bit for bit with the reference traversal, LT to the splat order,
BPT with its light cache bit for bit (light vertices keep their texture
coordinate) and the image to 1e-4."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W = H = 32
L = 4


def textured_cube(rt):
    P = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float32)
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    pos, nrm, tc, idx = [], [], [], []
    for f in faces:
        q = P[list(f)]
        n = -q.mean(0) / np.linalg.norm(q.mean(0))
        b = len(pos)
        for v, t in zip(q, [(0, 0), (2, 0), (2, 2), (0, 2)]):
            pos.append(v)
            nrm.append(n)
            tc.append(t)
        idx += [[b, b + 2, b + 1], [b, b + 3, b + 2]]
    m = rt.default_material("Tex", albedo=(0.5, 0.5, 0.5), specular=(0.2, 0.2, 0.2), emissivity=(0.5, 0.4, 0.3))
    return rt.HostScene.build(np.array(pos), np.array(idx), normals=np.array(nrm), texcoords=np.array(tc),
                              materials=[m])


def textures():
    rng = np.random.default_rng(8)
    diffuse8 = rng.integers(0, 256, size=(8, 8, 4), dtype=np.uint8)
    refl = np.ones((4, 4, 4), np.float32)
    refl[..., 3] = rng.uniform(2.0, 60.0, size=(4, 4))
    emis = rng.uniform(0.0, 0.8, size=(5, 3, 4)).astype(np.float32)
    return diffuse8, refl, emis


def to_linear(t8):
    c = np.arange(256, dtype=np.float64) / 255.0
    lut = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4).astype(np.float32)
    out = lut[t8].astype(np.float32)
    out[..., 3] = t8[..., 3].astype(np.float32) / np.float32(255)
    return out


@pytest.fixture(scope="module")
def scene(rt):
    return textured_cube(rt)


@pytest.fixture(scope="module")
def tctx(rt):
    c = rt.Context(0)  # its own context: the bound textures stay out of other modules' tests
    yield c
    c.close()


def setup(rt, ctx, hs):
    ctx.load_scene(hs)
    ctx.set_screen(W, H)
    ctx.set_camera(rt.camera([0.1, 0.2, 0.3], [0.5, -0.3, -1.0], [0, 1, 0], 60, W, H))
    ctx.set_max_path_length(L)
    ctx.set_integrator_kernel(rt.KERNEL_REFERENCE)
    d8, refl, emis = textures()
    ids = [ctx.add_texture(d8), ctx.add_texture(refl), ctx.add_texture(emis)]
    ctx.set_material_texture(0, rt.TEX_DIFFUSE, ids[0])
    ctx.set_material_texture(0, rt.TEX_REFLECTIVENESS, ids[1])
    ctx.set_material_texture(0, rt.TEX_EMISSIVITY, ids[2])
    return np.array([[0, -1, 1, 2]]), [to_linear(d8), refl, emis]


def close(img, ref):
    return np.abs(img - ref) <= 1e-4 * np.abs(ref) + 1e-6


@pytest.mark.parametrize("integrator", ["pt", "lt"])
def test_textured_integrators_match_oracle(rt, oracle, tctx, scene, integrator):
    ctx = tctx
    mat_tex, texs = setup(rt, ctx, scene)
    sc = oracle.Scene(scene.arrays())
    cam = rt.camera([0.1, 0.2, 0.3], [0.5, -0.3, -1.0], [0, 1, 0], 60, W, H)
    sampler = rt.LightSampler(scene, 123)
    acc = np.zeros((H, W, 4), np.float32)
    print("case: textured %s" % integrator, flush=True)
    oracle.set_textures(mat_tex, texs)
    try:
        for it in range(2):
            ls = sampler.generate(128 if integrator == "pt" else 256)
            ctx.set_light_samples(ls)
            ctx.draw(rt.PT if integrator == "pt" else rt.LT, rt.wanghash(it))
            if integrator == "pt":
                oracle.pathtrace(sc, ls, cam, W, H, rt.wanghash(it), L, acc)
            else:
                oracle.lighttrace(sc, ls, cam, W, H, rt.wanghash(it), L, acc)
    finally:
        oracle.set_textures(None, None)
    img, n = ctx.readback(W, H)
    assert n == 2 and acc[..., :3].mean() > 0.1
    if integrator == "pt":
        assert np.array_equal(img.view(np.uint32), acc.view(np.uint32))
    else:
        ok = close(img, acc)
        assert ok.all(), np.argwhere(~ok)[:5]


def test_textured_bpt_matches_oracle(rt, oracle, tctx, scene):
    ctx = tctx
    mat_tex, texs = setup(rt, ctx, scene)
    sc = oracle.Scene(scene.arrays())
    cam = rt.camera([0.1, 0.2, 0.3], [0.5, -0.3, -1.0], [0, 1, 0], 60, W, H)
    sampler = rt.LightSampler(scene, 123)
    print("case: textured bpt", flush=True)
    oracle.set_textures(mat_tex, texs)
    try:
        counts = []
        for k in range(8):
            ls = sampler.generate(128)
            ctx.set_light_samples(ls)
            g = ctx.bpt_warmup(rt.wanghash(k + 1))
            assert g == oracle.bpt_warmup(sc, ls, cam, W, H, rt.wanghash(k + 1), L)
            counts.append(g)
        mean = sum(c[0] for c in counts) // 8
        cap = mean + mean // 10
        ctx.bpt_set_light_cache(cap)
        ctx.reset_accum()
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
    finally:
        oracle.set_textures(None, None)
    img, n = ctx.readback(W, H)
    ok = close(img, acc)
    assert ok.all(), np.argwhere(~ok)[:5]


def test_textures_are_dropped_with_the_scene(rt, tctx, scene):
    ctx = tctx
    setup(rt, ctx, scene)
    sampler = rt.LightSampler(scene, 1)
    ctx.set_light_samples(sampler.generate(128))
    ctx.load_scene(scene)  # a scene load drops textures and bindings
    ctx.set_light_samples(sampler.generate(128))
    ctx.reset_accum()
    ctx.draw(rt.PT, 5)
    a, _ = ctx.readback(W, H)
    with pytest.raises(rt.GpugiError):
        ctx.set_material_texture(0, rt.TEX_DIFFUSE, 0)  # texture 0 no longer exists
    ctx.set_material_texture(0, rt.TEX_DIFFUSE, -1)
    ctx.reset_accum()
    ctx.draw(rt.PT, 5)
    b, _ = ctx.readback(W, H)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_json_textures_through_load_scene_and_script(rt, oracle, tctx, scene, tmp_path):
    """A scene whose material json names a TGA albedo texture: Context.load_scene
    uploads and binds it (PT bit for bit against the oracle given the decoded
    texels), and gpugi_render renders the same image from the files."""
    import json
    import os
    import subprocess
    from test_host_textures import tga
    scene.save(str(tmp_path / "cube.rawscene"), str(tmp_path / "cube.json"))
    rgb = np.random.default_rng(9).integers(0, 256, (6, 10, 3), dtype=np.uint8)
    tga(tmp_path / "albedo.tga", rgb)
    with open(tmp_path / "cube.json") as f:
        j = json.load(f)
    j["materials"]["Tex"]["albedo"] = "albedo.tga"
    with open(tmp_path / "cube.json", "w") as f:
        json.dump(j, f)
    hs = rt.HostScene.load(str(tmp_path / "cube.rawscene"), str(tmp_path / "cube.json"))
    ctx = tctx
    ctx.load_scene(hs)
    ctx.set_screen(W, H)
    cam = rt.camera([0.1, 0.2, 0.3], [0.5, -0.3, -1.0], [0, 1, 0], 60, W, H)
    ctx.set_camera(cam)
    ctx.set_max_path_length(L)
    ctx.set_integrator_kernel(rt.KERNEL_REFERENCE)
    sampler = rt.LightSampler(hs, 123)
    sc = oracle.Scene(hs.arrays())
    acc = np.zeros((H, W, 4), np.float32)
    print("case: json texture", flush=True)
    oracle.set_textures(hs.material_textures(), [to_linear(t) if t.dtype == np.uint8 else t for t in hs.textures()])
    try:
        ls = sampler.generate(128)
        ctx.set_light_samples(ls)
        ctx.draw(rt.PT, rt.wanghash(0))
        oracle.pathtrace(sc, ls, cam, W, H, rt.wanghash(0), L, acc)
    finally:
        oracle.set_textures(None, None)
    img, _ = ctx.readback(W, H)
    assert np.array_equal(img.view(np.uint32), acc.view(np.uint32))
    # the script front end: same scene files, same first iteration
    app = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpugi_amd", "bin", "gpugi_render")
    out = str(tmp_path / "s.pfm")
    lines = ["sceneFilename = " + str(tmp_path / "cube.rawscene"), "cameraPos = 0.1, 0.2, 0.3",
             "cameraLookAt = 0.6, -0.1, -0.7", "cameraFOV = 60", "resolution = %d, %d" % (W, H),
             "pathLength = %d" % L, "traversal = 1", "renderer = 0", "waitIterations = 1", "namedscreenshot = " + out,
             "exit"]
    args = [app]
    for l in lines:
        args += ["-e", l]
    r = subprocess.run(args, capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stdout + r.stderr
    with open(out, "rb") as f:
        f.readline(), f.readline(), f.readline()
        simg = np.frombuffer(f.read(), dtype="<f4").reshape(H, W, 3)
    assert simg.mean() > 0.1
    plain = rt.HostScene.load(str(tmp_path / "cube.rawscene"), None)
    assert plain.textures() == []
