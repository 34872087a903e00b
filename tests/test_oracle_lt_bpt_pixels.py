"""The light tracer and BPT oracles pinned per pixel by closed forms.

Two scenes with a closed-form image, both a Lambertian plane y = 0 (albedo a)
seen from above:

- under one point light of intensity I: L(x) = a/pi * I * cos / (r^2 + 1e-5);
- under one downward-facing square area light of radiance Le (outside the
  view): L(x) = a/pi * Le * integral over the light of cos_x cos_y / r^2 dA,
  taken by a 24 x 24 midpoint rule.

Every other path escapes to the black environment (the light has albedo 0),
so each pixel is the integral of L over its footprint (8 x 8 supersamples).
The light tracer must converge to that image pixel by pixel; so must BPT.
For BPT this holds because the reference's heuristic MIS weights of the three
ways of making a camera-plane-light path add up to one on a Lambertian plane:
with p = cos/pi at the plane vertex toward the light, the eye connection gets
p/(1+p) * l/(l+p) (l = cos_y/pi * d^2 at an area light, 1 for a point light),
the eye path hitting the light p/(1+p) * p/(p+l), and the light path's camera
splat 1/(1+p) (pathtracer_bidir.comp:62-157, screenreproject.glsl:150-170,
lightcache.glsl:61-89).  Dropping DistanceMISHeuristic from one side only
breaks that sum, and moving the splat's pixel breaks the image.

The checks compare each pixel, each row and column sum, the image centroid
and the image mean.  The first row and column are left out: ProjectToScreen
rounds the screen position to the nearest pixel, so they receive half a
footprint (screenreproject.glsl:94-95).  So is the pixel where
showlightcaches.comp splats the point light itself.

test_planted_* build the oracle with a known mistake (oracle.planted) and show
that the same checks reject it: a y-flip or a one-pixel shift of the splat, and
a dropped DistanceMISHeuristic on either side.
"""
import concurrent.futures
import os

import numpy as np
import pytest

from gpugi_amd import workloads

A = 0.6  # 153 / 255: exact through the UNORM8 quantization of the albedo
W = H = 32
POINT = np.array([0.3, 1.5, -0.2])
INTENSITY = np.array([2.0, 3.0, 4.0])
LE = np.array([3.0, 2.0, 1.0])
LIGHT_C = np.array([1.5, 1.1, -0.3])  # centre of the area light, outside the camera's view
LIGHT_HALF = 0.25
CAM = ([0.0, 3.0, 0.5], [0.0, 0.0, 0.0], [0, 1, 0], 40)
WORKERS = min(8, os.cpu_count() or 1)


def plane_material(rt):
    return rt.default_material("Plane", albedo=(A, A, A), specular=(0, 0, 0))


def point_scene(rt):
    pos = np.array([[-20, 0, -20], [20, 0, -20], [20, 0, 20], [-20, 0, 20]], np.float32)
    nrm = np.tile(np.array([[0, 1, 0]], np.float32), (4, 1))
    hs = rt.HostScene.build(pos, np.array([[0, 2, 1], [0, 3, 2]]), normals=nrm, materials=[plane_material(rt)])
    hs.add_point_light(POINT, INTENSITY)
    return hs


def area_scene(rt):
    x0, x1 = LIGHT_C[0] - LIGHT_HALF, LIGHT_C[0] + LIGHT_HALF
    z0, z1 = LIGHT_C[2] - LIGHT_HALF, LIGHT_C[2] + LIGHT_HALF
    y = LIGHT_C[1]
    pos = np.array([[-20, 0, -20], [20, 0, -20], [20, 0, 20], [-20, 0, 20],
                    [x0, y, z0], [x1, y, z0], [x1, y, z1], [x0, y, z1]], np.float32)
    nrm = np.array([[0, 1, 0]] * 4 + [[0, -1, 0]] * 4, np.float32)
    # the light quad is wound so that cross(v0 - v1, v0 - v2) points down: its emission side
    idx = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7]])
    light = rt.default_material("Light", albedo=(0, 0, 0), specular=(0, 0, 0), emissivity=tuple(LE))
    return rt.HostScene.build(pos, idx, normals=nrm, material_ids=np.array([0, 0, 1, 1]),
                              materials=[plane_material(rt), light])


def _plane_points(cam, sub):
    """Hit points on y = 0 of sub x sub supersamples per pixel (FP64)."""
    U, V, Wv = (np.array(getattr(cam, k), np.float64) for k in ("u", "v", "w"))
    o = np.array(cam.position[:], np.float64)
    ys, xs = np.mgrid[0:H, 0:W]
    for j in range(sub):
        for i in range(sub):
            sx = (xs + (i + 0.5) / sub - 0.5) / W * 2 - 1
            sy = (ys + (j + 0.5) / sub - 0.5) / H * 2 - 1
            d = sx[..., None] * U + sy[..., None] * V + Wv
            t = -o[1] / d[..., 1]
            yield o + t[..., None] * d


def point_closed_form(cam, sub=8):
    out = np.zeros((H, W, 3))
    for x in _plane_points(cam, sub):
        l = POINT - x
        r2 = (l * l).sum(-1) + 1e-5
        cos = np.clip(l[..., 1] / np.sqrt(r2), 0, None)
        out += (A / np.pi) * INTENSITY * (cos / r2)[..., None]
    return out / (sub * sub)


def area_closed_form(cam, sub=8, quad=24):
    g = (np.arange(quad) + 0.5) / quad * 2 * LIGHT_HALF - LIGHT_HALF
    lx, lz = np.meshgrid(LIGHT_C[0] + g, LIGHT_C[2] + g, indexing="ij")
    lx, lz = lx.reshape(-1), lz.reshape(-1)
    dA = (2 * LIGHT_HALF / quad) ** 2
    out = np.zeros((H, W, 3))
    for x in _plane_points(cam, sub):
        dx = lx - x[..., 0:1]
        dz = lz - x[..., 2:3]
        dy = LIGHT_C[1] - x[..., 1:2]
        r2 = dx * dx + dy * dy + dz * dz
        out += (A / np.pi) * LE * ((dy * dy / (r2 * r2)).sum(-1) * dA)[..., None]
    return out / (sub * sub)


def render(rt, oracle, hs, sc, cam, kind, iterations, light_samples, path_length=3, seed=321, size=(W, H)):
    """Mean image of `iterations` draws.  The light-sample batches and the BPT
    cache capacity are made in order first; the draws then run on a thread pool
    (the oracle's C calls release the GIL), each into its own buffer."""
    w, h = size
    sampler = rt.LightSampler(hs, seed)
    cap = None
    if kind == "bpt":
        counts = [oracle.bpt_warmup(sc, sampler.generate(light_samples), cam, w, h, rt.wanghash(k + 1), path_length)
                  for k in range(8)]
        mean = sum(c[0] for c in counts) // 8
        cap = mean + mean // 10
    batches = [sampler.generate(light_samples) for _ in range(iterations)]

    def draw(it):
        acc = np.zeros((h, w, 4), np.float32)
        if kind == "pt":
            oracle.pathtrace(sc, batches[it], cam, w, h, rt.wanghash(it), path_length, acc)
        elif kind == "lt":
            oracle.lighttrace(sc, batches[it], cam, w, h, rt.wanghash(it), path_length, acc)
        else:
            oracle.bpt_draw(sc, batches[it], cam, w, h, rt.wanghash(it), path_length, cap, acc)
        return acc[..., :3].astype(np.float64)

    with concurrent.futures.ThreadPoolExecutor(max_workers=WORKERS) as ex:
        return sum(ex.map(draw, range(iterations))) / iterations


def failures(got, want, keep, pixel_tol, line_tol, mean_tol, centroid_tol=0.05):
    """Which of the per-pixel, row/column-sum, centroid and mean checks the image fails."""
    out = []
    g = np.where(keep[..., None], got, 0.0).sum(-1)
    w = np.where(keep[..., None], want, 0.0).sum(-1)
    rel = np.abs(g[keep] / w[keep] - 1)
    if rel.max() >= pixel_tol:
        out.append("pixel %.4f" % rel.max())
    for axis, name in ((1, "row"), (0, "column")):
        gs, ws = g.sum(axis), w.sum(axis)
        r = np.abs(gs[ws > 0] / ws[ws > 0] - 1)
        if r.max() >= line_tol:
            out.append("%s sum %.4f" % (name, r.max()))
    ys, xs = np.mgrid[0:H, 0:W]
    for c, name in ((xs, "x"), (ys, "y")):
        d = (g * c).sum() / g.sum() - (w * c).sum() / w.sum()
        if abs(d) >= centroid_tol:
            out.append("centroid %s %.4f" % (name, d))
    m = abs(g.sum() / w.sum() - 1)
    if m >= mean_tol:
        out.append("mean %.4f" % m)
    return out


def base_keep():
    keep = np.ones((H, W), bool)
    keep[0, :] = False
    keep[:, 0] = False
    return keep


@pytest.fixture(scope="module")
def point(rt, oracle):
    hs = point_scene(rt)
    sc = oracle.Scene(hs.arrays())
    cam = rt.camera(*CAM, W, H)
    keep = base_keep()
    lx, ly = oracle.screen_pixel(cam, W, H, POINT)
    # showlightcaches.comp:57-58 truncates where ProjectToScreen rounds
    U, V, Wv = (np.array(getattr(cam, k), np.float64) for k in ("u", "v", "w"))
    c = np.array(cam.position[:], np.float64) - POINT
    tx = int((-U.dot(c) / -Wv.dot(c) / U.dot(U) * 0.5 + 0.5) * W)
    ty = int((-V.dot(c) / -Wv.dot(c) / V.dot(V) * 0.5 + 0.5) * H)
    keep[ty, tx] = False
    return hs, sc, cam, point_closed_form(cam), keep, (tx, ty)


@pytest.fixture(scope="module")
def area(rt, oracle):
    hs = area_scene(rt)
    sc = oracle.Scene(hs.arrays())
    cam = rt.camera(*CAM, W, H)
    return hs, sc, cam, area_closed_form(cam), base_keep()


# pixel, row/column and mean tolerances: about 6, 5 and 5 standard deviations
# of the sampling noise at these sample counts
LT_POINT = dict(iterations=64, light_samples=4096)
LT_POINT_TOL = dict(pixel_tol=0.15, line_tol=0.025, mean_tol=0.01)
BPT_POINT = dict(iterations=64, light_samples=1024)
BPT_POINT_TOL = dict(pixel_tol=0.25, line_tol=0.04, mean_tol=0.01)
LT_AREA = dict(iterations=48, light_samples=2048)
LT_AREA_TOL = dict(pixel_tol=0.25, line_tol=0.04, mean_tol=0.01)
BPT_AREA = dict(iterations=64, light_samples=1024)
BPT_AREA_TOL = dict(pixel_tol=0.5, line_tol=0.06, mean_tol=0.01)


def test_screen_pixel_inverts_primary_rays(rt, oracle):
    """A point on a primary ray of pixel (x, y) projects back into (x, y): the
    splat's rounding and orientation are those of the jittered primary rays
    (pathtracer.comp:39-46), whose image the path-tracer tests pin."""
    cam = rt.camera(*CAM, W, H)
    n = 4096
    rays = oracle.raygen_primary(cam, W, H, workloads.MASTER_SEED, n)
    idx = np.arange(n) % (W * H)
    px, py = idx % W, idx // W
    got = np.array([oracle.screen_pixel(cam, W, H, r[0:3] + 2.5 * r[4:7]) or (-1, -1) for r in rays])
    # the jitter is recovered from the direction to tell apart the rays that end
    # within float rounding of a pixel edge
    U, V, Wv = (np.array(getattr(cam, k), np.float64) for k in ("u", "v", "w"))
    d = rays[:, 4:7].astype(np.float64)
    sx = (d @ U) / (d @ Wv) / U.dot(U)
    sy = (d @ V) / (d @ Wv) / V.dot(V)
    fx = (sx * 0.5 + 0.5) * W - px
    fy = (sy * 0.5 + 0.5) * H - py
    edge = (np.abs(np.abs(fx) - 0.5) < 1e-3) | (np.abs(np.abs(fy) - 0.5) < 1e-3)
    # pixel 0's left half and row 0's lower half fall outside the screen window
    outside = (got[:, 0] < 0)
    assert np.all(outside <= ((px == 0) & (fx < 0) | (py == 0) & (fy < 0) | edge))
    ok = ~outside & ~edge
    assert np.array_equal(got[ok, 0], px[ok]) and np.array_equal(got[ok, 1], py[ok])
    assert ok.mean() > 0.95
    # the view direction itself lands in the centre pixel
    centre = np.array(cam.position[:]) + 2.0 * np.array(cam.w[:])
    assert oracle.screen_pixel(cam, W, H, centre) == (W // 2, H // 2)
    # behind the camera: no pixel
    assert oracle.screen_pixel(cam, W, H, np.array(cam.position[:]) - np.array(cam.w[:])) is None


def test_lt_point_light_per_pixel(rt, oracle, point):
    hs, sc, cam, want, keep, (tx, ty) = point
    got = render(rt, oracle, hs, sc, cam, "lt", **LT_POINT)
    assert got[ty, tx].min() > 100 * want[ty, tx].max()  # the light sample's own splat is there
    assert failures(got, want, keep, **LT_POINT_TOL) == []


def test_bpt_point_light_per_pixel(rt, oracle, point):
    hs, sc, cam, want, keep, _ = point
    got = render(rt, oracle, hs, sc, cam, "bpt", **BPT_POINT)
    assert failures(got, want, keep, **BPT_POINT_TOL) == []


def test_pt_area_light_per_pixel(rt, oracle, area):
    hs, sc, cam, want, keep = area
    got = render(rt, oracle, hs, sc, cam, "pt", iterations=256, light_samples=128)
    assert failures(got, want, keep, pixel_tol=0.15, line_tol=0.02, mean_tol=0.01) == []


def test_lt_area_light_per_pixel(rt, oracle, area):
    hs, sc, cam, want, keep = area
    got = render(rt, oracle, hs, sc, cam, "lt", **LT_AREA)
    assert failures(got, want, keep, **LT_AREA_TOL) == []


def test_bpt_area_light_per_pixel(rt, oracle, area):
    hs, sc, cam, want, keep = area
    got = render(rt, oracle, hs, sc, cam, "bpt", **BPT_AREA)
    assert failures(got, want, keep, **BPT_AREA_TOL) == []


@pytest.mark.parametrize("mistake", ["YFLIP", "XSHIFT"])
def test_planted_splat_mistakes_are_rejected(rt, oracle, point, mistake):
    hs, sc, cam, want, keep, _ = point
    with oracle.planted(mistake):
        got = render(rt, oracle, hs, sc, cam, "lt", **LT_POINT)
        bpt = render(rt, oracle, hs, sc, cam, "bpt", **BPT_POINT)
    assert failures(got, want, keep, **LT_POINT_TOL) != []
    assert failures(bpt, want, keep, **BPT_POINT_TOL) != []


@pytest.mark.parametrize("mistake", ["NO_LIGHT_DIST_MIS", "NO_HIT_DIST_MIS"])
def test_planted_mis_mistakes_are_rejected(rt, oracle, area, mistake):
    hs, sc, cam, want, keep = area
    with oracle.planted(mistake):
        got = render(rt, oracle, hs, sc, cam, "bpt", **BPT_AREA)
    assert failures(got, want, keep, **BPT_AREA_TOL) != []


def test_lt_and_bpt_match_pt_per_block_on_cornell(rt, oracle):
    """Low-pass per-pixel agreement on the Cornell room, where paths bounce up
    to five times: 4 x 4-pixel block sums of LT and BPT against a 1024-sample
    PT render, away from pixels that can see the emitter and the border (see
    test_oracle_integrators.emitter_mask).  At these sample counts the block
    noise is about 2 % for LT and BPT and 0.6 % for PT."""
    from test_oracle_integrators import emitter_mask
    hs = rt.HostScene.cornell()
    sc = oracle.Scene(hs.arrays())
    w = h = 32
    cam = workloads.cornell_camera(w)
    mask = ~emitter_mask(rt, oracle, sc, cam, w, h)
    full = mask.reshape(8, 4, 8, 4).sum((1, 3)) >= 8

    def blocks(kind, iterations, light_samples):
        img = render(rt, oracle, hs, sc, cam, kind, iterations, light_samples, path_length=5, seed=123, size=(w, h))
        return np.where(mask, img.sum(-1), 0).reshape(8, 4, 8, 4).sum((1, 3))[full]

    pt = blocks("pt", 1024, 128)
    for kind, its, nls in (("lt", 96, 1024), ("bpt", 384, 128)):
        b = blocks(kind, its, nls)
        rel = np.abs(b / pt - 1)
        assert rel.max() < 0.12, (kind, rel.max())
        assert abs(b.sum() / pt.sum() - 1) < 0.015, (kind, b.sum() / pt.sum())
