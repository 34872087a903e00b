"""The HierarchyImportance passes of the oracle (oracle_hier.c) pinned by their
definitions on small scenes.

- Acquisition (acquisition2.comp, counthits.glsl): every added value is a
  power of two (1 for a camera ray, halved at every vertex), so each node's
  and slot's sum is exact whatever the order; a brute-force restatement in
  numpy FP64 (every box the ray's segment enters, found by testing every
  node, and every triangle it hits before its end) must give the same
  counts on rays whose segments are not within float rounding of a box or
  triangle boundary.  The root is entered by every camera ray that starts
  inside or hits the scene box.
- Propagation (hierarchypropagation_*.comp): x is divided by the iteration
  count; a leaf node's subtree importance is what its two slots received
  (at least 1.19e-7); an inner node's is the sum of its children's; y is
  (x / subtree)^2 times a third of the box surface.  Checked in FP64 from the
  acquired values.
- Materials (sampleleafmaterials.comp, pullmaterials.comp): a scene with one
  material reproduces that material's constants at every node, to half
  precision; a two-material scene's root is the area-weighted mean; the
  stored areas add up the tree.
- Half conversion: against numpy's float16 (IEEE round to nearest even) on
  every float16 boundary and random values.
"""
import numpy as np
import pytest

from gpugi_amd import workloads


def parents_fp(sc):
    h = sc.a["hierarchy"]
    n = len(h)
    fc = h[:, 3].view(np.uint32)
    esc = h[:, 7].view(np.uint32)
    parent = -np.ones(n, np.int64)
    for p in range(n):
        if fc[p] & 0x80000000:
            continue
        end = n if esc[p] == 0 else int(esc[p])
        k = int(fc[p])
        while p < k < end:
            parent[k] = p
            if esc[k] == 0:
                break
            k = int(esc[k])
    return parent


def test_half_conversion_matches_ieee(oracle):
    rng = np.random.default_rng(1)
    vals = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(1000.0),
        rng.uniform(-1e-4, 1e-4, 5000).astype(np.float32),            # subnormal halves
        np.array([0.0, -0.0, 65504.0, 65520.0, 65519.99, 1e9, -1e9, 6.1035156e-05, 5.9604645e-08,
                  2.9802322e-08, 2.9802326e-08, np.inf, -np.inf], np.float32),
    ])
    # every exact midpoint between neighbouring halves
    h = np.arange(0, 0x7c00, dtype=np.uint16).view(np.float16).astype(np.float32)
    mids = ((h[:-1].astype(np.float64) + h[1:].astype(np.float64)) / 2).astype(np.float32)
    vals = np.concatenate([vals, mids, -mids, h])
    got = np.array([oracle.half_bits(float(v)) for v in vals], np.uint16)
    want = vals.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)
    back = np.array([oracle.half_value(int(b)) for b in want[:5000]], np.float32)
    assert np.array_equal(back, want[:5000].view(np.float16).astype(np.float32))


def box_hits_fp32(h, o, d, length):
    """IntersectBox (intersectiontests.glsl:9-20) of one ray against every node box, in float32."""
    o = o.astype(np.float32)
    inv = (np.float32(1.0) / d.astype(np.float32)).astype(np.float32)
    mn, mx = h[:, 0:3].astype(np.float32), h[:, 4:7].astype(np.float32)
    with np.errstate(invalid="ignore", over="ignore"):
        tbot = inv * (mn - o)
        ttop = inv * (mx - o)
    tmn, tmx = np.minimum(ttop, tbot), np.maximum(ttop, tbot)
    first = np.maximum(np.float32(0), np.maximum(np.maximum(tmn[:, 0], tmn[:, 1]), np.maximum(tmn[:, 0], tmn[:, 2])))
    last = np.minimum(np.minimum(tmx[:, 0], tmx[:, 1]), np.minimum(tmx[:, 0], tmx[:, 2]))
    return (first <= last) & (first <= np.float32(length))


@pytest.mark.parametrize("name", ["cornell", "hf16"])
def test_count_hits_equals_every_entered_box(rt, oracle, name):
    """TraceRayCountHits adds to every node whose box the segment enters: in a
    BVH a child box lies inside its parent's, so the walk reaches exactly the
    nodes a test of every node box accepts.  Triangle slots: every valid slot
    the segment hits (brute force over all slots)."""
    hs = rt.HostScene.cornell() if name == "cornell" else rt.HostScene.heightfield(16, workloads.MASTER_SEED)
    sc = oracle.Scene(hs.arrays())
    h = sc.a["hierarchy"]
    nn = len(h)
    rng = np.random.default_rng(3)
    n = 400
    lo, hi = h[0, 0:3], h[0, 4:7]
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = lo + (hi - lo) * rng.uniform(-0.2, 1.2, (n, 3))
    d = rng.standard_normal((n, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3] = np.where(rng.uniform(size=n) < 0.5, np.float32(np.inf), rng.uniform(0.1, 2.0, n)).astype(np.float32)
    w = np.float32(0.5) ** rng.integers(0, 8, n).astype(np.float32)
    got = oracle.count_hits(sc, rays, w)
    want_nodes = np.zeros(nn, np.float64)
    slots = sc.a["triangles"]
    valid = slots[:, 0] != slots[:, 1]
    tri_hit = oracle.slot_hits(sc, rays)  # (n, slots): hit at t <= length
    for i in range(n):
        want_nodes += w[i] * box_hits_fp32(h, rays[i, 0:3], rays[i, 4:7], rays[i, 3])
    want_slots = (tri_hit * w[:, None]).sum(0)
    assert np.array_equal(got[:nn, 0].astype(np.float64), want_nodes)
    assert np.array_equal(got[nn:, 0].astype(np.float64), want_slots)
    assert (got[nn:, 0][~valid] == 0).all() and not got[:, 1].any()
    assert want_nodes[0] > 0 and want_slots.sum() > 0


def plane_scene(rt):
    pos = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32)
    nrm = np.tile(np.array([[0, 1, 0]], np.float32), (4, 1))
    m = rt.default_material("Plane", albedo=(0.6, 0.6, 0.6), specular=(0, 0, 0))
    hs = rt.HostScene.build(pos, np.array([[0, 2, 1], [0, 3, 2]]), normals=nrm, materials=[m])
    hs.add_point_light([0.3, 1.5, -0.2], [2.0, 3.0, 4.0])
    return hs


def test_acquisition_closed_form_on_a_plane(rt, oracle):
    """One flat leaf seen by every camera ray, lit by a point light above it:
    the shadow rays are never blocked and the bounce rays leave the flat box at
    once, so per iteration the node counts exactly one unit per pixel and its
    two triangles share those units; propagation gives subtree = that count
    and y = a third of the box surface (sx * sz / 3: the box is flat)."""
    hs = plane_scene(rt)
    sc = oracle.Scene(hs.arrays())
    cam = rt.camera([0.0, 3.0, 0.0], [0.0, 0.0, 0.0], [0, 0, -1], 30, 16, 16)
    ls = rt.LightSampler(hs, 1).generate(128)
    N = 10
    imp = np.zeros((len(sc.a["hierarchy"]) + len(sc.a["triangles"]), 2), np.float32)
    for it in range(1, N + 1):
        oracle.hier_acquire(sc, ls, cam, 16, 16, it, imp)
    assert len(sc.a["hierarchy"]) == 1
    assert imp[0, 0] == 16 * 16 * N
    assert imp[1:3, 0].sum() == 16 * 16 * N
    imp, sub = oracle.hier_propagate(sc, N, imp)
    assert imp[0, 0] == 256 and sub[0] == 256
    h = sc.a["hierarchy"][0]
    assert np.isclose(imp[0, 1], (h[4] - h[0]) * (h[6] - h[2]) / 3.0, rtol=1e-6)


@pytest.mark.parametrize("name", ["cornell", "hf16"])
def test_propagation_invariants(rt, oracle, name):
    hs = rt.HostScene.cornell() if name == "cornell" else rt.HostScene.heightfield(16, workloads.MASTER_SEED)
    sc = oracle.Scene(hs.arrays())
    w = h = 16
    cam = workloads.cornell_camera(w) if name == "cornell" else workloads.heightfield_camera(w, h)
    ls = rt.LightSampler(hs, 7).generate(128)
    N = 4
    nn = len(sc.a["hierarchy"])
    raw = np.zeros((nn + len(sc.a["triangles"]), 2), np.float32)
    for it in range(1, N + 1):
        oracle.hier_acquire(sc, ls, cam, w, h, it, raw)
    imp, sub = oracle.hier_propagate(sc, N, raw.copy())
    hier = sc.a["hierarchy"]
    fc = hier[:, 3].view(np.uint32)
    parent = parents_fp(sc)
    assert parent[0] == -1 and (parent[1:] >= 0).all() and (parent[1:] < np.arange(1, nn)).all()
    # x per iteration; a child is entered only by rays that entered its parent
    assert np.allclose(imp[:nn, 0], raw[:nn, 0] / np.float32(N), rtol=1e-7, atol=0)
    assert (raw[1:nn, 0] <= raw[parent[1:], 0]).all()
    leaf = (fc & 0x80000000) != 0
    first = 2 * (fc & 0x7fffffff).astype(np.int64)
    for i in range(nn):
        if leaf[i]:
            c = raw[nn + first[i], 0] + raw[nn + first[i] + 1, 0]
            want = max(c / N, 1.19e-07)
            assert np.isclose(sub[i], want, rtol=1e-6), (i, sub[i], want)
            assert raw[nn + first[i], 0] <= raw[i, 0] and raw[nn + first[i] + 1, 0] <= raw[i, 0]
        else:
            kids = np.nonzero(parent == i)[0]
            assert len(kids) == 2 and kids[0] == fc[i]
            assert np.isclose(sub[i], sub[kids].astype(np.float64).sum(), rtol=1e-6)
        bmin, bmax = hier[i, 0:3].astype(np.float64), hier[i, 4:7].astype(np.float64)
        s = bmax - bmin
        area = (s[0] * s[1] + s[0] * s[2] + s[1] * s[2]) / 3.0
        if leaf[i]:  # the approximation divides by what the slots received, without the floor
            c = (np.float64(raw[nn + first[i], 0]) + raw[nn + first[i] + 1, 0]) / N
            a = imp[i, 0] / c if c > 0 else np.inf
        else:
            a = imp[i, 0] / np.float64(sub[i])
        if np.isfinite(a):
            assert np.isclose(imp[i, 1], a * a * area, rtol=1e-5, atol=1e-30), (i, imp[i, 1], a * a * area)
    # the root is entered by every camera ray (the camera sits inside or looks at the scene box)
    assert imp[0, 0] >= (w // 8 * 8) * (h // 8 * 8) - 1e-3


def test_materials_of_one_material_and_area_weights(rt, oracle):
    hs = rt.HostScene.cornell()
    sc = oracle.Scene(hs.arrays())
    words = oracle.hier_materials(sc)
    nn = len(sc.a["hierarchy"])
    vals = np.array([[oracle.half_value(int(w) & 0xffff), oracle.half_value(int(w) >> 16)] for w in words.reshape(-1)],
                    np.float32).reshape(nn, 12, 2)
    area = vals[:, 3, 1]
    # the stored areas add up the tree (half precision)
    parent = parents_fp(sc)
    for i in range(nn):
        kids = np.nonzero(parent == i)[0]
        if len(kids):
            assert np.isclose(area[i], area[kids].sum(), rtol=2e-3)
    # total area of the room's triangles
    pos = sc.a["positions"].astype(np.float64)
    t = sc.a["triangles"]
    t = t[t[:, 0] != t[:, 1]]
    tri_area = 0.5 * np.linalg.norm(np.cross(pos[t[:, 1]] - pos[t[:, 0]], pos[t[:, 2]] - pos[t[:, 0]]), axis=1)
    assert np.isclose(area[0], tri_area.sum(), rtol=2e-3)
    # the root's diffuse colour is the area-weighted mean of the materials
    mats = sc.mat_floats
    want = (mats[t[:, 3], 0:3].astype(np.float64) * tri_area[:, None]).sum(0) / tri_area.sum()
    got = np.array([vals[0, 4, 0], vals[0, 4, 1], vals[0, 5, 0]])
    assert np.allclose(got, want, rtol=3e-3), (got, want)
    em = (mats[t[:, 3], 10:13].astype(np.float64) * tri_area[:, None]).sum(0) / tri_area.sum()
    assert np.allclose([vals[0, 6, 0], vals[0, 6, 1], vals[0, 7, 0]], em, rtol=3e-3)


def test_materials_single_material_is_constant(rt, oracle):
    hs = plane_scene(rt)
    sc = oracle.Scene(hs.arrays())
    words = oracle.hier_materials(sc)
    m = sc.mat_floats[0]
    w = words.reshape(-1, 12)[0]
    half = lambda x: np.float32(np.float16(x))
    assert oracle.half_value(int(w[4]) & 0xffff) == half(m[0])
    assert oracle.half_value(int(w[9]) >> 16) == half(m[16])  # refraction index
    assert oracle.half_value(int(w[3]) >> 16) == half(4.0)    # area of the 2 x 2 plane
