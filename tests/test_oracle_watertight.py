"""Watertight triangle test of the oracle (Woop, Benthin and Wald 2013).

Pinned by FP64 geometry on interior rays, by the closed-mesh property (a ray
leaving the inside of a closed mesh must hit it, also through shared edges
and vertices) and by exact special cases.  The reference test
(intersectiontests.glsl:63-79) is not watertight; the tests count its gaps on
the same rays for comparison."""
import numpy as np

from gpugi_amd import rt as _rt  # noqa: F401  (library fixture loads it)


def rays_array(o, d, tmax=3.402823466e38):
    r = np.zeros((len(d), 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmax
    r[:, 4:7] = d
    return r


def icosphere_scene(rt):
    hs = rt.HostScene.clutter(1, 11)
    a = hs.arrays()
    tris = a["triangles"].reshape(-1, 4)
    tris = tris[(tris[:, 0] != tris[:, 1]) & (tris[:, 3] >= 2)]
    pos = a["positions"]
    centre = pos[np.unique(tris[:, :3])].mean(0)
    return hs, a, tris, centre


def test_watertight_matches_fp64_in_the_interior(rt, oracle):
    rng = np.random.default_rng(4)
    p = rng.normal(size=(400, 3, 3)).astype(np.float32)
    positions = p.reshape(-1, 3)
    idx = np.arange(len(positions)).reshape(-1, 3)
    hs = rt.HostScene.build(positions, idx)
    sc = oracle.Scene(hs.arrays())
    a = hs.arrays()
    tris = a["triangles"].reshape(-1, 4)
    slots = np.nonzero(tris[:, 0] != tris[:, 1])[0]
    # a ray per triangle towards an interior point (barycentrics >= 0.05), from a random origin
    k = len(slots)
    w = rng.dirichlet([2, 2, 2], size=k) * 0.85 + 0.05
    w /= w.sum(1, keepdims=True)
    P = a["positions"][tris[slots, :3]]
    target = (w[:, :, None] * P).sum(1)
    o = target + rng.normal(size=(k, 3)) * 3
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    wt = oracle.bruteforce(sc, rays_array(o, d), watertight=True)
    ref = oracle.bruteforce(sc, rays_array(o, d))
    # both tests find a triangle for every ray; they agree wherever the fp64 winner is unambiguous
    assert (wt[:, 0] >= 0).all() and (ref[:, 0] >= 0).all()
    t_wt = wt[:, 1].view(np.float32).astype(np.float64)
    for i in range(k):
        s = wt[i, 0]
        p0, p1, p2 = (a["positions"][tris[s, j]].astype(np.float64) for j in range(3))
        n = np.cross(p1 - p0, p2 - p0)
        t64 = np.dot(p0 - o[i].astype(np.float32), n) / np.dot(d[i].astype(np.float32), n)
        assert abs(t_wt[i] - t64) <= 1e-4 * abs(t64) + 1e-5
    assert (wt[:, 0] == ref[:, 0]).mean() > 0.99


def test_closed_icosphere_has_no_gaps(rt, oracle):
    hs, a, tris, centre = icosphere_scene(rt)
    sc = oracle.Scene(a)
    pos = a["positions"]
    verts = pos[np.unique(tris[:, :3])]
    edges = set()
    for t in tris[:, :3]:
        for i in range(3):
            edges.add(tuple(sorted((int(t[i]), int(t[(i + 1) % 3])))))
    mids = np.array([(pos[e0] + pos[e1]) * np.float32(0.5) for e0, e1 in sorted(edges)], np.float32)
    targets = np.concatenate([verts, mids])
    rng = np.random.default_rng(8)
    o = centre + (rng.random((len(targets), 3)).astype(np.float32) - 0.5) * 0.01
    d = targets - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rays_array(o, d)
    wt = oracle.bruteforce(sc, r, watertight=True)
    ref = oracle.bruteforce(sc, r)
    print("rays through vertices/edges: %d, watertight misses %d, reference misses %d"
          % (len(r), (wt[:, 0] < 0).sum(), (ref[:, 0] < 0).sum()))
    assert (wt[:, 0] >= 0).all()
    # the escape walk with the watertight test finds the same triangles
    walk = oracle.trace(sc, r, watertight=True)
    assert (walk[:, 0] >= 0).mean() > 0.999


def test_heightfield_vertices_and_edges_have_no_gaps(rt, oracle):
    cells = 16
    hs = rt.HostScene.heightfield(cells, 5)
    a = hs.arrays()
    sc = oracle.Scene(a)
    pos = a["positions"]
    n1 = cells + 1
    grid = pos.reshape(-1, 3)
    inner = [j * n1 + i for j in range(1, cells) for i in range(1, cells)]
    targets = [grid[k] for k in inner]
    for j in range(1, cells):
        for i in range(1, cells):
            k = j * n1 + i
            targets.append((grid[k] + grid[k + 1]) * np.float32(0.5))       # along x
            targets.append((grid[k] + grid[k + n1]) * np.float32(0.5))      # along z
            targets.append((grid[k] + grid[k + n1 + 1]) * np.float32(0.5))  # diagonal
    targets = np.array(targets, np.float32)
    rng = np.random.default_rng(2)
    o = targets + np.array([0.0, 1.0, 0.0], np.float32) + (rng.random((len(targets), 3)) - 0.5).astype(np.float32)
    d = targets - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rays_array(o, d)
    wt = oracle.bruteforce(sc, r, watertight=True)
    ref = oracle.bruteforce(sc, r)
    print("heightfield rays: %d, watertight misses %d, reference misses %d"
          % (len(r), (wt[:, 0] < 0).sum(), (ref[:, 0] < 0).sum()))
    assert (wt[:, 0] >= 0).all()


def test_watertight_special_cases(rt, oracle):
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    hs = rt.HostScene.build(pos, np.array([[0, 1, 2]]))
    sc = oracle.Scene(hs.arrays())
    o = np.array([[0.25, 0.25, 1.0], [0.25, 0.25, -1.0], [2, 2, 1], [0.25, 0.25, 1e-5], [0.5, 0.5, 1.0],
                  [0.0, 0.0, 1.0], [0.25, 0.25, 1.0]], np.float32)
    d = np.array([[0, 0, -1], [0, 0, 1], [0, 0, -1], [0, 0, -1], [0, 0, -1], [0, 0, -1], [1, 0, 0]], np.float32)
    h = oracle.bruteforce(sc, rays_array(o, d), watertight=True)
    t = h[:, 1].view(np.float32)
    by, bz = h[:, 2].view(np.float32), h[:, 3].view(np.float32)
    assert h[0, 0] == 0 and t[0] == 1.0 and by[0] == 0.25 and bz[0] == 0.25   # front
    assert h[1, 0] == 0 and t[1] == 1.0                                       # back face, no culling
    assert h[2, 0] == -1                                                      # outside
    assert h[3, 0] == -1                                                      # t <= 1e-4
    assert h[4, 0] == 0 and by[4] == 0.5 and bz[4] == 0.5                     # on the hypotenuse
    assert h[5, 0] == 0 and by[5] == 0.0 and bz[5] == 0.0                     # through a vertex
    assert h[6, 0] == -1                                                      # parallel
