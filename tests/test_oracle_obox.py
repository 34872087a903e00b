"""Oriented-box hierarchy (bvhType 1: TreeNode<OBox>, FetchIntersectOBoxNode,
intersectiontests.glsl:82-128) in the host fitter and the oracle walk.

- Every record keeps the node's firstChild / escape words, carries a unit
  quaternion, and its box holds every vertex of the node's subtree (checked in
  float64 with the rotation written here as a rotation matrix, not as the
  walk's quaternion formula).
- The oracle's walk over the oriented boxes equals brute force over all
  triangles except on rays the FP64 classifier marks as edge / tie cases.
- Geometry along a diagonal gets rotated boxes with less surface than the
  axis-aligned ones; axis-aligned geometry keeps boxes no larger than its
  axis-aligned boxes.
"""
import numpy as np
import pytest

from gpugi_amd import workloads
from test_oracle_geometry import classify_mismatch, make_rays


def subtree_vertices(a, n):
    H, T, P = a["hierarchy"], a["triangles"], a["positions"]
    fc = H[:, 3].view(np.uint32)
    esc = H[:, 7].view(np.uint32)
    out, stack = [], [n]
    while stack:
        m = stack.pop()
        c = int(fc[m])
        if c & 0x80000000:
            leaf = c & 0x7FFFFFFF
            for s in range(2):
                t = T[2 * leaf + s]
                if t[0] == t[1]:
                    break
                out.extend(P[t[:3]])
        else:
            stack.append(c)
            if esc[c] != 0 and esc[c] != esc[m]:
                stack.append(int(esc[c]))
    return np.array(out, np.float64)


def rotation_matrix(q):
    x, y, z, w = (float(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def sticks_scene(rt):
    """Thin boxes along random diagonal directions."""
    rng = np.random.default_rng(21)
    pos, idx = [], []
    cube = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], np.float64)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    for k in range(24):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        u = np.cross(d, [0.3, 0.5, 0.81])
        u /= np.linalg.norm(u)
        v = np.cross(d, u)
        c = rng.uniform(-3, 3, 3)
        corners = c + (cube[:, :1] - 0.5) * 2.0 * d + (cube[:, 1:2] - 0.5) * 0.1 * u + (cube[:, 2:] - 0.5) * 0.1 * v
        b = len(pos)
        pos.extend(corners)
        for q in quads:
            idx += [[b + q[0], b + q[1], b + q[2]], [b + q[0], b + q[2], b + q[3]]]
    return rt.HostScene.build(np.array(pos, np.float32), np.array(idx))


SCENES = {
    "cornell": lambda rt: rt.HostScene.cornell(),
    "hf16": lambda rt: rt.HostScene.heightfield(16, workloads.MASTER_SEED),
    "sticks": sticks_scene,
}


@pytest.mark.parametrize("name", list(SCENES))
def test_records_enclose_their_subtrees(rt, name):
    hs = SCENES[name](rt)
    a = hs.arrays()
    ob = hs.obox_hierarchy()
    H = a["hierarchy"]
    assert np.array_equal(ob[:, 3].view(np.uint32), H[:, 3].view(np.uint32))
    assert np.array_equal(ob[:, 7].view(np.uint32), H[:, 7].view(np.uint32))
    q = ob[:, 8:12].astype(np.float64)
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-6)
    for n in range(len(H)):
        v = subtree_vertices(a, n)
        if len(v) == 0:
            continue
        R = rotation_matrix(q[n])
        local = (v - ob[n, 0:3].astype(np.float64)) @ R.T
        half = ob[n, 4:7].astype(np.float64)
        assert (np.abs(local) <= half * (1 + 1e-6)).all(), n


@pytest.mark.parametrize("name", list(SCENES))
def test_obox_walk_equals_bruteforce(rt, oracle, name):
    hs = SCENES[name](rt)
    a = hs.arrays()
    lo, hi = a["positions"].min(0), a["positions"].max(0)
    a_ob = dict(a, obox=hs.obox_hierarchy())
    sc, sco = oracle.Scene(a), oracle.Scene(a_ob)
    rng = np.random.default_rng(12)
    o = rng.uniform(lo, hi, (6000, 3))
    # half the rays aim at a vertex (with a small offset), half go anywhere
    target = a["positions"][rng.integers(0, len(a["positions"]), 6000)] + rng.normal(scale=0.01, size=(6000, 3))
    d = np.where(np.arange(6000)[:, None] % 2 == 0, target - o, rng.normal(size=(6000, 3)))
    rays = make_rays(o, d)
    got = oracle.trace(sco, rays)
    ref = oracle.bruteforce(sc, rays)
    idx, flags = classify_mismatch(oracle, sc, rays, got, ref)
    assert (flags != 0).all(), [(i, got[i], ref[i]) for i, f in zip(idx, flags) if f == 0][:5]
    assert len(idx) <= 3
    assert (got[:, 0] >= 0).mean() > 0.2
    anyh = oracle.trace(sco, rays, any_hit=True)
    assert ((anyh[:, 0] == 1) == (got[:, 0] >= 0)).mean() > 0.999


def box_area(half):
    h = half.astype(np.float64)
    return (h[:, 0] * h[:, 1] + h[:, 1] * h[:, 2] + h[:, 2] * h[:, 0]).sum()


def test_rotated_boxes_are_tighter_on_diagonal_geometry(rt):
    hs = sticks_scene(rt)
    a = hs.arrays()
    ob = hs.obox_hierarchy()
    aabb_half = (a["hierarchy"][:, 4:7].astype(np.float64) - a["hierarchy"][:, 0:3]) / 2
    assert box_area(ob[:, 4:7]) < 0.8 * box_area(aabb_half)
    assert (np.abs(ob[:, 11]) < 0.999).mean() > 0.5  # most nodes use a rotated frame
    cornell = rt.HostScene.cornell()
    ca = cornell.arrays()
    cob = cornell.obox_hierarchy()
    c_half = (ca["hierarchy"][:, 4:7].astype(np.float64) - ca["hierarchy"][:, 0:3]) / 2
    assert box_area(cob[:, 4:7]) <= box_area(c_half) * (1 + 1e-4) + 1e-6
