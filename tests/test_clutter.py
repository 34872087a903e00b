"""The clutter room (integrator benchmark scene) on the host: counts, closed
icospheres, materials within the 256-entry table, light tables."""
import numpy as np


def test_clutter_counts_and_materials(rt):
    hs = rt.HostScene.clutter(40, 1)
    info = hs.info()
    assert info["num_triangles"] == 12 + 4 + 40 * 1280
    a = hs.arrays()
    assert len(a["materials"]) == 242 <= 256
    assert info["bbox_min"] == [-10.0, 0.0, -10.0] and info["bbox_max"] == [10.0, 8.0, 10.0]
    # every sphere triangle uses a Lambertian (2..121) or glossy (122..241) bin
    tris = a["triangles"].reshape(-1, 4)
    valid = tris[:, 0] != tris[:, 1]
    mats = tris[valid, 3]
    assert set(np.unique(mats)) <= set(range(242))
    assert (mats == 1).sum() == 4  # the area-light fan
    m = a["materials"]
    glossy = m[122:242]
    exps = glossy["reflectiveness"][:, 3]
    assert (exps > 9.0).all() and (exps < 1100.0).all()
    assert (m[2:122]["reflectiveness"][:, 3] < 2.0).all()  # roughness 1: exponent ~1


def test_clutter_spheres_are_closed(rt):
    hs = rt.HostScene.clutter(3, 7)
    a = hs.arrays()
    tris = a["triangles"].reshape(-1, 4)
    tris = tris[tris[:, 0] != tris[:, 1]]
    sph = tris[tris[:, 3] >= 2][:, :3]
    assert len(sph) == 3 * 1280
    edges = {}
    for t in sph:
        for i in range(3):
            e = tuple(sorted((int(t[i]), int(t[(i + 1) % 3]))))
            edges[e] = edges.get(e, 0) + 1
    assert set(edges.values()) == {2}


def test_clutter_is_deterministic(rt):
    a = rt.HostScene.clutter(5, 3).arrays()
    b = rt.HostScene.clutter(5, 3).arrays()
    c = rt.HostScene.clutter(5, 4).arrays()
    assert np.array_equal(a["positions"], b["positions"])
    assert not np.array_equal(a["positions"], c["positions"])
