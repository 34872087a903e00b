"""Host-side logic of the product: C-ABI exports, SAH sweep builder
invariants, .rawscene I/O, light sampler and camera (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from gpugi_amd import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "gpugi_rt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gpugi_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function(rt):
    lib = rt.lib()
    declared = header_functions()
    assert len(declared) >= 35
    for name in declared:
        assert hasattr(lib, name), name
    # and the binding declares a signature for each of them
    assert set(declared) == set(rt.exported_symbols())


def tree(a):
    H = a["hierarchy"]
    fc = H[:, 3].view(np.uint32)
    esc = H[:, 7].view(np.uint32)
    return H, fc, esc


def check_bvh_invariants(a):
    H, fc, esc = tree(a)
    tri = a["triangles"]
    pos = a["positions"]
    nn = len(H)
    assert esc[0] == 0
    seen = np.zeros(len(tri), int)
    for i in range(nn):
        lo, hi = H[i, 0:3], H[i, 4:7]
        assert (lo <= hi).all()
        if fc[i] & 0x80000000:
            leaf = fc[i] & 0x7FFFFFFF
            slots = tri[2 * leaf: 2 * leaf + 2]
            valid = slots[:, 0] != slots[:, 1]
            assert valid[0], "a leaf holds at least one triangle"
            assert not (valid[1] and not valid[0])
            for s in np.nonzero(valid)[0]:
                seen[2 * leaf + s] += 1
                p = pos[slots[s, :3]]
                assert (p >= lo).all() and (p <= hi).all(), "leaf box contains its triangles"
        else:
            l = fc[i]
            assert l == i + 1, "preorder: left child follows its parent"
            r = esc[l]
            assert r > l and esc[r] == esc[i], "left escapes to right, right to the parent's escape"
            for c in (l, r):
                assert (H[c, 0:3] >= lo).all() and (H[c, 4:7] <= hi).all()
                # the node box is the tight union of its children
            assert np.array_equal(np.minimum(H[l, 0:3], H[r, 0:3]), lo)
            assert np.array_equal(np.maximum(H[l, 4:7], H[r, 4:7]), hi)
    valid_slots = tri[:, 0] != tri[:, 1]
    assert (seen[valid_slots] == 1).all() and (seen[~valid_slots] == 0).all()
    return int(valid_slots.sum())


def test_builder_invariants_cornell_and_heightfield(rt):
    s = rt.HostScene.cornell()
    assert check_bvh_invariants(s.arrays()) == 36
    h = rt.HostScene.heightfield(32, workloads.MASTER_SEED)
    assert check_bvh_invariants(h.arrays()) == 2 * 32 * 32


def test_builder_random_soup_and_degenerate_input(rt):
    rng = np.random.default_rng(4)
    p = rng.normal(size=(300, 3)).astype(np.float32)
    idx = rng.integers(0, 300, size=(200, 3)).astype(np.uint32)
    keep = (idx[:, 0] != idx[:, 1]) & (idx[:, 1] != idx[:, 2]) & (idx[:, 0] != idx[:, 2])
    s = rt.HostScene.build(p, idx)
    assert check_bvh_invariants(s.arrays()) == int(keep.sum())  # repeated-index triangles are dropped
    one = rt.HostScene.build(p[:3], np.array([[0, 1, 2]], np.uint32))
    a = one.arrays()
    assert len(a["hierarchy"]) == 1 and a["hierarchy"][0, 3].view(np.uint32) == 0x80000000
    with pytest.raises(rt.GpugiError):
        rt.HostScene.build(p[:3], np.array([[0, 1, 7]], np.uint32))  # vertex out of range


def sah_costs(boxes):
    """Cost of every split of a sorted run, SA(side)/SA(node) * count, in float32."""
    def sa(lo, hi):
        e = (hi - lo).astype(np.float32)
        return np.float32(2.0) * (e[..., 0] * e[..., 1] + e[..., 1] * e[..., 2] + e[..., 2] * e[..., 0])
    lo, hi = boxes[:, 0:3], boxes[:, 3:6]
    n = len(boxes)
    plo, phi = np.minimum.accumulate(lo), np.maximum.accumulate(hi)
    slo, shi = np.minimum.accumulate(lo[::-1])[::-1], np.maximum.accumulate(hi[::-1])[::-1]
    parent = sa(plo[-1], phi[-1])
    k = np.arange(n - 1)
    cl = sa(plo[:-1], phi[:-1]) / parent * (k + 1).astype(np.float32)
    cr = sa(slo[1:], shi[1:]) / parent * (n - 1 - k).astype(np.float32)
    return cl + cr


def test_sah_split_is_cheapest_sweep_split(rt):
    """At every inner node the left child holds the first k+1 triangles of the
    node's sorted run, with k the first minimum of the sweep cost over k >= 1
    (bvhmake/buildmethods/sweep.cpp:217-251)."""
    h = rt.HostScene.heightfield(16, workloads.MASTER_SEED)
    a = h.arrays()
    H, fc, esc = tree(a)
    tri, pos = a["triangles"], a["positions"]
    # run of each node in leaf order = the sorted order of the builder
    order = []
    runs = {}

    def collect(i):
        start = len(order)
        if fc[i] & 0x80000000:
            leaf = fc[i] & 0x7FFFFFFF
            for s in range(2):
                t = tri[2 * leaf + s]
                if t[0] != t[1]:
                    p = pos[t[:3]]
                    order.append(np.concatenate([p.min(0), p.max(0)]))
        else:
            collect(fc[i])
            collect(esc[fc[i]])
        runs[i] = (start, len(order))

    collect(0)
    boxes = np.array(order, np.float32)
    checked = 0
    for i, (b, e) in runs.items():
        if fc[i] & 0x80000000 or e - b < 3:
            continue
        costs = sah_costs(boxes[b:e])
        k_left = runs[fc[i]][1] - b - 1
        cand = costs[1:]
        best = 1 + int(np.argmin(cand)) if np.isfinite(cand).any() else 0
        assert k_left == best, (i, k_left, best)
        checked += 1
    assert checked > 100


def test_rawscene_roundtrip_and_json(rt, tmp_path):
    s = rt.HostScene.cornell()
    s.add_point_light([0.1, 1.5, 0.2], [3.0, 2.0, 1.0])
    path, js = str(tmp_path / "c.rawscene"), str(tmp_path / "c.json")
    s.save(path, js)
    t = rt.HostScene.load(path, js)
    a, b = s.arrays(), t.arrays()
    for k in ("hierarchy", "triangles", "positions", "vertex_info"):
        assert np.array_equal(a[k], b[k]), k
    assert a["materials"].tobytes() == b["materials"].tobytes()
    # chunk layout: 40-byte NamedArray headers in bvhmake's export order
    raw = open(path, "rb").read()
    names = []
    off = 0
    while off < len(raw):
        name = raw[off:off + 32].split(b"\0")[0].decode()
        ne, es = np.frombuffer(raw[off + 32:off + 40], np.uint32)
        names.append((name, int(ne), int(es)))
        off += 40 + int(ne) * int(es)
    assert [n for n, _, _ in names] == ["materialref", "vertices", "bounding_aabox", "hierarchy", "triangles"]
    assert dict((n, e) for n, _, e in names) == {"materialref": 32, "vertices": 32, "bounding_aabox": 24,
                                                "hierarchy": 12, "triangles": 32}


def test_rawscene_load_errors(rt, tmp_path):
    s = rt.HostScene.cornell()
    path = str(tmp_path / "c.rawscene")
    s.save(path)
    raw = open(path, "rb").read()
    with pytest.raises(rt.GpugiError) as e:
        rt.HostScene.load(str(tmp_path / "missing.rawscene"))
    assert e.value.code == -4  # GPUGI_ERR_IO
    trunc = str(tmp_path / "t.rawscene")
    open(trunc, "wb").write(raw[: len(raw) - 7])
    with pytest.raises(rt.GpugiError) as e:
        rt.HostScene.load(trunc)
    assert e.value.code == -5  # GPUGI_ERR_FORMAT
    # a corrupt escape pointer is rejected by the sanity check
    bad = bytearray(raw)
    off = 0
    while off < len(bad):
        name = bytes(bad[off:off + 32]).split(b"\0")[0]
        ne, es = np.frombuffer(bytes(bad[off + 32:off + 40]), np.uint32)
        if name == b"hierarchy":
            bad[off + 40 + 8: off + 40 + 12] = np.uint32(10 ** 6).tobytes()  # node 0 escape
        off += 40 + int(ne) * int(es)
    badp = str(tmp_path / "b.rawscene")
    open(badp, "wb").write(bytes(bad))
    with pytest.raises(rt.GpugiError):
        rt.HostScene.load(badp)
    # an unknown chunk is skipped
    extra = str(tmp_path / "x.rawscene")
    hdr = b"approx_sggx".ljust(32, b"\0") + np.array([3, 12], np.uint32).tobytes() + bytes(36)
    open(extra, "wb").write(raw + hdr)
    assert rt.HostScene.load(extra).info()["num_triangles"] == 36


def test_material_resolution(rt):
    a = rt.HostScene.cornell().arrays()["materials"]
    # UNORM8 quantization of RGB8 constant textures (scene.cpp:164,213): 0.75 -> 191/255
    assert a["diffuse"][0, 0] == np.float32(191.0 / 255.0)
    assert np.allclose(a["diffuse"][1], np.float32([161, 15, 10]) / np.float32(255))
    assert (a["reflectiveness"][:, 3] == np.float32(1.0) / np.float32(1.0 + 1e-20)).all()
    assert (a["emissivity"][3] == np.float32([17, 12, 4])).all()
    assert (a["fresnel1"] == 1.0 - a["fresnel0"]).all()


def test_camera_closed_form(rt):
    cam = rt.camera([0.0, 1.0, 3.7], [0.0, 1.0, 0.0], [0.0, 1.0, 0.0], 40.0, 320, 200)
    U, V, W = (np.array(getattr(cam, k), np.float64) for k in ("u", "v", "w"))
    f = np.tan(np.radians(20.0))
    assert np.allclose(W, [0, 0, -1]) and abs(U @ W) < 1e-7 and abs(V @ W) < 1e-7
    assert np.isclose(np.linalg.norm(V), f, rtol=1e-6) and np.isclose(np.linalg.norm(U), f * 1.6, rtol=1e-6)
    assert np.isclose(cam.pixel_area, np.linalg.norm(U) * np.linalg.norm(V) * 4 / (320 * 200), rtol=1e-6)


def test_light_sampler_area_point_and_dummy(rt, oracle):
    s = rt.HostScene.cornell()
    samples = rt.LightSampler(s, 123).generate(128)
    # all on the ceiling quad, moved 1e-4 along its downward normal
    assert np.allclose(samples["position"][:, 1], np.float32(1.99) - np.float32(1e-4), atol=1e-6)
    assert (np.abs(samples["position"][:, [0, 2]]) <= 0.3 + 1e-6).all()
    assert np.allclose(samples["intensity"], np.float32([17, 12, 4]) * np.float32(0.36) / 128, rtol=1e-5)
    assert np.allclose(samples["normal_phi"], -np.pi / 2, atol=1e-6)
    # the xorshift64* stream of lightsampler.cpp: 3 draws per area sample
    st = ctypes.c_uint64(123)
    draws = [oracle.lib().or_xorshiftf(ctypes.byref(st)) for _ in range(3 * 128)]
    xi0 = np.sqrt(np.float32(draws[1::3]))
    xi1 = np.float32(draws[2::3])
    # sample k lies in triangle k of the area CDF; check barycentric recovery on the first triangle
    first_tri = np.float32(draws[0::3]) <= 0.5
    assert first_tri.any() and (~first_tri).any()
    v0 = np.float32([-0.3, 1.99, -0.3])
    v1 = np.float32([0.3, 1.99, -0.3])
    v2 = np.float32([0.3, 1.99, 0.3])
    pos = v0 * (1 - xi0)[:, None] + v1 * (xi0 * (1 - xi1))[:, None] + v2 * (xi0 * xi1)[:, None]
    assert np.allclose(samples["position"][first_tri][:, [0, 2]], pos[first_tri][:, [0, 2]], atol=1e-6)
    # point lights take their flux share of the samples (lightsampler.cpp:27-29)
    s.add_point_light([0.0, 1.0, 0.0], [50.0, 50.0, 50.0])
    mixed = rt.LightSampler(s, 123).generate(128)
    npoint = int((mixed["normal_theta_cos"] == 2.0).sum())
    area_flux = 0.36 * np.pi * (0.2126 * 17 + 0.7152 * 12 + 0.0722 * 4)
    point_flux = np.log(1 + 50 * 4 * np.pi)
    assert npoint == int(np.float32(point_flux / (point_flux + area_flux)) * 128)
    pts = mixed[mixed["normal_theta_cos"] == 2.0]
    assert np.allclose(pts["intensity"], 50.0 / npoint)
    # a scene without any light yields black dummy samples (lightsampler.cpp:21-31)
    h = rt.HostScene.heightfield(4, 1)
    dummy = rt.LightSampler(h).generate(16)
    assert (dummy.view(np.float32) == 0).all()


def test_errors_without_context(rt):
    lib = rt.lib()
    assert lib.gpugi_hostscene_cornell(None) == -1
    assert lib.gpugi_camera_compute(None, None, None, 40.0, 1, 1, None) == -1
    assert b"invalid" in lib.gpugi_last_error(None)


def test_next_row_entry_points_reject_null_arguments(rt):
    """The texture, environment, MSE, bvh-type, importance-break, pinned-memory and
    PT-schedule entry points fail with GPUGI_ERR_ARG (-1) on a null context, before any device call."""
    lib = rt.lib()
    null = None
    assert lib.gpugi_add_texture(null, 1, 1, 0, null, null) == -1
    assert lib.gpugi_set_material_texture(null, 0, 0, 0) == -1
    assert lib.gpugi_set_environment_map(null, 0, null) == -1
    assert lib.gpugi_set_reference_image(null, null, 1, 1) == -1
    assert lib.gpugi_compute_mse(null, 1, null) == -1
    assert lib.gpugi_set_bvh_type(null, 1) == -1
    assert lib.gpugi_trace_importance_break(null, 0, null, null, null, 0, null, 0, 0.0, 0.0) == -1
    assert lib.gpugi_host_alloc(null, 16, null) == -1
    assert lib.gpugi_host_free(null, null) == -1
    assert lib.gpugi_hostscene_obox_hierarchy(null, null, 0) == -1
    assert lib.gpugi_hostscene_num_textures(null, null) == -1
    assert lib.gpugi_set_pt_schedule(null, 1) == -1
    hs = rt.HostScene.cornell()
    assert hs.textures() == [] and (hs.material_textures() == -1).all()
    ob = hs.obox_hierarchy()
    short = np.zeros(ob.size - 1, np.float32)
    assert lib.gpugi_hostscene_obox_hierarchy(hs._h, short.ctypes.data, short.size) == -1
