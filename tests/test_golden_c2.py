"""The C2 heightfield ray sets pinned by the committed golden fixture
(tests/golden/c2_golden.npz, written by make_c2_golden.py): the oracle's ray
generators and escape walk reproduce it bit for bit, and the rays satisfy the
set definitions independently of the generator code (camera origin and unit
directions, cosine-hemisphere directions facing away from the hit surface,
shadow rays aimed at the seeded lights with tmax = distance - 2e-4)."""
import os

import numpy as np
import pytest

from gpugi_amd import workloads

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c2_golden.npz")


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def hf1024(rt):
    return rt.HostScene.heightfield(workloads.HF_CELLS, workloads.MASTER_SEED)


def test_oracle_reproduces_the_fixture(oracle, hf1024, golden):
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_c2_golden", os.path.join(os.path.dirname(GOLDEN), "make_c2_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    fresh = mod.generate()
    for k, v in golden.items():
        assert np.array_equal(np.asarray(fresh[k]).view(np.uint32), v.view(np.uint32)), k


def test_ray_sets_follow_their_definitions(hf1024, golden):
    cam = workloads.heightfield_camera()
    co = golden["coherent"]
    assert np.array_equal(co[:, 0:3], np.tile(np.float32(list(cam.position)), (len(co), 1)))
    assert np.allclose(np.linalg.norm(co[:, 4:7], axis=1), 1.0, atol=1e-5)
    hits = golden["coherent_hits"]
    assert (hits[:, 0] >= 0).mean() > 0.5
    # incoherent: unit cosine-hemisphere directions about the hit triangle's normal,
    # which faces the incoming primary ray
    a = hf1024.arrays()
    pos, tri = a["positions"].reshape(-1, 3), a["triangles"]
    hit_idx = np.nonzero(hits[:, 0] >= 0)[0]
    inc = golden["incoherent"]
    k = hit_idx[: len(inc)]
    t = tri[hits[k, 0]]
    p0, p1, p2 = pos[t[:, 0]], pos[t[:, 1]], pos[t[:, 2]]
    n = np.cross(p0 - p2, p1 - p0).astype(np.float64)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    din = co[k, 4:7].astype(np.float64)
    n[np.einsum("ij,ij->i", n, din) > 0] *= -1
    d = inc[:, 4:7].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    assert (np.einsum("ij,ij->i", d, n) > -1e-5).all()
    hitp = co[k, 0:3] + co[k, 4:7] * hits[k, 1].view(np.float32)[:, None]
    assert np.allclose(inc[:, 0:3], hitp + 1e-4 * inc[:, 4:7], atol=1e-5)
    # shadow ray i: from the same hit point toward light i mod 4
    # (lightsource.glsl:18-21,39-40): origin p + 1e-4 d, tmax = dist - 2e-4 with
    # dist = sqrt(|L - p|^2 + 1e-5)
    lights = np.asarray(workloads.heightfield_lights(), np.float64).reshape(-1, 3)
    sh = golden["shadow"].astype(np.float64)
    L = lights[np.arange(len(sh)) % len(lights)]
    to = L - hitp.astype(np.float64)
    dist = np.sqrt(np.einsum("ij,ij->i", to, to) + 1e-5)
    dirs = to / dist[:, None]
    assert np.allclose(sh[:, 4:7], dirs, atol=1e-5)
    assert np.allclose(sh[:, 0:3], hitp + 1e-4 * dirs, atol=1e-5)
    assert np.allclose(sh[:, 3], dist - 2e-4, atol=1e-5)
    occluded = golden["shadow_any"]
    assert set(np.unique(occluded)) <= {0, 1} and 0.01 < occluded.mean() < 0.99
