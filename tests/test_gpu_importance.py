"""The importance-break walk on the GPU (gpugi_trace_importance_break,
TRACERAY_IMPORTANCE_BREAK of traceray.glsl:48-112) against the oracle bit
for bit: random per-node importance and thresholds, closest hit with the node
outputs and any hit, over axis-aligned and oriented boxes."""
import numpy as np
import pytest

from gpugi_amd import workloads
from test_oracle_geometry import make_rays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ictx(rt):
    c = rt.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("bvh", ["aabox", "obox"])
def test_importance_break_matches_oracle(rt, oracle, ictx, bvh):
    hs = rt.HostScene.heightfield(32, workloads.MASTER_SEED)
    a = hs.arrays()
    if bvh == "obox":
        a = dict(a, obox=hs.obox_hierarchy())
    sc = oracle.Scene(a)
    ictx.set_bvh_type(rt.BVH_OBOX if bvh == "obox" else rt.BVH_AABOX)
    ictx.load_scene(hs)
    rng = np.random.default_rng(41)
    lo, hi = a["positions"].min(0), a["positions"].max(0)
    o = rng.uniform(lo - 0.3, hi + 0.3, (8192, 3))
    rays = make_rays(o, rng.uniform(lo, hi, (8192, 3)) - o)
    rays[::3, 3] = rng.uniform(0.05, 1.0, len(rays[::3])).astype(np.float32)  # some short rays
    nn = len(a["hierarchy"])
    imp = np.stack([rng.uniform(-0.2, 1.0, nn), rng.uniform(0.0, 2.0, nn)], 1).astype(np.float32)
    r = np.ascontiguousarray(rays).view(rt.RAY_DTYPE).reshape(-1)
    for threshold, ray_threshold in ((0.0, 1.5), (0.3, 0.0), (-1.0, 4.0)):
        print("case: importance break %s %g %g" % (bvh, threshold, ray_threshold), flush=True)
        gh, gn = ictx.trace_importance_break(r, imp, threshold, ray_threshold)
        wh, wn = oracle.trace_importance(sc, rays, imp, threshold, ray_threshold)
        assert np.array_equal(np.ascontiguousarray(gh).view(np.int32).reshape(-1, 4), wh)
        assert np.array_equal(np.ascontiguousarray(gn).view(np.int32).reshape(-1, 4)[:, :3], wn[:, :3])
        broke = (wh[:, 0] == -1) & (wn[:, 0] >= 0)
        assert 0.01 < broke.mean() < 0.99  # the break happens for some rays, not all
        ga = ictx.trace_importance_break(r, imp, threshold, ray_threshold, rt.ANY_HIT)
        wa = oracle.trace_importance(sc, rays, imp, threshold, ray_threshold, any_hit=True)
        assert np.array_equal(np.ascontiguousarray(ga).view(np.int32).reshape(-1, 4)[:, 0], wa[:, 0])
