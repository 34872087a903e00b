"""The device ray generators and traversals against the committed C2 fixtures:
the first 4,096 rays of each full 2^24-ray set equal tests/golden/c2_golden.npz
bit for bit, the bit-exact walk reproduces the fixture's hit records and
any-hit flags, the wide walk agrees except on FP64-classified edge/tie rays,
and the full sets hash to the digests in tests/golden/c2_sha256.json."""
import hashlib
import json
import os

import numpy as np
import pytest

from gpugi_amd import workloads

pytestmark = pytest.mark.gpu

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_full_sets_match_the_fixtures(rt, oracle, ctx):
    with np.load(os.path.join(GOLDEN_DIR, "c2_golden.npz"), allow_pickle=False) as z:
        golden = {k: z[k] for k in z.files}
    with open(os.path.join(GOLDEN_DIR, "c2_sha256.json")) as f:
        digests = json.load(f)
    hs = rt.HostScene.heightfield(workloads.HF_CELLS, workloads.MASTER_SEED)
    ctx.load_scene(hs)
    print("case: C2 full ray sets", flush=True)
    R = workloads.HeightfieldRays(ctx)
    try:
        assert R.n == digests["rays_per_set"]
        n = len(golden["coherent"])
        for name, buf in (("coherent", R.primary), ("incoherent", R.incoherent), ("shadow", R.shadow)):
            head = buf.download(np.float32, n * 8).reshape(n, 8)
            assert np.array_equal(head.view(np.uint32), golden[name].view(np.uint32)), name
            full = buf.download(np.uint8, R.n * 32)
            assert hashlib.sha256(full.tobytes()).hexdigest() == digests["sha256"][name], name
    finally:
        R.free()
    sc = oracle.Scene(hs.arrays())
    for name in ("coherent", "incoherent"):
        rays = np.ascontiguousarray(golden[name]).view(rt.RAY_DTYPE).reshape(-1)
        want = golden[name + "_hits"]
        got = ctx.trace(rays, rt.CLOSEST_HIT, rt.KERNEL_REFERENCE).view(np.int32).reshape(-1, 4)
        assert np.array_equal(got, want), name
        wide = ctx.trace(rays, rt.CLOSEST_HIT, rt.KERNEL_WIDE).view(np.int32).reshape(-1, 4)
        diff = np.nonzero((wide != want).any(1))[0]
        if len(diff):
            flags = oracle.classify(sc, golden[name][diff], want[diff, 0], wide[diff, 0])
            assert (flags != 0).all(), name
        assert len(diff) <= max(2, 1e-3 * len(want)), (name, len(diff))
    shadow = np.ascontiguousarray(golden["shadow"]).view(rt.RAY_DTYPE).reshape(-1)
    got = ctx.trace(shadow, rt.ANY_HIT, rt.KERNEL_REFERENCE).view(np.int32).reshape(-1, 4)[:, 0]
    assert np.array_equal(got, golden["shadow_any"])
