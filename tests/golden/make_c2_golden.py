"""Writes tests/golden/c2_golden.npz: the first 4,096 rays of the three C2
heightfield ray sets and their expected results, from the CPU oracle.

Scene: 1024 x 1024 cells (2,097,152 triangles), master seed 0x9E3779B9.
Coherent set: primary rays 0..4095 (pixel i mod 1920*1080, jitter from the
master seed).  Incoherent set: rays 0..4095 from the primary hits in order
(misses skipped; the first 8,192 primaries hold more than 4,096 hits, so this
is a prefix of the full set).  Shadow set: rays 0..4095 toward the four
seeded point lights.  Expected results: closest-hit records {slot, t, b.y,
b.z} of the oracle's escape walk for the coherent and incoherent rays,
any-hit flags for the shadow rays.  CPU only; takes well under a minute.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from gpugi_amd import rt, workloads  # noqa: E402
from oracle import oracle  # noqa: E402

N = 4096
N_PRIMARY = 8192


def generate():
    hs = rt.HostScene.heightfield(workloads.HF_CELLS, workloads.MASTER_SEED)
    sc = oracle.Scene(hs.arrays())
    cam = workloads.heightfield_camera()
    prim = oracle.raygen_primary(cam, workloads.HF_WIDTH, workloads.HF_HEIGHT, workloads.MASTER_SEED, N_PRIMARY)
    ph = oracle.trace(sc, prim)
    if (ph[:, 0] >= 0).sum() < N:
        raise SystemExit("fewer primary hits than incoherent rays")
    inc = oracle.raygen_diffuse(sc, prim, ph, (workloads.MASTER_SEED + 1) & 0xffffffff, N)
    shadow = oracle.raygen_shadow(prim, ph, workloads.heightfield_lights(), N)
    return {
        "coherent": prim[:N], "coherent_hits": ph[:N],
        "incoherent": inc, "incoherent_hits": oracle.trace(sc, inc),
        "shadow": shadow, "shadow_any": oracle.trace(sc, shadow, any_hit=True)[:, 0].astype(np.int32),
    }


def main():
    out = os.path.join(HERE, "c2_golden.npz")
    np.savez_compressed(out, **generate())
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
