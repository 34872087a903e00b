"""Writes tests/golden/c2_sha256.json: SHA-256 digests of the full C2 ray
sets (2^24 rays of 32 bytes each: coherent, incoherent, shadow) as the
library's device ray generators write them for the 1024 x 1024 heightfield
and the master seed.  Needs the GPU: run it on an MI355X after a change that
is meant to alter the ray sets, and commit the result together with that
change.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from gpugi_amd import rt, workloads  # noqa: E402


def digests():
    hs = rt.HostScene.heightfield(workloads.HF_CELLS, workloads.MASTER_SEED)
    ctx = rt.Context(0)
    try:
        ctx.load_scene(hs)
        R = workloads.HeightfieldRays(ctx)
        out = {}
        for name, buf in (("coherent", R.primary), ("incoherent", R.incoherent), ("shadow", R.shadow)):
            out[name] = hashlib.sha256(buf.download(np.uint8, R.n * 32).tobytes()).hexdigest()
        R.free()
        return {"rays_per_set": R.n, "bytes_per_ray": 32, "sha256": out}
    finally:
        ctx.close()


def main():
    out = os.path.join(HERE, "c2_sha256.json")
    d = digests()
    with open(out, "w") as f:
        json.dump(d, f, indent=1)
    print(json.dumps(d))


if __name__ == "__main__":
    main()
