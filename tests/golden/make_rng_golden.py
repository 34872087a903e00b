"""Writes tests/golden/rng_golden.json from the reference's own RNG.

The vectors come from oracle/_ref/librandom_ref.so, which oracle/Makefile
compiles from the reference's utilities/random.cpp (its only dependency is its
own header).  Run `make -C oracle ref` first; the fixture is committed so that
tests run where the reference tree is absent.
"""
import ctypes
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle  # noqa: E402


def main():
    ref = oracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/librandom_ref.so is missing: run `make -C oracle ref`")
    wang_in = list(range(64)) + [0x9E3779B9, 0xFFFFFFFF, 0x80000000, 123456789, 0xDEADBEEF]
    wang = [int(ref._Z8WangHashj(x)) for x in wang_in]
    # Xorshift32 chains: random.glsl RandomUInt == utilities/random.cpp Xorshift(uint32)
    xs32 = {}
    for s in (1, 0x9E3779B9, wang[0], 0xFFFFFFFF):
        seq, x = [], s
        for _ in range(32):
            x = int(ref._Z8Xorshiftj(x))
            seq.append(x)
        xs32[str(s)] = seq
    # xorshift64* and XorshiftF from the LightSampler seed 123 (lightsampler.cpp:8)
    xs64, x = [], 123
    for _ in range(32):
        x = int(ref._Z8Xorshiftm(x))
        xs64.append(str(x))
    xf = []
    state = ctypes.c_uint64(123)
    for _ in range(64):
        xf.append(float(ref._Z9XorshiftFRm(ctypes.byref(state))))
    out = {
        "source": "utilities/random.cpp of the reference, compiled by oracle/Makefile (target ref)",
        "wanghash": {"in": wang_in, "out": wang},
        "xorshift32": xs32,
        "xorshift64star_from_123": xs64,
        "xorshiftf_from_123": xf,
    }
    with open(os.path.join(HERE, "rng_golden.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
