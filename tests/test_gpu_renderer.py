"""The C++ renderer layer end to end: gpugi_render scripts against the CPU oracle.

RendererSystem reproduces the reference's per-iteration sequence
(renderersystem.cpp:83-116, renderersystem.hpp:73-86): a renderer's
constructor sets its initial light-sample count, which draws light-sample
batch 1 with FrameSeed = WangHash(0); iteration k then draws with
FrameSeed = WangHash(k) and batch k+1.  The bidirectional path tracer first
runs 8 warm-up light passes (FrameSeed WangHash(0..7), batches 1..8), then
restarts the iteration count without a new batch, so its iteration k uses
WangHash(k) and batch k+9 (bidirectionalpathtracer.cpp:65-132).  The PFM holds
the accumulated RGB divided by the iteration count (hdrimage.cpp:6-33).
"""
import os
import subprocess

import numpy as np
import pytest

from gpugi_amd import workloads

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "gpugi_amd", "bin", "gpugi_render")
W = H = 32
L = 5


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline() == b"PF\n"
        w, h = map(int, f.readline().split())
        assert f.readline() == b"-1.000000\n"
        data = np.frombuffer(f.read(), dtype="<f4")
    assert data.size == w * h * 3
    return data.reshape(h, w, 3)


def render(tmp_path, renderer, iterations, traversal=1):
    out = str(tmp_path / ("r%d.pfm" % renderer))
    lines = ["sceneFilename = cornell", "cameraPos = 0, 1, 3.7", "cameraLookAt = 0, 1, 0", "cameraFOV = 40",
             "resolution = %d, %d" % (W, H), "pathLength = %d" % L, "traversal = %d" % traversal,
             "renderer = %d" % renderer, "waitIterations = %d" % iterations, "namedscreenshot = " + out, "exit"]
    args = [APP]
    for l in lines:
        args += ["-e", l]
    print("gpugi_render renderer=%d iterations=%d" % (renderer, iterations), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stdout + r.stderr
    return read_pfm(out)


def close(img, ref):
    return np.abs(img - ref) <= 1e-4 * np.abs(ref) + 1e-6


def test_pathtracer_script_is_bitexact(rt, oracle, tmp_path):
    img = render(tmp_path, 0, 3)
    hs = rt.HostScene.cornell()
    sc = oracle.Scene(hs.arrays())
    cam = workloads.cornell_camera(W)
    sampler = rt.LightSampler(hs, 123)
    acc = np.zeros((H, W, 4), np.float32)
    for k in range(3):
        oracle.pathtrace(sc, sampler.generate(128), cam, W, H, rt.wanghash(k), L, acc)
    ref = acc[..., :3] / np.float32(3)
    assert img.mean() > 0.01
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


def test_lighttracer_script(rt, oracle, tmp_path):
    img = render(tmp_path, 1, 2)
    hs = rt.HostScene.cornell()
    sc = oracle.Scene(hs.arrays())
    cam = workloads.cornell_camera(W)
    sampler = rt.LightSampler(hs, 123)
    acc = np.zeros((H, W, 4), np.float32)
    for k in range(2):
        oracle.lighttrace(sc, sampler.generate(256), cam, W, H, rt.wanghash(k), L, acc)
    ref = acc[..., :3] / np.float32(2)
    assert ref.mean() > 0
    ok = close(img, ref)
    assert ok.all(), np.argwhere(~ok)[:5]


def test_bidirectional_script_warmup_sequence(rt, oracle, tmp_path):
    img = render(tmp_path, 2, 2)
    hs = rt.HostScene.cornell()
    sc = oracle.Scene(hs.arrays())
    cam = workloads.cornell_camera(W)
    sampler = rt.LightSampler(hs, 123)
    batches = [sampler.generate(128) for _ in range(10)]
    entries = 0
    for k in range(8):  # warm-up pass k: FrameSeed WangHash(k), batch k+1
        e, _ = oracle.bpt_warmup(sc, batches[k], cam, W, H, rt.wanghash(k), L)
        entries += e & 0xffffffff  # read back as a uint32 counter
    mean = entries // 8
    cap = mean + mean // 10
    acc = np.zeros((H, W, 4), np.float32)
    for k in range(2):  # iteration k: WangHash(k), batch k+9
        acc, _, _ = oracle.bpt_draw(sc, batches[8 + k], cam, W, H, rt.wanghash(k), L, cap, acc)
    ref = acc[..., :3] / np.float32(2)
    assert ref.mean() > 0
    ok = close(img, ref)
    assert ok.all(), np.argwhere(~ok)[:5]
