"""gpugi_compute_mse (TextureMSE::ComputeCurrentMSE, texturemse.cpp:39-66)
against the oracle on a Cornell render, and the gpugi_render MSE commands
(application.cpp:173-188, 346-376)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "gpugi_amd", "bin", "gpugi_render")
W, H = 48, 40


@pytest.fixture(scope="module")
def rendered(rt, ctx):
    hs = rt.HostScene.cornell()
    ctx.load_scene(hs)
    ctx.set_screen(W, H)
    ctx.set_camera(rt.camera((0.0, 1.0, 3.7), (0.0, 1.0, 0.0), (0.0, 1.0, 0.0), 40.0, W, H))
    ctx.set_max_path_length(5)
    ctx.set_integrator_kernel(rt.KERNEL_WIDE)
    sampler = rt.LightSampler(hs, 7)
    for k in range(4):
        ctx.set_light_samples(sampler.generate(128))
        ctx.draw(rt.PT, rt.wanghash(k))
    img, n = ctx.readback(W, H)
    assert n == 4 and img[..., :3].mean() > 0.01
    return img


def test_mse_matches_oracle(ctx, oracle, rendered):
    rng = np.random.default_rng(5)
    ref = (rendered[..., :3] / np.float32(4) + rng.normal(0, 0.05, (H, W, 3))).astype(np.float32)
    print("mse: random reference %dx%d" % (W, H), flush=True)
    ctx.set_reference_image(ref)
    got = ctx.compute_mse(4)
    want = oracle.image_mse(rendered, 4, ref)
    assert abs(got - want) <= 1e-6 * want, (got, want)


def test_mse_zero_for_the_image_itself_and_resampled_references(ctx, rendered):
    img = rendered[..., :3] / np.float32(4)
    ctx.set_reference_image(img)
    assert ctx.compute_mse(4) == 0.0
    ctx.set_reference_image(np.repeat(np.repeat(img, 2, axis=0), 2, axis=1))
    assert ctx.compute_mse(4) == 0.0
    ctx.set_reference_image(np.ascontiguousarray(img[1::2, 1::2]))
    assert ctx.compute_mse(4) == 0.0
    ctx.set_reference_image(np.ascontiguousarray(img[0::2, 0::2]))
    assert ctx.compute_mse(4) > 0.0


def test_mse_divisor_zero_and_no_reference(rt):
    c = rt.Context(0)
    try:
        c.set_screen(8, 8)
        assert c.compute_mse(1) == float("inf")  # no reference image
        c.set_reference_image(np.ones((8, 8, 3), np.float32))
        assert c.compute_mse(1) == 3.0  # cleared accumulation: (0 - 1)^2 * 3
        v = c.compute_mse(0)  # 0 / 0 in every channel
        assert v != v
    finally:
        c.close()


def _script(*lines, timeout=100):
    args = [APP]
    for l in lines:
        args += ["-e", l]
    print("gpugi_render:", lines[-3:], flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


SETUP = ("sceneFilename = cornell", "cameraPos = 0, 1, 3.7", "cameraLookAt = 0, 1, 0", "cameraFOV = 40",
         "resolution = 32, 32", "pathLength = 5", "renderer = 0")


def _read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline() == b"PF\n"
        w, h = map(int, f.readline().split())
        assert f.readline() == b"-1.000000\n"
        return np.frombuffer(f.read(), dtype="<f4").reshape(h, w, 3)


def _mse_lines(out):
    return {int(l.split()[1]): float(l.split()[3]) for l in out.splitlines() if l.startswith("Iteration ")}


def test_script_mse_checks_and_wait(tmp_path):
    ref = str(tmp_path / "ref.pfm")
    _script(*SETUP, "waitIterations = 8", "namedscreenshot = " + ref, "exit")
    four = str(tmp_path / "four.pfm")
    _script(*SETUP, "waitIterations = 4", "namedscreenshot = " + four, "exit")
    r = _script(*SETUP, "referenceImage = " + ref, "mseCheckInterval = 4", "waitIterations = 8", "exit")
    mse = _mse_lines(r.stdout)
    assert sorted(mse) == [4, 8]
    assert mse[8] == 0.0  # the same seeds reproduce the reference image exactly
    a, b = _read_pfm(four), _read_pfm(ref)
    d = a - b
    want = float(((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]).astype(np.float64).mean())
    assert abs(mse[4] - want) <= 1e-5 * want
    # waitMSE resumes at the first check below the goal
    r = _script(*SETUP, "referenceImage = " + ref, "mseCheckInterval = 2", "waitMSE = 1e9", "exit")
    assert sorted(_mse_lines(r.stdout)) == [2] and "Resuming script" in r.stdout
    # without a reference image the wait is refused and the script goes on
    r = _script(*SETUP, "mseCheckInterval = 2", "waitMSE = 1", "exit")
    assert "no reference image loaded" in r.stderr
