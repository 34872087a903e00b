"""Path-length isolation and the late Russian-roulette start in the oracle.

SHOW_SPECIFIC_PATHLENGTH (renderer.hpp:6-9; pathtracer.comp:74-84,
lighttracer.comp:133-157, pathtracer_bidir.comp:75-104) keeps only the paths
of one length k (bounces; 1 = direct lighting).  Pins:

- the white furnace (every wall emits E, albedo a): the paths of length k
  carry E * a^k, for PT, LT and BPT alike.  The light tracer's host pass still
  splats the light samples themselves (LightPathtracer::Draw calls
  showlightcaches unconditionally, lightpathtracer.cpp:48-50), which the test
  measures apart (a length no path reaches leaves only that splat) and
  subtracts; the first row and column get half a footprint from the splat's
  rounding and are left out.  The NEE estimator's 1/r^2 tail near the cube's
  edges makes PT and BPT converge slowly from below, hence their wider band.
- the point-light plane, where every path has length 1: k = 1 reproduces the
  full PT and LT renders bit for bit (the same random numbers are drawn) and
  k = 2 leaves PT black and LT with the light-sample splat only.

A late roulette start (no reference counterpart; the clutter-room
configuration starts it at bounce 3) must leave every integrator unbiased:
the furnace total E * (1 + a + ... + a^L) with roulette from bounce 3.
"""
import numpy as np
import pytest

from test_oracle_integrators import furnace_scene
from test_oracle_lt_bpt_pixels import render, point_scene, CAM, W, H

E = np.array([1.0, 2.0, 3.0])
S = 32
L = 4


@pytest.fixture(scope="module")
def furnace(rt, oracle):
    hs = furnace_scene(rt)
    sc = oracle.Scene(hs.arrays())
    cam = rt.camera([0.1, 0.2, 0.3], [0.5, -0.3, -1.0], [0, 1, 0], 60, S, S)
    a = float(np.float32(128) / np.float32(255))  # albedo 0.5 through UNORM8
    return hs, sc, cam, a


def furnace_mean(rt, oracle, furnace, kind, its, nls, spl=-1, rr=0, path_length=L):
    hs, sc, cam, _ = furnace
    with oracle.path_options(spl, rr):
        img = render(rt, oracle, hs, sc, cam, kind, its, nls, path_length=path_length, seed=123, size=(S, S))
    return img[1:, 1:].reshape(-1, 3).mean(0)


@pytest.mark.parametrize("kind,its,nls,ks,tol", [
    ("pt", 256, 128, (0, 1, 2, 3), 0.04),
    ("lt", 32, 1024, (1, 2, 3, 4), 0.015),
    ("bpt", 128, 128, (0, 1, 2, 3), 0.04),
])
def test_specific_path_length_furnace(rt, oracle, furnace, kind, its, nls, ks, tol):
    a = furnace[3]
    splat = furnace_mean(rt, oracle, furnace, kind, its, nls, spl=100) if kind == "lt" else 0.0
    if kind == "lt":
        assert splat.min() > 0  # the light-sample splat runs whatever the path length
    for k in ks:
        got = furnace_mean(rt, oracle, furnace, kind, its, nls, spl=k) - splat
        want = E * a ** k
        assert np.allclose(got, want, rtol=tol), (kind, k, got / want)


def test_specific_path_length_point_plane(rt, oracle):
    hs = point_scene(rt)
    sc = oracle.Scene(hs.arrays())
    cam = rt.camera(*CAM, W, H)
    ls = rt.LightSampler(hs, 5).generate(256)
    for draw in (oracle.pathtrace, oracle.lighttrace):
        full = draw(sc, ls, cam, W, H, 77, 3)
        with oracle.path_options(1):
            one = draw(sc, ls, cam, W, H, 77, 3)
        with oracle.path_options(2):
            two = draw(sc, ls, cam, W, H, 77, 3)
        assert full[..., :3].max() > 0
        assert np.array_equal(one.view(np.uint32), full.view(np.uint32))
        if draw is oracle.pathtrace:
            assert not two[..., :3].any()
        else:
            with oracle.path_options(100):
                splat = draw(sc, ls, cam, W, H, 77, 3)
            assert np.array_equal(two.view(np.uint32), splat.view(np.uint32))
            assert np.count_nonzero(splat[..., 0]) == 1  # the point light's own pixel


@pytest.mark.parametrize("kind,its,nls,tol", [("pt", 256, 128, 0.03), ("lt", 32, 1024, 0.03), ("bpt", 128, 128, 0.03)])
def test_late_roulette_start_is_unbiased(rt, oracle, furnace, kind, its, nls, tol):
    a = furnace[3]
    pl = 5
    got = furnace_mean(rt, oracle, furnace, kind, its, nls, rr=3, path_length=pl)
    if kind == "lt":
        got = got - furnace_mean(rt, oracle, furnace, kind, its, nls, spl=100, rr=3, path_length=pl)
        want = E * sum(a ** j for j in range(1, pl + 1))
    else:
        want = E * sum(a ** j for j in range(pl + 1))
    assert np.allclose(got, want, rtol=tol), (kind, got / want)
    # roulette from bounce 0 is the reference's default
    hs, sc, cam, _ = furnace
    ls = rt.LightSampler(hs, 9).generate(nls)
    base = oracle.pathtrace(sc, ls, cam, S, S, 11, pl)
    with oracle.path_options(-1, 0):
        same = oracle.pathtrace(sc, ls, cam, S, S, 11, pl)
    assert np.array_equal(base.view(np.uint32), same.view(np.uint32))
