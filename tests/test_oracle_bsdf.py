"""The oracle's BSDF sampling and evaluation (materials.glsl:43-391) pinned
against each other and against closed forms, branch by branch.

For one incident direction and a material that has only one lobe:
- the evaluation's pdf integrates to 1 over the sphere (quadrature);
- the sampler draws from that pdf: the probability of cones around the lobe
  axis matches the quadrature of the evaluated pdf;
- the mean sampled throughput equals the integral of bsdf * |cos_out| for the
  diffuse lobe and twice it for the Phong reflection and refraction lobes: the
  reference evaluates those lobes with (n+1) / (2 * PI_2 * |cos_out|)
  (materials.glsl:321) but samples them with the pdf (n+1) / PI_2 * cos^n
  (:323, :160-170), so its next-event estimate of a glossy lobe is half of the
  sampled one (a property of the reference that this build keeps);
- with roughness 0 (exponent 1e20) the sampled direction is the exact mirror
  direction, or obeys Snell's law in the plane of incidence for refraction.
"""
import numpy as np
import pytest

N = np.array([0.0, 0.0, 1.0], np.float32)
INC = np.array([0.3, 0.2, -1.0])
INC = (INC / np.linalg.norm(INC)).astype(np.float32)


def material(diffuse=(0.5, 0.6, 0.7), opacity=(1, 1, 1), exponent=20.0, fresnel0=(0, 0, 0), rindex=1.5):
    m = np.zeros(20, np.float32)
    m[0:3] = diffuse
    m[3:6] = opacity
    m[9] = exponent
    m[13:16] = fresnel0
    m[16] = rindex
    return m


LOBES = {
    "diffuse": material(),
    "reflect": material(fresnel0=(1, 1, 1)),
    "refract": material(opacity=(0, 0, 0)),
}


def sphere_grid(nc=3000, nphi=1200):
    c = (np.arange(nc) + 0.5) / nc * 2 - 1
    phi = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    s = np.sqrt(1 - c * c)
    d = np.stack([s[:, None] * np.cos(phi)[None], s[:, None] * np.sin(phi)[None],
                  np.broadcast_to(c[:, None], (nc, nphi))], -1).reshape(-1, 3)
    return d.astype(np.float32), 4 * np.pi / (nc * nphi)


@pytest.fixture(scope="module")
def grid():
    return sphere_grid()


def lobe_axis(oracle, m):
    d, _, _, _ = oracle.bsdf_samples(m.copy() if m[9] > 1e19 else np.where(np.arange(20) == 9, 1e20, m), INC, N, 1, 1)
    return d[0].astype(np.float64)


@pytest.mark.parametrize("lobe", list(LOBES))
def test_pdf_integrates_to_one(oracle, grid, lobe):
    d, dw = grid
    _, pdf = oracle.bsdf_evals(LOBES[lobe], INC, N, d)
    assert abs(pdf.astype(np.float64).sum() * dw - 1) < 2e-3


@pytest.mark.parametrize("lobe", list(LOBES))
def test_sampler_follows_the_evaluated_pdf(oracle, grid, lobe):
    m = LOBES[lobe]
    d, dw = grid
    _, pdf = oracle.bsdf_evals(m, INC, N, d)
    axis = N.astype(np.float64) if lobe == "diffuse" else lobe_axis(oracle, m)
    dirs, _, spdf, dif = oracle.bsdf_samples(m, INC, N, 12345, 400000)
    assert (dif == (lobe == "diffuse")).all()
    cg = d.astype(np.float64) @ axis
    cs = dirs.astype(np.float64) @ axis
    for c0 in (0.5, 0.8, 0.9, 0.95):
        want = pdf[cg > c0].astype(np.float64).sum() * dw
        got = (cs > c0).mean()
        assert abs(got - want) < 4e-3, (c0, got, want)
    # the pdf the sampler reports is the evaluated pdf at the sampled direction
    _, epdf = oracle.bsdf_evals(m, INC, N, dirs[:2000])
    assert np.allclose(spdf[:2000], epdf, rtol=1e-3, atol=1e-6)


@pytest.mark.parametrize("lobe,ratio", [("diffuse", 1.0), ("reflect", 2.0), ("refract", 2.0)])
def test_sampled_throughput_against_evaluated_bsdf(oracle, grid, lobe, ratio):
    m = LOBES[lobe]
    d, dw = grid
    f, _ = oracle.bsdf_evals(m, INC, N, d)
    cos_out = np.abs(d.astype(np.float64) @ N.astype(np.float64))
    integral = (f.astype(np.float64) * cos_out[:, None]).sum(0) * dw
    _, tp, _, _ = oracle.bsdf_samples(m, INC, N, 777, 400000)
    mean = tp.astype(np.float64).mean(0)
    assert np.allclose(mean, ratio * integral, rtol=5e-3), (mean, integral)
    if lobe == "diffuse":
        assert np.allclose(mean, m[0:3], rtol=1e-6)  # cosine sampling: throughput = albedo exactly


def test_mirror_direction(oracle):
    m = material(fresnel0=(1, 1, 1), exponent=1e20)
    dirs, tp, _, _ = oracle.bsdf_samples(m, INC, N, 5, 64)
    r = INC.astype(np.float64) - 2 * INC.astype(np.float64).dot(N) * N
    assert np.allclose(dirs, r, atol=1e-6)
    assert np.allclose(tp, 1.0, atol=1e-6)  # |N.out| / |N.rrd| = 1 and pr / avg(pr) = 1


@pytest.mark.parametrize("eta_in", [True, False])
def test_refraction_obeys_snell(oracle, eta_in):
    m = material(opacity=(0, 0, 0), exponent=1e20, rindex=1.5)
    inc = INC if eta_in else -INC  # from outside (cos < 0) or leaving the medium (cos > 0)
    dirs, tp, _, _ = oracle.bsdf_samples(m, inc, N, 9, 16)
    i = inc.astype(np.float64)
    t = dirs[0].astype(np.float64)
    n1, n2 = (1.0, 1.5) if eta_in else (1.5, 1.0)
    sin_i = np.linalg.norm(np.cross(i, N))
    sin_t = np.linalg.norm(np.cross(t, N))
    assert abs(n1 * sin_i - n2 * sin_t) < 1e-5
    assert abs(np.dot(np.cross(i, N), t)) < 1e-5  # in the plane of incidence
    assert np.sign(t[2]) == np.sign(i[2])  # continues through the surface
    # radiance scaling 1 / eta^2 (eta = n1 / n2) and the |N.out| / |N.rrd| = 1 factor
    assert np.allclose(tp[0], (n2 / n1) ** 2, rtol=1e-5)
