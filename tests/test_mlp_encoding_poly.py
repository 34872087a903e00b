"""The sin / cos of the radiance-cache input encoding (`sincospi_enc` and the
double-angle steps of `encode_axis` in gpugi_amd/csrc/device/mlp.hip).

The coefficients are read from the kernel source, the kernel's FP32 arithmetic
is emulated operation by operation in numpy (an FMA as an exact float64
product plus the addend, rounded once to FP32), and the results are checked
against float64 sin / cos: the error bounds the kernel's comments state, and
the values the reduction fixes exactly (integers and half-integers)."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "gpugi_amd", "csrc", "device", "mlp.hip")
f32 = np.float32


def _coefficients():
    text = open(SRC).read()
    body = text[text.index("void sincospi_enc("):]
    body = body[:body.index("\n}\n")]
    num = r"(-?\d+\.\d+)f"
    q3, q2 = re.search(r"float q = fmaf\(z, %s, %s\);" % (num, num), body).groups()
    q1, q0 = re.findall(r"\bq = fmaf\(z, q, %s\);" % num, body)
    r4, r3 = re.search(r"float r = fmaf\(z, %s, %s\);" % (num, num), body).groups()
    r2, r1, r0 = re.findall(r"\br = fmaf\(z, r, %s\);" % num, body)
    pi = re.search(r"fmaf\(t, %s, \(t \* z\) \* q\)" % num, body).group(1)
    one = re.search(r"const float c = fmaf\(z, r, %s\);" % num, body).group(1)
    c = lambda v: f32(float(v))
    return [c(q3), c(q2), c(q1), c(q0)], [c(r4), c(r3), c(r2), c(r1), c(r0)], c(pi), c(one)


def _fma(a, b, c):
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(f32)


def sincospi_enc(x):
    """The kernel's steps in FP32 for FP32 inputs x."""
    Q, R, pi, one = _coefficients()
    x = np.asarray(x, f32)
    k = np.rint(x).astype(f32)
    t = (x - k).astype(f32)
    z = (t * t).astype(f32)
    q = _fma(z, Q[0], Q[1])
    for c in Q[2:]:
        q = _fma(z, q, c)
    r = _fma(z, R[0], R[1])
    for c in R[2:]:
        r = _fma(z, r, c)
    s = _fma(t, pi, ((t * z).astype(f32) * q).astype(f32))
    co = _fma(z, r, one)
    odd = (k.astype(np.int64) & 1).astype(bool)
    return np.where(odd, -s, s).astype(f32), np.where(odd, -co, co).astype(f32)


def encode_axis(p):
    """sin / cos of 2^k pi p for k < 12, as the kernel builds them."""
    p = np.asarray(p, f32)
    sn, cs = [], []
    for a in range(0, 12, 3):
        s, c = sincospi_enc(np.ldexp(p, a).astype(f32))
        sn.append(s)
        cs.append(c)
        for _ in range(2):
            s, c = (((s + s).astype(f32) * c).astype(f32), ((c - s).astype(f32) * (c + s).astype(f32)).astype(f32))
            sn.append(s)
            cs.append(c)
    return np.stack(sn, 1), np.stack(cs, 1)


def test_reduced_range_error():
    t = np.linspace(-0.5, 0.5, 1_000_001).astype(f32)
    s, c = sincospi_enc(t)
    tt = np.pi * t.astype(np.float64)
    assert np.max(np.abs(s - np.sin(tt))) <= 1.2e-7
    assert np.max(np.abs(c - np.cos(tt))) <= 1.2e-7


def test_scaled_positions_error():
    # the arguments the encoding passes: positions in [0, 1] times 2^0, 2^3, 2^6, 2^9
    rng = np.random.default_rng(3)
    p = rng.random(500_000).astype(f32)
    for a in (0, 3, 6, 9):
        x = np.ldexp(p, a).astype(f32)
        s, c = sincospi_enc(x)
        xx = np.pi * x.astype(np.float64)
        assert np.max(np.abs(s - np.sin(xx))) <= 1.2e-7
        assert np.max(np.abs(c - np.cos(xx))) <= 1.2e-7


def test_exact_at_integers_and_half_integers():
    n = np.arange(-600, 601).astype(f32)
    s, c = sincospi_enc(n)
    assert np.all(s == 0.0)
    assert np.all(c == np.where(n.astype(np.int64) % 2 == 0, 1.0, -1.0))
    s, c = sincospi_enc(n + f32(0.5))
    assert np.max(np.abs(np.abs(s) - 1.0)) <= 1.2e-7
    assert np.max(np.abs(c)) <= 1.2e-7
    # sign: sin(pi (n + 1/2)) = (-1)^n
    assert np.all(np.sign(s) == np.where(n.astype(np.int64) % 2 == 0, 1.0, -1.0))


def test_encoded_features_error():
    rng = np.random.default_rng(4)
    edge = np.array([0.0, 1.0, 0.5, 0.25, 2.0 ** -12, 1.0 - 2.0 ** -24, 1.0 / 3.0], np.float32)
    p = np.concatenate([edge, np.linspace(0, 1, 400_001).astype(f32), rng.random(400_000).astype(f32)])
    sn, cs = encode_axis(p)
    x = np.pi * np.ldexp(p.astype(np.float64)[:, None], np.arange(12)[None, :])
    assert np.max(np.abs(sn - np.sin(x))) <= 7e-7
    assert np.max(np.abs(cs - np.cos(x))) <= 7e-7


@pytest.mark.parametrize("which", range(9))
def test_each_coefficient_matters(which):
    """A coefficient off by 1 % (a mistyped digit) breaks the reduced-range
    bound, so the bound pins every coefficient the source holds."""
    Q, R, pi, one = _coefficients()
    coeffs = Q + R
    t = np.linspace(-0.5, 0.5, 200_001).astype(f32)
    z = (t * t).astype(np.float64)
    tt = np.pi * t.astype(np.float64)
    bumped = [float(v) for v in coeffs]
    bumped[which] *= 1.0 + 1e-2
    q = np.polyval(bumped[:4], z)
    r = np.polyval(bumped[4:], z)
    err = max(np.max(np.abs(float(pi) * t + t * z * q - np.sin(tt))), np.max(np.abs(1.0 + z * r - np.cos(tt))))
    assert err > 1.2e-7
