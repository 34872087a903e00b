"""Oracle FP32 transcendentals pinned against FP64 libm within ulp bounds.

The renderer defines sin/cos/pow/atan2 by these Cephes-style kernels (the
device uses the same algorithm), so what must be pinned is their accuracy.
A dropped polynomial term, a wrong quadrant or a wrong reduction constant
breaks these bounds by orders of magnitude.
"""
import numpy as np


def ulp_err(got, want):
    got = np.float32(got)
    want64 = np.float64(want)
    w32 = np.float32(want64)
    if np.isinf(w32) or np.isnan(w32):
        return 0.0 if got == w32 else np.inf
    spacing = np.spacing(np.abs(w32)) if w32 != 0 else np.float32(1.4e-45)
    return float(abs(np.float64(got) - want64) / np.float64(spacing))


def test_sin_cos_accuracy(oracle):
    L = oracle.lib()
    xs = np.concatenate([np.linspace(-7.0, 7.0, 20001), np.linspace(0, 6.28318530717958, 4097),
                         np.float32([0.0, np.pi / 2, np.pi, 3 * np.pi / 2, 1e-6, -1e-6])]).astype(np.float32)
    worst_abs = 0.0
    for x in xs:
        s, c = L.or_sinf(float(x)), L.or_cosf(float(x))
        ws, wc = np.sin(np.float64(x)), np.cos(np.float64(x))
        worst_abs = max(worst_abs, abs(s - ws), abs(c - wc))
        # relative bound away from zeros, absolute bound near them
        if abs(ws) > 1e-3:
            assert ulp_err(s, ws) <= 4.0, (x, s, ws)
        if abs(wc) > 1e-3:
            assert ulp_err(c, wc) <= 4.0, (x, c, wc)
    assert worst_abs < 2e-7


def test_sin_cos_identity_and_symmetry(oracle):
    L = oracle.lib()
    for x in np.linspace(-6, 6, 999, dtype=np.float32):
        s, c = L.or_sinf(float(x)), L.or_cosf(float(x))
        assert abs(s * s + c * c - 1.0) < 1e-6
        assert L.or_sinf(float(-x)) == -s or abs(L.or_sinf(float(-x)) + s) < 1e-7


def test_log2_exp2_pow_accuracy(oracle):
    L = oracle.lib()
    xs = np.concatenate([np.geomspace(1e-30, 1e30, 4001), np.linspace(0.5, 2.0, 2001)]).astype(np.float32)
    for x in xs:
        got = L.or_log2f(float(x))
        want = np.log2(np.float64(x))
        assert abs(got - want) <= max(2e-7 * abs(want), 2e-7), (x, got, want)
    for y in np.linspace(-140.0, 127.5, 5001, dtype=np.float32):
        got = L.or_exp2f(float(y))
        want = np.exp2(np.float64(y))
        if want < 1.17549435e-38:  # denormal results: absolute bound
            assert abs(got - want) <= 2 * 1.4e-45 + 4e-7 * want
        else:
            assert ulp_err(got, want) <= 3.0, (y, got, want)
    # pow on the renderer's domain: x in [0,1], exponent in (0, 1000]
    rng = np.random.default_rng(0)
    for x, e in zip(rng.random(3000).astype(np.float32), np.geomspace(1e-3, 1000, 3000).astype(np.float32)):
        got = L.or_powf(float(x), float(e))
        want = np.float64(x) ** np.float64(e)
        if want > 1e-30:
            # error of exp2(e * log2 x) grows with |e * log2 x|
            tol = 4e-7 * (1.0 + abs(np.float64(e) * np.log2(np.float64(x)))) * want
            assert abs(got - want) <= tol + 1e-37, (x, e, got, want)
    assert L.or_powf(0.0, 3.0) == 0.0
    assert L.or_log2f(0.0) == -np.inf
    assert np.isnan(L.or_log2f(-1.0))
    assert L.or_exp2f(200.0) == np.inf and L.or_exp2f(-200.0) == 0.0


def test_atan2_accuracy_and_quadrants(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(3)
    pts = rng.normal(size=(5000, 2)).astype(np.float32)
    for y, x in pts:
        got = L.or_atan2f(float(y), float(x))
        want = np.arctan2(np.float64(y), np.float64(x))
        assert abs(got - want) <= 4e-7 * max(1.0, abs(want)), (y, x, got, want)
    assert L.or_atan2f(1.0, 0.0) == np.float32(np.pi / 2)
    assert L.or_atan2f(-1.0, 0.0) == np.float32(-np.pi / 2)
    assert L.or_atan2f(0.0, 0.0) == 0.0
