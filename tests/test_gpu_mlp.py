"""Radiance-cache MLP kernels (BF16 matrix cores) against the numpy oracle.

Encoding, inference and the backward pass agree with the oracle to BF16
precision (the kernels sum in FP32 in matrix-core order); the Adam update is
exact for a given gradient; training lowers the loss on a smooth target."""
import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


def features(n, seed):
    rng = np.random.default_rng(seed)
    f = np.zeros((n, 16), np.float32)
    f[:, 0:3] = rng.random((n, 3))
    d = rng.normal(size=(n, 3))
    f[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    f[:, 6:9] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    f[:, 9:12] = rng.random((n, 3))
    f[:, 12] = rng.random(n)
    return f


def smooth_target(f):
    t = np.zeros((len(f), 4), np.float32)
    t[:, 0] = 0.5 + 0.4 * np.sin(2 * np.pi * f[:, 0]) * f[:, 9]
    t[:, 1] = 0.5 + 0.4 * np.cos(2 * np.pi * f[:, 2]) * f[:, 10]
    t[:, 2] = 0.3 + 0.2 * f[:, 7] * f[:, 11]
    return t


@pytest.fixture(scope="module")
def mctx(rt):
    c = rt.Context(0)
    c.mlp_init(7)
    yield c
    c.close()


def test_encoding(mctx):
    f = features(1000, 1)
    # positions where sin or cos of 2^k pi p is 0 or +-1 for every k, or where the
    # reduction to [-1/2, 1/2] changes sides
    edge = np.array([0.0, 1.0, 0.5, 0.25, 2.0 ** -9, 2.0 ** -10, 2.0 ** -12, 1.0 - 2.0 ** -24, 2.0 ** -24,
                     1.0 / 3.0, 0.75 + 2.0 ** -12, 0.5 - 2.0 ** -11], np.float32)
    f[:len(edge), 0] = edge
    f[:len(edge), 1] = edge[::-1]
    f[len(edge):2 * len(edge), 2] = edge
    enc = mlp.from_bf16_bits(mctx.mlp_encode(f))
    ref = mlp.encode(f)
    # the device's FP32 sin / cos are within 7e-7 of the exact values (a reduced
    # polynomial and two double-angle steps): at most one BF16 step apart
    assert np.all(np.abs(enc - ref) <= 2.0 ** -7 * np.abs(ref) + 1e-6)
    assert (enc[:, 82:] == 0).all()


@pytest.mark.parametrize("n", [1, 16, 1000, 4096])
def test_inference_matches_oracle(mctx, n):
    f = features(n, 2 + n)
    w = mctx.mlp_get_weights()
    y = mctx.mlp_infer(f)
    _, ref = mlp.forward(w, mlp.encode(f))
    assert (y[:, 3] == 0).all()
    err = np.linalg.norm(y[:, :3] - ref[:, :3]) / np.linalg.norm(ref[:, :3])
    assert err < 1e-2, err


@pytest.mark.parametrize("n", [32, 1000, 8192])
def test_gradient_matches_oracle(mctx, n):
    f = features(n, 11 + n)
    t = smooth_target(f)
    w = mctx.mlp_get_weights()
    g, loss = mctx.mlp_gradient(f, t)
    lref, gref = mlp.loss_and_gradient(w, mlp.encode(f), t)
    assert loss == pytest.approx(lref, rel=1e-2)
    for l in range(5):
        a, b = g[mlp.W_OFF[l]:mlp.W_OFF[l + 1]], gref[mlp.W_OFF[l]:mlp.W_OFF[l + 1]]
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert err < 3e-2, (l, err)
    assert (g[mlp.W_OFF[4] + 3 * 64:] == 0).all()


def test_gradient_is_deterministic(mctx):
    f = features(3000, 5)
    t = smooth_target(f)
    g1, l1 = mctx.mlp_gradient(f, t)
    g2, l2 = mctx.mlp_gradient(f, t)
    assert np.array_equal(g1.view(np.uint32), g2.view(np.uint32)) and l1 == l2


def test_adam_step_is_exact(rt):
    c = rt.Context(0)
    try:
        c.mlp_init(3)
        f = features(2048, 8)
        t = smooth_target(f)
        w0 = c.mlp_get_weights()
        m = np.zeros_like(w0)
        v = np.zeros_like(w0)
        w = w0
        for step in range(1, 4):
            g, _ = c.mlp_gradient(f, t)
            c.mlp_train_step(f, t, lr=1e-3)
            w, m, v = mlp.adam(w, m, v, g, step, 1e-3)
            assert np.array_equal(c.mlp_get_weights().view(np.uint32), w.view(np.uint32))
    finally:
        c.close()


def test_training_lowers_the_loss(rt):
    c = rt.Context(0)
    try:
        c.mlp_init(11)
        f = features(16384, 21)
        t = smooth_target(f)
        first = c.mlp_train_step(f, t, lr=1e-3)
        for _ in range(300):
            last = c.mlp_train_step(f, t, lr=1e-3)
        assert last < 0.2 * first, (first, last)
        # held-out samples generalize
        fh = features(4096, 22)
        y = c.mlp_infer(fh)
        th = smooth_target(fh)
        rel = np.mean((y[:, :3] - th[:, :3]) ** 2 / (y[:, :3] ** 2 + 0.01))
        assert rel < 0.5 * first
    finally:
        c.close()
