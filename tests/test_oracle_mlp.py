"""The MLP oracle pinned by independent references: torch autograd for the
backward pass, torch.optim.Adam for the update, closed forms for the
encoding and the BF16 rounding."""
import numpy as np
import pytest
import torch

from oracle import mlp


def random_problem(n, seed):
    rng = np.random.default_rng(seed)
    f = np.zeros((n, 16), np.float32)
    f[:, 0:3] = rng.random((n, 3))
    d = rng.normal(size=(n, 3))
    f[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    f[:, 6:9] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    f[:, 9:12] = rng.random((n, 3))
    f[:, 12] = rng.random(n)
    w = np.zeros(mlp.PARAMS, np.float32)
    for l in range(5):
        fan = mlp.LAYER_IN[l]
        rows = 3 if l == 4 else mlp.LAYER_OUT[l]
        blk = (rng.random((rows, fan)) * 2 - 1) * np.sqrt(6.0 / fan)
        w[mlp.W_OFF[l]:mlp.W_OFF[l] + rows * fan] = blk.reshape(-1)
    t = np.zeros((n, 4), np.float32)
    t[:, :3] = rng.random((n, 3)) * 2
    return f, w, t


def test_bf16_rounding_closed_form():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 1e-30, 3.0e38], np.float32)
    r = mlp.bf16(x)
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1 + 2 ** -7) and r[3] == np.float32(1 + 2 ** -6)
    assert r[4] == -2.5
    assert abs(r[5] - 1e-30) <= 2 ** -8 * 1e-30
    assert (mlp.from_bf16_bits(mlp.bf16_bits(x)) == r).all()


def test_encoding_closed_form():
    f = np.zeros((1, 16), np.float32)
    f[0, :3] = (0.25, 0.5, 0.0)
    f[0, 3:13] = np.arange(1, 11) / 16.0
    e = mlp.encode(f, exact=True)[0]
    assert e[0] == pytest.approx(np.sin(np.pi / 4)) and e[36] == pytest.approx(np.cos(np.pi / 4))
    assert e[1] == pytest.approx(1.0) and abs(e[37]) < 1e-7        # k=1: sin(pi/2), cos(pi/2)
    assert e[12] == pytest.approx(1.0) and e[48] == pytest.approx(0.0, abs=1e-7)  # axis y, p = 1/2
    assert (e[24:36] == 0).all() and (e[60:72] == 1).all()           # axis z at 0
    assert (e[72:82] == f[0, 3:13]).all() and (e[82:] == 0).all()


def test_backward_matches_torch_autograd():
    f, w, t = random_problem(67, 3)
    x = mlp.encode(f, exact=True)
    loss, g = mlp.loss_and_gradient(w, x, t, exact=True)
    Wt = [torch.tensor(m, dtype=torch.float64, requires_grad=True) for m in mlp.split(w)]
    h = torch.tensor(x, dtype=torch.float64)
    for l in range(4):
        h = torch.relu(h @ Wt[l].T)
    y = (h @ Wt[4].T)[:, :3]
    tt = torch.tensor(t[:, :3], dtype=torch.float64)
    L = (((y - tt) ** 2) / (y.detach() ** 2 + 0.01)).mean()
    L.backward()
    gt = np.concatenate([m.grad.numpy().reshape(-1) for m in Wt])
    assert loss == pytest.approx(L.item(), rel=1e-12)
    assert np.allclose(g, gt, rtol=1e-9, atol=1e-15)
    # the padded output rows never receive a gradient
    assert (g[mlp.W_OFF[4] + 3 * 64:] == 0).all()


def test_bf16_emulation_stays_close_to_exact():
    f, w, t = random_problem(256, 5)
    l0, g0 = mlp.loss_and_gradient(w, mlp.encode(f, exact=True), t, exact=True)
    l1, g1 = mlp.loss_and_gradient(w, mlp.encode(f), t)
    assert l1 == pytest.approx(l0, rel=2e-2)
    assert np.linalg.norm(g1 - g0) <= 5e-2 * np.linalg.norm(g0)


def test_adam_matches_torch_optimizer():
    rng = np.random.default_rng(9)
    w = rng.normal(size=mlp.PARAMS).astype(np.float32)
    p = torch.tensor(w.copy(), requires_grad=True)
    opt = torch.optim.Adam([p], lr=1e-3, betas=(0.9, 0.99), eps=1e-8)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    for step in range(1, 6):
        g = rng.normal(size=mlp.PARAMS).astype(np.float32)
        p.grad = torch.tensor(g)
        opt.step()
        w, m, v = mlp.adam(w, m, v, g, step, 1e-3)
        assert np.allclose(w, p.detach().numpy(), rtol=0, atol=2e-6)
