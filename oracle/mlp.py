"""Radiance-cache MLP oracle -- TEST INFRASTRUCTURE ONLY.

numpy restatement of the cache network the HIP kernels run (96 encoded
inputs, 4 hidden ReLU layers of 64, 3 outputs padded to 16, no biases), its
relative-L2 loss, backward pass and Adam step.  The cache has no counterpart
in the reference; the kernels' arithmetic is what this restates:

- inputs, weights, activations and backpropagated gradients are rounded to
  BF16 (round to nearest even) at the points the kernels round them;
- products are summed in float64 here and in FP32 (matrix-core order) on the
  GPU, so the two agree to BF16 precision, not bit for bit;
- Adam runs in float32 with the kernels' operation order and matches them
  exactly for a given gradient.

`exact=True` turns every BF16 rounding off; in that mode the backward pass
is pinned to torch autograd (tests/test_oracle_mlp.py).
"""
import numpy as np

N_IN, N_HID, N_OUT, N_USED = 96, 64, 16, 3
LAYER_IN = (96, 64, 64, 64, 64)
LAYER_OUT = (64, 64, 64, 64, 16)
W_OFF = (0, 6144, 10240, 14336, 18432, 19456)
PARAMS = 19456
FREQS = 12


def bf16(x):
    """float32 values rounded to BF16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def bf16_bits(x):
    return (bf16(x).view(np.uint32) >> 16).astype(np.uint16)


def from_bf16_bits(b):
    return (np.asarray(b, np.uint32) << 16).view(np.float32)


def encode(features, exact=False):
    """n x 16 raw features -> n x 96 inputs (float32; BF16 values unless exact)."""
    f = np.asarray(features, np.float64).reshape(-1, 16)
    n = len(f)
    e = np.zeros((n, N_IN), np.float64)
    for a in range(3):
        for k in range(FREQS):
            x = np.pi * (2.0 ** k) * f[:, a]
            e[:, a * FREQS + k] = np.sin(x)
            e[:, 36 + a * FREQS + k] = np.cos(x)
    e[:, 72:81] = f[:, 3:12]
    e[:, 81] = f[:, 12]
    e = e.astype(np.float32)
    return e if exact else bf16(e)


def split(w):
    w = np.asarray(w, np.float32)
    return [w[W_OFF[l]:W_OFF[l + 1]].reshape(LAYER_OUT[l], LAYER_IN[l]) for l in range(5)]


def _rounder(exact):
    """float64 -> float64 through BF16 (the kernels' storage points), or unchanged."""
    if exact:
        return lambda a: np.asarray(a, np.float64)
    return lambda a: bf16(np.asarray(a, np.float64).astype(np.float32)).astype(np.float64)


def forward(w, x, exact=False):
    """Returns (activations [input, h1..h4], outputs n x 16 float64)."""
    rnd = _rounder(exact)
    W = [rnd(m) for m in split(w)]
    acts = [np.asarray(x, np.float32).astype(np.float64)]
    for l in range(4):
        acts.append(rnd(np.maximum(acts[-1] @ W[l].T, 0.0)))
    y = acts[-1] @ W[4].T
    return acts, y


def loss_and_gradient(w, x, targets, exact=False):
    """Relative L2 loss mean_{s, c<3} (y - t)^2 / (y^2 + 0.01), denominator held
    constant, and its gradient with respect to the 19,456 weights."""
    rnd = _rounder(exact)
    t = np.asarray(targets, np.float64).reshape(-1, 4)
    n = len(t)
    acts, y = forward(w, x, exact)
    W = [rnd(m) for m in split(w)]
    inv = 1.0 / (3.0 * n)
    d = y[:, :N_USED] - t[:, :N_USED]
    den = y[:, :N_USED] ** 2 + 0.01
    loss = float(np.sum(d * d / den) * inv)
    dy = np.zeros((n, N_OUT), np.float64)
    dy[:, :N_USED] = 2.0 * d / den * inv
    dy = rnd(dy)
    grads = [None] * 5
    for l in range(4, -1, -1):
        grads[l] = dy.T @ acts[l]
        if l > 0:
            dy = rnd(np.where(acts[l] > 0.0, dy @ W[l], 0.0))
    g = np.concatenate([gr.reshape(-1) for gr in grads])
    return loss, g


def adam(w, m, v, g, step, lr, beta1=0.9, beta2=0.99, eps=1e-8):
    """One Adam step in float32 with the kernel's operation order; returns (w, m, v)."""
    f = np.float32
    w, m, v, g = (np.asarray(a, np.float32) for a in (w, m, v, g))
    b1, b2 = f(beta1), f(beta2)
    m = b1 * m + (f(1) - b1) * g
    v = b2 * v + (f(1) - b2) * (g * g)
    # bias corrections: 1 - beta^step in double from the float32 betas, then rounded
    bc1 = f(1.0 - float(b1) ** step)
    bc2 = f(1.0 - float(b2) ** step)
    mhat = m / bc1
    vhat = v / bc2
    w = w - f(lr) * mhat / (np.sqrt(vhat) + f(eps))
    return w.astype(np.float32), m.astype(np.float32), v.astype(np.float32)
