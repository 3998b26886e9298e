"""The sweep surrogate MLP (8 -> 128 -> 128 -> 128 -> 2, SiLU).

CPU: a numpy FP64 restatement of the forward and backward pass is pinned by
central finite differences; a second numpy version rounds to BF16 at the
same points as the HIP kernels (weights, layer inputs, the output and hidden
gradients that enter products) and is checked against the FP64 one.
GPU bars:
  * forward: within 2e-3 (absolute, outputs of order 1) of the BF16 mirror;
  * gradient: relative 2-norm error <= 1e-2 against the BF16 mirror, and
    <= 5e-2 against FP64 (the BF16 rounding of the operands);
  * training with Adam lowers the loss on a smooth 8 -> 2 function by 20x.
"""
import numpy as np
import pytest

from gpucode_amd.surrogate import MLP, unpack

IN, H, OUT = 8, 128, 2
NP = H * IN + H + H * H + H + H * H + H + OUT * H + OUT


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def ref_loss_grad(p, x, t, rnd=lambda a: np.asarray(a, np.float64)):
    """Mean squared error and its gradient; rnd marks the BF16 rounding
    points of the kernels (identity for the FP64 restatement)."""
    W1, b1, W2, b2, W3, b3, W4, b4 = [np.asarray(a, np.float64) for a in unpack(p)]
    Ws = [rnd(W1), rnd(W2), rnd(W3), rnd(W4)]
    bs = [b1, b2, b3, b4]
    a = [rnd(x)]
    Z = []
    for l in range(3):
        z = a[-1] @ Ws[l].T + bs[l]
        Z.append(z)
        a.append(rnd(silu(z)))
    y = a[-1] @ Ws[3].T + bs[3]
    n2 = y.size
    loss = float(((y - t) ** 2).mean())
    dy = 2.0 * (y - t) / n2
    grads = [None] * 8
    dyb = rnd(dy)
    grads[6] = dyb.T @ a[3]
    grads[7] = dy.sum(0)
    dh = dyb @ Ws[3]
    for l in (2, 1, 0):
        dz = dh * dsilu(Z[l])
        dzb = rnd(dz)
        grads[2 * l] = dzb.T @ a[l]
        grads[2 * l + 1] = dz.sum(0)
        dh = dzb @ Ws[l]
    return y, loss, np.concatenate([g.ravel() for g in grads])


def random_params(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=NP) * 0.15
    return p.astype(np.float32)


def test_reference_gradient_by_finite_differences():
    rng = np.random.default_rng(0)
    p = random_params(1).astype(np.float64)
    x = rng.uniform(-1, 1, (16, IN))
    t = rng.normal(size=(16, OUT))
    _, _, g = ref_loss_grad(p, x, t)
    for k in rng.choice(NP, 60, replace=False):
        e = np.zeros(NP)
        h = 1e-6
        e[k] = h
        lp = ref_loss_grad(p + e, x, t)[1]
        lm = ref_loss_grad(p - e, x, t)[1]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-9, k


def test_bf16_mirror_close_to_fp64():
    rng = np.random.default_rng(2)
    p = random_params(3)
    x = rng.uniform(-1, 1, (64, IN)).astype(np.float32)
    t = rng.normal(size=(64, OUT)).astype(np.float32)
    y64, _, g64 = ref_loss_grad(p, x, t)
    yb, _, gb = ref_loss_grad(p, x, t, rnd=bf16)
    assert np.abs(y64 - yb).max() < 5e-2
    assert np.linalg.norm(g64 - gb) < 5e-2 * np.linalg.norm(g64)
    # the rounding itself: 1 + 2^-8 is a tie and rounds to even (1.0)
    assert bf16(np.float32(1.0 + 2 ** -8)) == 1.0
    assert bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6


def test_parameter_count():
    assert NP == 34434


@pytest.mark.gpu
def test_forward_and_gradient_match_reference():
    m = MLP(seed=5)
    assert m.nparams() == NP
    p = random_params(7)
    m.set_params(p)
    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 1, (300, IN)).astype(np.float32)
    t = rng.normal(size=(300, OUT)).astype(np.float32)
    y = m.forward(x)
    yb, lb, gb = ref_loss_grad(p, x, t, rnd=bf16)
    y64, l64, g64 = ref_loss_grad(p, x, t)
    assert np.abs(y - yb).max() < 2e-3
    loss, g = m.loss_grad(x, t)
    assert abs(loss - lb) < 1e-3 * lb
    assert np.linalg.norm(g - gb) <= 1e-2 * np.linalg.norm(gb)
    assert np.linalg.norm(g - g64) <= 5e-2 * np.linalg.norm(g64)
    # every layer's gradient, not just the norm of all of them
    for k, (a, b) in enumerate(zip(unpack(g), unpack(gb))):
        assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-7, k
    # a forward pass over many rows (several tiles in both directions)
    xs = rng.uniform(-1, 1, (5000, IN)).astype(np.float32)
    assert np.abs(m.forward(xs) - ref_loss_grad(p, xs, np.zeros((5000, OUT)), rnd=bf16)[0]).max() < 2e-3
    m.close()


@pytest.mark.gpu
def test_training_fits_a_smooth_function():
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, (4096, IN)).astype(np.float32)
    def f(x):
        return np.stack([np.sin(2 * x[:, 0]) + x[:, 1] ** 2 - 0.3 * x[:, 2] * x[:, 3],
                         np.cos(x[:, 4]) * x[:, 5] + 0.5 * x[:, 6] - x[:, 7] ** 3], 1)

    raw = f(x)
    mu, sd = raw.mean(0), raw.std(0)
    t = ((raw - mu) / sd).astype(np.float32)
    m = MLP(seed=1)
    l0, _ = m.loss_grad(x, t)
    loss = m.train(x, t, epochs=60, batch=256, lr=2e-3, seed=3)
    print(l0, loss)
    assert loss < l0 / 20
    # held-out points of the same function
    xv = rng.uniform(-1, 1, (1000, IN)).astype(np.float32)
    tv = ((f(xv) - mu) / sd).astype(np.float32)
    lv, _ = m.loss_grad(xv, tv)
    assert lv < l0 / 10
    m.close()
