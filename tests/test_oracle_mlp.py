"""Pins the network oracle (oracle/mlp_oracle.py) independently of its own
formulas, so the GPU training kernels that are checked against it inherit a
real pin:

  backward   central finite differences (FP64) of the mean-squared-error loss
             for every parameter block W1, b1, W2, b2, W3, b3
  gelu_grad  central differences of gelu itself
  gelu       the tanh form's defining identity x sigmoid(2 sqrt(2/pi)(x + 0.044715 x^3))
             and its limits (gelu(x) -> x, -> 0)
  adam       Adam's first step moves every parameter by lr sign(g) (bias
             correction: m_hat = g, v_hat = g^2), and a hand-iterated
             three-step run
  to_bf16    torch's float32 -> bfloat16 rounding (round to nearest even)
"""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo

BLOCKS = (("W1", mo.OFF_W1, mo.OFF_B1), ("b1", mo.OFF_B1, mo.OFF_W2), ("W2", mo.OFF_W2, mo.OFF_B2),
          ("b2", mo.OFF_B2, mo.OFF_W3), ("W3", mo.OFF_W3, mo.OFF_B3), ("b3", mo.OFF_B3, mo.NPARAMS))


def _params(rng):
    p = np.empty(mo.NPARAMS)
    p[mo.OFF_W1:mo.OFF_B1] = rng.normal(0, 1 / np.sqrt(5), 5 * 64)
    p[mo.OFF_B1:mo.OFF_W2] = rng.normal(0, 0.1, 64)
    p[mo.OFF_W2:mo.OFF_B2] = rng.normal(0, 1 / 8, 64 * 64)
    p[mo.OFF_B2:mo.OFF_W3] = rng.normal(0, 0.1, 64)
    p[mo.OFF_W3:mo.OFF_B3] = rng.normal(0, 1 / 8, 64 * 3)
    p[mo.OFF_B3:] = rng.normal(0, 0.1, 3)
    return p


def _bump(p, i, h):
    q = p.copy()
    q[i] += h
    return q


def _loss(p, x, t):
    y, _ = mo.forward(p, x)
    return ((y - t) ** 2).mean()


@pytest.mark.parametrize("name,lo,hi", BLOCKS)
def test_backward_matches_central_differences_of_the_loss(name, lo, hi):
    rng = np.random.default_rng(31)
    p = _params(rng)
    x = rng.normal(0, 1.5, (37, 5))
    t = rng.normal(0, 1, (37, 3))
    y, cache = mo.forward(p, x)
    dy = 2 * (y - t) / y.size  # d mean((y-t)^2) / dy
    g = mo.backward(p, x, cache, dy)
    idx = np.arange(lo, hi)
    if idx.size > 48:
        idx = rng.choice(idx, 48, replace=False)
    for i in idx:
        # Richardson-extrapolated central differences: truncation ~h^4, and
        # cancellation ~1e-16 |loss| / h ~ 1e-12 at h = 1e-4
        h = 1e-4 * max(1.0, abs(p[i]))
        d1 = (_loss(_bump(p, i, h), x, t) - _loss(_bump(p, i, -h), x, t)) / (2 * h)
        d2 = (_loss(_bump(p, i, h / 2), x, t) - _loss(_bump(p, i, -h / 2), x, t)) / h
        fd = (4 * d2 - d1) / 3
        assert abs(g[i] - fd) <= 1e-8 * abs(fd) + 1e-11, (name, i, g[i], fd)


def test_backward_pins_every_block_against_a_plausible_mistake():
    """A gradient with gelu_grad replaced by gelu'(x) ~ 1/2 + x/2 (the
    small-x slope) misses the central differences on every hidden block."""
    rng = np.random.default_rng(32)
    p = _params(rng)
    x = rng.normal(0, 1.5, (29, 5))
    t = rng.normal(0, 1, (29, 3))
    y, cache = mo.forward(p, x)
    dy = 2 * (y - t) / y.size
    orig = mo.gelu_grad
    try:
        mo.gelu_grad = lambda z: 0.5 + 0.5 * z
        bad = mo.backward(p, x, cache, dy)
    finally:
        mo.gelu_grad = orig
    for name, lo, hi in BLOCKS[:4]:
        i = lo + 3
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (_loss(pp, x, t) - _loss(pm, x, t)) / (2 * h)
        assert abs(bad[i] - fd) > 1e-4 * abs(fd), name


def test_gelu_grad_is_the_derivative_of_gelu():
    x = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (mo.gelu(x + h) - mo.gelu(x - h)) / (2 * h)
    assert np.abs(mo.gelu_grad(x) - fd).max() < 1e-8


def test_gelu_tanh_form_identities():
    x = np.linspace(-8, 8, 321)
    a = np.sqrt(2 / np.pi) * (x + 0.044715 * x ** 3)
    # (atol: the tanh form cancels to 0 where x sigmoid(.) is ~1e-21)
    assert np.allclose(mo.gelu(x), x / (1 + np.exp(-2 * a)), rtol=1e-14, atol=1e-15)
    assert mo.gelu(np.array([0.0]))[0] == 0.0
    assert abs(mo.gelu(np.array([10.0]))[0] - 10.0) < 1e-12
    assert abs(mo.gelu(np.array([-10.0]))[0]) < 1e-12
    # the tanh form is within 1e-3 of the exact x Phi(x) (Hendrycks & Gimpel)
    from scipy.special import erf
    assert np.abs(mo.gelu(x) - 0.5 * x * (1 + erf(x / np.sqrt(2)))).max() < 1e-3


def test_adam_first_step_and_short_run():
    rng = np.random.default_rng(33)
    n = 50
    p = rng.normal(size=n)
    g = rng.normal(size=n)
    lr = 1e-2
    p1, m, v = mo.adam(p, g, np.zeros(n), np.zeros(n), lr, 0.9, 0.999, 0.0, 1)
    assert np.allclose(p1, p - lr * np.sign(g), rtol=0, atol=1e-15)
    # three steps with the moment recursions iterated by hand
    gs = [rng.normal(size=n) for _ in range(3)]
    pk, mk, vk = p.copy(), np.zeros(n), np.zeros(n)
    q, mm, vv = p.copy(), np.zeros(n), np.zeros(n)
    for s, gk in enumerate(gs, start=1):
        pk, mk, vk = mo.adam(pk, gk, mk, vk, lr, 0.8, 0.99, 1e-8, s)
        mm = sum(0.2 * 0.8 ** (s - j) * gs[j - 1] for j in range(1, s + 1))
        vv = sum(0.01 * 0.99 ** (s - j) * gs[j - 1] ** 2 for j in range(1, s + 1))
        q = q - lr * (mm / (1 - 0.8 ** s)) / (np.sqrt(vv / (1 - 0.99 ** s)) + 1e-8)
    assert np.allclose(pk, q, rtol=1e-13, atol=1e-15)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(34)
    x = np.concatenate([rng.normal(0, 3, 5000), rng.uniform(-1e-3, 1e-3, 1000),
                        np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0, -0.0])]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mo.to_bf16(x), ref)
