"""The network oracle (oracle/unet_oracle.py) pinned by central differences
in float64, and its layer table against the library's parameter count."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import unet_oracle as U  # noqa: E402


def small_problem(c0=2, B=1, H=16, W=16, seed=0):
    rng = np.random.default_rng(seed)
    n = U.param_count(c0)
    p = rng.normal(0, 0.3, n)
    x = rng.normal(0, 1, (B, H, W, 3))
    t = rng.normal(0, 1, (B, H, W, 2))
    m = (rng.uniform(size=(B, H, W)) > 0.3).astype(np.float64)
    return p, x, t, m


def test_gradients_match_central_differences():
    c0 = 2
    p, x, t, m = small_problem(c0)
    _, loss, g = U.forward_backward(p, c0, x, t, m, use_bf16=False)
    rng = np.random.default_rng(1)
    idx = rng.choice(len(p), 40, replace=False)
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp = U.forward_backward(q, c0, x, t, m, use_bf16=False, need_grad=False)[1]
        q[i] -= 2 * h
        lm = U.forward_backward(q, c0, x, t, m, use_bf16=False, need_grad=False)[1]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (i, fd, g[i])


def test_loss_is_masked_mean_square():
    c0 = 2
    p, x, t, m = small_problem(c0)
    out, loss, _ = U.forward_backward(p, c0, x, t, m, use_bf16=False, need_grad=False)
    ref = (m[..., None] * (out - t) ** 2).sum() / (2 * m.sum())
    assert abs(loss - ref) <= 1e-12 * abs(ref)
    # a zero head reproduces the head bias everywhere
    q = p.copy()
    layers = U.layer_table(c0)
    off = sum(k * k * co * ci + co for ci, co, k, _ in layers[:-1])
    q[off:off + 2 * layers[-1][0]] = 0.0
    q[off + 2 * layers[-1][0]:] = [0.25, -0.5]
    out, _, _ = U.forward_backward(q, c0, x, t, m, use_bf16=False, need_grad=False)
    assert np.allclose(out[..., 0], 0.25) and np.allclose(out[..., 1], -0.5)


def test_bf16_rounding_is_nearest_even():
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38], np.float32)
    r = U.bf16(a)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == 1.0 and r[4] == -2.5
    assert np.all(np.abs(r - a) <= np.abs(a) * 2 ** -8)


def test_param_count_of_the_c0_24_network():
    # widths 24, 48, 96, 192: 1.1 M parameters
    assert U.param_count(24) == 1095578
