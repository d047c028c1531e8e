"""Pins of the surrogate-network oracle (CPU): exact float64 gradients against
central finite differences, and the BF16-rounded forms against the exact ones."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    _, n = mo.layout()
    p = np.empty(n)
    off, _ = mo.layout()
    for (w, b), fi, fo in zip(off, mo.FAN_IN, mo.FAN_OUT):
        a = np.sqrt(6.0 / fi) * scale
        p[w:b] = rng.uniform(-a, a, fi * fo)
        p[b:b + fo] = rng.uniform(-0.1, 0.1, fo)
    return p


def test_param_count():
    assert mo.layout()[1] == 5 * 128 + 128 + 3 * (128 * 128 + 128) + 128 + 1


def test_exact_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(1)
    X = rng.uniform(-1, 1, (16, 5))
    Y = rng.uniform(-1, 1, 16)
    g, _ = mo.gradients64(p, X, Y)
    idx = rng.choice(len(p), 200, replace=False)
    idx = np.concatenate([idx, [len(p) - 1, len(p) - 2, 5 * 128 + 3]])  # b5, a W5 entry, b1[3]
    eps = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp = np.mean((mo.forward64(q, X) - Y) ** 2)
        q[i] -= 2 * eps
        lm = np.mean((mo.forward64(q, X) - Y) ** 2)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.1415926, 0.0])
    r = mo.bf16(x)
    assert r[0] == 1.0
    # ties to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7 -> 1
    assert r[1] == 1.0
    # 1 + 3*2^-9 rounds up to 1 + 2^-7
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == -2.5 and r[5] == 0.0
    assert abs(r[4] - np.pi) < 2 ** -7 * 4


def test_rounded_forms_track_exact_forms():
    rng = np.random.default_rng(3)
    p = _params(2)
    X = rng.uniform(-1, 1, (64, 5))
    Y = rng.uniform(-1, 1, 64)
    T, _ = mo.forward(p, X)
    T64 = mo.forward64(p, X)
    assert np.max(np.abs(T - T64)) < 0.05 * (np.abs(T64).max() + 1)
    g, _ = mo.gradients(p, X, Y)
    g64, _ = mo.gradients64(p, X, Y)
    assert np.linalg.norm(g - g64) < 0.05 * np.linalg.norm(g64)


def test_scaled_form_cancels_its_scalings(monkeypatch):
    """forward_scaled (the fused kernel's rounding points) carries -log2(e) z
    through the hidden layers; with rounding switched off the scalings cancel
    and it is the exact network, and with rounding on it stays within the
    BF16 tolerance of it."""
    rng = np.random.default_rng(5)
    p = _params(3)
    X = rng.uniform(-1, 1, (256, 5))
    T64 = mo.forward64(p, X)
    Ts = mo.forward_scaled(p, X)
    assert np.max(np.abs(Ts - T64)) < 0.05 * (np.abs(T64).max() + 1)
    monkeypatch.setattr(mo, "bf16", lambda x: np.asarray(x, np.float64))
    monkeypatch.setattr(mo, "f32", lambda x: np.asarray(x, np.float64))
    np.testing.assert_allclose(mo.forward_scaled(p, X), T64, rtol=1e-12, atol=1e-12)
