"""Surrogate network kernels on the GPU against the CPU oracle: the BF16 MFMA
GEMM in every transpose form, the fused forward, and the backward pass."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("M,N,K", [(64, 32, 16), (200, 130, 70), (5, 128, 333), (257, 1, 128)])
def test_gemm_matches_bf16_reference(amca, ta, tb, M, N, K):
    from amca3d_amd.surrogate import gemm_selftest
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = rng.uniform(-1, 1, (K, M) if ta else (M, K))
    B = rng.uniform(-1, 1, (N, K) if tb else (K, N))
    got = gemm_selftest(A, B, ta, tb)
    a = mo.bf16(A)
    b = mo.bf16(B)
    ref = (a.T if ta else a) @ (b.T if tb else b)
    # FP32 accumulation of exact BF16 products
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * K + 1e-6)


def _params(seed):
    rng = np.random.default_rng(seed)
    off, n = mo.layout()
    p = np.empty(n)
    for (w, b), fi, fo in zip(off, mo.FAN_IN, mo.FAN_OUT):
        a = np.sqrt(6.0 / fi)
        p[w:b] = rng.uniform(-a, a, fi * fo)
        p[b:b + fo] = rng.uniform(-0.1, 0.1, fo)
    return p.astype(np.float32)


def test_forward_matches_oracle(amca):
    from amca3d_amd.surrogate import MLP
    m = MLP(1)
    p = _params(5)
    m.set_params(p)
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, (1000, 5)).astype(np.float32)
    t = m.forward(X)
    ref, _ = mo.forward(p.astype(np.float64), X)
    # the oracle rounds at the same points; FP32 accumulation order and the
    # occasional BF16 tie broken differently remain
    np.testing.assert_allclose(t, ref, rtol=0, atol=2e-2 * (np.abs(ref).max() + 1))
    assert np.corrcoef(t, ref)[0, 1] > 0.9999
    m.close()


@pytest.mark.parametrize("n", [1, 31, 1000, 70001])
def test_fused_forward_matches_oracle_and_layered(amca, n):
    """The fused kernel (weights in LDS, activations in registers, k order of
    each layer permuted to the accumulator layout, hidden layers carrying
    -log2(e) z) computes the network: its own rounding points (the oracle's
    forward_scaled) to FP32 accumulation order and the hardware exp2 and
    reciprocal, and the layered kernels (activations rounded unscaled) to
    BF16 rounding."""
    from amca3d_amd.surrogate import MLP
    m = MLP(1)
    p = _params(7)
    m.set_params(p)
    rng = np.random.default_rng(n)
    X = rng.uniform(-1, 1, (n, 5)).astype(np.float32)
    tf = m.forward_fused(X)
    tl = m.forward(X)
    m.close()
    ref = mo.forward_scaled(p.astype(np.float64), X)
    s = np.abs(ref).max() + 1
    np.testing.assert_allclose(tf, ref, rtol=0, atol=2e-2 * s)
    assert np.median(np.abs(tf - ref)) < 1e-5 * s
    scale = np.abs(tl).max() + 1
    np.testing.assert_allclose(tf, tl, rtol=0, atol=2e-2 * scale)
    if n >= 1000:
        # a BF16 place of the activations, spread over many rows
        assert np.median(np.abs(tf - tl)) < 1e-3 * scale
    if n <= 1000:
        ref, _ = mo.forward(p.astype(np.float64), X)
        np.testing.assert_allclose(tf, ref, rtol=0, atol=2e-2 * (np.abs(ref).max() + 1))


def test_fused_timing_counts_launches_and_rows(amca):
    """HIP-event totals of the fused kernel: one launch per call, its rows,
    a positive duration; enabling again resets them; off adds nothing."""
    from amca3d_amd.surrogate import MLP
    m = MLP(2)
    X = np.random.default_rng(0).uniform(-1, 1, (5000, 5)).astype(np.float32)
    m.enable_timing(True)
    m.forward_fused(X)
    m.forward_fused(X[:123])
    ms, launches, rows = m.fused_times()
    assert launches == 2 and rows == 5123 and ms > 0.0
    m.enable_timing(True)
    assert m.fused_times() == (0.0, 0, 0)
    m.enable_timing(False)
    m.forward_fused(X)
    assert m.fused_times() == (0.0, 0, 0)
    m.close()


def test_gradients_match_oracle(amca):
    from amca3d_amd.surrogate import MLP
    m = MLP(1)
    p = _params(7)
    m.set_params(p)
    rng = np.random.default_rng(4)
    X = rng.uniform(-1, 1, (777, 5)).astype(np.float32)
    Y = rng.uniform(-1, 1, 777).astype(np.float32)
    g, loss = m.gradients(X, Y)
    ref, rloss = mo.gradients(p.astype(np.float64), X, Y)
    assert loss == pytest.approx(rloss, rel=1e-3)
    off, _ = mo.layout()
    for l, ((w, b), fo) in enumerate(zip(off, mo.FAN_OUT)):
        for lo, hi, what in ((w, b, "W"), (b, b + fo, "b")):
            err = np.linalg.norm(g[lo:hi] - ref[lo:hi])
            assert err <= 2e-2 * np.linalg.norm(ref[lo:hi]) + 1e-7, (l, what, err)
    # and against the exact float64 gradient, to BF16 accuracy
    g64, _ = mo.gradients64(p.astype(np.float64), X, Y)
    assert np.linalg.norm(g - g64) < 5e-2 * np.linalg.norm(g64)
    m.close()
