"""The pre-sizing MLP's HIP kernels against the float64 CPU oracle
(oracle/mlp_ref.py): forward, gradient (pinned by finite differences first),
BF16 matrix-core inference and its MFMA operand layout, training."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_ref as R

N_PARAMS = 256 * 128 + 128 + 128 * 128 + 128 + 128 * 2 + 2


def _params(seed=0):
    rng = np.random.default_rng(seed)
    w = np.concatenate([rng.normal(0, (2.0 / 256) ** 0.5, 256 * 128), rng.normal(0, 0.05, 128),
                        rng.normal(0, (2.0 / 128) ** 0.5, 128 * 128), rng.normal(0, 0.05, 128),
                        rng.normal(0, (2.0 / 128) ** 0.5, 128 * 2), rng.normal(0, 0.05, 2)])
    return w.astype(np.float32)


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    w = _params(1).astype(np.float64)
    x = rng.normal(size=(16, 256))
    t = rng.normal(size=(16, 2))
    _, g = R.loss_grad(w, x, t)
    idx = np.concatenate([rng.integers(0, 256 * 128, 12), [256 * 128 + 3], 256 * 128 + 128 + rng.integers(0, 16384, 8),
                          [N_PARAMS - 1, N_PARAMS - 3, N_PARAMS - 100]])
    for i in idx:
        e = np.zeros_like(w)
        e[i] = 1e-6
        fd = (R.loss(w + e, x, t) - R.loss(w - e, x, t)) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(g[i]), (i, fd, g[i])


def test_reference_gradient_catches_a_dropped_term():
    """A backward pass that forgets the ReLU mask disagrees with the finite
    differences: the check above has teeth."""
    rng = np.random.default_rng(2)
    w = _params(2).astype(np.float64)
    x = rng.normal(size=(8, 256))
    t = rng.normal(size=(8, 2))
    W1, b1, W2, b2, W3, b3 = R.split(w)
    y, (xx, z1, a1, z2, a2) = R.forward(w, x)
    dy = 2.0 * (y - t) / y.size
    d2_wrong = dy @ W3.T                      # mask dropped
    gb2_wrong = d2_wrong.sum(0)
    i = 256 * 128 + 128 + 128 * 128           # first entry of b2
    e = np.zeros_like(w)
    e[i:i + 128] = 0
    diffs = []
    for j in range(128):
        e[:] = 0
        e[i + j] = 1e-6
        fd = (R.loss(w + e, x, t) - R.loss(w - e, x, t)) / 2e-6
        diffs.append(abs(fd - gb2_wrong[j]))
    assert max(diffs) > 1e-4


@pytest.mark.gpu
def test_hip_forward_and_gradient_match_oracle(gpu):
    from gkc_amd import learned
    rng = np.random.default_rng(3)
    m = learned.MLP(seed=5)
    w = _params(3)
    m.set_params(w)
    x = rng.normal(size=(300, 256)).astype(np.float32)
    t = rng.normal(size=(300, 2)).astype(np.float32)
    print("[gpu-case] mlp forward/grad", flush=True)
    y = m.forward(x)
    yr, _ = R.forward(w.astype(np.float64), x)
    np.testing.assert_allclose(y, yr, rtol=1e-4, atol=1e-4)
    L, g = m.loss_grad(x, t)
    Lr, gr = R.loss_grad(w.astype(np.float64), x, t)
    assert abs(L - Lr) <= 1e-4 * max(1.0, abs(Lr))
    np.testing.assert_allclose(g, gr, rtol=2e-3, atol=2e-6)
    m.close()


@pytest.mark.gpu
def test_mfma_tile_layout_exact(gpu):
    """Small-integer matrices are exact in bf16 and their products exact in
    fp32, so the matrix-core tile must reproduce numpy bit for bit."""
    import ctypes as C
    from gkc_amd import capi
    rng = np.random.default_rng(4)
    for K, N in [(16, 32), (128, 128), (256, 128), (128, 2)]:
        A = rng.integers(-4, 5, size=(32, K)).astype(np.float32)
        B = rng.integers(-4, 5, size=(K, N)).astype(np.float32)
        Cm = np.zeros((32, N), np.float32)
        print("[gpu-case] mfma tile", K, N, flush=True)
        f = C.POINTER(C.c_float)
        rc = capi.lib().gkc_bf16_matmul32(0, A.ctypes.data_as(f), B.ctypes.data_as(f), K, N, Cm.ctypes.data_as(f))
        assert rc == 0
        np.testing.assert_array_equal(Cm, A @ B)


@pytest.mark.gpu
def test_bf16_inference_close_to_fp32(gpu):
    from gkc_amd import learned
    rng = np.random.default_rng(5)
    m = learned.MLP(seed=6)
    m.set_params(_params(6))
    x = rng.normal(size=(1000, 256)).astype(np.float32)
    print("[gpu-case] mlp bf16", flush=True)
    y32 = m.forward(x)
    y16 = m.forward_bf16(x)
    # bf16 keeps 8 significant bits: relative error of the outputs ~1e-2
    scale = np.abs(y32).mean()
    assert np.abs(y16 - y32).mean() <= 0.02 * scale + 1e-3
    assert np.abs(y16 - y32).max() <= 0.15 * scale + 1e-2
    m.close()


@pytest.mark.gpu
def test_training_fits_a_known_function(gpu):
    from gkc_amd import learned
    rng = np.random.default_rng(7)
    x = rng.normal(size=(4096, 256)).astype(np.float32)
    A = rng.normal(size=(256, 2)).astype(np.float32) / 16
    t = np.tanh(x @ A).astype(np.float32)
    m = learned.MLP(seed=8)
    print("[gpu-case] mlp train", flush=True)
    losses = m.train(x, t, epochs=30, batch=256, lr=1e-3)
    assert losses[-1] < 0.1 * losses[0]
    m.close()
