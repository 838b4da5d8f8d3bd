"""Pins of the FP64 denoiser oracle (oracle/denoiser_ref.py): parameter
count, the convolution against a direct loop, and every gradient against
central finite differences."""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import denoiser_ref as D  # noqa: E402


def test_parameter_count():
    assert D.num_params() == 43270


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(1, 5, 5, 5, 3))
    W = rng.normal(size=(3, 3, 3, 3, 2))
    b = rng.normal(size=2)
    y = D._conv(x, W, b)
    for (z, yy, xx) in [(0, 0, 0), (2, 1, 0), (1, 2, 2)]:
        ref = b.copy()
        for kz in range(3):
            for ky in range(3):
                for kx in range(3):
                    ref += x[0, z + kz, yy + ky, xx + kx] @ W[kz, ky, kx]
        assert np.allclose(y[0, z, yy, xx], ref)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    p = D.init_params(3)
    p["W3"] *= 10.0  # make the head matter
    for k in ("b1", "b2", "b3"):
        p[k] = rng.normal(scale=0.05, size=p[k].shape)
    X = rng.uniform(-1, 1, size=(2, 10, 10, 10, 12))
    T = rng.uniform(-1, 1, size=(2, 8, 8, 8, 6))
    M = (rng.uniform(size=T.shape) > 0.3).astype(np.float64)
    L, g = D.loss_and_grads(p, X, T, M)
    eps = 1e-6
    for name in D.ORDER:
        flat = p[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            Lp, _ = D.loss_and_grads(p, X, T, M)
            flat[idx] = old - eps
            Lm, _ = D.loss_and_grads(p, X, T, M)
            flat[idx] = old
            fd = (Lp - Lm) / (2 * eps)
            an = g[name].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-6 + 1e-4 * abs(fd), (name, idx, fd, an)
