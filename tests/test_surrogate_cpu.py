"""CPU pins of the surrogate's FP64 reference (oracle/mlp_ref.py): analytic gradients against
central finite differences, periodicity of the features, and the feature-vector definition shared
with the product module."""
import numpy as np

import mlp_ref as R
from ephip import surrogate as S


def test_fourier_vectors_shared_definition():
    assert np.array_equal(R.fourier_vectors(32), S.fourier_vectors(32))
    m = R.fourier_vectors(32)
    assert len({tuple(x) for x in m} | {tuple(-x) for x in m}) == 64  # no +-m duplicates
    assert (m ** 2).sum(1).max() <= 6


def test_features_periodic():
    m = R.fourier_vectors(32)
    rng = np.random.default_rng(0)
    x = rng.random((5, 3))
    assert np.allclose(R.features(x, m), R.features(x + np.array([1, -2, 3]), m), atol=1e-12)


def test_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    m = R.fourier_vectors(8)
    p = R.init_params(rng, 16, 12, 5)
    for k in p:
        p[k] = p[k] + 0.1 * rng.standard_normal(p[k].shape)
    x = rng.random((7, 3))
    t = rng.standard_normal((7, 5))
    _, g = R.loss_and_grad(p, x, t, m)
    h = 1e-6
    for k in p:
        flat = p[k].reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp, _ = R.loss_and_grad(p, x, t, m)
            flat[idx] = old - h
            lm, _ = R.loss_and_grad(p, x, t, m)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[k].reshape(-1)[idx]) <= 1e-7 + 1e-6 * abs(fd), (k, idx)
