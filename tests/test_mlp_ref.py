"""Pins of the numpy cost-model reference (CPU): its analytic gradient against
central finite differences, softmax masking and the parameter count the C ABI
reports."""
import numpy as np
import pytest

from oracle import mlp_ref


def _problem(rows=7, seed=0):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, mlp_ref.NPARAM)
    x = rng.normal(0, 1, (rows, mlp_ref.F))
    mask = np.zeros((rows, mlp_ref.K))
    for r in range(rows):
        lo = 5 * (r % 3)
        mask[r, lo:lo + 5] = 1
    t = rng.random((rows, mlp_ref.K)) * mask
    t /= t.sum(1, keepdims=True)
    return p, x, t, mask


def test_param_count_matches_library(edpp):
    assert edpp.lib().edpp_mlp_num_params() == mlp_ref.NPARAM == 5967


def test_gradient_matches_finite_differences():
    p, x, t, mask = _problem()
    g = mlp_ref.grad(p, x, t, mask)
    rng = np.random.default_rng(1)
    for i in rng.choice(mlp_ref.NPARAM, 60, replace=False):
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (mlp_ref.loss(p + e, x, t, mask) - mlp_ref.loss(p - e, x, t, mask)) / 2e-6
        assert fd == pytest.approx(g[i], rel=1e-4, abs=1e-8), i


def test_masked_softmax():
    p, x, t, mask = _problem()
    prob, _ = mlp_ref.forward(p, x, mask)
    assert np.allclose(prob.sum(1), 1.0) and np.all(prob[mask == 0] == 0)
    # a mask with one candidate gives probability 1 and zero loss for that target
    m1 = np.zeros_like(mask)
    m1[:, 3] = 1
    t1 = m1.copy()
    assert mlp_ref.loss(p, x, t1, m1) == pytest.approx(0.0, abs=1e-12)
