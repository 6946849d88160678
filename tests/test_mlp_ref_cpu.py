"""Pin the CPU restatement of the cost model's backprop (oracle/mlp_ref.py)
against central finite differences in float64, and its bf16 rounding against
the definition of round-to-nearest-even."""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref  # noqa: E402


@pytest.mark.parametrize("dims", [(3, 4, 5), (7, 16, 8)])
def test_gradient_matches_finite_differences(dims):
    n_in, h1, h2 = dims
    rng = np.random.default_rng(1)
    p = rng.normal(size=mlp_ref.n_params(n_in, h1, h2)) * 0.5
    X = rng.normal(size=(11, n_in))
    y = rng.normal(size=11)
    loss, g = mlp_ref.loss_and_grad(p, X, y, n_in, h1, h2, bf16=False)
    eps = 1e-6
    for i in rng.choice(len(p), size=min(len(p), 40), replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (mlp_ref.loss_and_grad(pp, X, y, n_in, h1, h2, bf16=False)[0] -
              mlp_ref.loss_and_grad(pm, X, y, n_in, h1, h2, bf16=False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0, 3.0e38], dtype=np.float32)
    r = mlp_ref.bf16_round(x)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even (1.0)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 2 ** -6)
    assert r[3] == -2.5 and r[4] == 0.0
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)
