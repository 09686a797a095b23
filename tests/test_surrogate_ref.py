"""Pins of the numpy surrogate reference (the checker of the GPU training and inference kernels):
the gradient formula against central finite differences of the unrounded network, and the BF16
rounding against known bit patterns.  CPU only."""
import numpy as np

import surrogate_ref as R
from sdfibm_amd.surrogate import N_PARAMS, init_params


def test_bf16_rounding_bits():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1.0 + 2 ** -9 + 2 ** -20], dtype=np.float32)
    # 1+2^-8 is a tie between 1 and 1+2^-7: to even (1); 1+3*2^-8 ties to 1+2^-6 (even mantissa)
    np.testing.assert_array_equal(R.bf16(v), [1.0, 1.0, 1.0 + 2 ** -6, -2.5, 1.0])


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = init_params(3).astype(np.float64)
    X = rng.normal(0, 1, (40, 16))
    y = rng.uniform(0, 1, 40)
    g, _ = R.grad(p, X, y, rounding=False)
    assert g.size == N_PARAMS
    idx = rng.choice(N_PARAMS, 60, replace=False)
    h = 1e-6
    for k in list(idx) + [N_PARAMS - 1, 0]:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        lp = 0.5 * np.sum((R.forward(pp, X, False)[0] - y) ** 2)
        lm = 0.5 * np.sum((R.forward(pm, X, False)[0] - y) ** 2)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, g[k])
