"""The FP64 numpy reference of the warm-start MLP against finite
differences (CPU)."""
import numpy as np

import mlp_ref as M


def test_param_count_matches_header():
    from shapes_amd import engine
    assert M.NPARAM == engine.MLP_NPARAM == 35202


def test_gradient_matches_finite_differences():
    p = M.init_params(3).astype(np.float64)
    rng = np.random.default_rng(0)
    X = rng.normal(size=(29, 14))
    T = rng.normal(size=(29, 2))
    g, _ = M.gradient(p, X, T, emulate_bf16=False)
    eps = 1e-6
    for i in rng.choice(M.NPARAM, 40, replace=False).tolist() + [0, 35201, 35200]:
        pp = p.copy()
        pp[i] += eps
        pm = p.copy()
        pm[i] -= eps
        fd = (M.gradient(pp, X, T, False)[1] - M.gradient(pm, X, T, False)[1]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), i


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38], dtype=np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0          # tie to even
    assert r[2] == 1.0 + 2 ** -6                  # tie to even (up)
    assert r[3] == -2.5
