"""The learned part on the CPU: the numpy FP64 oracle of the MLP is pinned
by central finite differences, and the library's per-vertex features are
checked on meshes whose optical depths are known in closed form."""
import numpy as np
import pytest

import transfer_amd as ta
from oracle import mlp


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = mlp.init_params(3) + 0.05 * rng.standard_normal(mlp.NPARAM)
    x = rng.standard_normal((40, mlp.F))
    t = rng.uniform(0.05, 0.95, 40)
    mu, sd = np.zeros(mlp.F), np.ones(mlp.F)
    _, g = mlp.loss_and_grad(p, x, t, mu, sd, round_weights=False)
    idx = np.concatenate([rng.choice(mlp.NPARAM, 60, replace=False), [mlp.NPARAM - 1]])
    h = 1e-6
    for i in idx:
        e = np.zeros(mlp.NPARAM)
        e[i] = h
        lp, _ = mlp.loss_and_grad(p + e, x, t, mu, sd, round_weights=False)
        lm, _ = mlp.loss_and_grad(p - e, x, t, mu, sd, round_weights=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-5], np.float32)
    r = mlp.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
    assert abs(r[4] - 3.0e-5) <= 3.0e-5 * 2 ** -8


def test_features_optical_depth_and_contrast():
    n = 8
    m = ta.Mesh.box((n, n, n), (0, 0, 0), (1, 1, 1))
    kappa = 3.0
    m.set_fields(kappa=m.field("const", [kappa]), Ip=m.field("const", [1.0]))
    f = ta.mlp_features(m, 0.5)
    X = m.geometry()[0]
    bnd = np.any((X < 1e-12) | (X > 1 - 1e-12), axis=1)
    tau = np.exp(f[:, 3]) - 1e-3
    assert np.all(np.abs(tau[bnd]) < 1e-6)
    # the centre is 4 axis edges from the boundary: tau = kappa * 0.5 exactly
    c = np.argmin(np.linalg.norm(X - 0.5, axis=1))
    assert tau[c] == pytest.approx(kappa * 0.5, rel=1e-5)
    np.testing.assert_allclose(f[:, 0], np.log(kappa), rtol=1e-6)
    np.testing.assert_allclose(f[:, 6], 0.0, atol=1e-6)           # uniform field: no contrast
    np.testing.assert_allclose(f[:, 1], 0.5)
    np.testing.assert_allclose(f[:, 5], np.exp(-tau * np.sqrt(1.5)), rtol=1e-5, atol=1e-7)
