"""The CPU restatement of the surrogate MLP: analytic gradients against
central finite differences (float64, no bf16 rounding)."""
import numpy as np

from oracle import mlp as R


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    P = rng.normal(scale=0.3, size=R.N_PARAMS)
    X = rng.normal(size=(64, 3))
    y = rng.normal(size=64)
    ok = (rng.uniform(size=64) > 0.1).astype(float)
    _, g = R.loss_and_grad(P, X, y, ok, bf=False)
    n = X.shape[0]
    idx = np.concatenate([rng.choice(np.arange(R.OFF_W1, R.OFF_B1), 8), rng.choice(np.arange(R.OFF_B1, R.OFF_W2), 4),
                          rng.choice(np.arange(R.OFF_W2, R.OFF_B2), 16), rng.choice(np.arange(R.OFF_B2, R.OFF_W3), 4),
                          rng.choice(np.arange(R.OFF_W3, R.OFF_B3), 4), [R.OFF_B3]])
    for i in idx:
        e = np.zeros_like(P)
        e[i] = 1e-6
        lp, _ = R.loss_and_grad(P + e, X, y, ok, bf=False)
        lm, _ = R.loss_and_grad(P - e, X, y, ok, bf=False)
        fd = (lp - lm) / 2e-6 / n
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.0078125, 1.01171875, -3.14159265], np.float32)
    b = R.bf16(x)
    assert b[0] == 1.0 and b[1] == 1.0  # tie below rounds to even
    assert b[2] == 1.0078125 and b[3] == 1.015625  # tie above rounds to even
    assert abs(b[4] + 3.140625) < 1e-12
