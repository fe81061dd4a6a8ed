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


def test_cold_pressure_is_exact_at_rest_and_a_lower_bound():
    """p0 of the surrogate's input: the c2p root with W frozen at its p = 0 value.
    At rest (m = 0) W = 1 for every p, so p0 is the exact pressure; in motion
    W falls as p grows, so p0 <= p, and p/p0 -> 1 as the flow slows down."""
    import oracle as O

    rng = np.random.default_rng(11)
    gamma = 5.0 / 3.0
    Gg = gamma / (gamma - 1.0)
    n = 4000
    rho = 10 ** rng.uniform(-3, 2, n)
    p = rho * 10 ** rng.uniform(-5, 2, n)
    W = np.column_stack([rho, np.zeros((n, 3)), p])
    _, U = O.p2c(W, gamma=gamma)
    p0 = R.cold_pressure(Gg, U[:, 0], np.zeros(n), U[:, 4])
    assert np.allclose(p0, p, rtol=1e-9)
    lor = np.exp(rng.uniform(0, np.log(50), n))
    lor[: n // 10] = 1.0 + 1e-5 * rng.uniform(size=n // 10)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    W = np.column_stack([rho, d * np.sqrt(1 - 1 / lor**2)[:, None], p])
    _, U = O.p2c(W, gamma=gamma)
    p0 = R.cold_pressure(Gg, U[:, 0], np.linalg.norm(U[:, 1:4], axis=1), U[:, 4])
    r = p / p0
    assert (r >= 1 - 1e-6).all() and (r <= 5.0).all()
    slow = lor < 1.00001
    assert slow.sum() >= n // 10 and np.allclose(r[slow], 1.0, rtol=1e-3)
