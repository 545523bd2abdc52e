"""The float64 back-propagation reference of the frequency-map network
(tests/mlp_ref.py) against central finite differences of its own loss: every
bias, every W1 and W4 entry and 300 random W2/W3 entries, relative error
<= 1e-6.  This pins the reference that the GPU gradients are checked against."""
import numpy as np

import mlp_ref as R


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(41)
    p = rng.normal(0, 0.3, R.NP)
    X = rng.normal(size=(16, 4))
    T = rng.normal(size=(16, 3))
    loss, g = R.loss_and_grad(p, X, T)
    idx = list(range(R.OFF["W1"][0], R.OFF["W1"][0] + R.OFF["W1"][1]))
    for name in ("b1", "b2", "b3", "W4", "b4"):
        o, n = R.OFF[name]
        idx += list(range(o, o + n))
    for name in ("W2", "W3"):
        o, n = R.OFF[name]
        idx += list(o + rng.choice(n, 150, replace=False))
    worst = 0.0
    for k in idx:
        h = 1e-6 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (R.loss_and_grad(pp, X, T)[0] - R.loss_and_grad(pm, X, T)[0]) / (2 * h)
        worst = max(worst, abs(fd - g[k]) / max(abs(g[k]), 1e-3 * np.max(np.abs(g))))
    assert worst <= 1e-6, worst


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5, 3.14159265])
    got = R.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 (the even neighbour, 1, wins);
    # 1 + 3 2^-9 rounds up, 1 + 2^-9 down
    np.testing.assert_array_equal(got[:4], [1.0, 1.0, 1.0 + 2 ** -7, 1.0])
    assert got[4] == -2.5
    assert abs(got[5] - 3.140625) == 0.0
