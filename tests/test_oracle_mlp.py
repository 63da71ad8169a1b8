"""CPU: pinning the cost-model checker (oracle/mlp_ref.py) — BF16 rounding
against its definition, and the analytic gradient against central finite
differences in float64."""
import numpy as np

from oracle import mlp_ref as M


def test_bf16_round_to_nearest_even():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)  # BF16 spacing at 1.0
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -20, -2.5, 0.0, 3.0e38],
                    np.float32)
    got = M.bf16(vals)
    assert got[0] == one
    assert got[1] == one                      # tie -> even mantissa (1.0)
    assert got[2] == one + 2 * ulp            # tie -> even (1 + 2 ulp)
    assert got[3] == one + ulp                # above the tie -> up
    assert got[4] == np.float32(-2.5) and got[5] == 0.0
    assert np.isfinite(got[6])
    # idempotent, and the low 16 bits are zero
    assert np.array_equal(M.bf16(got), got)
    assert np.all((got.view(np.uint32) & 0xFFFF) == 0)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    d, n = 9, 13
    p = rng.normal(0, 0.3, M.n_params(d))
    x = rng.normal(0, 1, (n, d))
    t = rng.normal(0, 1, n)
    loss, g = M.loss_and_grad(p, x, t, quantize=False, dtype=np.float64)
    idx = rng.choice(len(p), size=60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = M.loss_and_grad(p + e, x, t, quantize=False, dtype=np.float64)
        lm, _ = M.loss_and_grad(p - e, x, t, quantize=False, dtype=np.float64)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_planted_gradient_mistake_is_detected():
    # a wrong sign on the b2 gradient must fail the finite-difference check
    rng = np.random.default_rng(1)
    d = 5
    p = rng.normal(0, 0.3, M.n_params(d))
    x, t = rng.normal(0, 1, (7, d)), rng.normal(0, 1, 7)
    _, g = M.loss_and_grad(p, x, t, quantize=False, dtype=np.float64)
    i = M.H * d + M.H + M.H * M.H + 3  # an entry of b2
    e = np.zeros_like(p)
    e[i] = 1e-6
    fd = (M.loss_and_grad(p + e, x, t, False, np.float64)[0] - M.loss_and_grad(p - e, x, t, False, np.float64)[0]) / 2e-6
    assert abs(fd) > 1e-6 and abs(fd - (-g[i])) > 1e-6 + 1e-5 * abs(fd)


def test_adam_step_definition():
    p, g = np.array([1.0, -2.0]), np.array([0.5, -0.25])
    p1, m, v = M.adam(p, g, np.zeros(2), np.zeros(2), 1, 0.1)
    # first step: m_hat = g, v_hat = g^2, so every parameter moves by lr * sign(g)
    assert np.allclose(p1, p - 0.1 * np.sign(g), atol=1e-6)
