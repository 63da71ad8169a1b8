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


class _Fixed:
    """A cost model stand-in returning fixed predictions (host-side test of
    the margin rule; the network itself is covered by the GPU tests)."""

    def __init__(self, pred):
        self.pred = np.asarray(pred, np.float32)

    def forward(self, x):
        return self.pred[: len(x)]


def test_margin_calibration_keeps_held_out_decisions_from_regressing():
    from vsgm_amd import learned as L

    # three groups of three orders; labels: log(time / heuristic time), the heuristic order has 0
    y = np.array([0.0, -0.7, 0.4,    # group 0: order 1 really faster
                  0.0, 0.5, 0.9,     # group 1: every other order slower
                  0.0, -0.2, 0.3], np.float32)
    group = np.repeat([0, 1, 2], 3)
    held = {"x": np.zeros((9, L.IN_DIM), np.float32), "y": y, "group": group}
    # predictions: group 0 predicts order 1 best by 0.6; group 1 (wrongly) order 1 best by 0.25;
    # group 2 predicts order 1 best by 0.1
    pred = np.array([0.0, -0.6, 0.2, 0.0, -0.25, 0.3, 0.0, -0.1, 0.5], np.float32)
    dec = L.decisions(pred, held)
    assert [round(p, 3) for p, _ in dec] == [0.6, 0.25, 0.1]
    assert [round(r, 3) for _, r in dec] == [0.7, -0.5, 0.2]
    margin, table = L.calibrate_margin(_Fixed(pred), held)
    # below 0.25 the wrong group-1 decision is taken; at 0.3 and 0.5 only group 0's (gain 0.7)
    assert margin == 0.3
    assert {t["margin"]: t["regressions"] for t in table}[0.2] == 1
    # when every finite margin lets a regression through, the heuristic plan is always kept
    bad = np.array([0.0, 5.0, 5.0, 0.0, -3.0, 0.0, 0.0, 5.0, 5.0], np.float32)
    assert L.calibrate_margin(_Fixed(bad), held)[0] == np.inf


def test_split_groups_keeps_groups_whole():
    from vsgm_amd import learned as L

    data = {"x": np.zeros((20, 3), np.float32), "y": np.arange(20, dtype=np.float32), "group": np.repeat(np.arange(5), 4)}
    tr, ho = L.split_groups(data, holdout=0.4, seed=1)
    assert len(tr["y"]) + len(ho["y"]) == 20
    assert not set(tr["group"].tolist()) & set(ho["group"].tolist())
    assert len(set(ho["group"].tolist())) == 2


def test_matrix_core_mode_stays_close_to_the_fp32_activation_model():
    # the matrix-core restatement rounds activations and deltas to BF16 (a
    # relative error of at most 2^-9 each); its outputs and gradient stay
    # within a few percent of the FP32-activation model and are not equal
    rng = np.random.default_rng(4)
    d, n = 40, 500
    p = (rng.normal(0, 1, M.n_params(d)) * 0.2).astype(np.float32)
    x = rng.normal(0, 1, (n, d)).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    y0, _ = M.forward(p, x)
    y1, _ = M.forward(p, x, mfma=True)
    assert np.max(np.abs(y1 - y0)) <= 0.03 * np.max(np.abs(y0)) and not np.array_equal(y0, y1)
    l0, g0 = M.loss_and_grad(p, x, t)
    l1, g1 = M.loss_and_grad(p, x, t, mfma=True)
    assert abs(l1 - l0) <= 0.03 * l0
    assert np.max(np.abs(g1 - g0)) <= 0.05 * np.max(np.abs(g0)) and not np.array_equal(g0, g1)


def test_sample_roots_is_seeded_sorted_and_sized():
    from vsgm_amd import learned as L

    a = L.sample_roots(1 << 16, seed=3)
    assert a.dtype == np.uint32 and len(a) == (1 << 16) // 64
    assert np.all(np.diff(a.astype(np.int64)) > 0) and a.max() < (1 << 16)
    assert np.array_equal(a, L.sample_roots(1 << 16, seed=3))
    assert not np.array_equal(a, L.sample_roots(1 << 16, seed=4))
    assert len(L.sample_roots(500, seed=1)) == 500  # fewer vertices than the minimum: all of them
    assert len(L.sample_roots(1 << 12, seed=1, fraction=8, minimum=16)) == 512
