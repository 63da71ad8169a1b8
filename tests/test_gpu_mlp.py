"""GPU: the cost-model kernels (forward, backward, Adam) against the CPU FP32
checker in oracle/mlp_ref.py, and training that fits a known function.

Tolerances: the kernels and the checker use the same BF16-rounded weights
and FP32 sums in a different order, so values agree to ~1e-5 relative;
gradients are summed with FP32 atomics in arbitrary order (1e-4 relative
plus a small absolute floor)."""
import numpy as np
import pytest

from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(gpu):
    from vsgm_amd.hip import DeviceContext

    c = DeviceContext(0)
    yield c
    c.close()


@pytest.mark.parametrize("mfma", [True, False])
@pytest.mark.parametrize("d,n", [(1, 1), (17, 5), (100, 1000), (128, 333)])
def test_forward_matches_checker(ctx, d, n, mfma):
    from vsgm_amd.mlp import CostModel

    m = CostModel(d, seed=d, ctx=ctx)
    m.use_mfma(mfma)
    rng = np.random.default_rng(d)
    p = m.get_params()
    assert p.size == M.n_params(d)
    p[M.H * d:M.H * d + M.H] = rng.normal(0, 0.1, M.H)  # non-zero biases
    p[-1] = 0.25
    m.set_params(p)
    assert np.array_equal(m.get_params(), p)
    x = rng.normal(0, 1, (n, d)).astype(np.float32)
    y = m.forward(x)
    want, _ = M.forward(p, x, mfma=mfma)
    assert np.allclose(y, want, rtol=1e-5, atol=1e-5), np.max(np.abs(y - want))
    m.close()


@pytest.mark.parametrize("mfma", [True, False])
@pytest.mark.parametrize("d,n", [(8, 3), (100, 777), (128, 4096)])
def test_backward_matches_checker(ctx, d, n, mfma):
    from vsgm_amd.mlp import CostModel

    m = CostModel(d, seed=3 + d, ctx=ctx)
    m.use_mfma(mfma)
    rng = np.random.default_rng(5)
    p = m.get_params()
    p[M.H * d:M.H * d + M.H] = rng.normal(0, 0.2, M.H)
    m.set_params(p)
    x = rng.normal(0, 1, (n, d)).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    loss, g = m.backward(x, t)
    want_loss, want_g = M.loss_and_grad(p, x, t, mfma=mfma)
    assert abs(loss - want_loss) <= 1e-5 * abs(want_loss) + 1e-6
    scale = np.max(np.abs(want_g))
    assert np.allclose(g, want_g, rtol=1e-4, atol=1e-5 * scale), np.max(np.abs(g - want_g))
    m.close()


def test_adam_step_matches_checker(ctx):
    from vsgm_amd import mlp as mlpmod
    from vsgm_amd.mlp import CostModel

    d, n = 20, 300
    m = CostModel(d, seed=9, ctx=ctx)
    rng = np.random.default_rng(7)
    x = rng.normal(0, 1, (n, d)).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    p = m.get_params().astype(np.float64)
    mm, vv = np.zeros_like(p), np.zeros_like(p)
    for step in (1, 2, 3):
        _, g = m.backward(x, t)
        ctx.check(mlpmod._lib().vsgm_mlp_adam_step(m.h, 1e-2, 0.9, 0.999, 1e-8), "adam")
        p, mm, vv = M.adam(p, g.astype(np.float64), mm, vv, step, 1e-2)
        got = m.get_params()
        assert np.allclose(got, p, rtol=1e-5, atol=1e-6), (step, np.max(np.abs(got - p)))
        p = got.astype(np.float64)  # follow the device's FP32 state
    m.close()


def test_training_fits_a_smooth_function(ctx):
    from vsgm_amd.mlp import CostModel

    d, n = 12, 4000
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, (n, d)).astype(np.float32)
    t = (np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2] + 0.5 * np.abs(x[:, 3])).astype(np.float32)
    m = CostModel(d, seed=1, ctx=ctx)
    before = float(np.mean((m.forward(x) - t) ** 2))
    losses = m.train(x, t, epochs=40, batch=128, lr=3e-3, seed=2)
    after = float(np.mean((m.forward(x) - t) ** 2))
    assert losses[-1] < losses[0]
    assert after < 0.1 * before and after < 0.02, (before, after)
    m.close()
