"""CPU pins of the numpy cost-model reference (tests/mlp_ref.py): the
analytic gradient against central finite differences, BF16 rounding known
answers."""
import numpy as np

import mlp_ref


def test_bf16_rounding_known_answers():
    v = np.array([1.0, 1.0 + 2**-8, 1.0 + 2**-7, 1.0 + 3 * 2**-8, -2.5, 0.0], np.float32)
    r = mlp_ref.bf16(v)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0);
    # 1 + 3*2^-8 is a tie between 1 + 2^-7 and 1 + 2^-6: rounds to even (1 + 2^-6)
    assert list(r) == [1.0, 1.0, 1.0 + 2**-7, 1.0 + 2**-6, -2.5, 0.0]


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = mlp_ref.init(3).astype(np.float64)
    x = rng.normal(size=(40, 8))
    t = rng.normal(size=40) * 0.3
    # float64, no BF16 rounding: the closed form must match the derivative
    loss, g = mlp_ref.loss_and_grad(p, x, t, dtype=np.float64, round_bf16=False)
    eps = 1e-6
    for k in rng.choice(len(p), 120, replace=False):
        pp = p.copy(); pp[k] += eps
        pm = p.copy(); pm[k] -= eps
        lp, _ = mlp_ref.loss_and_grad(pp, x, t, dtype=np.float64, round_bf16=False)
        lm, _ = mlp_ref.loss_and_grad(pm, x, t, dtype=np.float64, round_bf16=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_loss_decreases_along_negative_gradient():
    rng = np.random.default_rng(1)
    p = mlp_ref.init(4).astype(np.float64)
    x = rng.normal(size=(200, 8))
    t = np.tanh(x[:, 0] - 0.5 * x[:, 5])
    l0, g = mlp_ref.loss_and_grad(p, x, t, dtype=np.float64, round_bf16=False)
    l1, _ = mlp_ref.loss_and_grad(p - 1e-2 * g, x, t, dtype=np.float64, round_bf16=False)
    assert l1 < l0


def test_gate_rows_carry_no_gradient():
    """Rows whose history is unambiguous (feature 5 == feature 6 != 0) output
    feature 5 exactly and add nothing to the gradient; rows with ambiguous or
    missing history use the network (finite differences on a mixed batch)."""
    rng = np.random.default_rng(2)
    p = mlp_ref.init(5).astype(np.float64)
    x = rng.normal(size=(60, 8))
    x[:20, 6] = x[:20, 5]            # unambiguous history
    x[20:30, 5] = 0.0                # no history
    x[20:30, 6] = 0.0
    t = rng.normal(size=60) * 0.3
    y, _ = mlp_ref.forward(p, x, dtype=np.float64, round_bf16=False)
    assert np.array_equal(y[:20], x[:20, 5])
    _, g_all = mlp_ref.loss_and_grad(p, x, t, dtype=np.float64, round_bf16=False)
    _, g_rest = mlp_ref.loss_and_grad(p, x[20:], t[20:], dtype=np.float64, round_bf16=False)
    # the mean over 60 rows = (40 / 60) x the mean over the 40 gated rows
    assert np.allclose(g_all, g_rest * 40 / 60, rtol=1e-12, atol=1e-15)
    loss, g = mlp_ref.loss_and_grad(p, x, t, dtype=np.float64, round_bf16=False)
    eps = 1e-6
    for k in rng.choice(len(p), 60, replace=False):
        pp = p.copy(); pp[k] += eps
        pm = p.copy(); pm[k] -= eps
        lp, _ = mlp_ref.loss_and_grad(pp, x, t, dtype=np.float64, round_bf16=False)
        lm, _ = mlp_ref.loss_and_grad(pm, x, t, dtype=np.float64, round_bf16=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
