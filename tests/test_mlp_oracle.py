"""Pins for the FP64 MLP oracle (no GPU): the analytic gradient against central
finite differences, GELU against its integral definition, BF16 rounding."""
import numpy as np
import pytest
from scipy.integrate import quad


@pytest.fixture(scope="module")
def mo():
    import mlp_oracle
    return mlp_oracle


def test_gradient_matches_finite_differences(mo):
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.15, mo.NPAR)
    x = rng.uniform(-1, 1, (7, 4))
    t = rng.normal(0, 1, (7, 3))
    loss, g = mo.loss_grad(w, x, t)
    idx = np.concatenate([rng.choice(np.arange(o, o + int(np.prod(s))), min(6, int(np.prod(s))), replace=False)
                          for o, s in mo.OFF.values()])
    for q in idx:
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[q] += h
        wm[q] -= h
        fd = (mo.loss_grad(wp, x, t)[0] - mo.loss_grad(wm, x, t)[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-7 * max(1.0, abs(fd)), q


def test_gelu_and_derivative(mo):
    for z in (-3.0, -0.7, 0.0, 0.4, 2.5):
        phi = quad(lambda s: np.exp(-s * s / 2) / np.sqrt(2 * np.pi), -np.inf, z)[0]
        assert abs(mo.gelu(np.array(z)) - z * phi) < 1e-12
        h = 1e-6
        fd = (mo.gelu(np.array(z + h)) - mo.gelu(np.array(z - h))) / (2 * h)
        assert abs(mo.gelu_d(np.array(z)) - fd) < 1e-8


def test_bf16_rounding(mo):
    assert mo.bf16(np.array([1.0]))[0] == 1.0
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: round to even (1.0)
    assert mo.bf16(np.array([1.0 + 2 ** -8]))[0] == 1.0
    assert mo.bf16(np.array([1.0 + 3 * 2 ** -8]))[0] == 1.0 + 2 ** -6
    assert mo.bf16(np.array([-3.14159]))[0] == -3.140625
