"""The FP64 MLP oracle (oracle/mlp_oracle.py), pinned by central finite
differences of its own loss, and its bfloat16 rounding by exact cases."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def mlo():
    import mlp_oracle
    return mlp_oracle


def _params(mlo, rng):
    p = np.zeros(mlo.NTOTAL)
    o = 0
    for name, shape in mlo.SHAPES:
        k = int(np.prod(shape))
        fan = shape[1] if len(shape) == 2 else 1
        p[o:o + k] = rng.normal(0, 1.0 / np.sqrt(fan), k) if len(shape) == 2 else rng.normal(0, 0.1, k)
        o += k
    p[o:o + 16] = rng.normal(0, 1, 16)
    p[o + 16:o + 32] = rng.uniform(0.5, 2.0, 16)
    p[o + 32:o + 39] = rng.uniform(0.5, 2.0, 7)
    return p


def test_gradient_matches_finite_differences(mlo):
    rng = np.random.default_rng(0)
    p = _params(mlo, rng)
    X = rng.normal(0, 1, (9, 16))
    Y = rng.normal(0, 1, (9, 7))
    loss, g = mlo.loss_and_grad(p, X, Y)
    idx = np.concatenate([rng.choice(mlo.NPARAM, 60, replace=False),
                          [0, 2047, 2048, 2175, 2176, 18559, 35135, 35136, 36095, mlo.NPARAM - 1]])
    for i in idx:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlo.loss_and_grad(pp, X, Y)[0] - mlo.loss_and_grad(pm, X, Y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(fd)), (i, fd, g[i])


def test_forward_matches_loss(mlo):
    rng = np.random.default_rng(1)
    p = _params(mlo, rng)
    X = rng.normal(0, 1, (5, 16))
    y = mlo.forward(p, X)
    q = mlo.unpack(p)
    # loss at Y = forward(X) is zero; a unit shift of one normalised output costs 1/(5*7)
    assert mlo.loss_and_grad(p, X, y)[0] < 1e-20
    Y2 = y.copy()
    Y2[0, 3] += q["y_scale"][3]
    assert mlo.loss_and_grad(p, X, Y2)[0] == pytest.approx(1.0 / 35, rel=1e-9)


def test_bf16_rounding(mlo):
    # exactly representable values survive, halfway cases round to even
    assert mlo.bf16(np.array([1.0, -2.5, 0.0]))[0:3].tolist() == [1.0, -2.5, 0.0]
    one_ulp = 2.0 ** -7
    assert mlo.bf16(np.array([1.0 + one_ulp / 2]))[0] == 1.0
    assert mlo.bf16(np.array([1.0 + 3 * one_ulp / 2]))[0] == 1.0 + 2 * one_ulp
    x = np.random.default_rng(2).normal(size=1000)
    assert np.max(np.abs(mlo.bf16(x) - x) / np.abs(x)) <= 2.0 ** -8  # half an ulp of 8 bits
