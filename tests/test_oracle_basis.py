"""Oracle pins: float fld/cld and the 1D basis functions.

fld/cld are checked against exact rational arithmetic; the uGIMP weight is
checked against its definition as the particle-domain average of the linear
hat function (quadrature), plus partition of unity, linear reproduction and a
finite-difference derivative.
"""
import math
import random
from fractions import Fraction

import numpy as np
import pytest


def _exact_floor(x, y):
    return math.floor(Fraction(x) / Fraction(y))


def _exact_ceil(x, y):
    return math.ceil(Fraction(x) / Fraction(y))


def test_fld_cld_julia_manual_example(orc):
    # Julia manual: fld(6.0, 0.1) == 59.0 although 6.0 / 0.1 == 60.0
    assert 6.0 / 0.1 == 60.0
    assert orc.fld(6.0, 0.1) == 59.0
    assert orc.cld(6.0, 0.1) == 60.0
    assert orc.fld(0.3, 0.1) == 2.0 and orc.cld(0.3, 0.1) == 3.0


def test_fld_cld_exact_rational(orc):
    rng = random.Random(7)
    cases = []
    for _ in range(4000):
        y = rng.choice([0.1, 0.0025, 0.005, 0.5, 1.0 / 3.0, rng.uniform(1e-3, 2.0)])
        k = rng.randint(-400, 400)
        x = rng.choice([k * y, k * y + rng.uniform(-1e-12, 1e-12), rng.uniform(-3, 3),
                        float(Fraction(k) * Fraction(y))])
        cases.append((x, y))
    for x, y in cases:
        assert orc.fld(x, y) == _exact_floor(x, y), (x, y)
        assert orc.cld(x, y) == _exact_ceil(x, y), (x, y)
    for x, y in cases[:500]:
        assert orc.fld(x, -y) == _exact_floor(x, -y)
        assert orc.cld(x, -y) == _exact_ceil(x, -y)


def _ugimp_quadrature(d, h, lp, n=20001):
    """uGIMP weight by definition: mean over the particle domain
    [d - lp/2, d + lp/2] of the linear hat max(0, 1 - |s|/h)."""
    s = np.linspace(d - lp / 2, d + lp / 2, n)
    f = np.maximum(0.0, 1.0 - np.abs(s) / h)
    return np.trapezoid(f, s) / lp


@pytest.mark.parametrize("h,lp", [(1.0, 0.5), (0.0025, 0.00125), (0.02, 0.02 / 3), (1.0, 1.0)])
def test_ugimp_matches_domain_average(orc, h, lp):
    for d in np.linspace(-(h + lp / 2) * 1.1, (h + lp / 2) * 1.1, 301):
        N, _ = orc.ugimp_basis(d, h, lp)
        assert abs(N - _ugimp_quadrature(d, h, lp)) < 2e-6, d


@pytest.mark.parametrize("h,lp", [(1.0, 0.5), (0.005, 0.0025), (0.02, 0.01)])
def test_ugimp_derivative_and_branch_points(orc, h, lp):
    eps = 1e-7 * h
    pts = list(np.linspace(-(h + lp), h + lp, 257))
    for d in pts:
        N, dN = orc.ugimp_basis(d, h, lp)
        Np, _ = orc.ugimp_basis(d + eps, h, lp)
        Nm, _ = orc.ugimp_basis(d - eps, h, lp)
        # the weight is C1: the central difference matches everywhere
        assert abs((Np - Nm) / (2 * eps) - dN) < 1e-5 / h, d
    # continuity at every branch point and zero at/after the support edge
    for b in (lp / 2, h - lp / 2, h + lp / 2):
        for s in (1, -1):
            lo, _ = orc.ugimp_basis(s * b * (1 - 1e-12), h, lp)
            hi, _ = orc.ugimp_basis(s * b * (1 + 1e-12), h, lp)
            assert abs(lo - hi) < 1e-9
    assert orc.ugimp_basis(h + lp / 2, h, lp) == (0.0, 0.0)
    assert orc.ugimp_basis(-(h + lp / 2) * 1.01, h, lp) == (0.0, 0.0)
    # centre value: 1 - lp / (4 h)
    assert abs(orc.ugimp_basis(0.0, h, lp)[0] - (1 - lp / (4 * h))) < 1e-15


@pytest.mark.parametrize("h,lp", [(0.0025, 0.00125), (1.0, 0.5), (0.3, 0.1)])
def test_ugimp_partition_of_unity_and_reproduction(orc, h, lp):
    rng = np.random.default_rng(3)
    nodes = np.arange(-5, 6) * h
    for xp in rng.uniform(-2 * h, 2 * h, 200):
        w = np.array([orc.ugimp_basis(xp - xn, h, lp) for xn in nodes])
        assert abs(w[:, 0].sum() - 1.0) < 1e-13
        assert abs((w[:, 0] * nodes).sum() - xp) < 1e-13 * max(h, 1.0)
        assert abs(w[:, 1].sum()) < 1e-12 / h


def test_linear_basis(orc):
    h = 0.25
    nodes = np.arange(-4, 5) * h
    for xp in np.random.default_rng(1).uniform(-h, h, 100):
        # within the cell, the two corner weights are 1 - |d|/h and sum to one
        c0 = math.floor(xp / h) * h
        w = [orc.linear_basis(xp - xn, h) for xn in (c0, c0 + h)]
        assert abs(w[0][0] - (1 - (xp - c0) / h)) < 1e-15
        assert abs(w[0][0] + w[1][0] - 1.0) < 1e-15
        assert w[0][1] == -1 / h and w[1][1] == 1 / h
    assert orc.linear_basis(0.0, h) == (1.0, -0.0) or orc.linear_basis(0.0, h)[1] == 0.0
