"""Pins the oracle's Heaviside projection and its sensitivity
(PDEFilter::HeavisideFilter / ComputeProjectedSens, PDEFilter.cpp:242-341)
without restating their formula:

* special points: x~ = 1 maps to 1 and x~ = 0 to 0 (then the Xmin clamp);
* the beta -> 0 limit is the identity, a large beta gives the unit step at x0;
* symmetry about x0 = 0.5 (xp(x0 + t) + xp(x0 - t) = 1) and monotonicity;
* the chain-rule factor of ComputeProjectedSens equals a central finite
  difference of HeavisideFilter, and it scales dgdx like dfdx;
* the beta continuation over a run of calls itr = 0..100 against a schedule
  tabulated by hand from PDEFilter.cpp:259-269 (beta0 below itr 30, then
  +1 on every itr divisible by 10, capped at betamax), read through the
  projected value of x~ = 0.75, which for x0 = 0.5 is
  1/2 + tanh(beta/4) / (2 tanh(beta/2)) (tanh(a) + tanh(b) over a common
  denominator; the identity is checked on its own in mpmath).
A planted mistake (a shifted schedule, a dropped factor beta) fails these.
"""
import mpmath as mp
import numpy as np
import pytest


def _prob(O, **kw):
    prm = dict(Emax=1.0, proj=1, x0=0.5, beta0=1.0, betamax=8.0, Xmin=0.0)
    prm.update(kw)
    return O.cantilever(8, 4, 2, 2, **prm)


def test_special_points(O):
    for beta0 in (0.5, 1.0, 3.0, 8.0):
        P = _prob(O, beta0=beta0)
        xf = np.full(P.ne, 0.5)
        xf[0], xf[1] = 1.0, 0.0
        xp, beta = P.heaviside(xf, 0)
        assert beta == beta0
        assert xp[0] == 1.0
        assert abs(xp[1]) <= 1e-16
        assert abs(xp[2] - 0.5) <= 1e-16  # x~ = x0 = 0.5 maps to 1/2 for x0 = 0.5
    P = _prob(O, Xmin=0.01)
    xf = np.zeros(P.ne)
    xp, _ = P.heaviside(xf, 0)
    assert np.all(xp == 0.01)  # clamped to Xmin after the projection


def test_limits_symmetry_monotonicity(O):
    P = _prob(O, beta0=1e-4)
    rng = np.random.default_rng(1)
    xf = rng.uniform(0.0, 1.0, P.ne)
    xp, _ = P.heaviside(xf, 0)
    assert np.max(np.abs(xp - xf)) < 1e-8  # beta -> 0: identity (error O(beta^2))
    P = _prob(O, beta0=400.0)
    xf = np.where(np.arange(P.ne) % 2 == 0, 0.45, 0.55)
    xp, _ = P.heaviside(xf, 0)
    assert np.all(np.abs(xp - (xf > 0.5)) < 1e-12)  # large beta: unit step at x0
    P = _prob(O, beta0=3.0)
    t = np.linspace(0.0, 0.5, P.ne)
    a, _ = P.heaviside(0.5 + t, 0)
    b, _ = P.heaviside(0.5 - t, 0)
    assert np.max(np.abs(a + b - 1.0)) < 1e-15
    assert np.all(np.diff(a) > 0)


@pytest.mark.parametrize("beta0,x0", [(1.0, 0.5), (3.0, 0.5), (6.0, 0.3)])
def test_projected_sens_is_derivative(O, beta0, x0):
    P = _prob(O, beta0=beta0, x0=x0)
    rng = np.random.default_rng(2)
    xf = rng.uniform(0.05, 0.95, P.ne)
    h = 1e-6
    hp, _ = P.heaviside(xf + h, 0)
    hm, _ = P.heaviside(xf - h, 0)
    fd = (hp - hm) / (2 * h)
    P.heaviside(xf, 0)  # beta of the last call is the one the sensitivity uses
    df = rng.standard_normal(P.ne)
    dg = rng.uniform(0.5, 1.5, P.ne)
    rdf, rdg = P.projected_sens(xf, df, dg[None, :])
    assert np.max(np.abs(rdf - df * fd)) <= 1e-8 * np.max(np.abs(df * fd))
    assert np.max(np.abs(rdg[0] - dg * fd)) <= 1e-8 * np.max(np.abs(dg * fd))


def _beta_table(itr, beta0, betamax):
    """PDEFilter.cpp:259-269 tabulated by hand for a call at every itr."""
    if itr < 30:
        return beta0
    return min(betamax, beta0 + (itr // 10) - 2)


def test_beta_table_by_hand():
    assert [_beta_table(i, 1.0, 8.0) for i in (0, 29, 30, 39, 40, 49, 50, 89, 90, 100)] == \
        [1, 1, 2, 2, 3, 3, 4, 7, 8, 8]
    assert [_beta_table(i, 1.0, 4.5) for i in (59, 60, 70)] == [4, 4.5, 4.5]


def test_identity_at_three_quarters():
    mp.mp.dps = 30
    for b in (1, 2, 5, 8):
        b = mp.mpf(b)
        lhs = (mp.tanh(b / 2) + mp.tanh(b / 4)) / (mp.tanh(b / 2) + mp.tanh(b / 2))
        rhs = mp.mpf(1) / 2 + mp.tanh(b / 4) / (2 * mp.tanh(b / 2))
        assert abs(lhs - rhs) < mp.mpf(10) ** -25


@pytest.mark.parametrize("betamax", [8.0, 4.5])
def test_beta_continuation(O, betamax):
    P = _prob(O, beta0=1.0, betamax=betamax)
    xf = np.full(P.ne, 0.75)
    for itr in range(0, 101):
        xp, beta = P.heaviside(xf, itr)
        want = _beta_table(itr, 1.0, betamax)
        assert beta == want, (itr, beta, want)
        b = float(want)
        assert abs(xp[0] - (0.5 + np.tanh(b / 4) / (2 * np.tanh(b / 2)))) < 1e-15, itr


def test_proj_off_copies(O):
    P = _prob(O, proj=0, Xmin=0.01)
    xf = np.random.default_rng(3).uniform(0.0, 1.0, P.ne)
    xp, _ = P.heaviside(xf, 40)
    assert np.array_equal(xp, xf)  # VecCopy, no clamp (PDEFilter.cpp:288-291)
