"""The time-series tools behind the decorrelated-samples metric, on synthetic
AR(1) series whose integrated autocorrelation time is known in closed form:
x_t = phi x_{t-1} + e_t has rho(t) = phi^t, so tau_int = 1/2 + phi / (1 - phi)."""
import math

import numpy as np
import pytest

from patchyparticles_amd.stats import (autocorr, block_mean, replica_drift, stationary, tau_int, tau_int_err,
                                      tau_int_replicas)


def _ar1(phi, n, seed):
    rng = np.random.default_rng(seed)
    e = rng.normal(size=n)
    x = np.empty(n)
    x[0] = e[0] / math.sqrt(1 - phi * phi)
    for t in range(1, n):
        x[t] = phi * x[t - 1] + e[t]
    return x


@pytest.mark.parametrize("phi", [0.0, 0.5, 0.9, 0.98])
def test_tau_int_of_ar1(phi):
    x = _ar1(phi, 400000, seed=int(phi * 100))
    want = 0.5 + phi / (1 - phi)
    got = tau_int(x)
    # the estimator's relative error is about sqrt(2 (2W + 1) / n) with W ~ 5 tau
    assert abs(got - want) < 0.06 * want + 0.02, (phi, got, want)
    rho = autocorr(x)
    assert rho[0] == pytest.approx(1.0)
    assert abs(rho[1] - phi) < 0.01


def test_block_standard_error_of_ar1():
    phi, n = 0.9, 400000
    x = _ar1(phi, n, seed=3)
    tau = 0.5 + phi / (1 - phi)
    var = 1 / (1 - phi * phi)
    m, se = block_mean(x, nblocks=20)
    want = math.sqrt(2 * tau * var / n)
    assert abs(se - want) < 0.4 * want
    assert abs(m) < 5 * want


def test_tau_int_error_matches_the_spread_of_estimates():
    """Over 300 independent AR(1) series (phi = 0.9, tau = 9.5, n = 20000,
    about 2100 tau) the stated error is conservative, at most 1.5 times the
    spread of the estimates (the formula holds for windows much longer than
    tau; at W = 5 tau it is about 1.3 times the spread), and the estimates
    centre on the true tau."""
    phi, n = 0.9, 20000
    want = 0.5 + phi / (1 - phi)
    est, errs = [], []
    for k in range(300):
        t, e, W = tau_int_err(_ar1(phi, n, seed=1000 + k))
        assert W >= 5 * t
        est.append(t)
        errs.append(e)
    est = np.array(est)
    spread = est.std(ddof=1)
    assert spread <= np.mean(errs) <= 1.5 * spread, (np.mean(errs), spread)
    assert abs(est.mean() - want) < 4 * spread / math.sqrt(len(est)) + 0.02 * want


def test_constant_series():
    assert tau_int(np.ones(100)) == 0.5


def test_tau_int_of_replicas_pools_independent_series():
    """16 independent AR(1) series (phi = 0.95, tau = 19.5) of 5000 samples
    each: the pooled estimate centres on the closed form within its stated
    error (over 100 such sets the error is at least the estimates' spread),
    and one replica alone gives the single-series estimator's answer."""
    phi, n, R = 0.95, 5000, 16
    want = 0.5 + phi / (1 - phi)
    est, errs = [], []
    for k in range(100):
        X = np.stack([_ar1(phi, n, seed=10000 + 100 * k + j) for j in range(R)], 1)
        t, e, W = tau_int_replicas(X)
        assert W >= 5 * t
        est.append(t)
        errs.append(e)
    est = np.array(est)
    spread = est.std(ddof=1)
    assert spread <= np.mean(errs) <= 1.6 * spread, (np.mean(errs), spread)
    assert abs(est.mean() - want) < 4 * spread / math.sqrt(len(est)) + 0.02 * want
    x = _ar1(0.9, 20000, seed=5)
    a, b = tau_int_replicas(x[:, None])[:2], tau_int_err(x)[:2]
    assert a[0] == pytest.approx(b[0]) and a[1] == pytest.approx(b[1])


def test_stationarity_rule_flags_a_drift():
    """Stationary AR(1) replicas pass the rule; the same series plus a slow
    linear trend (a relaxation) fail it on the drift, and too short a run
    fails it on tau's error."""
    phi, n, R = 0.9, 20000, 8
    X = np.stack([_ar1(phi, n, seed=700 + j) for j in range(R)], 1)
    tau, err, _ = tau_int_replicas(X)
    d, dse = replica_drift(X)
    assert stationary(tau, err, d, dse)
    trend = np.linspace(0.0, -1.0, n)[:, None]
    tau2, err2, _ = tau_int_replicas(X + trend)
    d2, dse2 = replica_drift(X + trend)
    assert d2 < -0.4 and not stationary(tau2, err2, d2, dse2)
    short = X[:200]
    tau3, err3, _ = tau_int_replicas(short)
    d3, dse3 = replica_drift(short)
    assert not stationary(tau3, err3, d3, dse3, max_rel_err=0.1)
