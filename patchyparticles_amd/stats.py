"""Time-series statistics for Monte Carlo observables.

tau_int: integrated autocorrelation time with Sokal's automatic window
(the smallest W with W >= c * tau_int(W), c = 5); decorrelated samples per
unit time are then rate / (2 tau_int).  tau_int_err adds the estimator's
statistical error.  tau_int_replicas pools independent replicas' series of
one observable; replica_drift and stationary decide whether such a set of
series may be read as an equilibrium one at all.
"""
from __future__ import annotations

import numpy as np


def autocorr(x):
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    x = x - x.mean()
    m = 1 << int(np.ceil(np.log2(2 * n)))
    f = np.fft.rfft(x, m)
    acf = np.fft.irfft(f * np.conj(f), m)[:n]
    if acf[0] <= 0:
        return np.ones(1)
    return acf / acf[0]


def _tau_window(x, c):
    rho = autocorr(x)
    if len(rho) < 2:
        return 0.5, 0
    csum = 0.5 + np.cumsum(rho[1:])
    for W in range(1, len(csum) + 1):
        if W >= c * csum[W - 1]:
            return float(max(csum[W - 1], 0.5)), W
    return float(max(csum[-1], 0.5)), len(csum)


def tau_int(x, c=5.0):
    """Integrated autocorrelation time in units of the sampling interval
    (tau = 1/2 + sum_{t>=1} rho(t), so an uncorrelated series gives 1/2)."""
    return _tau_window(x, c)[0]


def tau_int_err(x, c=5.0):
    """tau_int with its statistical error and window: (tau, err, W), where
    err = tau sqrt(2 (2W + 1) / n) (Madras and Sokal, J. Stat. Phys. 50, 109
    (1988), eq. 2.25).  The formula is conservative: on AR(1) series with
    W = 5 tau it is about 1.3 times the actual spread of the estimates.  A
    series shorter than about 100 tau gives an error above 45 %."""
    n = len(x)
    tau, W = _tau_window(x, c)
    err = tau * float(np.sqrt(2 * (2 * W + 1) / n)) if n > 0 else float("inf")
    return tau, err, W


def block_mean(x, nblocks=20):
    """Mean and standard error from non-overlapping block averages."""
    x = np.asarray(x, dtype=np.float64)
    nb = min(nblocks, len(x))
    L = len(x) // nb
    b = x[: nb * L].reshape(nb, L).mean(axis=1)
    return float(b.mean()), float(b.std(ddof=1) / np.sqrt(nb)) if nb > 1 else float("nan")


def tau_int_replicas(X, c=5.0):
    """tau_int of R independent series of one observable, X of shape (n, R)
    (R replicas sampled at the same interval): each series centred on its own
    mean, their autocovariances summed and normalised by the summed variances,
    then Sokal's window on that pooled function.  Returns (tau, err, W) with
    the Madras-Sokal error for a total length of R n samples,
    err = tau sqrt(2 (2W + 1) / (R n))."""
    X = np.asarray(X, dtype=np.float64)
    if X.ndim == 1:
        X = X[:, None]
    n, R = X.shape
    Y = X - X.mean(axis=0)
    m = 1 << int(np.ceil(np.log2(2 * n)))
    f = np.fft.rfft(Y, m, axis=0)
    acov = np.fft.irfft(f * np.conj(f), m, axis=0)[:n].sum(axis=1)
    if acov[0] <= 0:
        return 0.5, 0.0, 0
    rho = acov / acov[0]
    csum = 0.5 + np.cumsum(rho[1:])
    tau, W = float(max(csum[-1], 0.5)) if len(csum) else 0.5, len(csum)
    for w in range(1, len(csum) + 1):
        if w >= c * csum[w - 1]:
            tau, W = float(max(csum[w - 1], 0.5)), w
            break
    err = tau * float(np.sqrt(2 * (2 * W + 1) / (R * n))) if n > 0 else float("inf")
    return tau, err, W


def replica_drift(X):
    """Drift of R series X (n, R) across their run: the mean over replicas of
    (mean of the second half - mean of the first half), and its standard
    error from the spread over the replicas (each replica one independent
    estimate).  A drift of several standard errors means the runs are still
    relaxing or aging."""
    X = np.asarray(X, dtype=np.float64)
    if X.ndim == 1:
        X = X[:, None]
    h = X.shape[0] // 2
    d = X[h:2 * h].mean(axis=0) - X[:h].mean(axis=0)
    se = float(d.std(ddof=1) / np.sqrt(len(d))) if len(d) > 1 else float("nan")
    return float(d.mean()), se


def stationary(tau, tau_err, drift, drift_se, max_rel_err=0.25, max_drift_se=2.0):
    """Whether a tau_int estimate describes an equilibrium: its relative error
    below max_rel_err and the drift within max_drift_se standard errors."""
    return bool(tau > 0 and tau_err / tau < max_rel_err and np.isfinite(drift_se) and
                abs(drift) < max_drift_se * drift_se)
