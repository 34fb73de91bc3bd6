"""Time-series statistics for Monte Carlo observables.

tau_int: integrated autocorrelation time with Sokal's automatic window
(the smallest W with W >= c * tau_int(W), c = 5); decorrelated samples per
unit time are then rate / (2 tau_int).  tau_int_err adds the estimator's
statistical error.
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
