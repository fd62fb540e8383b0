"""Effective sample size of walker energy traces.

The reference judges equilibration only with a Spearman rank test
(NewtonSteps.py:419-428); this module adds the integrated autocorrelation
time of the energy, the quantity behind the energy-ESS/s throughput figure.
"""
import numpy as np


def autocorr(E, tmax, mean=None):
    """Walker-averaged normalised autocorrelation rho(0..tmax) of traces E
    with shape (T, nwalkers), centred on the ensemble mean (all walkers are
    assumed to sample the same stationary distribution)."""
    E = np.asarray(E, dtype=np.float64)
    T = E.shape[0]
    mu = E.mean() if mean is None else mean
    x = E - mu
    var = np.mean(x * x)
    rho = np.empty(min(tmax, T - 1) + 1)
    rho[0] = 1.0
    for t in range(1, len(rho)):
        rho[t] = np.mean(x[:-t] * x[t:]) / var
    return rho


def tau_int(E, c=5.0):
    """Integrated autocorrelation time tau = 1/2 + sum_{t=1..W} rho(t) with
    Sokal's automatic window: the smallest W with W >= c * tau(W).
    Returns (tau, W).  tau is in units of the sampling interval, so that the
    variance of a T-sample mean is 2*tau*var/T."""
    E = np.asarray(E, dtype=np.float64)
    T = E.shape[0]
    x = E - E.mean()
    var = np.mean(x * x)
    if var == 0:
        return 0.5, 0
    tau = 0.5
    for W in range(1, T):
        tau += np.mean(x[:-W] * x[W:]) / var
        if W >= c * tau:
            return tau, W
    return tau, T - 1


def energy_ess(E, c=5.0):
    """ESS = nwalkers * T / (2 tau) of energy traces (T, nwalkers)."""
    tau, W = tau_int(E, c)
    T, n = np.asarray(E).shape
    return n * T / (2 * tau), tau, W
