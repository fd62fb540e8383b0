"""Gauge transforms of Potts parameters (changeGauge.py of the reference),
vectorised over pairs (the reference loops over pairs in Python,
changeGauge.py:132-134, which costs seconds per round at L=232).

h: (L, q) fields; J: (L(L-1)/2, q*q) packed couplings.
"""
import numpy as np

from mi3gpu.utils.potts_common import getL


def impute_params(hs, Js):
    if hs is None and Js is None:
        raise ValueError("need fields or couplings")
    if Js is None:
        L, q = hs.shape
        Js = np.zeros((L * (L - 1) // 2, q * q), dtype=hs.dtype)
    if hs is None:
        L = getL(Js.shape[0])
        q = int(round(np.sqrt(Js.shape[1])))
        hs = np.zeros((L, q), dtype=Js.dtype)
    L, q = hs.shape
    return L, q, hs, Js


def zeroJGauge(hs, Js, weights=None):
    """Couplings to the zero-mean gauge (sum over either letter of every block
    is 0) with fields absorbing the change, so sequence energies are kept
    (changeGauge.py:37-84).  With `weights`, weighted means are used."""
    L, q, hs, Js = impute_params(hs, Js)
    if np.any(np.isinf(Js)) or np.any(np.isinf(hs)):
        raise ValueError("Error: Cannot convert to zero gauge because of infinities")
    Jx = Js.reshape((-1, q, q))
    if weights is None:
        mJ1, mJ2 = Jx.mean(axis=1), Jx.mean(axis=2)
        mJ = mJ1.mean(axis=1)
    else:
        w = weights.reshape((-1, q, q))
        mJ1 = np.average(Jx, weights=w, axis=1)
        mJ2 = np.average(Jx, weights=w, axis=2)
        mJ = np.average(Jx, weights=w, axis=(1, 2))
    J0 = (Jx - mJ1[:, None, :] - mJ2[:, :, None] + mJ[:, None, None]).reshape((-1, q * q))
    h0 = hs - np.sum(mJ) / L
    i, j = np.triu_indices(L, k=1)
    np.add.at(h0, j, mJ1)
    np.add.at(h0, i, mJ2)
    return h0, J0


def zeroGauge(hs, Js, weights=None):
    """zeroJGauge plus zero-mean fields; energies shift by a constant
    (changeGauge.py:86-111)."""
    h0, J0 = zeroJGauge(hs, Js, weights)
    h0 = h0 - np.mean(h0, axis=1, keepdims=True)
    return h0, J0


def fieldlessGaugeDistributed(hs, Js, weights=None):
    """Fold each field h_i(a) evenly into the L-1 couplings touching site i
    (changeGauge.py:113-135); energies are kept."""
    L, q, hs, Js = impute_params(hs, Js)
    hd = hs / (L - 1)
    i, j = np.triu_indices(L, k=1)
    J0 = Js + (hd[i][:, :, None] + hd[j][:, None, :]).reshape((-1, q * q))
    return np.zeros(hs.shape), J0


def fieldlessGaugeEven(hs, Js, weights=None):
    """zeroGauge, then fold the fields evenly (changeGauge.py:137-155).  The
    result does not depend on the gauge of the input; energies shift by a
    sequence-independent constant."""
    return fieldlessGaugeDistributed(*zeroGauge(hs, Js, weights))


def fieldlessGauge(hs, Js, weights=None):
    """Move fields into couplings with site 0 (changeGauge.py:157-179)."""
    L, q, hs, Js = impute_params(hs, Js)
    J0 = Js.copy()
    J0[0, :] += np.repeat(hs[0, :], q)
    for i in range(L - 1):
        J0[i, :] += np.tile(hs[i + 1, :], q)
    return np.zeros(hs.shape), J0
