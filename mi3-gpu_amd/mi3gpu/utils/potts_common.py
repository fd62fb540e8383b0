"""Potts-model formats and marginal helpers (potts_common.py of the reference).

Packed format: couplings and bimarginals are (L(L-1)/2, q*q) float32 arrays,
pairs ordered [(i,j) for i in range(L-1) for j in range(i+1,L)] and letter
pairs [(a,b) for a in alpha for b in alpha] (UserGuide.md:176-180).
"""
import numpy as np

alpha20 = "ACDEFGHIKLMNPQRSTVWY"
alpha21 = '-' + alpha20


def getL(size):
    """L from the number of pairs L(L-1)/2 (potts_common.py:25-26)."""
    return int(((1 + np.sqrt(1 + 8 * size)) // 2) + 0.5)


def getLq(J):
    """(L, q) of a packed array (potts_common.py:28-29)."""
    return getL(J.shape[0]), int(np.sqrt(J.shape[1]) + 0.5)


def pairs(L):
    return ((i, j) for i in range(L - 1) for j in range(i + 1, L))


def pair_first_index(L):
    """Packed row of pair (i, i+1) for every i."""
    i = np.arange(L)
    return i * (2 * L - i - 1) // 2


def bimarg_to_unimarg(ff):
    """Univariate marginals: site 0 from row sums of pair (0,1), site j>0 from
    column sums of pair (0,j) (potts_common.py:35-40), renormalised."""
    L, q = getLq(ff)
    ff = np.asarray(ff).reshape((-1, q, q))
    marg = np.empty((L, q), dtype=ff.dtype)
    marg[0] = ff[0].sum(axis=1)
    marg[1:] = ff[:L - 1].sum(axis=1)
    return marg / marg.sum(axis=1, keepdims=True)


getUnimarg = bimarg_to_unimarg


def indep_bimarg(fab):
    """f_i(a) f_j(b) for every pair, from each pair's own row/column sums
    (potts_common.py:42-47)."""
    L, q = getLq(fab)
    x = np.asarray(fab).reshape((-1, q, q))
    fa, fb = x.sum(axis=2), x.sum(axis=1)
    return (fa[:, :, None] * fb[:, None, :]).reshape(-1, q * q)


indepF = indep_bimarg


def validate_bimarg(ff, eps=1e-5):
    """Check that the unimarginals implied by every pair agree and sum to 1
    (potts_common.py:50-76); raises ValueError naming the first offender."""
    L, q = getLq(ff)
    x = np.asarray(ff).reshape((-1, q, q))
    uni = np.empty((L, q), dtype=np.float64)
    uni[0] = x[0].sum(axis=1)
    uni[1:] = x[:L - 1].sum(axis=1)
    delta = np.abs(uni.sum(axis=1) - 1)
    if np.max(delta) > eps:
        raise ValueError("unimarg for pos {} does not sum to 1: delta {}".format(
            np.argmax(delta), np.max(delta)))
    i, j = np.triu_indices(L, k=1)
    di = np.abs(uni[i] - x.sum(axis=2))
    dj = np.abs(uni[j] - x.sum(axis=1))
    for d, side, site in ((di, 'i', i), (dj, 'j', j)):
        bad = np.where(d.max(axis=1) > eps)[0]
        if len(bad):
            n = bad[0]
            m = np.argmax(d[n])
            raise ValueError("inconsitent {} unimarg {:.5f} with delta {:.5f} for bimarg {} "
                             "({},{})".format(side, uni[site[n]][m], d[n][m], n, i[n], j[n]))


def getC(fab):
    """Connected correlations and their normalised form (potts_common.py:78-91)."""
    L, q = getLq(fab)
    x = np.asarray(fab).reshape((-1, q, q))
    fa, fb = x.sum(axis=2), x.sum(axis=1)
    C = fab - (fa[:, :, None] * fb[:, None, :]).reshape(-1, q * q)
    ss = ((fa * (1 - fa))[:, :, None] * (fb * (1 - fb))[:, None, :]).reshape(-1, q * q)
    return C, C / np.sqrt(ss)


def getM(x, diag_fill=0):
    """Pair array -> symmetric (L, L, ...) matrix (potts_common.py:93-101)."""
    L = getL(x.shape[0])
    M = np.zeros((L, L) + x.shape[1:], dtype=x.dtype)
    i, j = np.triu_indices(L, k=1)
    M[i, j, ...] = x
    M = M + M.swapaxes(0, 1)
    d = np.arange(L)
    M[d, d, ...] = diag_fill
    return M


def printsome(a, prec=4):
    """First five values of a flattened array, as one line."""
    return np.array2string(np.asarray(a).flatten()[:5], precision=prec, sign=' ')[1:-1]
