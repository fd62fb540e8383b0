"""Pseudocounts on bimarginals (pseudocount.py of the reference)."""
import numpy as np

from mi3gpu.utils.potts_common import getLq, getUnimarg


def mutation_pc(ff, N, mode='jeffreys'):
    """Correlation-preserving pseudocount: mu = q/(2N+q) (Jeffreys) or
    q/(N+q) (Bayes) (pseudocount.py:8-24)."""
    L, q = getLq(ff)
    mu = q / (2 * N + q) if mode == 'jeffreys' else q / (N + q)
    f = getUnimarg(ff)
    i, j = np.triu_indices(L, k=1)
    fifj = (f[i][:, :, None] + f[j][:, None, :]).reshape(-1, q * q)
    out = (1 - mu) ** 2 * ff + (1 - mu) * mu * fifj / q + (mu / q) ** 2
    return out / np.sum(out, axis=1, keepdims=True)
