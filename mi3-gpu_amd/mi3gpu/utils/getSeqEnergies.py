"""Host-side Potts energies (getSeqEnergies.py of the reference), a user
utility for small sequence sets; the inference round computes energies on the
GPU (mcmcGPU.MCMCGPU.calcEnergies)."""
import numpy as np

from mi3gpu.utils.potts_common import getLq


def E_potts(s, J):
    """E(s) = sum_{i<j} J[pair(i,j), q*s_i + s_j], float64 accumulation
    (getSeqEnergies.py:38-60)."""
    L, q = getLq(J)
    s = np.asarray(s)
    assert s.shape[-1] == L
    single = s.ndim == 1
    s = np.atleast_2d(s).astype(np.intp)
    J = np.asarray(J)
    i, j = np.triu_indices(L, k=1)
    n = np.arange(len(i))
    E = np.zeros(s.shape[0], dtype='f8')
    for start in range(0, len(i), 4096):
        sl = slice(start, start + 4096)
        E += J[n[sl], q * s[:, i[sl]] + s[:, j[sl]]].sum(axis=1, dtype='f8')
    return E[0] if single else E


def E_indep(s, h):
    L, q = h.shape
    return np.sum(h[np.arange(L), s], axis=-1)
