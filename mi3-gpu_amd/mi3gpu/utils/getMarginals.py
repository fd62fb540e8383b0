"""Host-side marginals of a sequence set (getMarginals.py of the reference),
a user utility; the inference round counts on the GPU
(mcmcGPU.MCMCGPU.calcBicounts)."""
import numpy as np


def getMarginals(seqs, q, weights=None, nrmlz=True, uni=True, bi=True):
    """Uni- and bimarginals by per-pair bincount (getMarginals.py:27-53)."""
    seqs = np.asarray(seqs).astype(np.intp)
    N, L = seqs.shape
    norm = (lambda x: x / np.sum(x, axis=-1, keepdims=True)) if nrmlz else (lambda x: x)
    ret = ()
    if uni:
        ret += (norm(np.array([np.bincount(seqs[:, i], weights, q) for i in range(L)])),)
    if bi:
        ret += (norm(np.array([np.bincount(q * seqs[:, i] + seqs[:, j], weights, q * q)
                               for i in range(L - 1) for j in range(i + 1, L)])),)
    return ret
