"""Many independent Potts models of uneven L and q swept together on one GPU.

Each model has its own couplings (the reference's packed (L(L-1)/2, q*q)
float32 layout), walkers and random streams: model m is keyed by seed + m
and runs exactly the chains a single-model MCMCGPU with that seed would.
A round (`run(nsweeps)`) is one replay of a HIP graph holding the site lists
of every model, one persistent sweep kernel per class of models that share
the register layout, on parallel branches, and the step-counter advance.
`tune()` picks the walkers per work-group tile of each class by timing.

The reference runs one model per process; this module has no counterpart
there.
"""
import ctypes

import numpy as np

from mi3gpu._lib import lib, MI3_EINVAL, MI3_EUNSUP, Mi3Error

_u32p = ctypes.POINTER(ctypes.c_uint32)


def _check(b, rc):
    if rc >= 0:
        return rc
    msg = lib().mi3_batch_last_error(b)
    msg = msg.decode(errors='replace') if msg else ''
    if rc in (MI3_EINVAL, MI3_EUNSUP):
        raise ValueError(msg)
    raise Mi3Error(rc, msg)


class PottsBatch:
    def __init__(self, Ls, qs, nwalkers, seed=0, device=0):
        self.L = [int(x) for x in Ls]
        self.q = [int(x) for x in qs]
        self.N = [int(x) for x in nwalkers]
        if not (len(self.L) == len(self.q) == len(self.N)):
            raise ValueError("Ls, qs and nwalkers must have one entry per model")
        n = len(self.L)
        arr = lambda v: (ctypes.c_uint32 * n)(*v)
        self._b = ctypes.c_void_p()
        _check(None, lib().mi3_batch_create(int(device), n, arr(self.L), arr(self.q),
                                            arr(self.N), int(seed), ctypes.byref(self._b)))
        self.seed = int(seed)

    def __len__(self):
        return len(self.L)

    def close(self):
        if getattr(self, '_b', None):
            lib().mi3_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_couplings(self, m, J):
        L, q = self.L[m], self.q[m]
        J = np.ascontiguousarray(J, dtype='<f4')
        if J.shape != (L * (L - 1) // 2, q * q):
            raise ValueError(f"model {m}: couplings must have shape "
                             f"{(L * (L - 1) // 2, q * q)}, got {J.shape}")
        _check(self._b, lib().mi3_batch_set_couplings(self._b, m, J.ctypes.data))

    def set_seqs(self, m, seqs):
        seqs = np.asarray(seqs)
        if seqs.shape != (self.N[m], self.L[m]) or seqs.dtype != np.uint8:
            raise ValueError(f"model {m}: sequences must be uint8 of shape "
                             f"{(self.N[m], self.L[m])}")
        dev = np.ascontiguousarray(seqs.T)
        _check(self._b, lib().mi3_batch_set_seqs(self._b, m, dev.ctypes.data))

    def get_seqs(self, m):
        out = np.empty((self.L[m], self.N[m]), dtype=np.uint8)
        _check(self._b, lib().mi3_batch_get_seqs(self._b, m, out.ctypes.data))
        return np.ascontiguousarray(out.T)

    def run(self, nsweeps=1):
        """nsweeps sweeps (nsweeps*L_m single-site updates per walker of model
        m) of every model, asynchronously; one graph launch."""
        _check(self._b, lib().mi3_batch_run(self._b, int(nsweeps)))

    def wait(self):
        _check(self._b, lib().mi3_batch_synchronize(self._b))

    def tune(self, nsweeps=1):
        """Time every allowed tile per class on nsweeps real sweeps (the
        walkers move on) and keep the fastest."""
        _check(self._b, lib().mi3_batch_tune(self._b, int(nsweeps)))

    def tile(self, m):
        w = ctypes.c_uint32(0)
        _check(self._b, lib().mi3_batch_tile(self._b, m, ctypes.byref(w)))
        return w.value

    def set_tile(self, m, wpb):
        _check(self._b, lib().mi3_batch_set_tile(self._b, m, int(wpb)))

    def step_count(self, m):
        s = ctypes.c_uint64(0)
        _check(self._b, lib().mi3_batch_step_count(self._b, m, ctypes.byref(s)))
        return s.value
