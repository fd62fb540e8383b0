"""ctypes loader for oracle/_build/liboracle.so (the C restatement in
oracle/oracle.c).  Test infrastructure only."""
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = _HERE / '_build' / 'liboracle.so'
_lib = None

_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags='C_CONTIGUOUS')
_f4p = np.ctypeslib.ndpointer(dtype=np.float32, flags='C_CONTIGUOUS')
_f8p = np.ctypeslib.ndpointer(dtype=np.float64, flags='C_CONTIGUOUS')
_u4p = np.ctypeslib.ndpointer(dtype=np.uint32, flags='C_CONTIGUOUS')
_c = ctypes


def build():
    subprocess.run(['make', '-s', '-C', str(_HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            build()
        L = ctypes.CDLL(str(_LIB))
        L.orc_philox4x32_10.argtypes = [_u4p, _u4p, _u4p]
        L.orc_positions.argtypes = [_c.c_uint64, _c.c_uint32, _c.c_uint64, _c.c_uint32, _u4p]
        L.orc_metropolis.argtypes = [_f4p, _c.c_uint32, _c.c_uint32, _u8p, _c.c_uint32,
                                     _c.c_uint64, _c.c_uint64, _c.c_uint32, _c.c_float,
                                     _c.c_uint32, _f8p, _c.c_int]
        L.orc_metropolis_f32.argtypes = [_f4p, _c.c_uint32, _c.c_uint32, _u8p, _c.c_uint32,
                                         _c.c_uint64, _c.c_uint64, _c.c_uint32, _c.c_float,
                                         _c.c_uint32, _c.c_int]
        L.orc_energies.argtypes = [_f4p, _c.c_uint32, _c.c_uint32, _u8p, _c.c_uint32, _f8p,
                                   _c.c_int]
        L.orc_bicounts.argtypes = [_u8p, _c.c_uint32, _c.c_uint32, _c.c_uint32, _u4p, _c.c_int]
        L.orc_weighted_marg.argtypes = [_u8p, _f4p, _c.c_uint32, _c.c_uint32, _c.c_uint32,
                                        _f8p, _c.c_int]
        L.orc_gen_indep.argtypes = [_f4p, _c.c_uint32, _c.c_uint32, _c.c_uint64, _c.c_uint32,
                                    _c.c_uint32, _c.c_uint32, _u8p]
        L.orc_num_threads.restype = _c.c_int
        _lib = L
    return _lib


def _Lq(J):
    nP, qq = J.shape
    L = int(((1 + np.sqrt(1 + 8 * nP)) // 2) + 0.5)
    return L, int(round(np.sqrt(qq)))


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.ascontiguousarray(ctr, np.uint32),
                            np.ascontiguousarray(key, np.uint32), out)
    return out


def positions(seed, L, t0, n):
    out = np.zeros(max(n, 1), np.uint32)
    lib().orc_positions(seed, L, t0, n, out)
    return out[:n]


def metropolis(seqs, J, seed, t0, nsteps, beta=1.0, walker0=0, nthreads=0):
    """float64-dE chain; returns (seqs, margins)."""
    J = np.ascontiguousarray(J, np.float32)
    L, q = _Lq(J)
    s = np.ascontiguousarray(seqs, np.uint8).copy()
    mg = np.zeros(s.shape[0], np.float64)
    rc = lib().orc_metropolis(J, L, q, s, s.shape[0], seed, t0, nsteps, beta, walker0, mg,
                              nthreads)
    assert rc == 0
    return s, mg


def metropolis_f32(seqs, J, seed, t0, nsteps, beta=1.0, walker0=0, nthreads=0):
    J = np.ascontiguousarray(J, np.float32)
    L, q = _Lq(J)
    s = np.ascontiguousarray(seqs, np.uint8).copy()
    rc = lib().orc_metropolis_f32(J, L, q, s, s.shape[0], seed, t0, nsteps, beta, walker0,
                                  nthreads)
    assert rc == 0
    return s


def energies(seqs, J, nthreads=0):
    J = np.ascontiguousarray(J, np.float32)
    L, q = _Lq(J)
    s = np.ascontiguousarray(seqs, np.uint8)
    E = np.zeros(s.shape[0], np.float64)
    lib().orc_energies(J, L, q, s, s.shape[0], E, nthreads)
    return E


def bicounts(seqs, q, nthreads=0):
    s = np.ascontiguousarray(seqs, np.uint8)
    N, L = s.shape
    out = np.zeros((L * (L - 1) // 2, q * q), np.uint32)
    lib().orc_bicounts(s, N, L, q, out, nthreads)
    return out


def weighted_marg(seqs, w, q, nthreads=0):
    s = np.ascontiguousarray(seqs, np.uint8)
    N, L = s.shape
    out = np.zeros((L * (L - 1) // 2, q * q), np.float64)
    lib().orc_weighted_marg(s, np.ascontiguousarray(w, np.float32), N, L, q, out, nthreads)
    return out


def gen_indep(cprob, seed, call, N, walker0=0):
    cprob = np.ascontiguousarray(cprob, np.float32)
    L, qm1 = cprob.shape
    out = np.zeros((N, L), np.uint8)
    lib().orc_gen_indep(cprob, L, qm1 + 1, seed, call, N, walker0, out)
    return out


def num_threads():
    return lib().orc_num_threads()
