"""Philox4x32-10 in numpy and the random-stream layout of the MI355X engine.

Philox4x32-10 follows the published algorithm (J. K. Salmon, M. A. Moraes,
R. O. Dror, D. E. Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11):
ten rounds of two 32x32->64 multiplies with the constants M0=0xD2511F53,
M1=0xCD9E8D57 and a Weyl key schedule W0=0x9E3779B9, W1=0xBB67AE85.
It is pinned by the known-answer vectors published with that paper's
reference implementation (tests/test_oracle.py::test_philox_kat).

Stream layout (replaces the MWC64X skip-ahead streams of the reference,
mcmcGPU.py:410-445, and the host position list, mcmcGPU.py:447-458):

* key        = the 64-bit context seed, (lo, hi).
* walker draw for global step t of walker w:
      r = philox((t>>1) lo, (t>>1) hi, w, TAG_WALK)
      (x, y) = (r0, r1) if t is even else (r2, r3)
      proposed residue b = mulhi32(x, q)            (mcmc.cl:441 uses x % q)
      acceptance uniform u = (y >> 8) * 2**-24      (mcmc.cl:29-31)
* site for global step t (shared by all walkers, as mcmcGPU.py:453-455):
      r = philox(t lo, t hi, 0, TAG_POS); pos = mulhi32(r0, L)
* independent reseed, call number g, walker w, sites 4k..4k+3:
      r = philox(k, g, w, TAG_INDEP); u_site = (r[site%4] >> 8) * 2**-24
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
MASK32 = np.uint64(0xFFFFFFFF)

TAG_WALK = 0x57414C4B   # 'WALK'
TAG_POS = 0x504F5349    # 'POSI'
TAG_INDEP = 0x494E4450  # 'INDP'


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10.  All arguments broadcast; returns 4 uint32
    arrays.  Round structure per Salmon et al. (2011), section 3."""
    c0, c1, c2, c3 = (np.asarray(x, dtype=np.uint32) for x in (c0, c1, c2, c3))
    c0, c1, c2, c3 = np.broadcast_arrays(c0, c1, c2, c3)
    c0, c1, c2, c3 = (x.astype(np.uint32).copy() for x in (c0, c1, c2, c3))
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    with np.errstate(over='ignore'):
        for _ in range(10):
            p0 = M0 * c0.astype(np.uint64)
            p1 = M1 * c2.astype(np.uint64)
            hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
            lo0 = (p0 & MASK32).astype(np.uint32)
            hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
            lo1 = (p1 & MASK32).astype(np.uint32)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
            k0 = np.uint32(k0 + W0)
            k1 = np.uint32(k1 + W1)
    return c0, c1, c2, c3


def split_seed(seed):
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return seed & 0xFFFFFFFF, seed >> 32


def mulhi32(x, n):
    """floor(x * n / 2**32) for uint32 x: a uniform draw in [0, n)."""
    return ((np.asarray(x, dtype=np.uint64) * np.uint64(n)) >> np.uint64(32)).astype(np.uint32)


def uniform24(y):
    """(y >> 8) * 2**-24 as float32, the reference's uniformMap (mcmc.cl:29-31)."""
    return ((np.asarray(y, dtype=np.uint32) >> np.uint32(8)).astype(np.float32)
            * np.float32(2.0 ** -24))


def site_positions(seed, L, t0, nsteps):
    """Shared site sequence for global steps t0 .. t0+nsteps-1."""
    k0, k1 = split_seed(seed)
    t = np.arange(t0, t0 + nsteps, dtype=np.uint64)
    r0, _, _, _ = philox4x32_10((t & MASK32).astype(np.uint32),
                                (t >> np.uint64(32)).astype(np.uint32),
                                0, TAG_POS, k0, k1)
    return mulhi32(r0, L)


def walker_draws(seed, q, walkers, t):
    """(b, u) for each walker in `walkers` at global step t."""
    k0, k1 = split_seed(seed)
    th = np.uint64(t) >> np.uint64(1)
    r0, r1, r2, r3 = philox4x32_10(np.uint32(th & MASK32), np.uint32(th >> np.uint64(32)),
                                   np.asarray(walkers, dtype=np.uint32), TAG_WALK, k0, k1)
    if int(t) & 1:
        x, y = r2, r3
    else:
        x, y = r0, r1
    return mulhi32(x, q), uniform24(y)


def indep_uniforms(seed, call, walkers, L):
    """(len(walkers), L) float32 uniforms used by the independent reseed."""
    k0, k1 = split_seed(seed)
    walkers = np.asarray(walkers, dtype=np.uint32)
    nk = (L + 3) // 4
    out = np.empty((walkers.size, 4 * nk), dtype=np.float32)
    for k in range(nk):
        r = philox4x32_10(np.uint32(k), np.uint32(call), walkers, TAG_INDEP, k0, k1)
        for e in range(4):
            out[:, 4 * k + e] = uniform24(r[e])
    return out[:, :L]
