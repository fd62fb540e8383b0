"""The allocation-free weighted-marginal entry point
(mi3_weighted_marg_noalloc) against mi3_weighted_marg, called through the C
ABI: bit-identical bins on two calls in a row, the second with fewer
sequences, at a size whose pair blocks take several runs of the library's
scratch accumulator (L=232, q=21) and at a small one; oversize requests are
refused before any launch."""
import ctypes

import numpy as np
import pytest

from oracle import cboracle

pytestmark = pytest.mark.gpu


def _ctx(tmp_path, L, q, N, seed=11):
    from mi3gpu.mcmcGPU import MCMCGPU
    g = MCMCGPU((0, 0, None, None), L, q, N, 512, tmp_path, 1024, seed)
    g.initJstep()
    return g


def _ptr(g, name):
    from mi3gpu._lib import lib, check, enc
    p = ctypes.c_void_p()
    check(g._ctx, lib().mi3_buf_device_ptr(g._ctx, enc(name), ctypes.byref(p)))
    return p


def _noalloc(g, nseq, out):
    from mi3gpu._lib import lib, check
    check(g._ctx, lib().mi3_weighted_marg_noalloc(
        g._ctx, _ptr(g, 'seq main'), g.nwalkers, _ptr(g, 'weights'), nseq, _ptr(g, out)))
    return _read(g, out)


def _read(g, name):
    from mi3gpu._lib import lib, check, enc
    n = g.nPairs * g.q * g.q
    a = np.empty(n, '<f4')
    t = ctypes.c_int()
    check(g._ctx, lib().mi3_buf_read_async(g._ctx, enc(name), a.nbytes, 0, ctypes.byref(t)))
    check(g._ctx, lib().mi3_ticket_wait(g._ctx, t.value, a.ctypes.data_as(ctypes.c_void_p),
                                        a.nbytes))
    return a.reshape(g.nPairs, g.q * g.q)


def _existing(g, nseq):
    from mi3gpu._lib import lib, check
    check(g._ctx, lib().mi3_weighted_marg(g._ctx, b'seq main', b'weights', nseq))
    return g.getBuf('bi').read()


@pytest.mark.parametrize("L,q,N,n2", [
    (232, 21, 8192, 5000),   # six runs of the accumulator; ragged second call
    (40, 4, 2048, 1040),     # one run
])
def test_weighted_marg_noalloc_matches(gpu_ok, tmp_path, L, q, N, n2):
    from mi3gpu._lib import lib, check
    print(f"weighted_marg_noalloc L={L} q={q} N={N} then {n2}", flush=True)
    g = _ctx(tmp_path, L, q, N)
    rng = np.random.default_rng(L)
    s = rng.integers(0, q, size=(N, L)).astype('<u1')
    w = rng.exponential(size=N).astype('<f4')
    g.setBuf('seq main', s)
    g.setBuf('weights', w)
    check(g._ctx, lib().mi3_buf_alloc(g._ctx, b'bi noalloc', g.nPairs * q * q * 4))
    for n in (N, n2):
        want = _existing(g, n)
        # poison the output so a bin the entry point failed to write shows
        check(g._ctx, lib().mi3_buf_fill_u32(g._ctx, b'bi noalloc', 0x7fc00000))
        got = _noalloc(g, n, 'bi noalloc')
        np.testing.assert_array_equal(got.view('<u4'), want.view('<u4'))
        Fx = cboracle.weighted_marg(s[:n], w[:n], q)
        occ = Fx > 0
        np.testing.assert_allclose(got[occ], Fx[occ], rtol=2.4e-7, atol=0)
        assert np.all(got[~occ] == 0)


def test_weighted_marg_noalloc_limits(gpu_ok, tmp_path):
    """Requests beyond the scratch limit and bad strides are refused on the
    host, before any launch."""
    from mi3gpu._lib import lib, MI3_EINVAL
    g = _ctx(tmp_path, 20, 4, 512)
    seq, w = _ptr(g, 'seq main'), _ptr(g, 'weights')
    check_bi = _ptr(g, 'bi')
    big = (1 << 20) + 16
    assert lib().mi3_weighted_marg_noalloc(g._ctx, seq, big, w, big, check_bi) == MI3_EINVAL
    assert b'exceeds the allocation-free limit' in lib().mi3_last_error(g._ctx)
    assert lib().mi3_weighted_marg_noalloc(g._ctx, seq, 520, w, 512, check_bi) == MI3_EINVAL
    assert lib().mi3_weighted_marg_noalloc(g._ctx, seq, 512, w, 513, check_bi) == MI3_EINVAL
    assert lib().mi3_weighted_marg_noalloc(g._ctx, None, 512, w, 512, check_bi) == MI3_EINVAL
