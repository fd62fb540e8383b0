"""Host-side logic of the drop-in package and the C ABI surface (CPU only)."""
import bz2
import io
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import potts_oracle as po

REPO = Path(__file__).resolve().parent.parent
GOLDEN = Path(__file__).resolve().parent / 'golden'


# --------------------------------------------------------------------- C ABI

def header_functions():
    text = (REPO / 'include' / 'mi3hip.h').read_text()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(mi3_[A-Za-z0-9_]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    from mi3gpu import _lib
    L = _lib.lib()
    declared = header_functions()
    assert len(declared) >= 35
    for name in declared:
        assert hasattr(L, name), name
    # the ctypes table covers the whole header, and nothing else
    assert sorted(_lib.SIGNATURES) == declared


def test_library_reports_no_device_cleanly():
    from mi3gpu import _lib
    if _lib.device_count() > 0:
        pytest.skip("a device is visible")
    import ctypes
    ctx = ctypes.c_void_p()
    rc = _lib.lib().mi3_create(0, 10, 4, 512, 1, 0, ctypes.byref(ctx))
    assert rc < 0
    assert _lib.lib().mi3_last_error(None)


def test_create_argument_validation_message():
    """Argument checks run before any device access."""
    import ctypes
    from mi3gpu import _lib
    ctx = ctypes.c_void_p()
    assert _lib.lib().mi3_create(0, 10, 4, 500, 1, 0, ctypes.byref(ctx)) == _lib.MI3_EINVAL
    assert b'multiple of 512' in _lib.lib().mi3_last_error(None)
    assert _lib.lib().mi3_create(0, 10, 40, 512, 1, 0, ctypes.byref(ctx)) == _lib.MI3_EUNSUP
    assert _lib.lib().mi3_create(0, 1025, 21, 512, 1, 0, ctypes.byref(ctx)) == _lib.MI3_EUNSUP
    assert b'L <= 1024' in _lib.lib().mi3_last_error(None)
    with pytest.raises(ValueError):
        _lib.check(None, _lib.MI3_EINVAL)


# --------------------------------------------------------------------- formats

def test_potts_common_matches_oracle():
    from mi3gpu.utils import potts_common as pc
    b = np.load(GOLDEN / 'hiv_bimarg_pc.npy')
    assert pc.getLq(b) == (93, 4)
    np.testing.assert_allclose(pc.getUnimarg(b), po.unimarg(b), rtol=1e-5)
    pc.validate_bimarg(b)
    ind = pc.indepF(b)
    x = b.reshape(-1, 4, 4).astype(np.float64)
    np.testing.assert_allclose(ind.reshape(-1, 4, 4),
                               x.sum(2)[:, :, None] * x.sum(1)[:, None, :], rtol=1e-5)


def test_validate_bimarg_rejects_inconsistent():
    from mi3gpu.utils.potts_common import validate_bimarg
    b = np.load(GOLDEN / 'hiv_bimarg_pc.npy').copy()
    b[5, :2] += [0.01, -0.01]   # keeps the row sum, breaks the marginal of site i
    with pytest.raises(ValueError, match='inconsitent'):
        validate_bimarg(b)


def test_gauges_match_oracle_loops():
    from mi3gpu.utils import changeGauge as cg
    rng = np.random.default_rng(1)
    L, q = 7, 5
    h = rng.normal(size=(L, q))
    J = rng.normal(size=(L * (L - 1) // 2, q * q))
    h1, J1 = cg.zeroJGauge(h, J)
    h2, J2 = po.zeroJGauge(h, J)
    np.testing.assert_allclose(h1, h2, atol=1e-12)
    np.testing.assert_allclose(J1, J2, atol=1e-12)
    np.testing.assert_allclose(cg.fieldlessGaugeEven(h, J)[1], po.fieldlessGaugeEven(h, J)[1],
                               atol=1e-12)
    # distributed gauge keeps energies exactly
    _, Jd = cg.fieldlessGaugeDistributed(h, J)
    s = rng.integers(0, q, size=(30, L))
    np.testing.assert_allclose(po.energies(s, Jd),
                               po.energies(s, J) + h[np.arange(L), s].sum(axis=1), atol=1e-10)


def test_pseudocount_matches_oracle():
    from mi3gpu.utils.pseudocount import mutation_pc
    b = np.load(GOLDEN / 'hiv_bimarg_pc.npy').astype(np.float64)
    np.testing.assert_allclose(mutation_pc(b, 100), po.jeffreys_pseudocount(b, 100), rtol=1e-12)


def test_host_energies_and_marginals_utils():
    from mi3gpu.utils.getSeqEnergies import E_potts
    from mi3gpu.utils.getMarginals import getMarginals
    rng = np.random.default_rng(2)
    L, q = 9, 3
    J = rng.normal(size=(L * (L - 1) // 2, q * q)).astype('f4')
    s = rng.integers(0, q, size=(40, L)).astype('u1')
    np.testing.assert_allclose(E_potts(s, J), po.energies(s, J), rtol=1e-12)
    assert np.isclose(E_potts(s[3], J), po.energies(s[3:4], J)[0])
    f, ff = getMarginals(s, q)
    np.testing.assert_allclose(ff, po.bicounts(s, q) / 40)


@pytest.mark.parametrize("zipf", [False, True])
def test_seqload_roundtrip(tmp_path, zipf):
    from mi3gpu.utils.seqload import loadSeqs, writeSeqs
    rng = np.random.default_rng(3)
    s = rng.integers(0, 4, size=(123, 17)).astype('u1')
    fn = tmp_path / 'seqs'
    writeSeqs(fn, s, 'ABCD', zipf=zipf)
    if zipf:
        assert fn.read_bytes()[:3] == b'BZh'
    got, ids, hdr = loadSeqs(fn, alpha='ABCD')
    np.testing.assert_array_equal(got, s)
    assert ids is None
    # ids and headers
    writeSeqs(fn, s[:3], 'ABCD', ids=['a', 'bb', 'ccc'], headers=['hello'], write_param=True)
    got, ids, hdr = loadSeqs(fn, alpha=None)
    np.testing.assert_array_equal(got, s[:3])
    assert list(ids) == [b'a', b'bb', b'ccc']
    assert hdr[0].startswith('PARAM') and hdr[1].strip() == 'hello'


def test_seqload_errors(tmp_path):
    from mi3gpu.utils.seqload import loadSeqs
    fn = tmp_path / 's'
    fn.write_bytes(b"ABCD\nABC\n")
    with pytest.raises(ValueError, match='wrong length'):
        loadSeqs(fn, alpha='ABCD')
    fn.write_bytes(b"ABCD\nABXD\n")
    with pytest.raises(ValueError, match="Invalid residue 'X'"):
        loadSeqs(fn, alpha='ABCD')
    fn.write_bytes(b"ABCD\nDCBA\n//\n")
    assert loadSeqs(fn, alpha='ABCD')[0].shape == (2, 4)


# --------------------------------------------------------------------- drop-in surface

def test_pack_seqs_reference_layout_roundtrip():
    from mi3gpu.mcmcGPU import MCMCGPU
    g = MCMCGPU.__new__(MCMCGPU)
    g.L, g.SWORDS, g.SBYTES = 10, 3, 12
    s = np.random.default_rng(4).integers(0, 21, size=(37, 10)).astype('u1')
    mem = g.packSeqs_4(s)
    assert mem.shape == (3, 37) and mem.dtype == np.dtype('<u4')
    # word w of sequence n holds residues 4w..4w+3, little endian
    assert mem[1, 5] == int.from_bytes(bytes(s[5, 4:8]), 'little')
    np.testing.assert_array_equal(g.unpackSeqs_4(mem), s)


def test_heuristics_match_reference_values():
    from mi3gpu.mcmcGPU import wgsize_heuristic, histogram_heuristic, unpackJ_CPU
    assert wgsize_heuristic(21) == 512 and wgsize_heuristic(4) == 256
    with pytest.raises(Exception):
        wgsize_heuristic(21, '256')
    assert histogram_heuristic(21) == (8, 128)
    assert histogram_heuristic(4) == (256, 512)
    J = np.random.default_rng(5).normal(size=(6, 9)).astype('f4')
    full = unpackJ_CPU(J).reshape(4, 4, 3, 3)
    np.testing.assert_array_equal(full, po.unpack_full(J, 4, 3).astype('f4'))


def test_option_registry_defaults():
    from mi3gpu.Mi3 import optionRegistry
    o = optionRegistry()
    assert o['nsteps']['default'] == 2048 and o['newtonsteps']['default'] == 1024
    assert np.isclose(o['gamma']['default'], 4e-4) and o['damping']['default'] == 0.001
    assert np.isclose(o['fracNeff']['default'], 0.9) and o['reseed']['default'] == 'single_indep'
    assert o['equiltime']['default'] == 'auto' and o['min_equil']['default'] == 64
    assert o['max_equil']['default'] == 32768 and o['mcsteps']['default'] == 64


def test_config_file_roundtrip(tmp_path):
    from mi3gpu.Mi3 import Parser
    p = Parser('t', 't')
    p.addopt(None, 'nwalkers gamma reg preopt outdir config alpha')
    a = p.parse_args(['--nwalkers', '1024', '--gamma', '0.001', '--reg', 'l1z:0.01',
                      '--preopt', '--alpha', 'ABCD'])
    fn = tmp_path / 'config.cfg'
    p.write_config_file(a, [fn])
    b = p.parse_args(['--config', str(fn), '--gamma', '0.002'])
    assert b.nwalkers == 1024 and b.reg == 'l1z:0.01' and b.preopt and b.alpha == 'ABCD'
    assert np.isclose(b.gamma, 0.002)     # command line wins over the file


def test_newton_args_validation(tmp_path):
    from mi3gpu.Mi3 import process_newton_args, Parser
    p = Parser('t', 't')
    p.addopt(None, 'bimarg mcsteps newtonsteps newton_delta fracNeff damping reg '
                   'distribute_jstep gamma preopt reseed')
    b = np.load(GOLDEN / 'hiv_bimarg_pc.npy')
    np.save(tmp_path / 'b8.npy', b.astype('f8'))
    with pytest.raises(Exception, match="f4"):
        process_newton_args(p.parse_args(['--bimarg', str(tmp_path / 'b8.npy')]), lambda *a: None)
    z = b.copy()
    z[0, 0] = 0
    np.save(tmp_path / 'z.npy', z)
    with pytest.raises(Exception, match="0 < f"):
        process_newton_args(p.parse_args(['--bimarg', str(tmp_path / 'z.npy')]), lambda *a: None)
    np.save(tmp_path / 'b.npy', b)
    a = process_newton_args(p.parse_args(['--bimarg', str(tmp_path / 'b.npy'), '--reg',
                                          'l1z:0.0001']), lambda *a: None)
    assert a.reg == 'l1z' and a.regarg == (0.0001,)
    def reg(spec):
        return process_newton_args(p.parse_args(['--bimarg', str(tmp_path / 'b.npy'), '--reg',
                                                 spec]), lambda *a: None)
    # argument formats and defaults of the reference's option parser
    assert reg('SCADJ:1').regarg == (1.0, 4.0)
    assert reg('SCADJ:1:2.5').regarg == (1.0, 2.5)
    with pytest.raises(Exception, match=">= 2.0"):
        reg('SCADJ:1:1.5')
    assert reg('X:0.01').regarg == (0.01,)
    s, r, a_ = reg('SCADX:10:0.1').regarg
    assert (r, a_) == (0.1, 4.0) and abs(s - 2 * 10 / (5 * 0.01)) < 1e-9
    assert reg('expX:0.001').regarg == (0.001,)
    assert reg('ddE:0.5').regarg == (0.5,)
    assert reg('SCADddE:0.5').regarg == (0.5, 0.5)
    assert reg('SCADddE:0.5:0.2').regarg == (0.5, 0.2)
    np.save(tmp_path / 'creg.npy', np.full(b.shape, 0.01, 'f4'))
    assert reg('Xij:' + str(tmp_path / 'creg.npy')).regarg.shape == b.shape
    np.save(tmp_path / 'bad.npy', np.zeros(3, 'f4'))
    with pytest.raises(Exception, match="wrong format"):
        reg('Xij:' + str(tmp_path / 'bad.npy'))
    with pytest.raises(Exception, match="must be one of"):
        process_newton_args(p.parse_args(['--bimarg', str(tmp_path / 'b.npy'), '--reg',
                                          'bogus:1']), lambda *a: None)


def test_independent_couplings_model():
    """'independent' initial couplings reproduce the target unimarginals:
    P(s) prod_i f_i(s_i) (Mi3.py:1177-1179)."""
    from mi3gpu.Mi3 import getCouplings, attrdict
    b = np.load(GOLDEN / 'hiv_bimarg_pc.npy')
    from mi3gpu.utils.potts_common import getUnimarg
    f = getUnimarg(b)[:5]
    L, q = 5, 4
    f = f / f.sum(axis=1, keepdims=True)
    args = attrdict(couplings='independent', init_model=None)
    J, L2, q2 = getCouplings(args, L, q, f, lambda *a: None)
    assert (L2, q2) == (L, q) and J.dtype == np.float32
    import itertools
    st = np.array(list(itertools.product(range(q), repeat=L)))
    E = po.energies(st, J)
    p = np.exp(-(E - E.min()))
    p /= p.sum()
    for i in range(L):
        m = np.bincount(st[:, i], weights=p, minlength=q)
        np.testing.assert_allclose(m, f[i], rtol=1e-5)


def test_newton_stats_helpers():
    import mi3gpu.NewtonSteps as NS
    t = np.array([[0.5, 0.5], [0.005, 0.995]])
    m = np.array([[0.4, 0.6], [0.005, 0.995]])
    ferr, ssr, maxd = NS.bimarg_stats(t, m)
    np.testing.assert_allclose([ferr, ssr, maxd], [np.mean([0.2, 0.2, 0.0]), 0.02, 0.1])
    assert NS.getNeff(np.ones(10)) == 10
    np.testing.assert_allclose(NS.getNeff([1.0, 0.0, 0.0]), 1.0)


def test_no_cpu_fallback_without_library(monkeypatch, tmp_path):
    from mi3gpu import _lib
    monkeypatch.setattr(_lib, '_lib', None)
    monkeypatch.setattr(_lib, 'LIBPATH', tmp_path / 'missing.so')
    with pytest.raises(ImportError, match='no CPU path'):
        _lib.lib()


def test_profile_log_summary():
    from mi3gpu.utils.analyze_gpu_profiling import summarize
    r = summarize(["EVT mcmc 0 2000000000 8000000000", "EVT mcmc 0 1000000000",
                   "noise", "EVT getBuf 5 10 100"])
    assert r['mcmc'][1] == 2 and abs(r['mcmc'][0] - 3.0) < 1e-9 and r['getBuf'][2] == 100


@pytest.mark.parametrize("phi", [0.0, 0.5, 0.8])
def test_tau_int_ar1_closed_form(phi):
    """AR(1) x_t = phi x_{t-1} + e_t has rho(t) = phi^t, so the integrated
    autocorrelation time is (1 + phi) / (2 (1 - phi)); a wrong window or a
    missing 1/2 misses it by far more than the 5% tolerance."""
    from mi3gpu.utils.ess import tau_int, energy_ess
    rng = np.random.default_rng(11)
    T, n = 2000, 400
    x = np.zeros((T, n))
    x[0] = rng.normal(size=n) / np.sqrt(1 - phi ** 2)
    for t in range(1, T):
        x[t] = phi * x[t - 1] + rng.normal(size=n)
    tau, W = tau_int(x)
    expect = (1 + phi) / (2 * (1 - phi))
    assert abs(tau - expect) < 0.05 * expect
    assert W >= 5 * tau
    ess = energy_ess(x)[0]
    assert abs(ess - n * T / (2 * expect)) < 0.06 * n * T / (2 * expect)


def test_parallel_bz2_is_standard_multistream(tmp_path):
    """Multi-stream output decompresses (bz2 module, the stream-by-stream
    decoder) to the exact text, and the seqs file reads back unchanged."""
    import bz2
    from mi3gpu.utils.seqload import bz2_compress_parallel, writeSeqs, loadSeqs
    rng = np.random.default_rng(4)
    seqs = rng.integers(0, 21, size=(60000, 232)).astype('u1')   # ~14 MB of text
    writeSeqs(tmp_path / 's', seqs, zipf=True)
    raw = (tmp_path / 's').read_bytes()
    assert raw[:3] == b'BZh' and raw.count(b'BZh91AY&SY') >= 2     # several streams
    dec = bz2.BZ2Decompressor()
    first = dec.decompress(raw)
    assert dec.eof and len(dec.unused_data) > 0                    # really multi-stream
    text = bz2.decompress(raw)
    assert text.startswith(first) and text.count(b'\n') == 60000
    back = loadSeqs(tmp_path / 's')[0]
    np.testing.assert_array_equal(back, seqs)
    small = b'ACD\n' * 10
    assert bz2.decompress(bz2_compress_parallel(small)) == small


def test_background_status_write_flushes(tmp_path):
    import mi3gpu.NewtonSteps as NS
    from mi3gpu.utils.seqload import loadSeqs
    L, q = 6, 4
    nP = L * (L - 1) // 2
    f = np.full((nP, q * q), 1.0 / (q * q), dtype='f4')
    seqs = np.random.default_rng(1).integers(0, q, size=(512, L)).astype('u1')
    (tmp_path / 'run_0').mkdir()
    NS.writeStatus('run_0', 0, f, None, f, np.zeros_like(f), seqs, np.zeros(512), 'ABCD',
                   None, None, tmp_path, lambda *a: None)
    NS.flush_writes()
    np.testing.assert_array_equal(loadSeqs(tmp_path / 'run_0' / 'seqs', alpha='ABCD')[0], seqs)
    assert (tmp_path / 'run_0' / 'info.txt').exists()
