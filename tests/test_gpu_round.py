"""Ensemble-level parity of the GPU sampler and end-to-end runs of the
drop-in command line on the GPU."""
import itertools
from pathlib import Path

import numpy as np
import pytest
from scipy import stats

from oracle import cboracle, potts_oracle as po

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / 'golden'


def make_gpu(tmp_path, L, q, N, seed, nsteps, wgsize=512):
    from mi3gpu.mcmcGPU import MCMCGPU
    g = MCMCGPU((0, 0, None, None), L, q, N, wgsize, tmp_path, 1024, seed)
    g.initMCMC(nsteps, 0, 2 ** 63)
    return g


def test_sampler_exact_enumeration(gpu_ok, tmp_path):
    """81-state model: the GPU ensemble matches exp(-E)/Z (chi-square)."""
    L, q, N = 4, 3, 2 ** 16
    J = np.random.default_rng(13).normal(0, 0.6, size=(6, 9)).astype('f4')
    states = np.array(list(itertools.product(range(q), repeat=L)))
    E = po.energies(states, J)
    p = np.exp(-(E - E.min()))
    p /= p.sum()
    g = make_gpu(tmp_path, L, q, N, 5, 400)
    g.setBuf('J', J)
    g.setBuf('seq main', np.random.default_rng(1).integers(0, q, size=(N, L)).astype('u1'))
    g.runMCMC()
    s = g.getBuf('seq main').read().astype(int)
    obs = np.bincount((s * q ** np.arange(L - 1, -1, -1)).sum(axis=1), minlength=q ** L)
    res = stats.chisquare(obs, p * N)
    print("chi2 p-value", res.pvalue)
    assert res.pvalue > 1e-4


def merged_chi2(c1, c2, minexp=5):
    """Two-sample chi-square over one pair's q*q bins, pooling sparse bins."""
    tot = c1 + c2
    order = np.argsort(tot)
    keep = tot[order] >= 2 * minexp
    a = np.append(c1[order][keep], c1[order][~keep].sum())
    b = np.append(c2[order][keep], c2[order][~keep].sum())
    m = (a + b) > 0
    table = np.vstack([a[m], b[m]])
    if table.shape[1] < 2:
        return 1.0
    return stats.chi2_contingency(table, correction=False)[1]


def test_ensemble_matches_cpu_port(gpu_ok, tmp_path):
    """Equilibrated GPU and CPU ensembles of one model agree within sampling
    error: SSR near its floor 2*sum f(1-f)/N and per-pair chi-square
    p-values uniform."""
    L, q, N = 24, 5, 8192
    rng = np.random.default_rng(2)
    J = rng.normal(0, 0.15, size=(L * (L - 1) // 2, q * q)).astype('f4')
    s0 = rng.integers(0, q, size=(N, L)).astype('u1')
    nsteps = 200 * L
    g = make_gpu(tmp_path, L, q, N, 11, nsteps)
    g.setBuf('J', J)
    g.setBuf('seq main', s0)
    g.runMCMC()
    g.calcBicounts('main')
    cg = g.getBuf('bicount').read().astype(np.int64)
    sc = cboracle.metropolis_f32(s0, J, 977, 0, nsteps)
    cc = cboracle.bicounts(sc, q).astype(np.int64)
    fg, fc = cg / N, cc / N
    ssr = np.sum((fg - fc) ** 2)
    f = (fg + fc) / 2
    floor = 2 * np.sum(f * (1 - f)) / N
    pv = np.array([merged_chi2(a, b) for a, b in zip(cg, cc)])
    print(f"SSR {ssr:.4g} floor {floor:.4g}; pair p<0.01: {np.mean(pv < 0.01):.3f}")
    assert ssr < 1.5 * floor
    assert np.mean(pv < 0.01) < 0.03
    assert stats.kstest(pv, 'uniform').pvalue > 1e-4


def test_infer_cli_hiv(gpu_ok, tmp_path):
    """Three rounds of `Mi3.py infer` on the reference's HIV fixture."""
    import mi3gpu.Mi3 as Mi3
    out = tmp_path / 'hiv'
    args = ['infer', '--init_model', 'independent', '--bimarg', str(GOLDEN / 'hiv_bimarg_pc.npy'),
            '--alpha', 'ABCD', '--nwalkers', '16384', '--damping', '0.01', '--mcsteps', '3',
            '--reg', 'l1z:0.0001', '--nsteps', str(93 * 4), '--equiltime', '16',
            '--newtonsteps', '256', '--rngseed', '7', '--outdir', str(out)]
    p = Mi3.main(args)
    for r in ('run_0', 'run_1', 'run_2'):
        for f in ('J.npy', 'info.txt', 'seqs', 'perturbedJ.npy', 'predictedBimarg.npy',
                  'nsteps', 'jstep', 'newtonsteps'):
            assert (out / r / f).exists(), (r, f)
    log = (out / 'log').read_text()
    assert 'Total MCMC running time' in log and 'Performed' in log
    ssr = [float(l.split('SSR:')[1].split()[0]) for l in log.splitlines() if 'Error: SSR' in l]
    print("SSR per round", ssr)
    assert len(ssr) == 3 and ssr[2] < ssr[0]
    assert (out / 'config.cfg').read_text().count('nwalkers') == 1
    J = np.load(out / 'run_2' / 'J.npy')
    assert J.shape == (4278, 16) and J.dtype == np.float32 and np.isfinite(J).all()


def test_gen_energies_benchmark_modes(gpu_ok, tmp_path):
    import mi3gpu.Mi3 as Mi3
    from mi3gpu.utils.seqload import loadSeqs
    L, q = 20, 4
    J = np.random.default_rng(3).normal(0, 0.2, size=(L * (L - 1) // 2, q * q)).astype('f4')
    np.save(tmp_path / 'J.npy', J)
    Mi3.main(['gen', '--init_model', 'zero', '--couplings', str(tmp_path / 'J.npy'),
              '--alpha', 'ABCD',
              '--nwalkers', '1024', '--nsteps', '200', '--equiltime', '4', '--seedseq', 'AAAA' * 5,
              '--outdir', str(tmp_path / 'gen'), '--rngseed', '1'])
    seqs = loadSeqs(tmp_path / 'gen' / 'seqs', alpha='ABCD')[0]
    assert seqs.shape == (1024, L)
    E = np.load(tmp_path / 'gen' / 'energies.npy')
    np.testing.assert_allclose(E, po.energies(seqs, J), atol=1e-4)
    es = Mi3.main(['energies', str(tmp_path / 'E.npy'), '--couplings', str(tmp_path / 'J.npy'),
                   '--alpha', 'ABCD', '--seqs', str(tmp_path / 'gen' / 'seqs'),
                   '--outdir', str(tmp_path / 'en')])
    np.testing.assert_allclose(np.load(tmp_path / 'E.npy'), po.energies(seqs, J), atol=1e-4)
    rate = Mi3.main(['benchmark', '--nloop', '3', '--couplings', str(tmp_path / 'J.npy'),
                     '--alpha', 'ABCD', '--nwalkers', '2048', '--nsteps', '100',
                     '--seedseq', 'ABCD' * 5, '--outdir', str(tmp_path / 'bench')])
    assert rate > 0
    assert 'MC steps per second' in (tmp_path / 'bench' / 'log').read_text()


@pytest.mark.parametrize("iterback", [False, True])
def test_subseq_mode(gpu_ok, tmp_path, iterback):
    """`Mi3.py subseq`: log relative subsequence frequencies from
    substituted energies, against the fp64 energies oracle."""
    import mi3gpu.Mi3 as Mi3
    from mi3gpu.utils.seqload import writeSeqs
    from scipy.special import logsumexp
    L, q = 10, 4
    rng = np.random.default_rng(6)
    J = rng.normal(0, 0.3, size=(L * (L - 1) // 2, q * q)).astype('f4')
    back = rng.integers(0, q, size=(700, L)).astype('u1')
    subs = rng.integers(0, q, size=(40, L)).astype('u1')
    np.save(tmp_path / 'J.npy', J)
    writeSeqs(tmp_path / 'back', back, alpha='ABCD')
    writeSeqs(tmp_path / 'subs', subs, alpha='ABCD')
    fix = [2, 5, 9]
    args = ['subseq', ','.join(map(str, fix)), str(tmp_path / 'back'), str(tmp_path / 'subs'),
            '--couplings', str(tmp_path / 'J.npy'), '--alpha', 'ABCD',
            '--outdir', str(tmp_path / ('out%d' % iterback))]
    if iterback:
        args.append('--iterbackgrounds')
    got = Mi3.main(args)
    Eb = po.energies(back, J)
    if not iterback:
        want = np.empty(len(subs))
        for n, s in enumerate(subs):
            h = back.copy()
            h[:, fix] = s[fix]
            want[n] = logsumexp(Eb - po.energies(h, J))
    else:
        want = np.full(len(subs), -np.inf)
        other = [i for i in range(L) if i not in fix]
        for n, b in enumerate(back):
            h = subs.copy()
            h[:, other] = b[other]
            want = np.logaddexp(want, Eb[n] - po.energies(h, J))
    np.testing.assert_allclose(got, want, atol=2e-4, rtol=1e-5)
    assert (tmp_path / ('out%d' % iterback) / 'probs.npy').exists()
