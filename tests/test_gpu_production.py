"""Parity on the paths that production sizes take, called through the C ABI
(via the mcmcGPU drop-in wrapper):

- bicounts with the sequence range split over several waves (partial counts
  added by integer atomics), bit-exact, up to 2^20 sequences at L=232 q=21;
- one sweep launch over the whole 512-work-group grid of 2^18 walkers;
- weighted marginals at 2^18 sequences with Newton-like weights, every bin
  within 1e-5 relative of the float64 oracle;
- the quasi-Newton loop at L=232 q=21 within 1e-5, beside the float64
  oracle's own error;
- the --beta reweighting branches of the driver and `equiltime auto`.
"""
import ctypes
import types

import numpy as np
import pytest

from oracle import cboracle, potts_oracle as po

pytestmark = pytest.mark.gpu


def model(L, q, seed, sj=0.1):
    rng = np.random.default_rng(seed)
    return rng.normal(0, sj, size=(L * (L - 1) // 2, q * q)).astype('<f4')


def seqs_rand(N, L, q, seed):
    return np.random.default_rng(seed).integers(0, q, size=(N, L)).astype('<u1')


def make_gpu(tmp_path, L, q, N, seed=1234, nsteps=None, wgsize=512):
    from mi3gpu.mcmcGPU import MCMCGPU
    g = MCMCGPU((0, 0, None, None), L, q, N, wgsize, tmp_path, 1024, seed)
    if nsteps is not None:
        g.initMCMC(nsteps, 0, 2 ** 63)
    return g


def ksplit_of(g, nseq):
    from mi3gpu._lib import lib, check
    k = ctypes.c_uint32()
    check(g._ctx, lib().mi3_bicount_split(g._ctx, nseq, ctypes.byref(k)))
    return k.value


@pytest.mark.parametrize("L,q,N,nbuf,split", [
    (50, 21, 8192, None, 4),            # split in 4 parts
    (50, 21, 8192, 8192 + 17, 4),       # ragged last step in the last part
    (99, 4, 2 ** 16, None, 32),         # the largest split, q = 4
    (232, 21, 16384, None, 5),
    (232, 21, 2 ** 18, None, 5),        # production walker count
    (232, 21, 2 ** 20, None, 5),        # standalone-kernel workload
])
def test_bicounts_split_bit_exact(gpu_ok, tmp_path, L, q, N, nbuf, split):
    print(f"bicounts L={L} q={q} N={N} nbuf={nbuf}", flush=True)
    g = make_gpu(tmp_path, L, q, N)
    if nbuf is None:
        s = seqs_rand(N, L, q, 5)
        g.setBuf('seq main', s)
        buf, n = 'main', N
    else:
        g.initLargeBufs(nbuf)
        s = seqs_rand(nbuf, L, q, 6)
        g.setBuf('seq large', s)
        buf, n = 'large', nbuf
    assert ksplit_of(g, n) == split
    g.calcBicounts(buf)
    c = g.getBuf('bicount').read()
    want = cboracle.bicounts(s, q)
    np.testing.assert_array_equal(c, want)
    assert np.all(c.sum(axis=1, dtype=np.uint64) == n)


def test_sweep_production_grid(gpu_ok, tmp_path):
    """Two 48-step launches on 2^18 walkers in 512-walker tiles (512
    work-groups), walker for walker against the oracle chain."""
    L, q, N, nsteps, seed = 232, 21, 2 ** 18, 48, 31
    J = model(L, q, 7, sj=0.1)
    s0 = seqs_rand(N, L, q, 8)
    g = make_gpu(tmp_path, L, q, N, seed=seed, nsteps=nsteps, wgsize=512)
    g.setBuf('J', J)
    g.setBuf('seq main', s0)
    g.runMCMC()
    g.runMCMC()
    got = g.getBuf('seq main').read()
    want, m1 = cboracle.metropolis(s0, J, seed, 0, nsteps)
    want, m2 = cboracle.metropolis(want, J, seed, nsteps, nsteps)
    margin = np.minimum(m1, m2)
    same = np.all(got == want, axis=1)
    safe = margin > 1e-3
    blocks = same.reshape(-1, 512).mean(axis=1)
    print(f"{same.mean() * 100:.3f}% identical, {safe.mean() * 100:.2f}% with margin > 1e-3, "
          f"worst work-group {blocks.min() * 100:.1f}%, changed {np.mean(got != s0):.4f}")
    assert np.all(same[safe])
    assert same.mean() > 0.99
    assert blocks.min() > 0.9
    assert np.mean(got != s0) > 0.05


@pytest.mark.parametrize("beta", [0.5, 1.7])
def test_sweep_beta_matches_oracle(gpu_ok, tmp_path, beta):
    """Acceptance exp(-beta dE) > u at beta != 1 (the reference's BETA,
    mcmc.cl:445-451)."""
    L, q, N, nsteps, seed = 99, 4, 1024, 198, 12
    J = model(L, q, 3, sj=0.3)
    s0 = seqs_rand(N, L, q, 4)
    g = make_gpu(tmp_path, L, q, N, seed=seed, nsteps=nsteps)
    g.beta = beta
    g.setBuf('J', J)
    g.setBuf('seq main', s0)
    g.runMCMC()
    got = g.getBuf('seq main').read()
    want, margin = cboracle.metropolis(s0, J, seed, 0, nsteps, beta=beta)
    same = np.all(got == want, axis=1)
    print(f"beta={beta}: {same.mean() * 100:.2f}% identical")
    assert np.all(same[margin > 1e-3])
    assert same.mean() > 0.98
    # a different beta gives different chains
    other, _ = cboracle.metropolis(s0, J, seed, 0, nsteps, beta=1.0)
    assert np.mean(np.all(other == got, axis=1)) < 0.5


def newton_like_weights(s, q, neff_frac, seed):
    """Weights exp(-(dE - min dE)) of a random coupling perturbation dJ scaled
    so that Neff / N is about neff_frac (Zwanzig weights late in a Newton
    phase)."""
    N, L = s.shape
    rng = np.random.default_rng(seed)
    dJ = rng.normal(0, 1, size=(L * (L - 1) // 2, q * q)).astype('<f4')
    e = cboracle.energies(s, dJ)
    sd = np.sqrt(np.log(1 / neff_frac)) / e.std()
    dE = e * sd
    w = np.exp(-(dE - dE.min())).astype('<f4')
    return w


def seqs_skewed(N, L, q, seed):
    """Residues drawn from p(a) ~ exp(-a/2) under a random relabelling per
    site, so that pair bins range from ~2^-2 N down to single sequences."""
    rng = np.random.default_rng(seed)
    p = np.exp(-0.5 * np.arange(q))
    p /= p.sum()
    s = rng.choice(q, size=(N, L), p=p).astype('<u1')
    perm = np.argsort(rng.random((L, q)), axis=1).astype('<u1')
    return perm[np.arange(L)[None, :], s]


def test_weighted_marg_production_rtol(gpu_ok, tmp_path):
    """2^18 sequences, L=100 q=21, skewed residues (bins from a quarter of
    the sequences down to single ones), Newton-like weights (Neff ~ 0.9 N):
    every bin holding at least one sequence within 1e-5 relative of the
    float64 sums, bins holding none exactly 0."""
    L, q, N = 100, 21, 2 ** 18
    s = seqs_skewed(N, L, q, 21)
    w = newton_like_weights(s, q, 0.9, 22)
    neff = w.astype(np.float64).sum() ** 2 / (w.astype(np.float64) ** 2).sum()
    g = make_gpu(tmp_path, L, q, N)
    g.initJstep()
    g.setBuf('seq main', s)
    g.setBuf('weights', w)
    g.weightedMarg('main')
    F = g.getBuf('bi').read().astype(np.float64)
    Fx = cboracle.weighted_marg(s, w, q)
    cnt = cboracle.bicounts(s, q)
    occ = cnt > 0
    rel = np.abs(F - Fx)[occ] / Fx[occ]
    one = cnt == 1
    rel_one = (np.abs(F - Fx)[one] / Fx[one]).max()
    print(f"Neff/N {neff / N:.3f}; max rel err {rel.max():.3g} over {occ.sum()} bins, "
          f"single-sequence bins {rel_one:.3g} ({one.sum()} bins), "
          f"weight span {w.min():.3g}..{w.max():.3g}")
    assert 0.85 < neff / N < 0.95
    assert one.sum() > 1000
    assert np.all(F[~occ] == 0)
    assert rel.max() < 1e-5


def _newton_problem(L, q, N, seed):
    """Sequences near equilibrium of a model, a target from an independent
    sample of a slightly different model (with pseudocount)."""
    J = model(L, q, seed, 0.05)
    s = cboracle.metropolis(seqs_rand(N, L, q, seed + 1), J, seed, 0, 4 * L)[0]
    J2 = (J + model(L, q, seed + 2, 0.01)).astype('<f4')
    t = cboracle.metropolis(seqs_rand(N, L, q, seed + 3), J2, seed + 4, 0, 4 * L)[0]
    bt = po.jeffreys_pseudocount(cboracle.bicounts(t, q).astype(np.float64) / N, N)
    return J, s, bt.astype('<f4')


def _run_iter_newton(tmp_path, J, s, bt, gamma, pc, steps, frac, reg=None, regarg=None,
                     beta=None):
    import mi3gpu.NewtonSteps as NS
    from mi3gpu.node_manager import GPU_node
    from mi3gpu.Mi3 import attrdict
    N, L = s.shape
    q = int(round(np.sqrt(J.shape[1])))
    g = make_gpu(tmp_path, L, q, N)
    g.initJstep()
    g.beta = beta
    gpus = GPU_node([g])
    gpus.setBuf('J', J)
    gpus.setBuf('bi target', bt)
    gpus.setSeqs('main', s)
    p = attrdict(bimarg=bt, gamma0=gamma, pcdamping=pc, newtonSteps=steps, fracNeff=frac,
                 nwalkers=N, reg=reg, regarg=regarg, distribute_jstep='all', beta=beta,
                 outdir=tmp_path)
    return NS.iterNewton(p, None, gpus, lambda *a: None)


def test_iter_newton_production_rtol(gpu_ok, tmp_path):
    """25 quasi-Newton iterations at L=232 q=21 on 8192 sequences against the
    float64 oracle: J + dJ and the predicted bimarginals within 1e-5
    relative.  The oracle's own error is estimated by running it again on
    the sequences in reverse order (every sum in another order)."""
    L, q, N = 232, 21, 8192
    J, s, bt = _newton_problem(L, q, N, 40)
    gamma, pc, steps, frac = 4e-4, 0.01, 25, 0.5
    i, trialJ, bim = _run_iter_newton(tmp_path, J, s, bt, gamma, pc, steps, frac)

    def oracle(seqs):
        return po.iter_newton(
            seqs, J, bt, gamma, pc, steps, frac,
            energies_fn=lambda ss, dj: cboracle.energies(ss, dj.astype('f4')),
            wmarg_fn=lambda ss, w: cboracle.weighted_marg(ss, w.astype('f4'), q))

    oi, oJ, obim, ow, neffs = oracle(s)
    ri, rJ, rbim, _, _ = oracle(s[::-1].copy())
    dJ_scale = np.abs(oJ - J).max()
    err_J = np.abs(trialJ - oJ) / (np.abs(oJ) + dJ_scale)
    occ = obim > 0
    err_b = np.abs(bim - obim)[occ] / obim[occ]
    own_J = np.abs(rJ - oJ) / (np.abs(oJ) + dJ_scale)
    own_b = np.abs(rbim - obim)[occ] / obim[occ]
    print(f"iterations {i} (oracle {oi}), Neff/N {neffs[-1] / N:.3f}, max |dJ| {dJ_scale:.3g}; "
          f"J+dJ rel err {err_J.max():.3g} (oracle's own {own_J.max():.3g}); "
          f"bimarg rel err {err_b.max():.3g} (oracle's own {own_b.max():.3g})")
    assert i == oi == steps - 1
    assert err_J.max() < 1e-5
    assert err_b.max() < 1e-5
    assert np.all(bim[~occ] == 0)


def test_iter_newton_beta_matches_oracle(gpu_ok, tmp_path):
    """--beta: sequences sampled at beta are reweighted to beta = 1 inside the
    Newton loop (NewtonSteps.py:206-214, 223-224)."""
    L, q, N, beta = 40, 4, 4096, 0.8
    J, s, bt = _newton_problem(L, q, N, 60)
    gamma, pc, steps, frac = 4e-4, 0.01, 12, 0.3
    i, trialJ, bim = _run_iter_newton(tmp_path, J, s, bt, gamma, pc, steps, frac, beta=beta)
    oi, oJ, obim, ow, neffs = po.iter_newton(
        s, J, bt, gamma, pc, steps, frac, beta=beta,
        energies_fn=lambda ss, dj: cboracle.energies(ss, dj.astype('f4')),
        wmarg_fn=lambda ss, w: cboracle.weighted_marg(ss, w.astype('f4'), q))
    dJ_scale = np.abs(oJ - J).max()
    err_J = np.abs(trialJ - oJ) / (np.abs(oJ) + dJ_scale)
    occ = obim > 0
    err_b = np.abs(bim - obim)[occ] / obim[occ]
    print(f"beta {beta}: iterations {i} (oracle {oi}), J+dJ err {err_J.max():.3g}, "
          f"bimarg err {err_b.max():.3g}")
    assert i == oi
    assert err_J.max() < 1e-5
    assert err_b.max() < 1e-5


def _mcmc_param(tmp_path, **kw):
    from mi3gpu.Mi3 import attrdict
    p = dict(equiltime=2, trackequil=0, outdir=tmp_path, beta=None, min_equil=64,
             max_equil=32768, tracked=None)
    p.update(kw)
    return attrdict(p)


def test_run_mcmc_beta_reweighting(gpu_ok, tmp_path):
    """runMCMC with --beta: bimarginals of the walkers reweighted by
    exp((beta-1)(E - min E)) (NewtonSteps.py:468-481, mcmc.cl:635-644)."""
    import mi3gpu.NewtonSteps as NS
    from mi3gpu.node_manager import GPU_node
    L, q, N, beta = 30, 5, 4096, 0.7
    J = model(L, q, 5, 0.2)
    g = make_gpu(tmp_path, L, q, N, seed=3, nsteps=4 * L)
    g.initJstep()
    g.beta = beta
    gpus = GPU_node([g])
    gpus.setSeqs('main', seqs_rand(N, L, q, 6))
    bim, bicount, es, e_rho, _, step = NS.runMCMC(gpus, J, 'run_0', _mcmc_param(
        tmp_path, beta=beta), lambda *a: None)
    assert bicount is None and step == 2
    s = g.getBuf('seq main').read()
    E = cboracle.energies(s, J)
    np.testing.assert_allclose(es, E, rtol=0, atol=1e-5 * np.abs(E).max())
    w = np.exp((beta - 1) * (E - E.min()))
    F = cboracle.weighted_marg(s, w.astype('f4'), q)
    want = F / F.sum(axis=1, keepdims=True)
    occ = want > 0
    err = np.abs(bim - want)[occ] / want[occ]
    print(f"beta reweighted bimarg: max rel err {err.max():.3g}")
    assert err.max() < 1e-5
    assert np.all(bim[~occ] == 0)


def test_run_mcmc_auto_equilibration(gpu_ok, tmp_path):
    """equiltime 'auto' (NewtonSteps.py:395-442): 8 calls, then doubling
    batches with an energy/bimarginal snapshot after each, until the Spearman
    tests pass and at least min_equil calls ran; snapshots are saved with
    --trackequil."""
    import mi3gpu.NewtonSteps as NS
    from mi3gpu.node_manager import GPU_node
    L, q, N = 20, 4, 4096
    J = model(L, q, 9, 0.1)
    g = make_gpu(tmp_path, L, q, N, seed=4, nsteps=2 * L)
    gpus = GPU_node([g])
    gpus.setSeqs('main', seqs_rand(N, L, q, 7))
    logs = []
    (tmp_path / 'run_0').mkdir()
    bim, bicount, es, e_rho, _, step = NS.runMCMC(
        gpus, J, 'run_0', _mcmc_param(tmp_path, equiltime='auto', trackequil=1,
                                      tracked=['bim', 'E']), lambda *a: logs.append(a))
    text = '\n'.join(' '.join(map(str, a)) for a in logs)
    print(text)
    assert step >= 64 and 'Equilibrated.' in text
    # snapshots after 16, 32, 64, ... calls
    for k in (16, 32, 64):
        assert (tmp_path / 'run_0' / 'equilibration' / f'energies_{k}.npy').exists()
        assert (tmp_path / 'run_0' / 'equilibration' / f'bimarg_{k}.npy').exists()
    assert len(e_rho) >= 3
    np.testing.assert_array_equal(bicount, cboracle.bicounts(g.getBuf('seq main').read(), q))
    assert g.mcmc_steps_done() == step * 2 * L


def test_infer_cli_auto_equiltime(gpu_ok, tmp_path):
    """One `Mi3.py infer` round on the HIV fixture with the reference's
    default `--equiltime auto`."""
    from pathlib import Path
    import mi3gpu.Mi3 as Mi3
    golden = Path(__file__).resolve().parent / 'golden'
    out = tmp_path / 'hiv'
    Mi3.main(['infer', '--init_model', 'independent', '--bimarg',
              str(golden / 'hiv_bimarg_pc.npy'), '--alpha', 'ABCD', '--nwalkers', '8192',
              '--damping', '0.01', '--mcsteps', '1', '--reg', 'l1z:0.0001', '--nsteps', '186',
              '--newtonsteps', '64', '--rngseed', '3', '--outdir', str(out)])
    log = (out / 'log').read_text()
    assert 'Equilibrated.' in log or 'Reached Max Steps' in log
    assert (out / 'run_0' / 'J.npy').exists()
