"""Parity of every HIP kernel of the inference round against the CPU oracle,
called through the C ABI (via the mcmcGPU drop-in wrapper)."""
import numpy as np
import pytest

from oracle import cboracle, potts_oracle as po, philox

pytestmark = pytest.mark.gpu


def model(L, q, seed, sj=0.1):
    rng = np.random.default_rng(seed)
    return rng.normal(0, sj, size=(L * (L - 1) // 2, q * q)).astype('<f4')


def seqs_rand(N, L, q, seed):
    return np.random.default_rng(seed).integers(0, q, size=(N, L)).astype('<u1')


def make_gpu(tmp_path, L, q, N, seed=1234, nsteps=None, profile=False, wgsize=None):
    from mi3gpu.mcmcGPU import MCMCGPU, wgsize_heuristic
    g = MCMCGPU((0, 0, None, None), L, q, N, wgsize or wgsize_heuristic(q), tmp_path, 1024,
                seed, profile=profile)
    if nsteps is not None:
        g.initMCMC(nsteps, 0, 2 ** 63)
    return g


SHAPES = [(99, 4), (232, 21), (37, 7), (5, 3), (300, 21), (64, 21), (33, 2), (450, 4),
          (700, 21), (1000, 4), (400, 21), (500, 17)]


@pytest.mark.parametrize("L,q", SHAPES)
def test_energies_match_oracle(gpu_ok, tmp_path, L, q):
    N = 512
    J = model(L, q, 1)
    s = seqs_rand(N, L, q, 2)
    g = make_gpu(tmp_path, L, q, N)
    g.setBuf('J', J)
    g.setBuf('seq main', s)
    g.calcEnergies('main')
    E = g.getBuf('E main').read()
    Ex = cboracle.energies(s, J)
    scale = cboracle.energies(s, np.abs(J))
    err = np.abs(E - Ex) / scale
    print(f"L={L} q={q} max rel err (vs sum|terms|) {err.max():.3g}")
    assert err.max() < 1e-5


def test_energies_four_words_ring_many_groups(gpu_ok, tmp_path):
    """Four state words at q=21 (the energy kernel's overlapping chunk ring)
    over five work-groups."""
    L, q, N = 420, 21, 2560
    J = model(L, q, 4)
    s = seqs_rand(N, L, q, 5)
    g = make_gpu(tmp_path, L, q, N)
    g.setBuf('J', J)
    g.setBuf('seq main', s)
    g.calcEnergies('main')
    E = g.getBuf('E main').read()
    Ex = cboracle.energies(s, J)
    scale = cboracle.energies(s, np.abs(J))
    err = np.abs(E - Ex) / scale
    print(f"L={L} q={q} N={N} max rel err (vs sum|terms|) {err.max():.3g}")
    assert err.max() < 1e-5


def test_wm_atomic_probe_runs_and_checks_sizes(gpu_ok, tmp_path):
    """The weighted marginals' atomic-ceiling probe runs in bounds (q=21, the
    largest histogram) and refuses sizes outside its documented limits."""
    from mi3gpu._lib import lib
    g = make_gpu(tmp_path, 40, 21, 512, profile=True)
    assert lib().mi3_wm_atomic_probe(g._ctx, 8, 16) == 0
    g.wait()
    ms, n = g.kernel_ms('wmAtomicProbe')
    assert n == 1 and ms > 0
    assert lib().mi3_wm_atomic_probe(g._ctx, 0, 16) != 0
    assert lib().mi3_wm_atomic_probe(g._ctx, 8, 0) != 0
    assert lib().mi3_wm_atomic_probe(g._ctx, 65537, 16) != 0
    assert lib().mi3_wm_atomic_probe(g._ctx, 8, (1 << 20) + 1) != 0


def test_energies_special_cases(gpu_ok, tmp_path):
    L, q, N = 40, 5, 512
    g = make_gpu(tmp_path, L, q, N)
    s = seqs_rand(N, L, q, 3)
    g.setBuf('seq main', s)
    g.setBuf('J', np.zeros((L * (L - 1) // 2, q * q), '<f4'))
    g.calcEnergies('main')
    assert np.all(g.getBuf('E main').read() == 0)
    g.setBuf('J', np.full((L * (L - 1) // 2, q * q), 0.5, '<f4'))
    g.calcEnergies('main')
    np.testing.assert_allclose(g.getBuf('E main').read(), 0.5 * L * (L - 1) / 2, rtol=1e-6)
    # a single nonzero entry: pair (3, 17), letters (2, 4)
    J = np.zeros((L * (L - 1) // 2, q * q), '<f4')
    n = po.pair_list(L).index((3, 17))
    J[n, 2 * q + 4] = 1.5
    g.setBuf('J', J)
    g.calcEnergies('main')
    want = np.where((s[:, 3] == 2) & (s[:, 17] == 4), 1.5, 0.0)
    np.testing.assert_array_equal(g.getBuf('E main').read(), want.astype('f4'))


@pytest.mark.parametrize("L,q,N", [(99, 4, 512), (232, 21, 1024), (37, 7, 512), (5, 3, 512),
                                   (300, 21, 512), (20, 25, 1024), (700, 21, 512)])
def test_bicounts_bit_exact(gpu_ok, tmp_path, L, q, N):
    s = seqs_rand(N, L, q, 5)
    g = make_gpu(tmp_path, L, q, N)
    g.setBuf('seq main', s)
    g.calcBicounts('main')
    c = g.getBuf('bicount').read()
    want = cboracle.bicounts(s, q)
    assert c.dtype == np.uint32
    np.testing.assert_array_equal(c, want)
    assert np.all(c.sum(axis=1) == N)
    g.bicounts_to_bimarg('main')
    np.testing.assert_array_equal(g.getBuf('bi').read(), po.bicounts_to_bimarg(want, N))


def test_bicounts_ragged_large_buffer(gpu_ok, tmp_path):
    L, q, N = 50, 21, 512
    g = make_gpu(tmp_path, L, q, N)
    g.initLargeBufs(1000)
    s = seqs_rand(777, L, q, 6)
    g.setBuf('seq large', s)
    g.calcBicounts('large')
    np.testing.assert_array_equal(g.getBuf('bicount').read(), cboracle.bicounts(s, q))
    g.calcEnergies('large')
    E = g.getBuf('E large').read()
    assert E.shape == (777,)
    J = np.zeros((L * (L - 1) // 2, q * q), '<f4')
    np.testing.assert_array_equal(E, 0)


@pytest.mark.parametrize("L,q,N,nsteps,wg", [(99, 4, 512, 198, 128), (99, 4, 1024, 198, 256),
                                             (232, 21, 512, 64, 512), (232, 21, 1024, 64, 128),
                                             (37, 7, 512, 111, 256), (300, 21, 512, 30, 512),
                                             (8, 3, 512, 81, 512), (256, 5, 512, 40, 512),
                                             (300, 21, 1024, 30, 256), (400, 4, 512, 20, 256),
                                             (500, 4, 512, 12, 128), (350, 21, 1024, 20, 512),
                                             (384, 4, 512, 16, 512), (450, 21, 512, 12, 512),
                                             (420, 5, 512, 12, 512), (600, 21, 512, 24, 512),
                                             (700, 21, 512, 24, 256), (1000, 4, 512, 40, 128)])
def test_sweep_matches_oracle_trajectories(gpu_ok, tmp_path, L, q, N, nsteps, wg):
    J = model(L, q, 7, sj=0.3 if q < 10 else 0.1)
    s0 = seqs_rand(N, L, q, 8)
    seed = 99
    g = make_gpu(tmp_path, L, q, N, seed=seed, nsteps=nsteps, wgsize=wg)
    g.setBuf('J', J)
    g.setBuf('seq main', s0)
    g.runMCMC()
    g.runMCMC()
    got = g.getBuf('seq main').read()
    t1 = nsteps + (nsteps & 1)
    want, m1 = cboracle.metropolis(s0, J, seed, 0, nsteps)
    want, m2 = cboracle.metropolis(want, J, seed, t1, nsteps)
    margin = np.minimum(m1, m2)
    same = np.all(got == want, axis=1)
    safe = margin > 1e-3
    print(f"L={L} q={q}: {same.mean() * 100:.2f}% identical, {safe.mean() * 100:.1f}% "
          f"walkers with margin > 1e-3, changed sites {np.mean(got != s0):.3f}")
    assert np.all(same[safe])
    assert same.mean() > 0.98
    assert np.mean(got != s0) > 0.01


def test_gen_indep_bit_exact(gpu_ok, tmp_path):
    L, q, N = 61, 21, 1024
    rng = np.random.default_rng(3)
    f = rng.dirichlet(np.ones(q), size=L)
    g = make_gpu(tmp_path, L, q, N, seed=77, nsteps=10)
    g.prepare_indep(f)
    cp = g.getBuf('cprob').read()
    np.testing.assert_array_equal(cp, po.prepare_indep(f))
    g.gen_indep('main')
    a = g.getBuf('seq main').read()
    g.gen_indep('main')
    b = g.getBuf('seq main').read()
    np.testing.assert_array_equal(a, cboracle.gen_indep(cp, 77, 0, N))
    np.testing.assert_array_equal(b, cboracle.gen_indep(cp, 77, 1, N))


def test_fill_seqs_and_roundtrip(gpu_ok, tmp_path):
    L, q, N = 45, 6, 512
    g = make_gpu(tmp_path, L, q, N)
    s = seqs_rand(N, L, q, 1)
    g.setBuf('seq main', s)
    np.testing.assert_array_equal(g.getBuf('seq main').read(), s)
    g.fillSeqs(s[3])
    np.testing.assert_array_equal(g.getBuf('seq main').read(), np.tile(s[3], (N, 1)))


def _newton_setup(tmp_path, L, q, N, seed=4):
    rng = np.random.default_rng(seed)
    J = model(L, q, seed, 0.05)
    s = seqs_rand(N, L, q, seed + 1)
    bt = rng.dirichlet(np.ones(q * q), size=L * (L - 1) // 2).astype('<f4')
    bi = rng.dirichlet(np.ones(q * q), size=L * (L - 1) // 2).astype('<f4')
    dJ = (rng.normal(0, 0.01, size=J.shape)).astype('<f4')
    g = make_gpu(tmp_path, L, q, N)
    g.initJstep()
    g.setBuf('J', J)
    g.setBuf('seq main', s)
    g.setBuf('bi target', bt)
    g.setBuf('bi', bi)
    g.setBuf('dJ', dJ)
    return g, J, s, bt, bi, dJ


@pytest.mark.parametrize("L,q", [(30, 21), (99, 4), (12, 25)])
def test_updateJ_bit_exact(gpu_ok, tmp_path, L, q):
    g, J, s, bt, bi, dJ = _newton_setup(tmp_path, L, q, 1024)
    g.updateJ(4e-4, 1e-3)
    got = g.getBuf('dJ').read()
    # float32 arithmetic in the reference's operation order (mcmc.cl:842)
    want = dJ - np.float32(4e-4) * (bt - bi) / (bi + np.float32(1e-3))
    np.testing.assert_array_equal(got, want.astype('<f4'))
    np.testing.assert_allclose(got, po.updatedJ(bt, bi, 4e-4, 1e-3, dJ), rtol=1e-5, atol=1e-9)


@pytest.mark.parametrize("reg", ['l1z', 'l2z'])
@pytest.mark.parametrize("fused", [True, False])
def test_update_reg_matches_oracle(gpu_ok, tmp_path, reg, fused):
    L, q, lam = 30, 21, 0.05
    g, J, s, bt, bi, dJ = _newton_setup(tmp_path, L, q, 512)
    if fused:
        g.updateJ(4e-4, 1e-3)
        getattr(g, 'reg_' + reg)(4e-4, 1e-3, lam)
        d1 = po.updatedJ(bt, bi, 4e-4, 1e-3, dJ)
    else:
        getattr(g, 'reg_' + reg)(4e-4, 1e-3, lam)
        d1 = dJ.astype(np.float64)
    got = g.getBuf('dJ').read()
    fn = po.reg_l1z if reg == 'l1z' else po.reg_l2z
    want, amb = fn(bi, 4e-4, 1e-3, lam, J, d1)
    ok = ~amb
    # relative to the magnitude of the pair block (J + dJ), since single
    # entries of a zero-gauged block can be arbitrarily close to 0
    scale = (np.abs(J) + np.abs(want)).max(axis=1, keepdims=True) * np.ones_like(want)
    err = np.abs(got - want)[ok] / scale[ok]
    print(f"{reg} fused={fused}: max rel err {err.max():.3g}, ambiguous {amb.mean():.4f}")
    assert err.max() < 1e-5
    assert amb.mean() < 0.01


@pytest.mark.parametrize("L,q,N", [(70, 21, 1024), (130, 4, 1536)])
def test_weighted_marg_and_renorm(gpu_ok, tmp_path, L, q, N):
    g, J, s, bt, bi, dJ = _newton_setup(tmp_path, L, q, N)
    w = np.random.default_rng(1).exponential(size=N).astype('<f4')
    g.setBuf('weights', w)
    g.weightedMarg('main')
    F = g.getBuf('bi').read()
    Fx = cboracle.weighted_marg(s, w, q)
    occ = Fx > 0
    # exact 64-bit fixed-point sums rounded once to float: within an ulp
    np.testing.assert_allclose(F[occ], Fx[occ], rtol=2.4e-7, atol=0)
    assert np.all(F[~occ] == 0)
    g.renormalize_bimarg()
    np.testing.assert_allclose(g.getBuf('bi').read(), po.renormalize(Fx), rtol=1e-5, atol=1e-7)
    # unit weights reduce to the counts exactly (the bin sums are c * 2^62/N
    # to within 2^-53 relative, which rounds to the float of the count c)
    g.setBuf('weights', np.ones(N, '<f4'))
    g.weightedMarg('main')
    counts = cboracle.bicounts(s, q).astype('f4')
    np.testing.assert_array_equal(g.getBuf('bi').read(), counts)


def test_weighted_marg_ragged_large(gpu_ok, tmp_path):
    """A sequence count that is not a multiple of the kernel's groups (the
    tail path), through the large buffers."""
    L, q, N, n = 70, 21, 512, 1000
    g, J, s0, bt, bi, dJ = _newton_setup(tmp_path, L, q, N)
    g.initLargeBufs(n)
    s = seqs_rand(n, L, q, 9)
    g.setBuf('seq large', s)
    w = np.random.default_rng(3).exponential(size=n).astype('<f4')
    g.setBuf('weights large', w)
    g.weightedMarg('large')
    F = g.getBuf('bi').read()
    Fx = cboracle.weighted_marg(s, w, q)
    occ = Fx > 0
    np.testing.assert_allclose(F[occ], Fx[occ], rtol=2.4e-7, atol=0)
    assert np.all(F[~occ] == 0)


@pytest.mark.parametrize("L", [2, 3, 32, 33, 95])
def test_weighted_marg_unit_edges(gpu_ok, tmp_path, L):
    """Work units of 32 consecutive pairs: one pair (L=2, 31 idle columns), a
    unit spanning every row (L=3), units ending exactly on a row (L=32, 33
    with 496 and 528 pairs) and the short last rows of L=95; 991 sequences
    leave a tail of 31 after the last whole group of 32."""
    q, n = 5, 991
    g = make_gpu(tmp_path, L, q, 512)
    g.initJstep()
    g.initLargeBufs(n)
    s = seqs_rand(n, L, q, 11)
    g.setBuf('seq large', s)
    w = np.random.default_rng(12).exponential(size=n).astype('<f4')
    g.setBuf('weights large', w)
    g.weightedMarg('large')
    F = g.getBuf('bi').read()
    Fx = cboracle.weighted_marg(s, w, q)
    occ = Fx > 0
    np.testing.assert_allclose(F[occ], Fx[occ], rtol=2.4e-7, atol=0)
    assert np.all(F[~occ] == 0)
    # unit weights: the counts exactly
    g.setBuf('weights large', np.ones(n, '<f4'))
    g.weightedMarg('large')
    np.testing.assert_array_equal(g.getBuf('bi').read(), cboracle.bicounts(s, q).astype('f4'))


def test_min_weights_stats(gpu_ok, tmp_path):
    L, q, N = 20, 4, 4096
    g = make_gpu(tmp_path, L, q, N)
    g.initJstep()
    rng = np.random.default_rng(2)
    E = rng.normal(0, 2, size=N).astype('<f4')
    g.setBuf('E main', E)
    g.min_buf('E main')
    assert g.getBuf('minout').read()[0] == np.float32(po.min_floats(E))
    g.setBuf('E main', np.abs(E) + 1)
    g.min_buf('E main')
    assert g.getBuf('minout').read()[0] == 0.0   # the reference's min(0, .) quirk
    g.setBuf('E main', E)
    g.dE_to_weights('main', -0.25)
    w = g.getBuf('weights').read()
    np.testing.assert_allclose(w, po.dE_to_weights(E, -0.25), rtol=2e-6)
    g.weight_statistics('main')
    st = g.getBuf('weightstats').read()
    s1, s2 = po.weight_stats(w)
    np.testing.assert_allclose(st, [s1, s2], rtol=1e-6)


@pytest.mark.parametrize("reg", [None, 'l1z'])
def test_iter_newton_matches_oracle(gpu_ok, tmp_path, reg):
    """Several full quasi-Newton iterations through the drop-in driver."""
    import mi3gpu.NewtonSteps as NS
    from mi3gpu.node_manager import GPU_node
    from mi3gpu.Mi3 import attrdict
    L, q, N = 24, 4, 2048
    J = model(L, q, 11, 0.2)
    s = cboracle.metropolis(seqs_rand(N, L, q, 1), J, 5, 0, 400)[0]
    bt = po.jeffreys_pseudocount(po.bicounts_to_bimarg(cboracle.bicounts(
        seqs_rand(4096, L, q, 9), q), 4096).astype(np.float64), 4096).astype('<f4')
    g = make_gpu(tmp_path, L, q, N)
    g.initJstep()
    gpus = GPU_node([g])
    gpus.setBuf('J', J)
    gpus.setBuf('bi target', bt)
    gpus.setSeqs('main', s)
    p = attrdict(bimarg=bt, gamma0=0.004, pcdamping=0.01, newtonSteps=25, fracNeff=0.5,
                 nwalkers=N, reg=reg, regarg=(0.01,) if reg else None,
                 distribute_jstep='all', beta=None, outdir=tmp_path)
    logs = []
    i, trialJ, bim = NS.iterNewton(p, None, gpus, lambda *a: logs.append(a))
    oi, oJ, obim, ow, neffs = po.iter_newton(
        s, J, bt, 0.004, 0.01, 25, 0.5, reg=reg, reg_lambda=0.01,
        energies_fn=lambda ss, dj: cboracle.energies(ss, dj.astype('f4')),
        wmarg_fn=lambda ss, w: cboracle.weighted_marg(ss, w.astype('f4'), q))
    print("iterations", i, oi, "neff", neffs[-1])
    assert i == oi
    dJ_scale = np.abs(oJ - J).max()
    err_J = np.abs(trialJ - oJ) / (np.abs(oJ) + dJ_scale)
    occ = obim > 0
    err_b = np.abs(bim - obim)[occ] / obim[occ]
    print(f"J+dJ rel err {err_J.max():.3g}, bimarg rel err {err_b.max():.3g}")
    assert err_J.max() < 1e-5
    assert err_b.max() < 1e-5


REG_CASES = [
    ('SCADJ', (0.05, 3.0)),
    ('X', (0.01,)),
    ('X', (1e3,)),          # clipped at |X|/(Xnorm gamma) everywhere
    ('Xij', None),
    ('SCADX', (1.0, 0.002, 3.0)),
    ('expX', (0.002,)),
    ('ddE', (0.01,)),
    ('SCADddE', (0.01, 0.05)),
]


@pytest.mark.parametrize("L,q", [(12, 21), (20, 4)])
@pytest.mark.parametrize("reg,arg", REG_CASES)
def test_other_regularisers_match_oracle(gpu_ok, tmp_path, L, q, reg, arg):
    gamma, pc = 4e-4, 1e-3
    g, J, s, bt, bi, dJ = _newton_setup(tmp_path, L, q, 512)
    nP = J.shape[0]
    lam = None
    if reg == 'Xij':
        creg = np.zeros_like(J)
        lam = np.random.default_rng(9).uniform(0.001, 0.05, nP).astype('<f4')
        creg.ravel()[:nP] = lam       # the kernel reads lambdas[pair]
        g.setBuf('Creg', creg)
        g.reg_Xij(gamma, pc)
    else:
        getattr(g, 'reg_' + reg)(gamma, pc, *arg)
    got = g.getBuf('dJ').read().astype(np.float64)
    d0 = dJ.astype(np.float64)
    amb = np.zeros(J.shape, bool)
    if reg == 'SCADJ':
        want, amb = po.reg_SCADJ(bi, gamma, pc, *arg, J, d0)
    elif reg == 'X':
        want = po.reg_X(bi, gamma, pc, arg[0], J, d0)
    elif reg == 'Xij':
        want = po.reg_X(bi, gamma, pc, None, J, d0, lambdas=lam)
    elif reg == 'SCADX':
        want = po.reg_SCADX(bi, gamma, pc, *arg, J, d0)
    elif reg == 'expX':
        want, amb = po.reg_expX(bi, gamma, pc, arg[0], J, d0)
    elif reg == 'ddE':
        want, amb = po.reg_ddE(bi, gamma, pc, arg[0], J, d0)
    else:
        want, amb = po.reg_ddE(bi, gamma, pc, arg[0], J, d0, scad_r=arg[1])
    ok = ~amb
    step_w, step_g = want - d0, got - d0
    assert np.abs(step_w).max() > 0, "the case takes no step"
    bscale = np.maximum(np.abs(want), np.abs(d0)).max(axis=1, keepdims=True) * np.ones_like(want)
    sscale = np.abs(step_w).max(axis=1, keepdims=True) * np.ones_like(want) + 1e-30
    err = (np.abs(got - want) / bscale)[ok]
    # the step relative to the block's largest step, beyond the float32
    # resolution of the stored dJ (4 ulp of |dJ|)
    ulp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
    serr = (np.maximum(np.abs(step_g - step_w) - 4 * ulp, 0) / sscale)[ok]
    print(f"{reg}{arg}: max err {err.max():.3g} (block scale), step err {serr.max():.3g}, "
          f"ambiguous {amb.mean():.4f}")
    assert err.max() < 1e-5
    assert serr.max() < 2e-3
    assert amb.mean() < 0.02


def test_store_restore_and_copy_subseq(gpu_ok, tmp_path):
    """storeSeqs (device and host sources), restoreSeqs and copySubseq on the
    site-major buffers, against numpy."""
    L, q, N = 37, 5, 512
    g = make_gpu(tmp_path, L, q, N)
    g.initLargeBufs(3 * N + 100)
    g.initSubseq()
    a = seqs_rand(N, L, q, 1)
    b = seqs_rand(N, L, q, 2)
    g.setBuf('seq main', a)
    g.storeSeqs()                        # device copy of main
    g.storeSeqs(b)                       # host sequences
    g.setBuf('seq main', b)
    g.storeSeqs()
    big = g.getBuf('seq large').read()
    assert big.shape == (3 * N, L)
    np.testing.assert_array_equal(big, np.concatenate([a, b, b]))
    with pytest.raises(Exception, match="past end"):
        g.storeSeqs()
    g.setBuf('seq main', seqs_rand(N, L, q, 3))
    g.restoreSeqs()
    np.testing.assert_array_equal(g.getBuf('seq main').read(), a)
    marks = np.zeros(L, 'u1')
    marks[[0, 7, 36]] = 1
    g.markPos(marks)
    g.copySubseq(5)
    big2 = g.getBuf('seq large').read()
    want = big.copy()
    want[:, [0, 7, 36]] = a[5, [0, 7, 36]]
    np.testing.assert_array_equal(big2, want)
