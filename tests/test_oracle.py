"""Pinning of the CPU oracle (oracle/) against published known answers, the
reference's own fixture, closed forms, invariants and brute force -- each
check chosen so that a plausible slip (a dropped term, a transposed block, a
wrong sign or index) fails it.  CPU only."""
import itertools
import json
from pathlib import Path

import numpy as np
import pytest
from scipy import stats

from oracle import cboracle, philox, potts_oracle as po

GOLDEN = Path(__file__).resolve().parent / 'golden'


def rnd_model(L, q, seed, sd=0.3):
    return np.random.default_rng(seed).normal(0, sd, size=(L * (L - 1) // 2, q * q)).astype('f4')


# --------------------------------------------------------------------- RNG

def test_philox_kat_numpy_and_c():
    """Published Philox4x32-10 known-answer vectors (Random123)."""
    kat = json.loads((GOLDEN / 'philox4x32_10_kat.json').read_text())
    for v in kat:
        got = [int(x) for x in philox.philox4x32_10(*v['ctr'], *v['key'])]
        assert got == v['out']
        assert list(cboracle.philox(v['ctr'], v['key'])) == v['out']


def test_stream_layout_c_matches_numpy():
    np.testing.assert_array_equal(philox.site_positions(17, 232, 1000, 64),
                                  cboracle.positions(17, 232, 1000, 64))
    pos = philox.site_positions(3, 7, 0, 20000)
    assert pos.min() == 0 and pos.max() == 6
    # uniform site choice
    assert stats.chisquare(np.bincount(pos, minlength=7)).pvalue > 1e-4


def test_uniform24_is_reference_uniformmap():
    # (x >> 8) * 2^-24 (mcmc.cl:29-31): exact float32, in [0, 1)
    x = np.array([0, 255, 256, 0xFFFFFFFF], dtype=np.uint32)
    np.testing.assert_array_equal(philox.uniform24(x),
                                  np.array([0, 0, 2 ** -24, 1 - 2 ** -24], dtype=np.float32))


# --------------------------------------------------------------------- formats

def test_hiv_fixture_is_consistent():
    """The reference's fixture: L=93, q=4, rows sum to 1, all unimarginals
    implied by different pairs agree (validate_bimarg, potts_common.py:50-76)."""
    b = np.load(GOLDEN / 'hiv_bimarg_pc.npy')
    assert b.dtype == np.float32 and b.shape == (4278, 16)
    assert po.getLq(b) == (93, 4)
    np.testing.assert_allclose(b.sum(axis=1), 1, atol=1e-5)
    assert b.min() > 0
    f = po.unimarg(b)
    L, q = 93, 4
    x = b.reshape(-1, q, q).astype(np.float64)
    for n, (i, j) in enumerate(po.pair_list(L)):
        np.testing.assert_allclose(x[n].sum(axis=1), f[i], atol=2e-5)
        np.testing.assert_allclose(x[n].sum(axis=0), f[j], atol=2e-5)


def test_unpack_pack_roundtrip_and_transpose():
    L, q = 6, 3
    J = rnd_model(L, q, 1)
    F = po.unpack_full(J, L, q)
    np.testing.assert_array_equal(po.pack_full(F), J.astype(np.float64))
    # the (j,i) block is the transpose of the (i,j) block
    for (i, j) in po.pair_list(L):
        np.testing.assert_array_equal(F[j, i], F[i, j].T)
    assert np.all(F[np.arange(L), np.arange(L)] == 0)


# --------------------------------------------------------------------- energies

def brute_energy(s, J, L, q):
    """Triple loop straight from the definition E = sum_{i<j} J_ij(s_i, s_j)."""
    e = 0.0
    n = 0
    for i in range(L - 1):
        for j in range(i + 1, L):
            e += float(J[n, q * int(s[i]) + int(s[j])])
            n += 1
    return e


def test_energies_brute_force_and_c():
    L, q = 7, 4
    J = rnd_model(L, q, 2)
    s = np.random.default_rng(3).integers(0, q, size=(40, L)).astype('u1')
    want = np.array([brute_energy(x, J, L, q) for x in s])
    np.testing.assert_allclose(po.energies(s, J), want, rtol=1e-12)
    np.testing.assert_allclose(cboracle.energies(s, J), want, rtol=1e-12)


def test_energies_quadratic_form():
    """E = 1/2 sum_{i != j} F[i, j, s_i, s_j] with the symmetric unpacked form
    (catches a transposed (j,i) block)."""
    L, q = 9, 5
    J = rnd_model(L, q, 4)
    F = po.unpack_full(J, L, q)
    s = np.random.default_rng(5).integers(0, q, size=(30, L))
    i, j = np.meshgrid(np.arange(L), np.arange(L), indexing='ij')
    qf = np.array([0.5 * F[i, j, x[i], x[j]].sum() for x in s])
    np.testing.assert_allclose(po.energies(s, J), qf, rtol=1e-12)


def test_energies_special_cases():
    L, q = 8, 3
    s = np.random.default_rng(6).integers(0, q, size=(20, L))
    Z = np.zeros((L * (L - 1) // 2, q * q), 'f4')
    assert np.all(po.energies(s, Z) == 0)
    np.testing.assert_allclose(po.energies(s, Z + 0.25), 0.25 * L * (L - 1) / 2)
    J = Z.copy()
    n = po.pair_list(L).index((2, 6))
    J[n, 1 * q + 2] = 1.0
    np.testing.assert_array_equal(po.energies(s, J), ((s[:, 2] == 1) & (s[:, 6] == 2)) * 1.0)


def test_delta_energy_is_energy_difference():
    L, q = 10, 4
    J = rnd_model(L, q, 7)
    F = po.unpack_full(J, L, q)
    rng = np.random.default_rng(8)
    for _ in range(20):
        s = rng.integers(0, q, size=L)
        pos, b = rng.integers(L), rng.integers(q)
        s2 = s.copy()
        s2[pos] = b
        dE = po.delta_energy(s, F, pos, b)
        np.testing.assert_allclose(dE, po.energies(s2[None], J)[0] - po.energies(s[None], J)[0],
                                   atol=1e-12)


# --------------------------------------------------------------------- counts

def test_bicounts_brute_force():
    L, q, N = 6, 4, 300
    s = np.random.default_rng(9).integers(0, q, size=(N, L))
    want = np.zeros((L * (L - 1) // 2, q * q), np.int64)
    for n, (i, j) in enumerate(po.pair_list(L)):
        for x in s:
            want[n, q * x[i] + x[j]] += 1
    np.testing.assert_array_equal(po.bicounts(s, q), want)
    np.testing.assert_array_equal(cboracle.bicounts(s.astype('u1'), q), want)
    assert np.all(want.sum(axis=1) == N)


def test_bicounts_to_bimarg_float32_division():
    c = np.array([[0, 1, 3, 7, 2 ** 20 - 1]], np.uint32)
    np.testing.assert_array_equal(po.bicounts_to_bimarg(c, 2 ** 20),
                                  c.astype(np.float32) / np.float32(2 ** 20))


def test_weighted_marg_reduces_and_is_linear():
    L, q, N = 5, 3, 200
    rng = np.random.default_rng(10)
    s = rng.integers(0, q, size=(N, L))
    np.testing.assert_array_equal(po.weighted_marg(s, np.ones(N), q), po.bicounts(s, q))
    w1, w2 = rng.random(N), rng.random(N)
    np.testing.assert_allclose(po.weighted_marg(s, 2 * w1 + w2, q),
                               2 * po.weighted_marg(s, w1, q) + po.weighted_marg(s, w2, q))
    np.testing.assert_allclose(cboracle.weighted_marg(s.astype('u1'), w1.astype('f4'), q),
                               po.weighted_marg(s, w1.astype('f4'), q), rtol=1e-12)


# --------------------------------------------------------------------- sampler

def exact_distribution(J, L, q, beta=1.0):
    states = np.array(list(itertools.product(range(q), repeat=L)))
    E = po.energies(states, J)
    p = np.exp(-beta * (E - E.min()))
    return states, p / p.sum()


def test_metropolis_numpy_and_c_agree():
    L, q = 6, 3
    J = rnd_model(L, q, 11, 0.5)
    s = np.random.default_rng(12).integers(0, q, size=(64, L)).astype('u1')
    a, ma = po.metropolis(s, J, 5, 10, 120)
    b, mb = cboracle.metropolis(s, J, 5, 10, 120)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_allclose(ma, mb, atol=1e-12)


def test_metropolis_exact_enumeration_chi2():
    """81-state model: the sampled ensemble matches exp(-E)/Z."""
    L, q, N = 4, 3, 60000
    J = rnd_model(L, q, 13, 0.6)
    states, p = exact_distribution(J, L, q)
    s0 = np.random.default_rng(14).integers(0, q, size=(N, L)).astype('u1')
    s, _ = cboracle.metropolis(s0, J, 21, 0, 400)
    idx = (s.astype(int) * (q ** np.arange(L - 1, -1, -1))).sum(axis=1)
    obs = np.bincount(idx, minlength=q ** L)
    res = stats.chisquare(obs, p * N)
    assert res.pvalue > 1e-4, res
    # a wrong sign in dE would produce exp(+E) instead
    bad = stats.chisquare(obs, exact_distribution(-J, L, q)[1] * N)
    assert bad.pvalue < 1e-20


def test_metropolis_beta_scales_distribution():
    L, q, N = 3, 3, 40000
    J = rnd_model(L, q, 15, 0.8)
    states, p = exact_distribution(J, L, q, beta=0.5)
    s0 = np.random.default_rng(16).integers(0, q, size=(N, L)).astype('u1')
    s, _ = cboracle.metropolis(s0, J, 3, 0, 300, beta=0.5)
    idx = (s.astype(int) * (q ** np.arange(L - 1, -1, -1))).sum(axis=1)
    assert stats.chisquare(np.bincount(idx, minlength=q ** L), p * N).pvalue > 1e-4


def test_metropolis_f32_port_tracks_f64_chain():
    L, q = 30, 5
    J = rnd_model(L, q, 17, 0.1)
    s = np.random.default_rng(18).integers(0, q, size=(256, L)).astype('u1')
    a, m = cboracle.metropolis(s, J, 9, 0, 300)
    b = cboracle.metropolis_f32(s, J, 9, 0, 300)
    same = np.all(a == b, axis=1)
    assert np.all(same[m > 1e-3]) and same.mean() > 0.95


def test_gen_indep_distribution_and_c():
    L, q, N = 5, 4, 40000
    f = np.random.default_rng(19).dirichlet(np.ones(q), size=L)
    cp = po.prepare_indep(f)
    s = cboracle.gen_indep(cp, 7, 0, N)
    np.testing.assert_array_equal(s[:500], po.gen_indep(cp, 7, 0, 500))
    for i in range(L):
        assert stats.chisquare(np.bincount(s[:, i], minlength=q), f[i] * N).pvalue > 1e-4
    # sites independent: pair (0,1) matches the product
    pair = np.bincount(q * s[:, 0].astype(int) + s[:, 1], minlength=q * q)
    assert stats.chisquare(pair, np.outer(f[0], f[1]).ravel() * N).pvalue > 1e-4


# --------------------------------------------------------------------- Newton phase

def test_updatedJ_closed_form():
    bt = np.array([[0.5, 0.25]])
    bi = np.array([[0.25, 0.25]])
    got = po.updatedJ(bt, bi, 0.1, 0.25, np.zeros((1, 2)))
    # dJ - gamma*(ft - f)/(f + pc): -0.1*0.25/0.5 = -0.05, and 0 where f = ft
    np.testing.assert_allclose(got, [[-0.05, 0.0]])


def test_zero_gauge_blocks_invariants():
    rng = np.random.default_rng(20)
    q = 4
    J = rng.normal(size=(7, q * q))
    Z = po.zero_gauge_blocks(J).reshape(-1, q, q)
    np.testing.assert_allclose(Z.sum(axis=1), 0, atol=1e-12)
    np.testing.assert_allclose(Z.sum(axis=2), 0, atol=1e-12)
    # idempotent, and blind to row/column constants
    np.testing.assert_allclose(po.zero_gauge_blocks(Z.reshape(7, -1)).reshape(Z.shape), Z,
                               atol=1e-12)
    shift = rng.normal(size=(7, q, 1)) + rng.normal(size=(7, 1, q))
    np.testing.assert_allclose(po.zero_gauge_blocks((J.reshape(-1, q, q) + shift).reshape(7, -1)),
                               Z.reshape(7, -1), atol=1e-12)


def test_reg_l1z_hand_example():
    """2x2 block whose zero-gauge form is J0 = [[1,-1],[-1,1]]*0.01."""
    J = np.array([[0.01, -0.01, -0.01, 0.01]])
    dJ = np.zeros_like(J)
    bi = np.full_like(J, 0.25)
    gamma, pc = 0.1, 0.0
    # R = -lam*sign(J0)*gamma/(f+pc); lam=0.001 -> |R| = 4e-4 < |J0|: step taken
    out, amb = po.reg_l1z(bi, gamma, pc, 0.001, J, dJ)
    np.testing.assert_allclose(out, -4e-4 * np.sign(J))
    assert not amb.any()
    # lam=0.1 -> |R| = 0.04 > |J0| = 0.01: sign flip is clipped, J0 set to 0
    out, _ = po.reg_l1z(bi, gamma, pc, 0.1, J, dJ)
    np.testing.assert_allclose(out, -J)
    # l2: R = -lam*J0*gamma/(f+pc)
    out, _ = po.reg_l2z(bi, gamma, pc, 0.5, J, dJ)
    np.testing.assert_allclose(out, -0.5 * J * 0.1 / 0.25)


def test_min_weights_stats_quirks():
    assert po.min_floats(np.array([3.0, 2.0])) == 0.0
    assert po.min_floats(np.array([3.0, -2.0])) == -2.0
    w = po.dE_to_weights(np.array([0.0, np.log(2)]), 0.0)
    np.testing.assert_allclose(w, [1.0, 0.5])
    assert po.weight_stats([1.0, 0.5]) == (1.5, 1.25)


def test_iter_newton_first_iteration_by_hand():
    """One iteration on a tiny model, written out step by step."""
    L, q, N = 3, 2, 6
    s = np.array([[0, 0, 1], [1, 1, 0], [0, 1, 1], [1, 0, 0], [0, 0, 0], [1, 1, 1]])
    J = rnd_model(L, q, 21, 0.1).astype(np.float64)
    bt = np.full((3, 4), 0.25)
    bt[0] = [0.4, 0.1, 0.1, 0.4]
    i, trialJ, bi, w, neffs = po.iter_newton(s, J, bt, 0.5, 0.1, 1, 0.0)
    f = po.bicounts(s, q) / N
    dJ = -0.5 * (bt - f) / (f + 0.1)
    E = po.energies(s, dJ)
    ww = np.exp(-E)
    F = po.weighted_marg(s, ww, q)
    np.testing.assert_allclose(trialJ, J + dJ, rtol=1e-12)
    np.testing.assert_allclose(bi, F / F.sum(axis=1, keepdims=True), rtol=1e-12)
    np.testing.assert_allclose(neffs[0], ww.sum() ** 2 / (ww ** 2).sum())


def test_iter_newton_stops_on_neff():
    L, q, N = 5, 3, 400
    s = np.random.default_rng(22).integers(0, q, size=(N, L))
    J = rnd_model(L, q, 23, 0.1).astype(np.float64)
    bt = po.jeffreys_pseudocount(po.bicounts(np.random.default_rng(24).integers(
        0, q, size=(50, L)), q) / 50, 50)
    i, _, _, _, neffs = po.iter_newton(s, J, bt, 0.05, 0.01, 500, 0.8)
    assert neffs[-1] < 0.8 * N and all(n >= 0.8 * N for n in neffs[:-1])
    assert i == len(neffs) - 1


# --------------------------------------------------------------------- gauges

def test_fieldless_gauge_even_shifts_energy_by_constant():
    L, q = 6, 4
    rng = np.random.default_rng(25)
    h = rng.normal(size=(L, q))
    J = rng.normal(size=(L * (L - 1) // 2, q * q))
    _, Jf = po.fieldlessGaugeEven(h, J)
    s = rng.integers(0, q, size=(50, L))
    E0 = po.energies(s, J) + h[np.arange(L), s].sum(axis=1)
    d = po.energies(s, Jf) - E0
    np.testing.assert_allclose(d, d[0], atol=1e-10)
    # fully constrained: re-gauging the input does not change the output
    h2, J2 = po.zeroJGauge(h, J)
    np.testing.assert_allclose(po.fieldlessGaugeEven(h2, J2)[1], Jf, atol=1e-12)


def test_jeffreys_pseudocount():
    b = np.load(GOLDEN / 'hiv_bimarg_pc.npy').astype(np.float64)
    out = po.jeffreys_pseudocount(b, 1e12)
    np.testing.assert_allclose(out, b, atol=1e-9)
    out = po.jeffreys_pseudocount(b, 10)
    np.testing.assert_allclose(out.sum(axis=1), 1)
    assert out.min() > b.min()


def test_jeffreys_pseudocount_hand_example():
    """One pair, q=2, N=1 written out by hand from pseudocount.py:8-24:
    mu = q/(2N+q) = 1/2; f_i = (0.7, 0.3) (row sums), f_j = (0.6, 0.4)
    (column sums); fifj = f_i(a) + f_j(b) = (1.3, 1.1, 0.9, 0.7);
    (1-mu)^2 ff + (1-mu) mu fifj / q + (mu/q)^2
      = (0.125, 0.05, 0.025, 0.05) + (0.1625, 0.1375, 0.1125, 0.0875) + 0.0625
      = (0.35, 0.25, 0.2, 0.2), which already sums to 1.  A wrong mu (the
    Bayes q/(N+q)) or a product f_i f_j in place of the outer sum gives other
    values."""
    ff = np.array([[0.5, 0.2, 0.1, 0.2]])
    want = np.array([[0.35, 0.25, 0.2, 0.2]])
    np.testing.assert_allclose(po.jeffreys_pseudocount(ff, 1), want, rtol=1e-14)
    from mi3gpu.utils.pseudocount import mutation_pc
    np.testing.assert_allclose(mutation_pc(ff, 1), want, rtol=1e-14)
    # the Bayes form: mu = 2/3 -> (1/9) ff + (2/9) fifj/2 + 1/9, renormalised
    bayes = ff / 9 + fifj_sum(ff) / 9 + 1 / 9
    np.testing.assert_allclose(mutation_pc(ff, 1, mode='bayes'), bayes / bayes.sum(),
                               rtol=1e-14)


def fifj_sum(ff):
    return np.array([[1.3, 1.1, 0.9, 0.7]])


def test_Xijab_contraction_matches_quadruple_loop():
    """getXijab's literal q^4 loop (mcmc.cl:1052-1067) on random blocks."""
    rng = np.random.default_rng(21)
    q, nP = 3, 4
    Jt = rng.normal(size=(nP, q * q))
    fi, fj = rng.random((nP, q)), rng.random((nP, q))
    got = po.block_Xijab(Jt, fi, fj, q)
    for p in range(nP):
        B = Jt[p].reshape(q, q)
        for c in range(q):
            for d in range(q):
                x = 0.0
                for a in range(q):
                    for b in range(q):
                        x += B[a, b] * (fi[p, a] - (a == c)) * (fj[p, b] - (b == d))
                assert abs(got[p, c, d] - x) < 1e-12


def _normalised_blocks(rng, nP, q):
    f = rng.random((nP, q * q)) + 0.05
    return f / f.sum(axis=1, keepdims=True)


def test_X_family_hand_example():
    """q=2: f = [[.4,.1],[.1,.4]] -> fi = fj = (.5,.5), C = f - .25,
    J = [[1,0],[0,0]] -> X = 0.15, Xnorm = 0.5625 (pc = 0)."""
    f = np.array([[0.4, 0.1, 0.1, 0.4]])
    J = np.array([[1.0, 0.0, 0.0, 0.0]])
    dJ = np.zeros_like(J)
    C, X, Xn, fi, fj = po.block_X(f, J, 2, 0.0)
    assert abs(X[0] - 0.15) < 1e-12 and abs(Xn[0] - 0.5625) < 1e-12
    # reg_X, lambda 1: min(1, 0.15/(0.5625*0.01)) = 1 -> dJ = -0.01*C/f
    got = po.reg_X(f, 0.01, 0.0, 1.0, J, dJ)
    np.testing.assert_allclose(got, -0.01 * np.array([[0.375, -1.5, -1.5, 0.375]]), rtol=1e-12)
    # lambda above the clip: lX' = |X|/(Xnorm*gamma) = 26.666..
    got = po.reg_X(f, 0.01, 0.0, 100.0, J, dJ)
    np.testing.assert_allclose(got, -0.01 * (0.15 / 0.005625) * np.array(
        [[0.15 / 0.4, -1.5, -1.5, 0.15 / 0.4]]), rtol=1e-12)
    # Xijab = (.5-[c=0])(.5-[d=0]); grad = C + f (X - Xijab) = +-0.11
    np.testing.assert_allclose(po.block_Xijab(J, fi, fj, 2)[0],
                               [[0.25, -0.25], [-0.25, 0.25]], atol=1e-15)
    got, amb = po.reg_expX(f, 0.01, 0.0, 1.0, J, dJ)
    grad = np.array([0.11, -0.11, -0.11, 0.11])
    np.testing.assert_allclose(got[0], -0.01 * np.exp(-0.15) * grad / f[0], rtol=1e-12)
    assert not amb.any()


def test_X_family_is_gauge_invariant():
    """Adding u(a) + v(b) to a block leaves X and X - Xijab unchanged when
    the block's bimarginals sum to 1, so every X-type step is unchanged."""
    rng = np.random.default_rng(5)
    q, nP = 4, 6
    f = _normalised_blocks(rng, nP, q)
    J = rng.normal(0, 0.3, (nP, q * q))
    dJ = rng.normal(0, 0.05, (nP, q * q))
    g = (rng.normal(size=(nP, q, 1)) + rng.normal(size=(nP, 1, q))).reshape(nP, q * q)
    args = (1e-3, 1e-3)
    pairs = [(po.reg_X(f, *args, 0.5, J, dJ), po.reg_X(f, *args, 0.5, J + g, dJ)),
             (po.reg_SCADX(f, *args, 2.0, 0.3, 3.0, J, dJ),
              po.reg_SCADX(f, *args, 2.0, 0.3, 3.0, J + g, dJ)),
             (po.reg_expX(f, *args, 0.2, J, dJ)[0], po.reg_expX(f, *args, 0.2, J + g, dJ)[0])]
    for a, b in pairs:
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12)


def test_SCADJ_piecewise_by_hand():
    """Zero-gauge 2x2 block [[x,-x],[-x,x]], f = 1/4, pc = 0, gamma = .1,
    r = .5, a = 3: R = min(r, |x|/gamma), (a r - |x|)/(a-1), or 0, and the
    step is R gamma sign(x)/f."""
    f = np.full((1, 4), 0.25)
    for x, R in [(0.01, 0.1), (0.04, 0.4), (1.0, 0.25), (-1.2, 0.15), (2.0, 0.0)]:
        J = np.array([[x, -x, -x, x]])
        got, amb = po.reg_SCADJ(f, 0.1, 0.0, 0.5, 3.0, J, np.zeros_like(J))
        exp = -R * 0.1 * np.sign(J) / 0.25
        np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-15)
        assert amb.all() == (abs(x) < 0.0505)


def test_ddE_gauge_invariance_and_q2_closed_form():
    rng = np.random.default_rng(8)
    q, nP = 3, 5
    f = _normalised_blocks(rng, nP, q)
    J = rng.normal(0, 0.3, (nP, q * q))
    dJ = np.zeros_like(J)
    g = (rng.normal(size=(nP, q, 1)) + rng.normal(size=(nP, 1, q))).reshape(nP, q * q)
    a, amb_a = po.reg_ddE(f, 0.01, 1e-3, 0.5, J, dJ)
    b, amb_b = po.reg_ddE(f, 0.01, 1e-3, 0.5, J + g, dJ)
    np.testing.assert_allclose(a, b, atol=1e-12)
    # q = 2: one double difference per entry, dR = sign(ddE)
    J2 = np.array([[0.3, -0.1, 0.2, 0.5]])
    dd = 0.3 - 0.2 - (-0.1) + 0.5     # J(a,b) - J(a',b) - J(a,b') + J(a',b'), a=b=0
    f2 = np.full((1, 4), 0.25)
    out, amb = po.reg_ddE(f2, 0.1, 0.0, 2.0, J2, np.zeros_like(J2))
    assert abs(out[0, 0] - (-2.0 * np.sign(dd) * 0.1 / 0.25)) < 1e-12
    # SCADddE below r: R = 1 -> same as lam * sign; above 4r: nothing
    out2, _ = po.reg_ddE(f2, 0.1, 0.0, 2.0, J2, np.zeros_like(J2), scad_r=10.0)
    np.testing.assert_allclose(out2, out, atol=1e-12)
    out3, _ = po.reg_ddE(f2, 0.1, 0.0, 2.0, J2, np.zeros_like(J2), scad_r=0.01)
    np.testing.assert_allclose(out3, 0.0, atol=1e-15)
