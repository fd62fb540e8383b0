"""Learned proposal on the GPU: the bf16 GEMM, the MADE forward/backward
against the oracle, the sampler against the network's own probabilities,
the mixed sampler's stationary distribution against exact enumeration,
and what the trained network gains over single-site sweeps alone.
Device memory comes from the library (no other GPU runtime is loaded)."""
import itertools

import numpy as np
import pytest

from oracle import made_oracle as mo, potts_oracle as po

pytestmark = pytest.mark.gpu

NAMES = ['W1t', 'b1', 'W2', 'b2', 'W3', 'b3']


def _to_bf16_bits(x):
    return (mo.bf16(x).view(np.uint32) >> 16).astype(np.uint16)


def _from_bf16_bits(u):
    return (u.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_bf16_layouts_exact(gpu_ok, ta, tb):
    """Small integers are exact in bf16 and in the float32 accumulation, so
    every layout must reproduce numpy bit for bit (incl. bias, relu, beta)."""
    from mi3gpu.made import DeviceArrays
    from mi3gpu._lib import lib
    M, N, K = 136, 72, 48
    rng = np.random.default_rng(ta * 2 + tb)
    A = rng.integers(-4, 5, (M, K)).astype(np.float32)
    B = rng.integers(-4, 5, (K, N)).astype(np.float32)
    bias = rng.integers(-3, 4, N).astype(np.float32)
    C0 = rng.integers(-5, 6, (M, N)).astype(np.float32)
    d = DeviceArrays()
    Ah = A.T.copy() if ta else A
    Bh = B.T.copy() if tb else B
    d.alloc('A', Ah.shape, 'u2')
    d.alloc('B', Bh.shape, 'u2')
    d.alloc('bias', (N,), 'f4')
    d.alloc('C', (M, N), 'f4')
    d.alloc('Cb', (M, N), 'u2')
    d.write('A', _to_bf16_bits(Ah))
    d.write('B', _to_bf16_bits(Bh))
    d.write('bias', bias)
    d.write('C', C0)
    lda = M if ta else K
    ldb = K if tb else N
    rc = lib().mi3_gemm_bf16(d.stream, ta, tb, M, N, K, 1.0, d['A'], lda, d['B'], ldb, 1.0,
                             d['C'], N, d['Cb'], N, d['bias'], 1)
    assert rc == 0
    d.sync()
    want = np.maximum(A @ B + C0 + bias[None, :], 0)
    np.testing.assert_array_equal(d.read('C'), want)
    np.testing.assert_array_equal(_from_bf16_bits(d.read('Cb')), mo.bf16(want))
    d.close()


def _perturb(m, rng, scale):
    for name in NAMES:
        v = m.get(name)
        m.set(name, v + rng.normal(0, scale.get(name, 0.0), v.shape))


def _params(m):
    return [m.get(k).astype(np.float64) for k in NAMES]


def test_made_forward_and_gradients_match_oracle(gpu_ok):
    from mi3gpu.made import MADE, made_masks
    L, q, H, B = 10, 4, 32, 64
    m = MADE(L, q, H, max_batch=B, seed=3)
    _perturb(m, np.random.default_rng(7), {k: 0.05 for k in NAMES})
    s = np.random.default_rng(1).integers(0, q, (B, L)).astype(np.uint8)
    m.mem.alloc('seqs', (L, B), 'u1')
    m.mem.write('seqs', s.T)
    m.loss_and_grads(m.mem['seqs'], B, B)
    nll = m.get('nll').astype(np.float64)
    P = _params(m)
    Ms = made_masks(L, q, H)
    g, nll_ref = mo.grads(s, *P, q, *Ms, bf=True)
    ref_each = mo.forward(s, *P, q, bf=True)['nll']
    print("nll", nll.mean(), nll_ref)
    assert abs(nll.mean() - nll_ref) < 1e-4 * abs(nll_ref)
    np.testing.assert_allclose(nll, ref_each, rtol=0, atol=5e-3)
    mask = {'W1t': Ms[0], 'W2': Ms[1], 'W3': Ms[2]}
    for name in NAMES:
        got = m.get('g_' + name).astype(np.float64) * mask.get(name, 1.0)
        ref = g[name]
        err = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-30)
        print(f"{name}: relative error {err:.2e}")
        assert err < 1e-3, name
    m.close()


def test_made_adam_matches_restatement(gpu_ok):
    """One Adam step from the stored gradients equals the textbook update
    (Kingma & Ba 2015) with the mask applied to gradient and result."""
    from mi3gpu.made import MADE, made_masks
    L, q, H, B = 6, 4, 16, 32
    m = MADE(L, q, H, max_batch=B, seed=2)
    s = np.random.default_rng(3).integers(0, q, (B, L)).astype(np.uint8)
    m.mem.alloc('seqs', (L, B), 'u1')
    m.mem.write('seqs', s.T)
    m.loss_and_grads(m.mem['seqs'], B, B)
    before = {k: m.get(k).astype(np.float64) for k in NAMES}
    grads = {k: m.get('g_' + k).astype(np.float64) for k in NAMES}
    lr, b1, b2, eps = 1e-2, 0.9, 0.999, 1e-8
    m.adam(lr, b1, b2, eps)
    Ms = dict(zip(('W1t', 'W2', 'W3'), made_masks(L, q, H)))
    for k in NAMES:
        gk = grads[k] * Ms.get(k, 1.0)
        mt = (1 - b1) * gk
        vt = (1 - b2) * gk * gk
        want = (before[k] - lr * (mt / (1 - b1)) / (np.sqrt(vt / (1 - b2)) + eps)) * \
            Ms.get(k, 1.0)
        np.testing.assert_allclose(m.get(k), want, rtol=1e-5, atol=1e-6, err_msg=k)
    # the bf16 copies follow the float32 parameters
    np.testing.assert_array_equal(_from_bf16_bits(m.get('W3b')),
                                  mo.bf16(m.get('W3').astype(np.float32)))
    m.close()


def test_made_sampler_matches_its_probabilities(gpu_ok):
    """Samples follow the network's own distribution (exact enumeration of
    the 4^8 states), and the returned nll equals the teacher-forced one."""
    from mi3gpu.made import MADE
    from scipy import stats
    L, q, H, B = 8, 4, 32, 2 ** 16
    m = MADE(L, q, H, max_batch=B, seed=5)
    _perturb(m, np.random.default_rng(9), {'b3': 1.0, 'W3': 0.5, 'W1t': 0.3})
    m.mem.alloc('out', (L, B), 'u1')
    m.mem.alloc('nll_s', (B,), 'f4')
    m.mem.alloc('nll_tf', (B,), 'f4')
    m.sample(B, m.mem['out'], B, m.mem['nll_s'], seed=11, counter=1)
    m.nll(m.mem['out'], B, B, m.mem['nll_tf'])
    s = m.get('out').T
    np.testing.assert_allclose(m.get('nll_s'), m.get('nll_tf'), rtol=0, atol=2e-3)
    # a second draw with another counter differs; the same counter repeats
    m.mem.alloc('out2', (L, B), 'u1')
    m.sample(B, m.mem['out2'], B, m.mem['nll_tf'], seed=11, counter=1)
    np.testing.assert_array_equal(m.get('out2'), m.get('out'))
    m.sample(B, m.mem['out2'], B, m.mem['nll_tf'], seed=11, counter=2)
    assert np.mean(m.get('out2') != m.get('out')) > 0.2
    states, lp = mo.log_prob_all(*_params(m), L, q, bf=True)
    p = np.exp(lp)
    p /= p.sum()
    key = lambda x: x[:, 0].astype(int) * 16 + x[:, 1] * 4 + x[:, 7]
    exp = np.bincount(key(states), weights=p, minlength=64) * B
    obs = np.bincount(key(s), minlength=64)
    res = stats.chisquare(obs, exp)
    print("chi2 p", res.pvalue)
    assert res.pvalue > 1e-4
    m.close()


def _ferro(L, q, c):
    """Fully connected ferromagnetic Potts couplings J_ij(a,b) = -c [a=b]."""
    J = np.zeros((L * (L - 1) // 2, q * q), np.float32)
    J[:, [a * q + a for a in range(q)]] = -c
    return J


def _exact_bimarg(J, L, q):
    states = np.array(list(itertools.product(range(q), repeat=L)), dtype=np.uint8)
    E = po.energies(states, J)
    w = np.exp(-(E - E.min()))
    w /= w.sum()
    return po.weighted_marg(states, w, q)


def _ssr(seqs, J, L, q):
    f = po.bicounts(seqs, q) / len(seqs)
    fe = _exact_bimarg(J, L, q)
    return np.sum((f - fe) ** 2), 2 * np.sum(fe * (1 - fe)) / len(seqs)


def test_mixed_sampler_exact_and_gains(gpu_ok, tmp_path):
    """Ordered ferromagnetic model (four degenerate phases): single-site
    sweeps stay in the phase they start in, so their bimarginals miss the
    exact ones by far; the mixed sampler with a MADE trained on exact
    samples moves between phases, its ensemble matches exact enumeration
    (SSR near the sampling floor), and the trained network is accepted far
    more often than an untrained one."""
    from mi3gpu.mcmcGPU import MCMCGPU
    from mi3gpu.made import MADE, MixedSampler
    L, q, H, N = 8, 4, 64, 8192
    J = _ferro(L, q, 0.55)
    states = np.array(list(itertools.product(range(q), repeat=L)), dtype=np.uint8)
    E = po.energies(states, J)
    w = np.exp(-(E - E.min()))
    w /= w.sum()
    rng = np.random.default_rng(0)
    train = states[rng.choice(len(states), size=32768, p=w)]
    s0 = np.zeros((N, L), np.uint8)

    def chain(made, cycles):
        g = MCMCGPU((0, 0, None, None), L, q, N, 512, tmp_path, 1024, 7)
        g.initMCMC(L, 0, 2 ** 63)
        g.setBuf('J', J)
        g.setBuf('seq main', s0)
        if made is None:
            for _ in range(cycles):
                g.runMCMC()
            return g.getBuf('seq main').read(), None
        ms = MixedSampler(g, made, seed=3)
        for _ in range(cycles):
            ms.step()
        out = g.getBuf('seq main').read()
        return out, ms.accepted() / (N * cycles)

    plain, _ = chain(None, 40)
    untrained = MADE(L, q, H, max_batch=1024, seed=1)
    _, acc_untrained = chain(untrained, 10)
    made = MADE(L, q, H, max_batch=1024, seed=1)
    made.mem.alloc('train', (L, len(train)), 'u1')
    made.mem.write('train', train.T)
    hist = made.train(made.mem['train'], len(train), len(train), steps=400, batch=1024, lr=3e-3)
    mixed, acc = chain(made, 40)
    ssr_p, floor = _ssr(plain, J, L, q)
    ssr_m, _ = _ssr(mixed, J, L, q)
    print(f"nll {hist[0][1]:.3f} -> {hist[-1][1]:.3f}; acceptance trained {acc:.3f} "
          f"untrained {acc_untrained:.4f}; SSR plain {ssr_p:.4g} mixed {ssr_m:.4g} "
          f"floor {floor:.4g}")
    assert hist[-1][1] < hist[0][1]
    assert acc > 0.2 and acc > 5 * acc_untrained
    assert ssr_m < 3 * floor
    assert ssr_p > 20 * floor


def test_gemm_bf16_hidden_product_at_size(gpu_ok):
    """The sampler's hidden product at its production shape (2^16 walkers x
    512 x 512, A row-major, B transposed as W2 is stored): small integers
    make bf16 inputs and float32 sums exact, so the result must equal numpy
    bit for bit, as must its bf16 copy."""
    from mi3gpu.made import DeviceArrays
    from mi3gpu._lib import lib
    M, N, K = 2 ** 16, 512, 512
    rng = np.random.default_rng(17)
    A = rng.integers(-3, 4, (M, K)).astype(np.float32)
    Bt = rng.integers(-3, 4, (N, K)).astype(np.float32)
    bias = rng.integers(-8, 9, N).astype(np.float32)
    d = DeviceArrays()
    d.alloc('A', (M, K), 'u2')
    d.alloc('B', (N, K), 'u2')
    d.alloc('bias', (N,), 'f4')
    d.alloc('C', (M, N), 'f4')
    d.alloc('Cb', (M, N), 'u2')
    d.write('A', _to_bf16_bits(A))
    d.write('B', _to_bf16_bits(Bt))
    d.write('bias', bias)
    rc = lib().mi3_gemm_bf16(d.stream, 0, 1, M, N, K, 1.0, d['A'], K, d['B'], K, 0.0,
                             d['C'], N, d['Cb'], N, d['bias'], 1)
    assert rc == 0
    d.sync()
    want = np.maximum(A @ Bt.T + bias[None, :], 0)
    np.testing.assert_array_equal(d.read('C'), want)
    np.testing.assert_array_equal(_from_bf16_bits(d.read('Cb')), mo.bf16(want))
    d.close()


def test_made_production_size_matches_oracle(gpu_ok):
    """The learned proposal's network at its production size (L=232, q=21,
    two hidden layers of 512) on a 256-sequence batch: per-sequence NLL and
    every gradient against the oracle with the GPU's bf16 rounding, then the
    sampler at that size: the NLL it accumulates while drawing equals the
    teacher-forced NLL of what it drew."""
    from mi3gpu.made import MADE, made_masks
    L, q, H, B = 232, 21, 512, 256
    m = MADE(L, q, H, max_batch=2048, seed=4)
    _perturb(m, np.random.default_rng(8), {'W1t': 0.05, 'W2': 0.02, 'W3': 0.05, 'b3': 0.5})
    s = np.random.default_rng(2).integers(0, q, (B, L)).astype(np.uint8)
    m.mem.alloc('seqs', (L, B), 'u1')
    m.mem.write('seqs', s.T)
    m.loss_and_grads(m.mem['seqs'], B, B)
    nll = m.get('nll')[:B].astype(np.float64)
    P = _params(m)
    Ms = made_masks(L, q, H)
    g, nll_ref = mo.grads(s, *P, q, *Ms, bf=True)
    ref_each = mo.forward(s, *P, q, bf=True)['nll']
    rel = np.abs(nll - ref_each) / np.abs(ref_each)
    print(f"nll mean {nll.mean():.4f} (oracle {nll_ref:.4f}), max rel {rel.max():.2e}")
    assert rel.max() < 1e-4
    mask = {'W1t': Ms[0], 'W2': Ms[1], 'W3': Ms[2]}
    for name in NAMES:
        got = m.get('g_' + name).astype(np.float64) * mask.get(name, 1.0)
        ref = g[name]
        err = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-30)
        print(f"{name}: relative error {err:.2e}")
        assert err < 1e-3, name
    # sampling at this size
    n = 2048
    m.mem.alloc('out', (L, n), 'u1')
    m.mem.alloc('nll_s', (n,), 'f4')
    m.mem.alloc('nll_tf', (n,), 'f4')
    m.sample(n, m.mem['out'], n, m.mem['nll_s'], seed=5, counter=1)
    m.nll(m.mem['out'], n, n, m.mem['nll_tf'])
    a, b = m.get('nll_s').astype(np.float64), m.get('nll_tf').astype(np.float64)
    drawn = m.get('out').T
    print(f"sampled nll {a.mean():.3f}, teacher-forced {b.mean():.3f}, "
          f"max diff {np.abs(a - b).max():.2e}")
    assert drawn.max() < q
    np.testing.assert_allclose(a, b, rtol=1e-4, atol=0)
    # the drawn sequences' nll under the oracle network, and by the training
    # path's forward pass (embedding, two GEMMs, softmax)
    ref_drawn = mo.forward(drawn[:256], *P, q, bf=True)['nll']
    np.testing.assert_allclose(b[:256], ref_drawn, rtol=1e-4)
    m.mem.alloc('nll_fw', (n,), 'f4')
    m.nll_forward(m.mem['out'], n, n, m.mem['nll_fw'])
    np.testing.assert_allclose(m.get('nll_fw').astype(np.float64), b, rtol=1e-4)
    m.close()
