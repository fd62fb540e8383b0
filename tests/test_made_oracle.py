"""Pins of the learned-proposal oracle: masks give the autoregressive
property, backward equals finite differences, sampling probabilities sum
to one."""
import numpy as np

from oracle import made_oracle as mo


def _masks(L, q, H):
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parents[1] / 'mi3-gpu_amd'))
    from mi3gpu.made import made_masks
    return made_masks(L, q, H)


def _params(L, q, H, seed=0):
    rng = np.random.default_rng(seed)
    M1t, M2, M3 = _masks(L, q, H)
    return (rng.normal(0, 0.5, (L * q, H)) * M1t, rng.normal(0, 0.1, H),
            rng.normal(0, 0.3, (H, H)) * M2, rng.normal(0, 0.1, H),
            rng.normal(0, 0.3, (L * q, H)) * M3, rng.normal(0, 0.1, L * q)), (M1t, M2, M3)


def test_masks_are_autoregressive():
    """Changing s_j changes only the logits of sites after j."""
    L, q, H = 6, 3, 16
    P, _ = _params(L, q, H)
    rng = np.random.default_rng(1)
    s = rng.integers(0, q, (1, L))
    base = mo.forward(s, *P, q, bf=False)['logits'].reshape(L, q)
    for j in range(L):
        t = s.copy()
        t[0, j] = (t[0, j] + 1) % q
        z = mo.forward(t, *P, q, bf=False)['logits'].reshape(L, q)
        changed = np.abs(z - base).max(axis=1) > 1e-12
        assert not changed[:j + 1].any()
    # site 0's distribution is the bias alone
    np.testing.assert_allclose(base[0], P[5][:q])


def test_probabilities_sum_to_one():
    L, q, H = 4, 3, 8
    P, _ = _params(L, q, H, seed=2)
    states, lp = mo.log_prob_all(*P, L, q, bf=False)
    assert abs(np.exp(lp).sum() - 1) < 1e-10


def test_backward_equals_finite_differences():
    L, q, H, B = 5, 3, 8, 7
    P, M = _params(L, q, H, seed=3)
    s = np.random.default_rng(4).integers(0, q, (B, L))
    g, _ = mo.grads(s, *P, q, *M, bf=False)
    names = ['W1t', 'b1', 'W2', 'b2', 'W3', 'b3']
    rng = np.random.default_rng(5)
    for k, name in enumerate(names):
        mask = M[[0, None, 1, None, 2, None][k]] if name.startswith('W') else None
        for _ in range(6):
            idx = tuple(rng.integers(0, n) for n in P[k].shape)
            if mask is not None and mask[idx] == 0:
                continue
            eps = 1e-6
            Pp = [x.copy() for x in P]
            Pm = [x.copy() for x in P]
            Pp[k][idx] += eps
            Pm[k][idx] -= eps
            fd = (mo.forward(s, *Pp, q, bf=False)['nll'].mean() -
                  mo.forward(s, *Pm, q, bf=False)['nll'].mean()) / (2 * eps)
            assert abs(fd - g[name][idx]) < 1e-6 * max(1.0, abs(fd)), (name, idx)


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 0.0], np.float32)
    r = mo.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0          # tie to even
    assert r[2] == np.float32(1.0078125)         # rounds up
    assert abs(r[3] - x[3]) <= 2 ** -8 * abs(x[3])
