"""Batched uneven models: each model's chains equal the single-model oracle
chains for its seed (seed + model index), whatever the tile choice, across
graph replays."""
import numpy as np
import pytest

from oracle import cboracle

pytestmark = pytest.mark.gpu

# (L, q, walkers): register classes of one to five state words (L <= 128, <= 256,
# <= 384, <= 512, > 512) x q classes; the four-word models cover the top of
# the batched workload's L range (50..400)
MODELS = [(20, 4, 512), (70, 21, 1024), (140, 4, 512), (33, 21, 512), (260, 4, 512),
          (99, 21, 512), (600, 21, 512), (400, 21, 1024), (390, 4, 512)]


def _setup(seed=40):
    from mi3gpu.batch import PottsBatch
    L, q, N = zip(*MODELS)
    b = PottsBatch(L, q, N, seed=seed)
    Js, S0 = [], []
    for m, (Lm, qm, Nm) in enumerate(MODELS):
        rng = np.random.default_rng(100 + m)
        J = rng.normal(0, 0.3 if qm < 10 else 0.1, size=(Lm * (Lm - 1) // 2, qm * qm))
        J = J.astype('<f4')
        s0 = rng.integers(0, qm, size=(Nm, Lm)).astype('u1')
        b.set_couplings(m, J)
        b.set_seqs(m, s0)
        Js.append(J)
        S0.append(s0)
    return b, Js, S0


def _check_model(got, s0, J, seed, steps):
    """`steps`: list of (t0, nsteps) launches."""
    want, margin = s0, None
    for t0, n in steps:
        want, mg = cboracle.metropolis(want, J, seed, t0, n)
        margin = mg if margin is None else np.minimum(margin, mg)
    same = np.all(got == want, axis=1)
    safe = margin > 1e-3
    assert np.all(same[safe]), f"{(~same[safe]).sum()} walkers differ"
    assert same.mean() > 0.98
    return same.mean()


def test_batch_chains_match_single_model_oracle(gpu_ok):
    b, Js, S0 = _setup()
    b.run(1)
    b.run(1)                      # graph replay: step counters moved on
    b.wait()
    for m, (L, q, N) in enumerate(MODELS):
        got = b.get_seqs(m)
        n = L
        t1 = n + (n & 1)
        frac = _check_model(got, S0[m], Js[m], 40 + m, [(0, n), (t1, n)])
        assert b.step_count(m) == 2 * t1
        print(f"model {m} L={L} q={q}: {frac * 100:.2f}% identical, tile {b.tile(m)}")
        assert np.mean(got != S0[m]) > 0.01
    b.close()


@pytest.mark.parametrize("wpb", [128, 256, 512])
def test_batch_tiles_give_identical_chains(gpu_ok, wpb):
    b1, Js, S0 = _setup()
    b2, _, _ = _setup()
    for m, (L, q, N) in enumerate(MODELS):
        if wpb <= 256 or L <= 512:      # 512-walker tiles hold at most 512 sites
            b2.set_tile(m, wpb)
    b1.run(2)
    b2.run(2)
    for m in range(len(MODELS)):
        np.testing.assert_array_equal(b1.get_seqs(m), b2.get_seqs(m))
    b1.close()
    b2.close()


def test_batch_tune_and_errors(gpu_ok):
    from mi3gpu.batch import PottsBatch
    b, Js, S0 = _setup()
    b.tune(1)
    for m, (L, q, N) in enumerate(MODELS):
        assert b.tile(m) in (128, 256, 512)
        assert b.step_count(m) > 0
    b.run(1)
    b.wait()
    with pytest.raises(ValueError):
        b.set_tile(4, 1024)           # 128, 256 or 512 walkers per tile
    with pytest.raises(ValueError):
        b.set_seqs(0, np.full((512, 20), 9, 'u1'))
    with pytest.raises(ValueError):
        b.set_tile(0, 64)
    with pytest.raises(ValueError):
        b.set_tile(6, 512)            # L = 600: 128 or 256 walkers per tile
    with pytest.raises(ValueError):
        PottsBatch([10], [4], [100])
    b.close()
