"""Learned proposal for the Potts sampler: a masked autoregressive MLP
(MADE, Germain et al. 2015) over sequence positions, trained on sequences
the MCMC produced and used as an independence Metropolis-Hastings proposal
mixed with single-site sweeps.  Acceptance uses the exact Potts energy, so
the stationary distribution is unchanged.

Forward and backward passes, the Adam step and the sampler are the
library's HIP kernels (bf16 matrix-core GEMMs with float32 accumulation).
Device memory is held in named buffers of a small library context of its
own; kernels run on a given HIP stream (by default that context's), so
they can be ordered with an engine context's work without extra waits.
The reference has no counterpart.
"""
import ctypes

import numpy as np

from mi3gpu._lib import lib, check, MI3_EINVAL, MI3_EUNSUP, Mi3Error


def _mk(rc):
    if rc >= 0:
        return rc
    msg = lib().mi3_made_last_error()
    msg = msg.decode(errors='replace') if msg else ''
    if rc in (MI3_EINVAL, MI3_EUNSUP):
        raise ValueError(msg)
    raise Mi3Error(rc, msg)


def made_masks(L, q, H):
    """Degree-based MADE masks for inputs/outputs ordered by site: input
    site i has degree i+1, hidden unit k of both layers degree
    1 + k mod (L-1), output site j sees hidden units of degree <= j.
    Returns (M1t (L*q, H), M2 (H, H), M3 (L*q, H)) as float32 0/1 arrays."""
    deg_in = np.repeat(np.arange(1, L + 1), q)              # (L*q,)
    m = 1 + (np.arange(H) % max(L - 1, 1))                 # (H,)
    M1t = (m[None, :] >= deg_in[:, None]).astype('f4')     # input (L*q) -> hidden
    M2 = (m[:, None] >= m[None, :]).astype('f4')           # hidden -> hidden
    out_site = np.repeat(np.arange(L), q)
    M3 = (m[None, :] <= out_site[:, None]).astype('f4')    # hidden -> output
    return M1t, M2, M3


def _addr(p, offset=0):
    return ctypes.c_void_p(p.value + offset)


class DeviceArrays:
    """Named device arrays in a private library context (its own stream and
    memory), addressed by raw pointer."""

    def __init__(self, device=0):
        self.ctx = ctypes.c_void_p()
        check(None, lib().mi3_create(device, 2, 2, 512, 0, 0, ctypes.byref(self.ctx)))
        s = ctypes.c_void_p()
        check(self.ctx, lib().mi3_stream(self.ctx, ctypes.byref(s)))
        self.stream = s
        self.meta = {}
        self.ptrs = {}

    def alloc(self, name, shape, dtype):
        if name in self.meta:
            raise ValueError(f"array {name!r} exists")
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        check(self.ctx, lib().mi3_buf_alloc(self.ctx, name.encode(), max(nbytes, 16)))
        self.meta[name] = (tuple(shape), np.dtype(dtype))
        p = ctypes.c_void_p()
        check(self.ctx, lib().mi3_buf_device_ptr(self.ctx, name.encode(), ctypes.byref(p)))
        self.ptrs[name] = p
        return p

    def __getitem__(self, name):
        return self.ptrs[name]

    def write(self, name, arr):
        shape, dt = self.meta[name]
        a = np.ascontiguousarray(arr, dtype=dt).reshape(shape)
        check(self.ctx, lib().mi3_buf_write(self.ctx, name.encode(), a.ctypes.data, a.nbytes, 0))

    def read(self, name, count=None):
        """Contents of an array (or its first `count` elements), in the
        private stream's order."""
        shape, dt = self.meta[name]
        out = np.empty(shape if count is None else (count,), dtype=dt)
        t = ctypes.c_int(-1)
        check(self.ctx, lib().mi3_buf_read_async(self.ctx, name.encode(), out.nbytes, 0,
                                                 ctypes.byref(t)))
        check(self.ctx, lib().mi3_ticket_wait(self.ctx, t.value, out.ctypes.data, out.nbytes))
        return out

    def sync(self):
        check(self.ctx, lib().mi3_synchronize(self.ctx))

    def close(self):
        if getattr(self, 'ctx', None):
            lib().mi3_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


PARAMS = ['W1t', 'b1', 'W2', 'b2', 'W3', 'b3']
MASK_OF = {'W1t': 'M1t', 'W2': 'M2', 'W3': 'M3'}
BF16_OF = {'W2': 'W2b', 'W3': 'W3b'}


class MADE:
    """MADE over L sites x q letters with two hidden layers of H units, for
    batches of at most `max_batch` sequences (workspace sized once).

    Parameters (float32, device): W1t (L*q, H), b1 (H), W2 (H, H), b2 (H),
    W3 (L*q, H), b3 (L*q), kept masked, so masked entries stay zero under
    training; bf16 copies W2b, W3b feed the matrix cores.  Sequences are
    site-major u8 (L rows of `ld` bytes), as the sampler stores walkers."""

    def __init__(self, L, q, H=512, max_batch=1024, seed=0, unimarg=None, device=0):
        if (L * q) % 8 or H % 8 or max_batch % 8 or max_batch <= 0:
            raise ValueError("L*q, H and max_batch must be positive multiples of 8")
        self.L, self.q, self.H, self.max_batch = L, q, H, max_batch
        self.mem = m = DeviceArrays(device)
        g = np.random.default_rng(seed)
        M1t, M2, M3 = made_masks(L, q, H)
        Lq = L * q
        init = {
            'W1t': g.normal(0, 1 / np.sqrt(L), (Lq, H)) * M1t,
            'b1': np.zeros(H),
            'W2': g.normal(0, np.sqrt(2 / H), (H, H)) * M2,
            'b2': np.zeros(H),
            'W3': g.normal(0, 1 / np.sqrt(H), (Lq, H)) * M3,
            'b3': (np.log(np.clip(np.asarray(unimarg, np.float64), 1e-6, None)).ravel()
                   if unimarg is not None else np.zeros(Lq)),
        }
        self.shape = {k: np.shape(v) for k, v in init.items()}
        for k, v in init.items():
            for pre in ('', 'g_', 'm_', 'v_'):
                m.alloc(pre + k, self.shape[k], 'f4')
            m.write(k, v)
        for k, M in (('M1t', M1t), ('M2', M2), ('M3', M3)):
            m.alloc(k, M.shape, 'f4')
            m.write(k, M)
        m.alloc('W2b', (H, H), 'u2')
        m.alloc('W3b', (Lq, H), 'u2')
        B = max_batch
        for name, shape, dt in (('h1', (B, H), 'f4'), ('h1b', (B, H), 'u2'),
                                ('h2', (B, H), 'f4'), ('h2b', (B, H), 'u2'),
                                ('logits', (B, Lq), 'f4'), ('nll', (B,), 'f4'),
                                ('dlb', (B, Lq), 'u2'), ('dlf', (B, Lq), 'f4'),
                                ('dh2', (B, H), 'f4'), ('dh2p', (B, H), 'f4'),
                                ('dh2pb', (B, H), 'u2'), ('dh1', (B, H), 'f4'),
                                ('dh1p', (B, H), 'f4')):
            m.alloc(name, shape, dt)
        self.t = 0
        st = m.stream
        _mk(lib().mi3_to_bf16(st, H * H, m['W2'], m['W2b']))
        _mk(lib().mi3_to_bf16(st, Lq * H, m['W3'], m['W3b']))
        m.sync()

    # -- helpers ---------------------------------------------------------------

    def _st(self, stream):
        return stream if stream is not None else self.mem.stream

    def _check_batch(self, B, ld):
        if B <= 0 or B > self.max_batch or B % 8:
            raise ValueError(f"batch of {B}: need a multiple of 8 in [8, {self.max_batch}]")
        if ld < B:
            raise ValueError("row stride smaller than the batch")

    def _gemm(self, st, ta, tb, M, N, K, A, lda, B, ldb, C=None, ldc=0, Cb=None, ldcb=0,
              bias=None, relu=0):
        _mk(lib().mi3_gemm_bf16(st, ta, tb, M, N, K, 1.0, A, lda, B, ldb, 0.0, C, ldc, Cb, ldcb,
                                bias, relu))

    def get(self, name):
        """A parameter, gradient ('g_' + name) or workspace array as numpy,
        read in the private stream's order: synchronise a foreign stream
        first."""
        self.mem.sync()
        return self.mem.read(name)

    def set(self, name, value):
        """Overwrite a parameter (masked as stored; bf16 copies refreshed)."""
        m = self.mem
        v = np.asarray(value, np.float32).reshape(self.shape[name])
        if name in MASK_OF:
            v = v * dict(zip(('W1t', 'W2', 'W3'), made_masks(self.L, self.q, self.H)))[name]
        m.write(name, v)
        if name in BF16_OF:
            _mk(lib().mi3_to_bf16(m.stream, v.size, m[name], m[BF16_OF[name]]))
        m.sync()

    # -- forward / backward ----------------------------------------------------

    def _forward(self, seq, ld, B, st):
        L, q, H, Lq, m = self.L, self.q, self.H, self.L * self.q, self.mem
        _mk(lib().mi3_made_embed(st, seq, ld, B, L, q, H, m['W1t'], m['b1'], m['h1'], m['h1b']))
        self._gemm(st, 0, 1, B, H, H, m['h1b'], H, m['W2b'], H, C=m['h2'], ldc=H, Cb=m['h2b'],
                   ldcb=H, bias=m['b2'], relu=1)
        self._gemm(st, 0, 1, B, Lq, H, m['h2b'], H, m['W3b'], H, C=m['logits'], ldc=Lq,
                   bias=m['b3'])

    def loss_and_grads(self, seq, ld, B, stream=None):
        """Negative log-likelihoods of B site-major sequences at device
        address `seq` (row stride ld) into the 'nll' workspace, and the
        gradients of their mean into the 'g_*' arrays (masked by Adam)."""
        self._check_batch(B, ld)
        L, q, H, Lq, m = self.L, self.q, self.H, self.L * self.q, self.mem
        st = self._st(stream)
        self._forward(seq, ld, B, st)
        _mk(lib().mi3_made_xent(st, m['logits'], seq, ld, B, L, q, 1.0 / B, m['nll'], m['dlb'],
                                m['dlf']))
        # output layer
        self._gemm(st, 1, 0, Lq, H, B, m['dlb'], Lq, m['h2b'], H, C=m['g_W3'], ldc=H)
        _mk(lib().mi3_colsum(st, B, Lq, m['dlf'], m['g_b3']))
        self._gemm(st, 0, 0, B, H, Lq, m['dlb'], Lq, m['W3b'], H, C=m['dh2'], ldc=H)
        _mk(lib().mi3_relu_grad(st, B * H, m['dh2'], m['h2'], m['dh2p'], m['dh2pb']))
        # hidden layer
        self._gemm(st, 1, 0, H, H, B, m['dh2pb'], H, m['h1b'], H, C=m['g_W2'], ldc=H)
        _mk(lib().mi3_colsum(st, B, H, m['dh2p'], m['g_b2']))
        self._gemm(st, 0, 0, B, H, H, m['dh2pb'], H, m['W2b'], H, C=m['dh1'], ldc=H)
        _mk(lib().mi3_relu_grad(st, B * H, m['dh1'], m['h1'], m['dh1p'], None))
        # one-hot input layer
        _mk(lib().mi3_made_embed_grad(st, seq, ld, B, L, q, H, m['dh1p'], m['g_W1t']))
        _mk(lib().mi3_colsum(st, B, H, m['dh1p'], m['g_b1']))

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        st, m = self._st(stream), self.mem
        self.t += 1
        for k in PARAMS:
            n = int(np.prod(self.shape[k]))
            mask = m[MASK_OF[k]] if k in MASK_OF else None
            pb = m[BF16_OF[k]] if k in BF16_OF else None
            _mk(lib().mi3_adam_step(st, n, m[k], m['g_' + k], m['m_' + k], m['v_' + k], mask, lr,
                                    beta1, beta2, eps, self.t, pb))

    def train(self, seq, ld, N, steps=200, batch=1024, lr=1e-3, seed=0, log=None,
              log_every=25):
        """Adam on minibatches of `batch` consecutive columns (at random
        offsets) of N site-major sequences at device address `seq` (row
        stride ld), on the private stream.  The caller orders the sequences'
        producer before this call (synchronise its stream).  Returns
        [(step, mean nll)] every log_every steps and at the end."""
        batch = min(batch, self.max_batch, N - N % 8)
        self._check_batch(batch, ld)
        rng = np.random.default_rng(seed)
        hist = []
        for s in range(steps):
            off = int(rng.integers(0, N - batch + 1))
            self.loss_and_grads(_addr(seq, off), ld, batch)
            self.adam(lr)
            if s % log_every == 0 or s == steps - 1:
                v = float(np.mean(self.mem.read('nll', batch)))
                hist.append((s, v))
                if log:
                    log(f"MADE step {s}: nll {v:.4f}")
        self.mem.sync()
        return hist

    # -- sampling ------------------------------------------------------------

    def sample(self, B, out, ld, nll_out, seed, counter, stream=None, walker_base=0):
        """Draw B sequences autoregressively into the site-major buffer at
        device address `out` (row stride ld); the float32 array at nll_out
        (B) receives their negative log probabilities.  Randomness from
        Philox((site, walker_base + b, counter, 'MADE')) keyed by seed.  One
        fused launch over all sites (each hidden unit formed once, at its
        degree)."""
        self._fused(B, out, ld, nll_out, seed, counter, walker_base, 1, stream)

    def nll(self, seq, ld, B, out, stream=None):
        """Writes -log q of B site-major sequences at `seq` to the float32
        array at device address `out`: the sampling pass with the letters
        read instead of drawn, so a drawn sequence scores exactly the nll
        its draw returned."""
        self._fused(B, seq, ld, out, 0, 0, 0, 0, stream)

    def _fused(self, B, seq, ld, nll_out, seed, counter, walker_base, draw, stream):
        self._check_batch(B, ld)
        m = self.mem
        _mk(lib().mi3_made_fused(self._st(stream), self.L, self.q, self.H, B, m['W1t'], m['b1'],
                                 m['W2b'], m['b2'], m['W3b'], m['b3'], int(seed),
                                 int(counter) & 0xffffffff, int(walker_base), seq, ld, nll_out,
                                 int(draw)))

    def nll_forward(self, seq, ld, B, out, stream=None):
        """-log q of B sequences by the training path's forward pass (the
        embedding, the two GEMMs and the per-sequence softmax)."""
        self._check_batch(B, ld)
        st, m = self._st(stream), self.mem
        self._forward(seq, ld, B, st)
        _mk(lib().mi3_made_xent(st, m['logits'], seq, ld, B, self.L, self.q, 1.0, out, None,
                                None))

    def close(self):
        self.mem.close()


class MixedSampler:
    """Cycles of single-site sweeps (engine launches of nsteps updates per
    walker) followed by one independence Metropolis-Hastings step with MADE
    proposals, on an MCMCGPU context's walkers ('seq main', energies 'E
    main' under 'J').  Proposals are drawn and scored in tiles of the
    MADE's max_batch walkers; every kernel runs on the engine's stream."""

    def __init__(self, gpu, made, seed=1):
        self.g, self.made = gpu, made
        self.seed = int(seed)
        self.counter = 0
        L, N = gpu.L, gpu.nseq['main']
        if made.L != L or made.q != gpu.q:
            raise ValueError("MADE and sampler sizes differ")
        self.L, self.N = L, N
        self.tile = min(made.max_batch, N)
        if N % self.tile:
            raise ValueError("walker count must be a multiple of the MADE batch")
        ctx = gpu._ctx
        for name, nbytes in ((b'seq prop', L * N), (b'E prop', 4 * N), (b'made nll cur', 4 * N),
                             (b'made nll prop', 4 * N), (b'made naccept', 16)):
            if not lib().mi3_buf_exists(ctx, name):
                check(ctx, lib().mi3_buf_alloc(ctx, name, nbytes))
        self.ptr = {}
        for name in (b'seq main', b'seq prop', b'E main', b'E prop', b'made nll cur',
                     b'made nll prop', b'made naccept'):
            p = ctypes.c_void_p()
            check(ctx, lib().mi3_buf_device_ptr(ctx, name, ctypes.byref(p)))
            self.ptr[name] = p
        s = ctypes.c_void_p()
        check(ctx, lib().mi3_stream(ctx, ctypes.byref(s)))
        self.stream = s

    def propose(self):
        """One MH step with MADE proposals for every walker."""
        g, made, P = self.g, self.made, self.ptr
        ctx, N, L, st, T = g._ctx, self.N, self.L, self.stream, self.tile
        g.calcEnergies('main')
        self.counter += 1
        for w0 in range(0, N, T):
            made.nll(_addr(P[b'seq main'], w0), N, T, _addr(P[b'made nll cur'], 4 * w0),
                     stream=st)
            made.sample(T, _addr(P[b'seq prop'], w0), N, _addr(P[b'made nll prop'], 4 * w0),
                        self.seed, self.counter, stream=st, walker_base=w0)
        check(ctx, lib().mi3_calc_energies(ctx, b'seq prop', b'J', b'E prop', N))
        _mk(lib().mi3_mh_accept(st, N, L, P[b'seq main'], N, P[b'seq prop'], N, P[b'E main'],
                                P[b'E prop'], P[b'made nll cur'], P[b'made nll prop'], self.seed,
                                self.counter, 0, P[b'made naccept']))
        check(ctx, lib().mi3_buf_mark_written(ctx, b'seq main'))

    def step(self, launches=1):
        for _ in range(launches):
            self.g.runMCMC()
        self.propose()

    def accepted(self):
        """MH moves accepted so far (synchronises)."""
        out = np.zeros(1, np.uint32)
        t = ctypes.c_int(-1)
        ctx = self.g._ctx
        check(ctx, lib().mi3_buf_read_async(ctx, b'made naccept', 4, 0, ctypes.byref(t)))
        check(ctx, lib().mi3_ticket_wait(ctx, t.value, out.ctypes.data, 4))
        return int(out[0])
