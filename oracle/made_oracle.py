"""numpy restatement of the learned-proposal network (test infrastructure).

A masked autoregressive MLP (MADE, Germain et al. 2015) over L sites x q
letters with two relu hidden layers, written from its definition:
  h1 = relu(b1 + sum_i W1t[q*i + s_i]),  h2 = relu(W2 h1 + b2),
  logits = W3 h2 + b3,  nll = -sum_j log softmax(logits_j)[s_j],
and its gradients by the chain rule.  With bf16=True the operands are
rounded to bfloat16 where the GPU path rounds them (activations and
weights entering the matrix-core products, the output gradient), so that
the GPU kernels can be compared to it closely; with bf16=False everything
is float64, which the finite-difference test pins.

No function here is used by the product path.
"""
import numpy as np


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    b = x.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def _r(x, on):
    return bf16(x).astype(np.float64) if on else np.asarray(x, np.float64)


def forward(seqs, W1t, b1, W2, b2, W3, b3, q, bf=True):
    """seqs (B, L) -> dict of activations, logits (B, L*q) and nll (B,)."""
    seqs = np.asarray(seqs)
    B, L = seqs.shape
    cols = q * np.arange(L)[None, :] + seqs                       # (B, L)
    h1 = b1[None, :].astype(np.float64) + np.asarray(W1t, np.float64)[cols].sum(axis=1)
    if bf:
        h1 = h1.astype(np.float32).astype(np.float64)
    h1b = _r(np.maximum(h1, 0), bf)
    W2b, W3b = _r(W2, bf), _r(W3, bf)
    h2 = np.maximum(h1b @ W2b.T + b2[None, :], 0)
    if bf:
        h2 = h2.astype(np.float32).astype(np.float64)
    h2b = _r(h2, bf)
    logits = h2b @ W3b.T + b3[None, :]
    z = logits.reshape(B, L, q)
    mx = z.max(axis=2, keepdims=True)
    lse = (mx + np.log(np.exp(z - mx).sum(axis=2, keepdims=True)))[..., 0]
    zs = np.take_along_axis(z, seqs[..., None].astype(np.int64), axis=2)[..., 0]
    nll = (lse - zs).sum(axis=1)
    return dict(h1=h1, h1b=h1b, h2=h2, h2b=h2b, W2b=W2b, W3b=W3b, logits=logits, lse=lse,
                nll=nll, cols=cols)


def grads(seqs, W1t, b1, W2, b2, W3, b3, q, M1t, M2, M3, bf=True):
    """Gradients of mean(nll) with respect to every parameter, masked."""
    seqs = np.asarray(seqs)
    B, L = seqs.shape
    a = forward(seqs, W1t, b1, W2, b2, W3, b3, q, bf)
    z = a['logits'].reshape(B, L, q)
    p = np.exp(z - a['lse'][..., None])
    onehot = np.zeros_like(p)
    np.put_along_axis(onehot, seqs[..., None].astype(np.int64), 1.0, axis=2)
    dl = ((p - onehot) / B).reshape(B, L * q)
    dlb = _r(dl, bf)
    gW3 = (dlb.T @ a['h2b']) * M3
    gb3 = dl.sum(axis=0)
    dh2 = dlb @ a['W3b']
    dh2p = dh2 * (a['h2'] > 0)
    dh2pb = _r(dh2p, bf)
    gW2 = (dh2pb.T @ a['h1b']) * M2
    gb2 = dh2p.sum(axis=0)
    dh1 = dh2pb @ a['W2b']
    dh1p = dh1 * (a['h1'] > 0)
    gW1t = np.zeros_like(np.asarray(W1t, np.float64))
    np.add.at(gW1t, a['cols'].ravel(), np.repeat(dh1p, L, axis=0))
    gW1t *= M1t
    gb1 = dh1p.sum(axis=0)
    return dict(W1t=gW1t, b1=gb1, W2=gW2, b2=gb2, W3=gW3, b3=gb3), a['nll'].mean()


def log_prob_all(W1t, b1, W2, b2, W3, b3, L, q, bf=True):
    """log q(x) of every sequence of a tiny model (q**L states)."""
    import itertools
    states = np.array(list(itertools.product(range(q), repeat=L)), dtype=np.uint8)
    return states, -forward(states, W1t, b1, W2, b2, W3, b3, q, bf)['nll']
