"""numpy restatement of the Mi3 inference-round kernels (test infrastructure).

This module is the checker for the HIP engine.  Every function cites the
reference file:line whose behaviour it restates.  Floating-point work is done
in float64 (the reference kernels use float32); where the reference's float32
arithmetic is part of the contract (bicounts_to_bimarg) the float32 result is
reproduced exactly.

Pinning (see tests/test_oracle.py):
  pair_list / unpack_full   file-format statement (UserGuide.md:176-180),
                            round trip pack(unpack(J)) == J
  energies                  brute force on tiny models, J=0, constant J
                            (c*L(L-1)/2), single nonzero entry, fieldless-gauge
                            shift invariance
  bicounts                  independent brute-force triple loop, row sums = N,
                            unimarginal consistency (validate_bimarg)
  delta_energy              E(s') - E(s) from `energies`
  metropolis                exact enumeration chi-square at L=4, q=3; closed
                            form independent-site marginals
  gen_indep                 chi-square against the product distribution
  updatedJ / reg_*          closed forms on hand-built 2x2 / 3x3 blocks,
                            zero-gauge invariants (row/col sums 0)
  min/weights/stats         numpy min / sum on the same vectors
  weighted_marg             w = 1 reduces to bicounts; linearity in w
  iter_newton               one iteration by hand; Neff stop rule

No function here is used by the product path.
"""
import numpy as np

from . import philox


# ----------------------------------------------------------------------------
# formats


def nP_of(L):
    return L * (L - 1) // 2


def getLq(J):
    """L, q from a packed (L(L-1)/2, q*q) array (potts_common.py:25-29)."""
    L = int(((1 + np.sqrt(1 + 8 * J.shape[0])) // 2) + 0.5)
    q = int(np.sqrt(J.shape[1]) + 0.5)
    return L, q


def pair_list(L):
    """Pair order of the packed format: [(i,j) for i<j] (UserGuide.md:176-180,
    potts_common.py:31-32)."""
    return [(i, j) for i in range(L - 1) for j in range(i + 1, L)]


def unpack_full(J, L, q):
    """(L, L, q, q) symmetric form: F[i,j,a,b] = J[pair(i,j), a*q+b],
    F[j,i,b,a] = the same value, zero diagonal blocks.  Restates unpackfV
    (mcmc.cl:43-65) and the CPU sampler's expansion
    (extras/mcmcCPUgenThreaded.c:193-206); the reference leaves the diagonal
    unfilled and skips it (mcmc.cl:391), which is the same as zero."""
    F = np.zeros((L, L, q, q), dtype=np.float64)
    i, j = np.triu_indices(L, k=1)
    blocks = np.asarray(J, dtype=np.float64).reshape(-1, q, q)
    F[i, j] = blocks
    F[j, i] = blocks.transpose(0, 2, 1)
    return F


def pack_full(F):
    L, _, q, _ = F.shape
    i, j = np.triu_indices(L, k=1)
    return F[i, j].reshape(-1, q * q)


# ----------------------------------------------------------------------------
# energies, counts, marginals


def energies(seqs, J):
    """E(s) = sum_{i<j} J[pair(i,j), q*s_i + s_j], accumulated in float64.
    Restates getSeqEnergies.E_potts (getSeqEnergies.py:38-60); the GPU
    kernel it checks is getEnergies (mcmc.cl:249-345)."""
    seqs = np.asarray(seqs)
    L, q = getLq(J)
    assert seqs.shape[-1] == L
    s = seqs.astype(np.int64)
    J = np.asarray(J, dtype=np.float64)
    E = np.zeros(s.shape[0], dtype=np.float64)
    n = 0
    for i in range(L - 1):
        qsi = q * s[:, i]
        for j in range(i + 1, L):
            E += J[n, qsi + s[:, j]]
            n += 1
    return E


def bicounts(seqs, q):
    """count_ij(a,b) = #{s: s_i=a, s_j=b}, pairs in packed order.
    Restates BiCounter (getMarginals.py:55-87, np.bincount per pair); the
    GPU kernel it checks is countBivariate (mcmc.cl:464-516)."""
    seqs = np.asarray(seqs).astype(np.int64)
    N, L = seqs.shape
    out = np.zeros((nP_of(L), q * q), dtype=np.int64)
    n = 0
    for i in range(L - 1):
        qsi = q * seqs[:, i]
        for j in range(i + 1, L):
            out[n] = np.bincount(qsi + seqs[:, j], minlength=q * q)
            n += 1
    return out


def bicounts_to_bimarg(counts, nseq):
    """float32(count) / float32(nseq), bit-for-bit (mcmc.cl:518-528)."""
    return np.asarray(counts).astype(np.float32) / np.float32(nseq)


def unimarg(bimarg):
    """Univariate marginals from packed bimarginals (potts_common.py:35-40)."""
    L, q = getLq(bimarg)
    ff = np.asarray(bimarg, dtype=np.float64).reshape(-1, q, q)
    pl = pair_list(L)
    f = np.zeros((L, q))
    f[0] = ff[0].sum(axis=1)
    for n, (i, j) in enumerate(pl):
        if i == 0:
            f[j] = ff[n].sum(axis=0)
    return f / f.sum(axis=1, keepdims=True)


# ----------------------------------------------------------------------------
# sampler


def delta_energy(seq, F, pos, b):
    """dE = sum_{m != pos} F[pos,m,b,s_m] - F[pos,m,a,s_m] with a = s_pos
    (mcmc.cl:360-413, extras/mcmcCPUgenThreaded.c:68-84)."""
    L = F.shape[0]
    m = np.arange(L)
    a = seq[pos]
    sm = np.asarray(seq, dtype=np.int64)
    return float(np.sum(F[pos, m, b, sm]) - np.sum(F[pos, m, a, sm]))


def metropolis(seqs, J, seed, t0, nsteps, beta=1.0, walker0=0):
    """Single-site Metropolis, all walkers sharing one random site per step.

    Restates metropolis (mcmc.cl:415-458): the proposal residue is uniform over
    all q letters (including the current one), acceptance iff
    exp(-beta*dE) > u with u a 24-bit uniform.  The random streams are the
    counter-based ones of oracle.philox.  Returns (new seqs, min acceptance
    margin per walker), where the margin is min |exp(-beta*dE) - u| over the
    walker's steps (used to skip walkers whose decisions are within float32
    round-off of the threshold)."""
    seqs = np.array(seqs, dtype=np.uint8, copy=True)
    N, L = seqs.shape
    Lq = getLq(J)
    q = Lq[1]
    F = unpack_full(J, L, q)
    pos_list = philox.site_positions(seed, L, t0, nsteps)
    walkers = np.arange(walker0, walker0 + N, dtype=np.uint32)
    margin = np.full(N, np.inf)
    rows = np.arange(N)
    for t in range(nsteps):
        pos = int(pos_list[t])
        b, u = philox.walker_draws(seed, q, walkers, t0 + t)
        s = seqs.astype(np.int64)
        a = s[:, pos]
        Fp = F[pos]                                  # (L, q, q)
        m = np.arange(L)
        dE = (Fp[m[None, :], b.astype(np.int64)[:, None], s].sum(axis=1)
              - Fp[m[None, :], a[:, None], s].sum(axis=1))
        p = np.exp(-beta * dE)
        acc = p > u.astype(np.float64)
        margin = np.minimum(margin, np.abs(p - u))
        seqs[rows[acc], pos] = b[acc]
    return seqs, margin


def prepare_indep(unimarg_):
    """Cumulative probabilities (L, q-1) used by the independent reseed
    (mcmcGPU.py:375-378)."""
    c = np.cumsum(np.asarray(unimarg_, dtype=np.float64), axis=1)
    return (c[:, :-1] / c[:, -1:]).astype(np.float32)


def gen_indep(cprob, seed, call, N, walker0=0):
    """Independent-site sequences: residue = #{k: u >= cprob[i,k]}, the
    result of the binary search in gen_indep (mcmc.cl:170-244) for a
    monotone cprob row."""
    L = cprob.shape[0]
    u = philox.indep_uniforms(seed, call, np.arange(walker0, walker0 + N), L)
    return (u[:, :, None] >= cprob[None, :, :]).sum(axis=2).astype(np.uint8)


# ----------------------------------------------------------------------------
# quasi-Newton phase


def updatedJ(bimarg_target, bimarg, gamma, pc, Ji):
    """Jo = Ji - gamma*(f_target - f)/(f + pc) (mcmc.cl:829-843), float64."""
    bt = np.asarray(bimarg_target, dtype=np.float64)
    bi = np.asarray(bimarg, dtype=np.float64)
    return np.asarray(Ji, dtype=np.float64) - gamma * (bt - bi) / (bi + pc)


def zero_gauge_blocks(J):
    """Per q x q block: J - rowmean - colmean + mean (mcmc.cl:764-822;
    changeGauge.zeroJGauge's coupling part, changeGauge.py:64-71)."""
    nP, qq = J.shape
    q = int(round(np.sqrt(qq)))
    B = np.asarray(J, dtype=np.float64).reshape(nP, q, q)
    hi = B.mean(axis=2, keepdims=True)
    hj = B.mean(axis=1, keepdims=True)
    mu = B.mean(axis=(1, 2), keepdims=True)
    return (B - hi - hj + mu).reshape(nP, qq)


def _reg_z(bimarg, gamma, pc, lJ, J, dJ, kind):
    J = np.asarray(J, dtype=np.float64)
    dJ = np.asarray(dJ, dtype=np.float64)
    bi = np.asarray(bimarg, dtype=np.float64)
    J0 = zero_gauge_blocks(J + dJ)
    if kind == 'l1':
        R = -lJ * np.sign(J0) * gamma / (bi + pc)
    else:
        R = -lJ * J0 * gamma / (bi + pc)
    flip = np.sign(J0) != np.sign(J0 + R)
    out = np.where(flip, dJ - J0, dJ + R)
    # decisions within float32 reach of the sign boundary
    scale = np.abs(J) + np.abs(dJ) + 1e-30
    ambiguous = (np.abs(J0) < 1e-5 * scale) | (np.abs(J0 + R) < 1e-5 * scale)
    return out, ambiguous


def reg_l1z(bimarg, gamma, pc, lJ, J, dJ):
    """L1 step on zero-gauged J+dJ with sign-flip clipping (mcmc.cl:845-872).
    Returns (new dJ, ambiguous-decision mask)."""
    return _reg_z(bimarg, gamma, pc, lJ, J, dJ, 'l1')


def reg_l2z(bimarg, gamma, pc, lJ, J, dJ):
    """L2 analogue (mcmc.cl:874-902).  Returns (new dJ, ambiguous mask)."""
    return _reg_z(bimarg, gamma, pc, lJ, J, dJ, 'l2')


def _blocks(x, q):
    return np.asarray(x, dtype=np.float64).reshape(-1, q, q)


def block_unimarg(bimarg, q):
    """Row sums fi(a) = sum_b f(a,b) and column sums fj(b) of every pair
    block (getUnimarg, mcmc.cl:934-970)."""
    B = _blocks(bimarg, q)
    return B.sum(axis=2), B.sum(axis=1)


def block_X(bimarg, Jt, q, pc):
    """Per block: C = f - fi fj, X = sum Jt*C, Xnorm = sum C^2/(f + pc)
    (reg_X / reg_Xij / reg_SCADX / reg_expX, mcmc.cl:972-1150)."""
    F = _blocks(bimarg, q)
    T = _blocks(Jt, q)
    fi, fj = F.sum(axis=2), F.sum(axis=1)
    C = F - fi[:, :, None] * fj[:, None, :]
    X = np.sum(T * C, axis=(1, 2))
    Xn = np.sum(C * C / (F + pc), axis=(1, 2))
    return C, X, Xn, fi, fj


def block_Xijab(Jt, fi, fj, q):
    """Xijab(c,d) = sum_ab Jt(a,b) (fi(a) - [a=c]) (fj(b) - [b=d])
    (getXijab, mcmc.cl:1052-1067), as one contraction per block."""
    T = _blocks(Jt, q)
    I = np.eye(q)
    Fi = fi[:, :, None] - I[None]          # [p, a, c]
    Fj = fj[:, :, None] - I[None]          # [p, b, d]
    return np.einsum('pab,pac,pbd->pcd', T, Fi, Fj)


def reg_SCADJ(bimarg, gamma, pc, lJ, a, J, dJ):
    """SCAD step on the zero-gauge J0 of J+dJ (reg_SCADJ, mcmc.cl:904-932):
    R = min(lJ, |J0|/gamma) for |J0| < lJ, (a lJ - |J0|)/(a-1) for
    |J0| < a lJ, else 0; dJ -= R gamma sign(J0)/(f + pc).  The step is
    continuous in J0, but for |J0| < lJ gamma it equals J0/(f + pc), which
    magnifies the float32 rounding of J0 by 1/(f + pc): those entries are
    returned as ambiguous.  Returns (new dJ, ambiguous mask)."""
    J, dJ = np.asarray(J, np.float64), np.asarray(dJ, np.float64)
    bi = np.asarray(bimarg, np.float64)
    J0 = zero_gauge_blocks(J + dJ)
    aJ = np.abs(J0)
    R = np.where(aJ < lJ, np.minimum(lJ, aJ / gamma),
                 np.where(aJ < a * lJ, (a * lJ - aJ) / (a - 1), 0.0))
    amb = aJ < 1.01 * lJ * gamma
    return dJ - R * gamma * np.sign(J0) / (bi + pc), amb


def reg_X(bimarg, gamma, pc, lX, J, dJ, lambdas=None):
    """X penalty (reg_X, mcmc.cl:1004-1028; with per-pair lambdas, reg_Xij
    972-1002): lX' = min(lX, |X|/(Xnorm gamma)),
    dJ -= gamma lX' C sign(X)/(f + pc)."""
    J, dJ = np.asarray(J, np.float64), np.asarray(dJ, np.float64)
    bi = np.asarray(bimarg, np.float64)
    nP, qq = bi.shape
    q = int(round(np.sqrt(qq)))
    C, X, Xn, _, _ = block_X(bi, J + dJ, q, pc)
    lam = np.full(nP, float(lX)) if lambdas is None else np.asarray(lambdas, np.float64)
    with np.errstate(divide='ignore', invalid='ignore'):
        lXe = np.fmin(lam, np.abs(X) / (Xn * gamma))
    R = (lXe * np.sign(X))[:, None] * C.reshape(nP, qq)
    return dJ - gamma * R / (bi + pc)


def reg_SCADX(bimarg, gamma, pc, s, r, a, J, dJ):
    """SCAD penalty on X (reg_SCADX, mcmc.cl:1077-1124).  Continuous in X."""
    J, dJ = np.asarray(J, np.float64), np.asarray(dJ, np.float64)
    bi = np.asarray(bimarg, np.float64)
    nP, qq = bi.shape
    q = int(round(np.sqrt(qq)))
    Jt = J + dJ
    C, X, Xn, fi, fj = block_X(bi, Jt, q, pc)
    Xab = block_Xijab(Jt, fi, fj, q).reshape(nP, qq)
    aX = np.abs(X)
    with np.errstate(divide='ignore', invalid='ignore'):
        rr = np.fmin(r, aX / (Xn * gamma * s))
    R = np.where(aX < r, rr * np.sign(X),
                 np.where(aX < a * r, (a * r * np.sign(X) - X) / (a - 1), 0.0))
    grad = C.reshape(nP, qq) + bi * (X[:, None] - Xab)
    return dJ - R[:, None] * gamma * s * grad / (bi + pc)


def reg_expX(bimarg, gamma, pc, lam, J, dJ):
    """exp(-|X|/lam) penalty (reg_expX, mcmc.cl:1126-1150).  Returns
    (new dJ, ambiguous mask): blocks with X ~ 0, where sign(X) jumps."""
    J, dJ = np.asarray(J, np.float64), np.asarray(dJ, np.float64)
    bi = np.asarray(bimarg, np.float64)
    nP, qq = bi.shape
    q = int(round(np.sqrt(qq)))
    Jt = J + dJ
    C, X, Xn, fi, fj = block_X(bi, Jt, q, pc)
    Xab = block_Xijab(Jt, fi, fj, q).reshape(nP, qq)
    grad = C.reshape(nP, qq) + bi * (X[:, None] - Xab)
    R = (np.sign(X) * np.exp(-np.abs(X) / lam))[:, None] * grad
    scale = np.sum(np.abs(Jt.reshape(nP, qq) * C.reshape(nP, qq)), axis=1)
    amb = np.repeat((np.abs(X) <= 1e-5 * scale)[:, None], qq, axis=1)
    return dJ - gamma * R / (bi + pc), amb


def _double_diffs(Jt, q):
    """dd[p, a, b, a2, b2] = Jt(a,b) - Jt(a2,b) - Jt(a,b2) + Jt(a2,b2)."""
    T = _blocks(Jt, q)
    a_b = T[:, :, :, None, None]                          # Jt(a, b)
    a2_b = T.transpose(0, 2, 1)[:, None, :, :, None]       # Jt(a2, b)
    a_b2 = T[:, :, None, None, :]                         # Jt(a, b2)
    a2_b2 = T[:, None, None, :, :]                        # Jt(a2, b2)
    return a_b - a2_b - a_b2 + a2_b2


def reg_ddE(bimarg, gamma, pc, lam, J, dJ, scad_r=None):
    """Double-difference penalty (reg_ddE, mcmc.cl:1152-1187; with scad_r,
    reg_SCADddE 1191-1243): over all a2 != a, b2 != b,
      ddE = Jt(a,b) - Jt(a2,b) - Jt(a,b2) + Jt(a2,b2),
      ddE:     dR = sum sign(ddE) / (q-1)^2,       dJ -= lam dR gamma/(f+pc)
      SCADddE: dR = sum lam sign(ddE) R(|ddE|) / (q-1)^2, dJ -= gamma dR/(f+pc)
    with R = 1 below r, (4r - |ddE|)/(3r) below 4r, else 0.  Returns
    (new dJ, ambiguous mask): entries with some |ddE| ~ 0 (sign jump)."""
    J, dJ = np.asarray(J, np.float64), np.asarray(dJ, np.float64)
    bi = np.asarray(bimarg, np.float64)
    nP, qq = bi.shape
    q = int(round(np.sqrt(qq)))
    dd = _double_diffs(J + dJ, q)                       # [p, a, b, a2, b2]
    off = (1 - np.eye(q))[:, None, :, None] * (1 - np.eye(q))[None, :, None, :]
    sg = np.sign(dd) * off[None]
    if scad_r is None:
        dR = sg.sum(axis=(3, 4)) / (q - 1) ** 2
        step = lam * dR.reshape(nP, qq) * gamma / (bi + pc)
    else:
        r = scad_r
        ad = np.abs(dd)
        R = np.where(ad < r, 1.0, np.where(ad < 4 * r, (4 * r - ad) / (4 * r - r), 0.0))
        dR = (lam * sg * R).sum(axis=(3, 4)) / (q - 1) ** 2
        step = gamma * dR.reshape(nP, qq) / (bi + pc)
    scale = np.abs(_blocks(J + dJ, q)).max(axis=(1, 2))[:, None, None, None, None]
    amb = ((np.abs(dd) <= 1e-5 * scale) & (off[None] > 0)).any(axis=(3, 4)).reshape(nP, qq)
    return dJ - step, amb


def min_floats(x):
    """min(0, min x): the accumulator of minFloats starts at 0 (mcmc.cl:568)."""
    return min(0.0, float(np.min(x)))


def dE_to_weights(dE, ref_E):
    """w = exp(-dE + ref_E) (mcmc.cl:624-633)."""
    return np.exp(-np.asarray(dE, dtype=np.float64) + ref_E)


def weight_stats(w):
    """(sum w, sum w^2) (mcmc.cl:586-622)."""
    w = np.asarray(w, dtype=np.float64)
    return float(w.sum()), float((w * w).sum())


def weighted_marg(seqs, w, q):
    """F_ij(a,b) = sum_s w_s [s_i=a][s_j=b], unnormalised (mcmc.cl:646-718;
    the weighted branch of BiCounter, getMarginals.py:72-86)."""
    seqs = np.asarray(seqs).astype(np.int64)
    w = np.asarray(w, dtype=np.float64)
    N, L = seqs.shape
    out = np.zeros((nP_of(L), q * q))
    n = 0
    for i in range(L - 1):
        qsi = q * seqs[:, i]
        for j in range(i + 1, L):
            out[n] = np.bincount(qsi + seqs[:, j], weights=w, minlength=q * q)
            n += 1
    return out


def renormalize(bi):
    """Divide each pair's q*q block by its sum (mcmc.cl:729-759)."""
    bi = np.asarray(bi, dtype=np.float64)
    return bi / bi.sum(axis=1, keepdims=True)


def iter_newton(seqs, J, bimarg_target, gamma, pc, newton_steps, frac_neff,
                reg=None, reg_lambda=0.0, energies_fn=None, wmarg_fn=None, beta=None):
    """The Zwanzig-reweighting quasi-Newton loop of one round on one GPU
    (NewtonSteps.py:159-268 with merge_bimarg = renormalize on one device,
    node_manager.py:283-287).

    With `beta`, the sequences were sampled at inverse temperature beta and
    are reweighted to beta = 1 (NewtonSteps.py:206-214, 223-224): every
    iteration's energies gain -(beta-1)*E(s) under the starting couplings,
    the first reference energy is the minimum of that term, and the stop
    rule compares Neff with the Neff of the reweighting alone.

    Returns (iterations i, trialJ = J + dJ, predicted bimarg, weights,
    Neff history).  `energies_fn` / `wmarg_fn` can substitute a faster
    implementation of the same mathematics (the C oracle)."""
    seqs = np.asarray(seqs)
    N, L = seqs.shape
    q = getLq(J)[1]
    efn = energies_fn or energies
    wfn = wmarg_fn or (lambda s, w: weighted_marg(s, w, q))
    bi = renormalize(bicounts_to_bimarg(bicounts(seqs, q), N).astype(np.float64))
    dJ = np.zeros_like(np.asarray(J, dtype=np.float64))
    N0 = N
    ref_dE = 0.0
    beta_mod = None
    if beta is not None:
        beta_mod = -(beta - 1) * np.asarray(efn(seqs, J), dtype=np.float64)
        ref_dE = float(np.min(beta_mod))
        w0 = np.exp(-(beta_mod - ref_dE))
        N0 = np.sum(w0) ** 2 / np.sum(w0 ** 2)
    neffs = []
    w = None
    i = 0
    for i in range(newton_steps):
        dJ = updatedJ(bimarg_target, bi, gamma, pc, dJ)
        if reg == 'l1z':
            dJ, _ = reg_l1z(bi, gamma, pc, reg_lambda, J, dJ)
        elif reg == 'l2z':
            dJ, _ = reg_l2z(bi, gamma, pc, reg_lambda, J, dJ)
        dE = efn(seqs, dJ)
        if beta_mod is not None:
            dE = dE + beta_mod
        mindE = min_floats(dE)
        w = dE_to_weights(dE, ref_dE)
        s1, s2 = weight_stats(w)
        bi = renormalize(wfn(seqs, w))
        neff = s1 * s1 / s2
        neffs.append(neff)
        if neff < frac_neff * N0:
            break
        ref_dE = mindE
    trialJ = np.asarray(J, dtype=np.float64) + dJ
    return i, trialJ, bi, w, neffs


# ----------------------------------------------------------------------------
# gauges (host-side helpers of the round, changeGauge.py:37-155)


def zeroJGauge(h, J):
    """changeGauge.zeroJGauge (changeGauge.py:37-84), unweighted."""
    L, q = h.shape
    Jx = np.asarray(J, dtype=np.float64).reshape(-1, q, q)
    mJ1, mJ2 = Jx.mean(axis=1), Jx.mean(axis=2)
    mJ = mJ1.mean(axis=1)
    J0 = (Jx - mJ1[:, None, :] - mJ2[:, :, None] + mJ[:, None, None]).reshape(-1, q * q)
    h0 = np.asarray(h, dtype=np.float64) - mJ.sum() / L
    for n, (i, j) in enumerate(pair_list(L)):
        h0[j] += mJ1[n]
        h0[i] += mJ2[n]
    return h0, J0


def fieldlessGaugeEven(h, J):
    """zeroGauge then distribute fields evenly (changeGauge.py:86-155),
    written with explicit loops as an independent check of the vectorised
    product version."""
    L, q = h.shape
    h0, J0 = zeroJGauge(h, J)
    h0 = h0 - h0.mean(axis=1, keepdims=True)
    out = J0.copy()
    hd = h0 / (L - 1)
    for n, (i, j) in enumerate(pair_list(L)):
        out[n] += np.repeat(hd[i], q)
        out[n] += np.tile(hd[j], q)
    return np.zeros_like(h0), out


def jeffreys_pseudocount(ff, N):
    """C-preserving Jeffreys pseudocount mu = q/(2N+q) (pseudocount.py:8-24)."""
    L, q = getLq(ff)
    mu = q / (2.0 * N + q)
    f = unimarg(ff)
    fifj = np.array([np.add.outer(f[i], f[j]).ravel() for i, j in pair_list(L)])
    out = (1 - mu) ** 2 * ff + (1 - mu) * mu * fifj / q + (mu / q) ** 2
    return out / out.sum(axis=1, keepdims=True)
