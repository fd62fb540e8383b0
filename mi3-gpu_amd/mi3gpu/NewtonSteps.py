# This file keeps the interface and behaviour of the file of the same name
# in Mi3-GPU, Copyright 2020 Allan Haldane, and is a derived work of it.
# Mi3-GPU is free software: you can redistribute it and/or modify it under
# the terms of the GNU General Public License as published by the Free
# Software Foundation, version 3 of the License; this file is distributed
# under the same licence, WITHOUT ANY WARRANTY; without even the implied
# warranty of MERCHANTABILITY or FITNESS FOR A PARTICULAR PURPOSE.  See the
# GNU General Public License <http://www.gnu.org/licenses/> for details.
"""Inference driver: MCMC rounds and the quasi-Newton (Zwanzig reweighting)
coupling updates (NewtonSteps.py of the reference).

Same functions, arguments, log lines and output files as the reference;
all device work goes through the `gpus` interface (node_manager.GPU_node).
Parallel tempering (runMCMC_tempered, broken in the reference) is not part of
this build.
"""
import time

import numpy as np
from scipy.stats import spearmanr

import mi3gpu
from mi3gpu.utils.changeGauge import fieldlessGaugeEven
from mi3gpu.utils.seqload import writeSeqs
from mi3gpu.utils.potts_common import printsome, indepF


################################################################################
# status output


def bimarg_stats(bimarg_target, bimarg_model):
    """(rel. error over f_target > 0.01, SSR, max |delta|) (NewtonSteps.py:33-41)."""
    d = bimarg_target - bimarg_model
    big = bimarg_target > 0.01
    with np.errstate(divide='ignore', invalid='ignore'):
        ferr = np.mean(np.abs(d[big]) / bimarg_target[big])
    return ferr, np.sum(d ** 2), np.max(np.abs(d))


def printstats(name, jstep, bicount, bimarg_target, bimarg_model, couplings, energies, e_rho,
               ptinfo):
    """Round summary lines: SSR, rel%, max%, covariance energy X
    (NewtonSteps.py:43-74)."""
    ferr, ssr, maxd = bimarg_stats(bimarg_target, bimarg_model)
    X = np.sum(couplings * (bimarg_model - indepF(bimarg_model)))
    Xo = np.sum(couplings * (bimarg_target - indepF(bimarg_target)))
    rhostr = '(none)'
    if e_rho is not None:
        rho = np.array([r[0] for r in e_rho])
        rhostr = np.array2string(rho, precision=2, floatmode='fixed', suppress_small=True)[1:-1]
    disp = (f"{name} {jstep: 6d} Error: SSR:{ssr:6.3f}   rel%:{ferr * 100:5.2f}   "
            f"max%:{maxd * 100:5.2f}   X:{X: 6.1f} ({Xo: 6.1f})\n"
            f"{name} bimarg: {printsome(bimarg_model)} ...\n"
            f"{name}      J: {printsome(couplings)} ...\n"
            f"{name} min(E) = {np.min(energies):.4f}     mean(E) = {np.mean(energies):.4f}"
            f"     std(E) = {np.std(energies):.4f}\n"
            f"{name} E Autocorr vs time: {rhostr}")
    if ptinfo is not None:
        disp += f"\n{name} PT swap rate: {ptinfo[1]}"
    return disp


class _BackgroundWriter:
    """One worker thread for the per-round sequence files, so that the bzip2
    compression of N*L bytes overlaps the next GPU phase.  `flush()` waits for
    every pending write and re-raises the first failure."""

    def __init__(self):
        self._ex = None
        self._pending = []

    def submit(self, fn, *args, **kw):
        from concurrent.futures import ThreadPoolExecutor
        if self._ex is None:
            self._ex = ThreadPoolExecutor(max_workers=1)
        self._pending.append(self._ex.submit(fn, *args, **kw))

    def flush(self):
        pending, self._pending = self._pending, []
        for f in pending:
            f.result()


background_writer = _BackgroundWriter()


def flush_writes():
    """Wait for the background sequence writes of earlier rounds."""
    background_writer.flush()


def writeStatus(name, Jstep, bimarg_target, bicount, bimarg_model, couplings, seqs, energies,
                alpha, e_rho, ptinfo, outdir, log, write_seqs=True, background=True):
    """info.txt + bzip2 seqs of the round (NewtonSteps.py:76-94).  The seqs
    file is written by a background thread unless background=False; the
    array must not be modified afterwards (collect() returns a fresh one)."""
    disp = printstats(name, Jstep, bicount, bimarg_target, bimarg_model, couplings, energies,
                      e_rho, ptinfo)
    with open(outdir / name / 'info.txt', 'wt') as f:
        f.write(disp)
    if write_seqs:
        if background:
            background_writer.flush()   # at most one round's file in flight
            background_writer.submit(writeSeqs, outdir / name / 'seqs', seqs, alpha, zipf=True)
        else:
            writeSeqs(outdir / name / 'seqs', seqs, alpha, zipf=True)
    log(disp)


################################################################################
# quasi-Newton phase


def getNeff(w):
    """Effective sample size (sum w)^2 / sum w^2 (NewtonSteps.py:126-130)."""
    w = np.asarray(w, dtype=np.float64)
    return np.sum(w) ** 2 / np.sum(w ** 2)


def NewtonStatus(n, trialJ, weights, bimarg_model, bimarg_target, log):
    ferr, ssr, maxd = bimarg_stats(bimarg_target, bimarg_model)
    w = weights
    log("Predicted statistics after perturbing J:")
    log(f"    SSR:{ssr:6.3f}   rel%:{ferr * 100:5.2f}   max%:{maxd * 100:5.2f}")
    log(f"    Neff: {getNeff(w):.1f}    wspan: {np.min(w):.3e}:{np.max(w):.3e}")
    log("  trialJ:", printsome(trialJ), '...')
    log("  bimarg:", printsome(bimarg_model), '...')
    log(" weights:", printsome(weights), '...')


def iterNewton(param, bimarg_model, gpus, log):
    """Zwanzig-reweighting coupling updates on the current sequences until
    Neff < fracNeff * N (NewtonSteps.py:159-268).  Returns (steps, trialJ,
    predicted bimarg)."""
    bimarg_target = param.bimarg
    gamma, pc = param.gamma0, param.pcdamping
    newtonSteps, Nfrac = param.newtonSteps, param.fracNeff
    N = param.nwalkers
    log("")
    log(f"Perturbation optimization for up to {newtonSteps} steps:")
    if param.reg:
        log(f"Using regularization {param.reg} ({param.regarg})")
    t0 = time.time()

    seqbuf, wbuf, ebuf, etmp = 'main', 'weights', 'E main', 'E tmp'
    if param.distribute_jstep == 'head_node':
        raise Exception('"head_node" option only makes sense when using MPI')

    gpus.calcEnergies(seqbuf)
    gpus.calcBicounts(seqbuf)
    gpus.bicounts_to_bimarg(seqbuf)
    gpus.merge_bimarg()
    gpus.fillBuf('dJ', 0)

    N0 = N
    ref_dE = 0.0
    if param.beta is not None:
        E = gpus.readBufs(ebuf)
        mod = [-(param.beta - 1) * x for x in E]
        dEB = np.concatenate(mod)
        ref_dE = np.min(dEB)
        N0 = getNeff(np.exp(-(dEB - ref_dE)))
        gpus.setBuf(etmp, [m.astype('<f4') for m in mod])
        log(f"Temperature reweight decreases Neff from {N} to {N0}")

    lastNeff = 2 * N0
    i = 0
    for i in range(newtonSteps):
        gpus.updateJ(gamma, pc)
        if param.reg is not None:
            # Xij takes its lambdas from the 'Creg' buffer set once per run
            extra = () if param.reg == 'Xij' else tuple(param.regarg)
            gpus.reg(param.reg, (gamma, pc) + extra)
        gpus.calcEnergies(seqbuf, 'dJ')
        if param.beta is not None:
            gpus.addFloatBuf(ebuf, etmp)
        gpus.min_buf(ebuf)
        mindE_fut = gpus.getBuf('minout')
        gpus.dE_to_weights(seqbuf, ref_dE)      # ref_dE lags one iteration
        gpus.weight_statistics(seqbuf)
        w_fut = gpus.getBuf('weightstats')
        gpus.weightedMarg(seqbuf)
        gpus.merge_bimarg()

        stats = [w.read() for w in w_fut]
        wsum = sum(float(s[0]) for s in stats)
        wsum2 = sum(float(s[1]) for s in stats)
        Neff = wsum ** 2 / wsum2
        if i % 64 == 0 or abs(lastNeff - Neff) / N0 > 0.05 or Neff < Nfrac * N0:
            log(f"J-step {i: 5d}   Neff: {Neff:.1f}   ({Neff / N0 * 100:.1f}% of {N0})")
            lastNeff = Neff
        if Neff < Nfrac * N0:
            log(f"Ending coupling updates because Neff/N < {Nfrac:.2f}")
            break
        ref_dE = min(float(x.read()[0]) for x in mindE_fut)
    log(f"Performed {i} coupling update steps")

    bi, J, dJ = gpus.head_gpu.readBufs(['bi', 'J', 'dJ'])
    bimarg_model, trialJ = bi[0], J[0] + dJ[0]
    weights = gpus.collect(wbuf)
    gpus.setBuf('J', trialJ)
    NewtonStatus(i, trialJ, weights, bimarg_model, bimarg_target, log)
    tot = np.sum(weights)
    if not np.isfinite(tot) or tot == 0:
        raise Exception("Error: Divergence. Decrease gamma or increase pc-damping")
    log(f"Total Newton-step running time: {time.time() - t0:.1f} s")
    gpus.logProfile()
    return i, trialJ, bimarg_model


def singleNewton(bimarg, gamma, param, gpus):
    """One coupling update from the given bimarginals (NewtonSteps.py:115-123,
    which calls wrappers that do not exist; here the update is applied to J
    itself)."""
    gpus.setBuf('bi', bimarg)
    gpus.updateJ(gamma, param.pcdamping, 'J')
    return gpus.head_gpu.getBuf('J')[0].read()


def NewtonSteps(runName, param, bimarg_model, gpus, log):
    outdir = param.outdir
    if param.newtonSteps != 1:
        Jsteps, newJ, bimarg_p = iterNewton(param, bimarg_model, gpus, log)
        np.save(outdir / runName / 'predictedBimarg', bimarg_p)
        np.save(outdir / runName / 'perturbedJ', newJ)
    else:
        log("Performing single newton update step")
        newJ = singleNewton(bimarg_model, param.gamma0, param, gpus)
        Jsteps = 1
    return Jsteps, newJ


def preOpt(param, gpus, log):
    """Newton phase on the loaded sequences before the first MCMC round
    (NewtonSteps.py:273-298)."""
    J, outdir = param.couplings, param.outdir
    log("Pre-Optimization:")
    gpus.setBuf('J', J)
    gpus.calcBicounts('main')
    gpus.calcEnergies('main')
    bicount, es, seqs = gpus.collect(['bicount', 'E main', 'seq main'])
    bimarg = bicount.astype(np.float32) / np.float32(np.sum(bicount[0, :]))
    (outdir / 'preopt').mkdir(parents=True, exist_ok=True)
    writeStatus('preopt', 0, param.bimarg, bicount, bimarg, J, seqs, es, param.alpha, None, None,
                outdir, log)
    Jsteps, newJ = NewtonSteps('preopt', param, bimarg, gpus, log)
    return newJ, Jsteps


################################################################################
# MCMC phase


def track_main_bufs(param, gpus, savedir=None, step=None):
    """Bicounts and energies of the walkers; optionally saved
    (NewtonSteps.py:367-383)."""
    gpus.calcBicounts('main')
    gpus.calcEnergies('main')
    bicounts, energies = gpus.collect(['bicount', 'E main'])
    bimarg_model = bicounts.astype('f4') / np.float32(np.sum(bicounts[0, :]))
    if savedir:
        tracked = param.tracked or []
        if 'bim' in tracked:
            np.save(savedir / f'bimarg_{step}', bimarg_model)
        if 'E' in tracked:
            np.save(savedir / f'energies_{step}', energies)
        if 'seq' in tracked:
            writeSeqs(savedir / f'seqs_{step}', gpus.collect('seq main'), param.alpha, zipf=True)
    return energies, bimarg_model


def runMCMC(gpus, couplings, runName, param, log):
    """Equilibrate the walkers under `couplings` and measure them
    (NewtonSteps.py:385-486).  equiltime 'auto': 8 calls, then doubling
    batches until the Spearman correlation of energies with the snapshots one
    and two batches back has p > 0.02 (and at least min_equil calls)."""
    nloop, trackequil, outdir = param.equiltime, param.trackequil, param.outdir
    gpus.setBuf('J', couplings)

    if nloop == 'auto':
        equil_dir = None
        if trackequil != 0:
            equil_dir = outdir / runName / 'equilibration'
            equil_dir.mkdir(parents=True, exist_ok=True)
        loops = 8
        for _ in range(loops):
            gpus.runMCMC()
        step = loops
        equil_e = []
        while True:
            for _ in range(loops):
                gpus.runMCMC()
            step += loops
            energies, _ = track_main_bufs(param, gpus, equil_dir, step)
            equil_e.append(energies)
            rstr = f"Step {step} <E>={np.mean(energies):.2f}. "
            if len(equil_e) >= 3:
                r1, p1 = spearmanr(equil_e[-1], equil_e[-2])
                r2, p2 = spearmanr(equil_e[-1], equil_e[-3])
                rstr += f"r=({r1:.3f}, {p1:.2g}) prev:({r2:.3f}, {p2:.2g}) "
                if p1 > 0.02 and p2 > 0.02 and step >= param.min_equil:
                    log(rstr + "Equilibrated.")
                    break
            if step >= param.max_equil:
                log(rstr + "Reached Max Steps. Stopping")
                break
            log(rstr + "Continuing.")
            loops *= 2
        e_rho = [spearmanr(ei, equil_e[-1]) for ei in equil_e]
    elif trackequil == 0:
        for _ in range(nloop):
            gpus.runMCMC()
        step = nloop
        e_rho = None
    else:
        equil_dir = outdir / runName / 'equilibration'
        equil_dir.mkdir(parents=True, exist_ok=True)
        equil_e = []
        for k in range(nloop // trackequil):
            for _ in range(trackequil):
                gpus.runMCMC()
            energies, _ = track_main_bufs(param, gpus, equil_dir, k * trackequil)
            equil_e.append(energies)
        step = nloop
        e_rho = [spearmanr(ei, equil_e[-1]) for ei in equil_e]

    if param.beta is None:
        gpus.calcBicounts('main')
        gpus.calcEnergies('main')
        bicount, es = gpus.collect(['bicount', 'E main'])
        bimarg_model = (bicount / np.sum(bicount[0, :])).astype('f4')
    else:
        log("Reweighting to account for modified temperature.")
        gpus.calcEnergies('main')
        es = gpus.collect('E main')
        gpus.fixed_beta_weights(np.min(es), 'main')
        gpus.weightedMarg('main')
        bimarg_model = gpus.collect('bi')
        bimarg_model = bimarg_model / np.sum(bimarg_model, axis=1, keepdims=True)
        bicount = None
    gpus.logProfile()
    return bimarg_model, bicount, es, e_rho, None, step


def MCMCstep(runName, Jstep, couplings, param, gpus, log):
    """One round: re-gauge J, run MCMC, report, adapt newtonSteps, Newton
    phase (NewtonSteps.py:610-687)."""
    outdir = param.outdir
    L, q = param.L, param.q
    bimarg_target = param.bimarg
    log("")
    log(f"Gradient Descent step {runName}")
    log("---------------------------")
    log(f"Total J update step {Jstep}")
    log("(Re-zeroing gauge of couplings)")
    tstart = time.time()
    couplings = fieldlessGaugeEven(np.zeros((L, q)), couplings)[1].astype('<f4')

    rundir = outdir / runName
    rundir.mkdir(parents=True, exist_ok=True)
    np.save(rundir / 'J', couplings)
    (rundir / 'newtonsteps').write_text(str(param.newtonSteps))
    (rundir / 'jstep').write_text(str(Jstep))

    if param.tempering is not None:
        raise Exception("parallel tempering is not part of this build")
    t0 = time.time()
    log("Running MCMC chains...")
    bimarg_model, bicount, energies, e_rho, ptinfo, equilsteps = runMCMC(
        gpus, couplings, runName, param, log)
    dt = time.time() - t0
    MC_s = equilsteps * param.nsteps * np.float64(gpus.nwalkers) / dt
    log(f"Total MCMC running time: {dt:.1f} s    ({MC_s:.3g} MC/s)")
    param.last_mcmc_time = dt
    param.last_mc_per_s = MC_s

    seqs = gpus.collect('seq main')
    tw = time.time()
    writeStatus(runName, Jstep, bimarg_target, bicount, bimarg_model, couplings, seqs, energies,
                param.alpha, e_rho, ptinfo, outdir, log,
                write_seqs=not getattr(param, 'skip_seq_write', False))
    t_write = time.time() - tw

    delta = param.newton_delta
    ssr = np.sum((bimarg_target - bimarg_model) ** 2)
    if param.last_ssr is not None and ssr > (param.last_ssr + param.min_ssr) / 2:
        param.newtonSteps = max(delta, param.newtonSteps - int(1.5 * delta))
        log(f"SSR increased over min. Decreasing newtonsteps to {param.newtonSteps}")
    param.last_ssr = ssr
    param.min_ssr = min(ssr, param.min_ssr)

    t1 = time.time()
    Jsteps, newJ = NewtonSteps(runName, param, bimarg_model, gpus, log)
    param.last_newton_time = time.time() - t1
    param.newtonSteps = min(2048, Jsteps + delta)
    log(f"Increasing newtonsteps to {param.newtonSteps}")
    (outdir / runName / 'nsteps').write_text(str(Jsteps))
    param.last_round_ssr = ssr
    # per-round timing record (wall s; not part of the reference's outputs)
    if 'round_stats' in param:
        wall = time.time() - tstart
        ferr, _, _ = bimarg_stats(bimarg_target, bimarg_model)
        X = np.sum(couplings * (bimarg_model - indepF(bimarg_model)))
        Xo = np.sum(couplings * (bimarg_target - indepF(bimarg_target)))
        param.round_stats.append({'run': runName, 'wall_s': wall,
                                  'mcmc_s': dt, 'status_write_s': t_write,
                                  'newton_s': param.last_newton_time, 'newton_steps': Jsteps,
                                  'jstep': int(Jstep), 'equil_launches': int(equilsteps),
                                  'ssr': float(ssr), 'ferr': float(ferr), 'X': float(X),
                                  'Xo': float(Xo)})
        if round_hook is not None:
            round_hook(param.round_stats[-1])
    return Jstep + Jsteps, seqs, energies, newJ


# Optional callable given each round's timing record as it completes (a
# caller that runs many rounds, such as a benchmark, can report progress).
round_hook = None


def newtonMCMC(param, gpus, start_run, Jstep, log):
    """All rounds from start_run to param.mcmcsteps (NewtonSteps.py:689-766)."""
    J = param.couplings
    gpus.setBuf('bi target', param.bimarg)
    if param.tempering is not None:
        raise Exception("parallel tempering is not part of this build")
    if param.reg == 'Xij':
        gpus.setBuf('Creg', param.regarg)

    Jsteps = 0
    if param.preopt:
        J, Jsteps = preOpt(param, gpus, log)
    else:
        log("No Pre-optimization")

    seqs = param.seqs
    seedseq = param.seedseq
    es = None
    if seqs is not None and param.reseed == 'single_best':
        gpus.calcEnergies('main')
        es = gpus.collect('E main')

    param.max_newtonSteps = param.newtonSteps
    param.min_ssr = np.inf
    param.last_ssr = None
    Jstep += Jsteps
    width = max(1, int(np.ceil(np.log10(max(param.mcmcsteps, 2)))))
    for i in range(start_run, param.mcmcsteps):
        runname = f'run_{i:0{width}d}'
        seed = None
        if seedseq is not None:
            seed, seedseq = seedseq, None
        elif param.reseed == 'single_indep':
            seed = mi3gpu.Mi3.generateSequences('independent', param.L, param.q, 1, log,
                                                param.unimarg)[0]
        elif param.reseed == 'single_random':
            seed = seqs[np.random.randint(0, seqs.shape[0])]
        elif param.reseed == 'single_best':
            seed = seqs[np.argmin(es)]

        rundir = param.outdir / runname
        rundir.mkdir(parents=True, exist_ok=True)
        if seed is not None:
            (rundir / 'seedseq').write_text("".join(param.alpha[c] for c in seed))
            gpus.fillSeqs(seed)
        elif param.reseed == 'independent':
            gpus.gen_indep('main')
        elif param.reseed == 'msa':
            gpus.setSeqs('main', param.seedmsa)

        Jstep, seqs, es, J = MCMCstep(runname, Jstep, J, param, gpus, log)
    flush_writes()
    return J
