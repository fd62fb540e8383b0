#!/usr/bin/env python3
# This file keeps the interface and behaviour of the file of the same name
# in Mi3-GPU, Copyright 2020 Allan Haldane, and is a derived work of it.
# Mi3-GPU is free software: you can redistribute it and/or modify it under
# the terms of the GNU General Public License as published by the Free
# Software Foundation, version 3 of the License; this file is distributed
# under the same licence, WITHOUT ANY WARRANTY; without even the implied
# warranty of MERCHANTABILITY or FITNESS FOR A PARTICULAR PURPOSE.  See the
# GNU General Public License <http://www.gnu.org/licenses/> for details.
"""Mi3 command line: `Mi3.py {infer,gen,energies,benchmark,subseq} [options]`
(Mi3.py of the reference).

Same modes, option names, defaults, log lines and output directory layout.
Option parsing uses argparse with the reference's `--config` file support
(configargparse's "key = value" format, read and written) built in, because
configargparse is not available.  MPI (`--mpi`) is not part of this build;
`benchmark` works (the reference's crashes at Mi3.py:667).
"""
import atexit
import datetime
import os
import signal
import socket
import sys
import time
from pathlib import Path
from shlex import quote as cmd_quote

import numpy as np
from numpy.random import randint, rand

import mi3gpu
import mi3gpu.NewtonSteps
from mi3gpu.utils.seqload import loadSeqs, writeSeqs
from mi3gpu.utils.changeGauge import fieldlessGaugeEven
from mi3gpu.utils.potts_common import printsome, getLq, getUnimarg, validate_bimarg
from mi3gpu.mcmcGPU import setup_GPU_context, initGPU, wgsize_heuristic, printGPUs
from mi3gpu.node_manager import GPU_node

progname = 'Mi3.py'
version = mi3gpu.__version__


class attrdict(dict):
    def __getattr__(self, attr):
        if attr.startswith('_'):
            return super().__getattr__(attr)
        return self.get(attr, None)

    def __setattr__(self, k, v):
        self[k] = v


################################################################################
# options


def _flag(v):
    return str(v).lower() in ('1', 'true', 'yes', 'on')


def optionRegistry():
    """Option names, types and defaults of the reference (Mi3.py:94-189)."""
    o = {}

    def add(name, **k):
        o[name] = k
    add('init_model', default='independent',
        help="One of 'zero', 'independent', or a directory name.")
    add('outdir', type=Path, default='output', help='Output Directory')
    add('finish', type=Path, help='Dir. of an unfinished run to finish')
    add('config', help='config file to load arguments from')
    add('rngseed', type=np.uint32, help='random seed')
    add('nwalkers', type=np.uint32, help="Number of MC walkers")
    add('nsteps', type=np.uint32, default=2048, help="number of mc steps per kernel call")
    add('wgsize', default=512, help="GPU workgroup size")
    add('gpus', help="GPUs to use (comma-sep list of platforms #s, eg '0,0')")
    add('profile', action='store_true', help="enable kernel profiling")
    add('nlargebuf', type=np.uint32, default=1,
        help='size of large seq buffer, in multiples of nwalkers')
    add('measurefperror', action='store_true', help="enable fp error calculation")
    add('beta', type=np.float32, help="beta at which to generate sequences")
    add('bimarg', help="Target bivariate marginals (npy file)")
    add('mcsteps', type=np.uint32, default=64, help="Number of rounds of MCMC generation")
    add('newtonsteps', default=1024, type=np.uint32,
        help="Initial number of newton steps per round.")
    add('newton_delta', default=32, type=np.uint32, help="Newton step number tuning scale")
    add('fracNeff', type=np.float32, default=0.9,
        help="stop coupling updates after Neff/N = fracNeff")
    add('gamma', type=np.float32, default=0.0004, help="Initial step size")
    add('damping', default=0.001, type=float, help="Damping parameter")
    add('reg', default=None, help="regularization format")
    add('preopt', action='store_true',
        help="Perform a round of newton steps before first MCMC run")
    add('reseed', choices=['none', 'single_best', 'single_random', 'single_indep',
                           'independent', 'uniform', 'msa'],
        default='single_indep', help="Strategy to reset walkers after each MCMC round")
    add('seedmsa', default=None, help="seed used of reseed=msa")
    add('distribute_jstep', choices=['head_gpu', 'head_node', 'all'], default='all',
        help="how to split newton step computation across GPUs")
    add('alpha', help="Alphabet, a sequence of letters")
    add('couplings', help="One of 'zero', 'independent', or a filename")
    add('L', help="sequence length", type=int)
    add('seedseq', help="Starting sequence. May be 'rand'")
    add('seqs', help="File containing sequences to pre-load to GPU")
    add('seqs_large', help="File containing sequences to pre-load to GPU")
    add('indep_marg', help="marg used to generate site-independent sequences")
    add('equiltime', default='auto', help="Number of MC kernel calls to equilibrate")
    add('min_equil', default=64, type=int,
        help="minimum MC calls to equilibrate when using 'equiltime=auto'")
    add('max_equil', default=32768, type=int,
        help="maximum MC calls to equilibrate when using 'equiltime=auto'")
    add('trackequil', type=np.uint32, default=0,
        help='Save "--tracked" data every TRACKEQUIL steps during mcmc')
    add('tracked', default='bim,E', help='Data saved when trackequil is enabled')
    add('tempering', help='optional inverse Temperature schedule')
    add('nswaps_temp', type=np.uint32, default=128, help='optional number of pt swaps')
    return o


_OPTIONS = optionRegistry()


class Parser:
    """argparse with configargparse-style `--config FILE` and
    write_config_file (the subset the reference uses)."""

    def __init__(self, prog, description):
        import argparse
        self.p = argparse.ArgumentParser(prog=prog, description=description)
        self.names = []

    def addopt(self, groupname, optstring):
        g = self.p.add_argument_group(groupname) if groupname else self.p
        for name in optstring.split():
            g.add_argument('--' + name, **_OPTIONS[name])
            self.names.append(name)

    def add_argument(self, *a, **k):
        return self.p.add_argument(*a, **k)

    @staticmethod
    def read_config(fn):
        out = []
        for line in Path(fn).read_text().splitlines():
            line = line.strip()
            if not line or line.startswith(('#', ';', '[')):
                continue
            sep = '=' if '=' in line else ':'
            k, _, v = line.partition(sep)
            out.append((k.strip(), v.strip()))
        return out

    def parse_args(self, args):
        args = list(args)
        if '--config' in args and 'config' in self.names:
            fn = args[args.index('--config') + 1]
            given = {a[2:].split('=')[0] for a in args if a.startswith('--')}
            extra = []
            for k, v in self.read_config(fn):
                if k in given or k not in self.names:
                    continue
                if _OPTIONS[k].get('action') == 'store_true':
                    if _flag(v):
                        extra.append('--' + k)
                else:
                    extra.extend(['--' + k, v])
            args = extra + args
        return self.p.parse_args(args)

    def write_config_file(self, ns, files):
        lines = []
        for k in self.names:
            v = getattr(ns, k, None)
            if k in ('config', 'finish') or v is None or v is False:
                continue
            lines.append(f"{k} = {'true' if v is True else v}")
        for fn in files:
            Path(fn).write_text("\n".join(lines) + "\n")


def requireargs(args, required):
    for r in required.split():
        if getattr(args, r) is None:
            raise Exception(f"error: argument --{r} is required")


def setup_seed(args, p, log):
    seed = args.rngseed if args.rngseed is not None else \
        np.frombuffer(os.urandom(4), dtype='u4')[0]
    np.random.seed(seed)
    p['rngseed'] = int(seed) + 1
    log(f"Using random seed {p.rngseed}")
    log("")


def print_node_startup(log, orig_args):
    log(f"Hostname:   {socket.gethostname()}")
    log(f"Start Time: {datetime.datetime.now()}")
    log("")
    log("Command line arguments:")
    log(" ".join(cmd_quote(a) for a in orig_args))
    log("")
    log("Mi3-GPU (MI355X engine) Version:", version)
    log("")


def setup_exit_hook(log):
    atexit.register(lambda: log(f"Exited at {datetime.datetime.now()}"))
    signal.signal(signal.SIGTERM, lambda signum, frame: sys.exit(1))


def divideWalkers(nwalkers, ngpus, log, wgsize=None):
    n_max = (nwalkers - 1) // ngpus + 1
    out = [n_max] * (ngpus - 1) + [nwalkers - (ngpus - 1) * n_max]
    if wgsize is not None and nwalkers % (ngpus * wgsize) != 0:
        log("Warning: number of MCMC walkers is not a multiple of wgsize*ngpus, so there are "
            "idle work units.")
    return out


def setup_GPUs(p, log, splitwalkers=True):
    """One GPU_node over the selected devices (Mi3.py:322-347)."""
    clinfo, devs, info = setup_GPU_context(None, None, p, log)
    (p.outdir / 'device_info').write_text(info[0] + "\n" + info[1])
    ngpus = len(devs)
    walkers = divideWalkers(p.nwalkers, ngpus, log, p.wgsize) if splitwalkers else \
        [p.nwalkers] * ngpus
    log(f"Found {ngpus} GPUs")
    log("GPU Initialization:")
    gpus = GPU_node([initGPU(n, clinfo, d, w, p, log)
                     for n, (d, w) in enumerate(zip(devs, walkers))])
    log('Running on GPUs:\n' + "\n".join(f'    {n}   ({w} walkers)'
                                        for n, w in zip(gpus.gpu_list, walkers)))
    return gpus


def _open_log(path):
    f = open(path, 'wt')
    return f, (lambda *s, **k: print(*s, file=f, flush=True, **k))


################################################################################
# argument processing


def process_GPU_args(args, L, q, outdir, log):
    log("GPU setup")
    log("---------")
    p = attrdict({'nsteps': args.nsteps, 'wgsize': args.wgsize, 'nwalkers': args.nwalkers,
                  'beta': getattr(args, 'beta', None), 'gpuspec': args.gpus,
                  'profile': args.profile, 'fperror': args.measurefperror,
                  'L': L, 'q': q, 'outdir': outdir})
    p['wgsize'] = wgsize_heuristic(q, p.wgsize)
    log(f"Total GPU walkers: {p.nwalkers}")
    log(f"Work Group Size: {p.wgsize}")
    log(f"{p.nsteps} MC steps per MCMC kernel call")
    if p.profile:
        log("Profiling Enabled")
    return p


REG_TYPES = ['l2z', 'l1z', 'SCADJ', 'X', 'Xij', 'SCADX', 'expX', 'ddE', 'SCADddE']


def parse_reg(spec, bimarg, log):
    """--reg TYPE:ARGS -> {'reg': TYPE, 'regarg': tuple or array}, with the
    argument formats and defaults of the reference (Mi3.py:1038-1113)."""
    rtype, _, rarg = spec.partition(':')
    if rtype not in REG_TYPES:
        raise Exception(f"reg must be one of {str(REG_TYPES)}")
    if rtype == 'ddE':
        lam = float(rarg)
        log(f"Regularizing using {rtype} with lambda = {lam}")
        return {'reg': rtype, 'regarg': (lam,)}
    if rtype == 'SCADddE':
        lam, _, r = rarg.partition(':')
        lam = float(lam)
        r = lam if r == '' else float(r)
        log(f"Regularizing using {rtype} with lambda = {lam}, r = {r}")
        return {'reg': rtype, 'regarg': (lam, r)}
    if rtype in ('l2z', 'l1z'):
        try:
            lJ = float(rarg)
        except ValueError:
            raise Exception(f"{rtype} specifier must be of form '{rtype}:lJ', eg "
                            f"'{rtype}:0.01'. Got '{spec}'")
        log(f"Regularizing using {rtype} norm with lambda_J = {lJ}")
        return {'reg': rtype, 'regarg': (lJ,)}
    if rtype == 'SCADJ':
        try:
            r, _, a = rarg.partition(':')
            r = float(r)
            a = float(a) if a != '' else 4.0
        except ValueError:
            raise Exception(f"{rtype} specifier must be of form '{rtype}:r:a', eg "
                            f"'{rtype}:10:0.1'. Got '{spec}'")
        if a < 2.0:
            raise Exception("SCADJ a parameter must be >= 2.0")
        log(f"Regularizing using SCADJ with r={r} a={a}")
        return {'reg': rtype, 'regarg': (r, a)}
    if rtype == 'X':
        try:
            lX = float(rarg)
        except ValueError:
            raise Exception(f"{rtype} specifier must be of form 'X:lX', eg 'X:0.01'. "
                            f"Got '{spec}'")
        log(f"Regularizing X with lambda_X = {lX}")
        return {'reg': rtype, 'regarg': (lX,)}
    if rtype == 'Xij':
        log(f"Regularizing with Xij from file {rarg}")
        lam = np.load(rarg)
        if lam.shape != bimarg.shape:
            raise Exception("Xij in wrong format")
        return {'reg': rtype, 'regarg': lam.astype('<f4')}
    if rtype == 'SCADX':
        try:
            d, _, r = rarg.partition(':')
            r, _, a = r.partition(':')
            d, r = float(d), float(r)
            a = float(a) if a != '' else 4.0
        except ValueError:
            raise Exception(f"{rtype} specifier must be of form '{rtype}:d:r:a' or "
                            f"{rtype}:d:r', eg '{rtype}:10:0.1'. Got '{spec}'")
        if a < 2.0:
            raise Exception("SCADX a parameter must be >= 2.0")
        s = 2 * d / ((1 + a) * r * r)   # cost d for large |X|
        log(f"Regularizing using SCADX with d={d} r={r} a={a}")
        return {'reg': rtype, 'regarg': (s, r, a)}
    # expX
    try:
        lam = float(rarg)
    except ValueError:
        raise Exception(f"{rtype} specifier must be of form 'expX:l', eg 'expX:0.001'. "
                        f"Got '{spec}'")
    log(f"Regularizing using expX with  lam = {lam}")
    return {'reg': rtype, 'regarg': (lam,)}


def process_newton_args(args, log):
    """Newton options and the target bimarginals (Mi3.py:1004-1116)."""
    log("Newton Solver Setup")
    log("-------------------")
    p = attrdict({'mcmcsteps': args.mcsteps, 'newtonSteps': args.newtonsteps,
                  'newton_delta': args.newton_delta, 'fracNeff': args.fracNeff,
                  'gamma0': args.gamma, 'pcdamping': args.damping, 'reseed': args.reseed,
                  'preopt': args.preopt, 'distribute_jstep': args.distribute_jstep})
    log(f"Running {p.mcmcsteps} Newton-MCMC rounds")
    log(f"Updating J locally with gamma={p.gamma0}, and pc-damping {p.pcdamping}")
    log(f"Running {p.newtonSteps} Newton update steps per round.")
    log(f"Using {p.distribute_jstep}-GPU mode for Newton-step calculations.")
    log(f"Reading target marginals from file {args.bimarg}")
    bimarg = np.load(args.bimarg)
    if bimarg.dtype != np.dtype('<f4'):
        raise Exception("Bimarg must be in 'f4' format")
    if np.any((bimarg <= 0) | (bimarg > 1)):
        raise Exception("All bimarg must be 0 < f < 1")
    validate_bimarg(bimarg)
    log("Target Marginals: " + printsome(bimarg) + "...")
    p['bimarg'] = bimarg
    if args.reg is not None:
        p.update(parse_reg(args.reg, bimarg, log))
    log("")
    return p


def updateLq(L, q, newL, newq, name):
    if newL is not None:
        if L is not None and L != newL:
            raise Exception(f"L from {name} ({newL}) inconsitent with previous value ({L})")
        L = newL
    if newq is not None:
        if q is not None and q != newq:
            raise Exception(f"q from {name} ({newq}) inconsitent with previous value ({q})")
        q = newq
    return L, q


def _load_f4(fn, what):
    arr = np.load(fn)
    if arr.dtype != np.dtype('<f4'):
        raise Exception(f"{what} must be in 'f4' format")
    return arr


GENERATED = ('uniform', 'independent')


def getCouplings(args, L, q, unimarg, log):
    """Initial couplings (Mi3.py:1158-1203).  --couplings wins over
    --init_model; either may be 'uniform' (all zero) or 'independent' (the
    fieldless form of fields -log f_i); otherwise --couplings is a .npy file
    and --init_model a run directory holding J.npy."""
    from_dir = not args.couplings
    source = args.init_model if from_dir else args.couplings
    if source is None:
        raise Exception("didn't get couplings or init_model")
    if source in GENERATED:
        args.couplings = source
        if L is None:
            raise Exception("Need L to generate couplings")
        zero = np.zeros((L * (L - 1) // 2, q * q), dtype='<f4')
        if source == 'uniform':
            log("Setting Initial couplings to uniform frequencies")
            couplings = zero
        else:
            log("Setting Initial couplings to independent model")
            if unimarg is None:
                raise Exception("Need univariate marginals to generate independent model "
                                "couplings")
            couplings = fieldlessGaugeEven(-np.log(unimarg), zero)[1].astype('<f4')
    else:
        fn = Path(source, 'J.npy') if from_dir else source
        if from_dir and not fn.is_file():
            raise Exception(f"could not find file {fn}")
        log(f"Reading couplings from file {fn}")
        couplings = _load_f4(fn, "Couplings")
    L2, q2 = getLq(couplings)
    L, q = updateLq(L, q, L2, q2, 'couplings')
    return couplings, L, q


def process_potts_args(args, L, q, unimarg, log):
    log("Potts Model Setup")
    log("-----------------")
    alpha = args.alpha.strip()
    L, q = updateLq(getattr(args, 'L', None), len(alpha), L, q, 'bimarg')
    couplings, L, q = getCouplings(args, L, q, unimarg, log)
    log(f"alphabet: {alpha}")
    log(f"q {q}  L {L}")
    log("Couplings: " + printsome(couplings) + "...")
    log("")
    return attrdict({'L': L, 'q': q, 'alpha': alpha, 'couplings': couplings})


def repeatseqs(seqs, n):
    return np.repeat(seqs, (n - 1) // seqs.shape[0] + 1, axis=0)[:n, :]


def generateSequences(gentype, L, q, nseqs, log, unimarg=None):
    """Host-side random or independent-model sequences (Mi3.py:1272-1284)."""
    if gentype in ('zero', 'uniform'):
        log(f"Generating {nseqs} random sequences...")
        return randint(0, q, size=(nseqs, L)).astype('<u1')
    if gentype == 'independent':
        log(f"Generating {nseqs} independent-model sequences...")
        if unimarg is None:
            raise Exception("marg must be provided to generate sequences")
        cp = np.cumsum(unimarg, axis=1)
        cp = cp / cp[:, -1:]
        return np.array([np.searchsorted(c, rand(nseqs)) for c in cp], dtype='<u1').T
    raise Exception(f"Unknown sequence generation mode '{gentype}'")


def loadseedseq(fn, alpha, log):
    log(f"Reading seedseq from file {fn}")
    with open(fn) as f:
        s = f.readline().strip()
    return np.array([alpha.index(c) for c in s], dtype='<u1')


def loadSequenceFile(sfile, alpha, log):
    log(f"Loading sequences from file {sfile}")
    seqs = loadSeqs(sfile, alpha=alpha)[0].astype('<u1')
    log(f"Found {seqs.shape[0]} sequences")
    return seqs


def loadSequenceDir(sdir, bufname, alpha, log):
    log(f"Loading {bufname} sequences from dir {sdir}")
    return loadSequenceFile(Path(sdir) / 'seqs', alpha, log)


def _first_source(candidates):
    """The first (kind, value) whose value is given and is a generator name
    exactly when kind is 'gen'."""
    for kind, v in candidates:
        if v is not None and (kind == 'gen') == (v in GENERATED):
            return kind, v
    return None, None


def _fit_rows(seqs, n, log):
    """Repeat or truncate `seqs` to n rows."""
    have = seqs.shape[0]
    if n > have:
        log(f"Repeating {have} sequences to make {n}")
        return repeatseqs(seqs, n)
    if n < have:
        log(f"Truncating {have} sequences to make {n}")
        return seqs[:n]
    return seqs


def process_sequence_args(args, L, alpha, log, unimarg=None, nseqs=None, needseed=False):
    """Initial walker sequences and/or a seed sequence (Mi3.py:1205-1270).
    Walkers: a generator named by --seqs, then one named by --init_model,
    then a --seqs file, then the `seqs` of an --init_model directory.  Seed:
    a generator or letters/file given by --seedseq, then an --init_model
    generator or the `seedseq` file of its directory."""
    if not nseqs and not needseed:
        return {}
    log("Sequence Setup")
    log("--------------")
    q = len(alpha)
    seedseq, seqs = None, None
    if nseqs is not None:
        kind, v = _first_source([('gen', args.seqs), ('gen', args.init_model),
                                 ('file', args.seqs), ('dir', args.init_model)])
        if kind == 'gen':
            seqs = generateSequences(v, L, q, nseqs, log, unimarg)
        elif kind == 'file':
            seqs = loadSequenceFile(v, alpha, log)
        elif kind == 'dir':
            seqs = loadSequenceDir(v, '', alpha, log)
        if seqs is None:
            raise Exception(f"Did not find requested {nseqs} sequences")
        seqs = _fit_rows(seqs, nseqs, log)
    if needseed:
        kind, v = _first_source([('gen', args.seedseq), ('given', args.seedseq),
                                 ('gen', args.init_model), ('dir', args.init_model)])
        origin = v if kind == 'gen' else 'from file'
        if kind == 'gen':
            seedseq = generateSequences(v, L, q, 1, log, unimarg)[0]
        elif kind == 'given':
            if all(c in alpha for c in v):
                seedseq = np.array([alpha.index(c) for c in v], dtype='<u1')
                origin = 'supplied'
            else:
                seedseq = loadseedseq(v, alpha, log)
        elif kind == 'dir':
            seedseq = loadseedseq(Path(v, 'seedseq'), alpha, log)
        if seedseq is not None:
            log(f"Seed seq ({origin}): {''.join(alpha[x] for x in seedseq)}")
    log("")
    return attrdict({'seedseq': seedseq, 'seqs': seqs})


def process_sample_args(args, log):
    p = attrdict({'equiltime': args.equiltime, 'min_equil': args.min_equil,
                  'max_equil': args.max_equil, 'trackequil': args.trackequil,
                  'tracked': args.tracked.split(',')})
    bad = [x for x in p.tracked if x not in ['bim', 'E', 'seq']]
    if bad:
        raise ValueError(f'Invalid "--tracked" values: {bad}')
    if p.equiltime != 'auto':
        p['equiltime'] = int(p.equiltime)
    if getattr(args, 'tempering', None):
        raise Exception("parallel tempering is not part of this build")
    log("MCMC Sampling Setup")
    log("-------------------")
    if p.equiltime == 'auto':
        log('Using "auto" equilibration')
    else:
        log(f"In each MCMC round, running {p.equiltime} GPU MCMC kernel calls")
    if p.equiltime != 'auto' and p.trackequil != 0:
        if p.equiltime % p.trackequil != 0:
            raise Exception("Error: trackequil must be a divisor of equiltime")
        log(f"Tracking equilibration every {p.trackequil} loops.")
    log("")
    return p


################################################################################
# modes


def inverseIsing(orig_args, infer_args, log):
    """`infer`: fit couplings to target bimarginals (Mi3.py:351-547)."""
    parser = Parser(progname + ' inverseIsing',
                    'Inverse Ising inference using a quasi-Newton MCMC algorithm on the GPU')
    parser.addopt('GPU options', 'nwalkers nsteps wgsize gpus profile beta measurefperror')
    parser.addopt('Sequence Options', 'seedseq seqs seqs_large')
    parser.addopt('Newton Step Options', 'bimarg mcsteps newtonsteps newton_delta fracNeff '
                                         'damping reg distribute_jstep gamma preopt reseed '
                                         'seedmsa')
    parser.addopt('Sampling Options', 'equiltime min_equil max_equil trackequil tracked '
                                      'tempering nswaps_temp')
    parser.addopt('Potts Model Options', 'alpha couplings L')
    parser.addopt(None, 'init_model outdir rngseed config finish')
    args = parser.parse_args(infer_args)

    if args.finish:
        outdir = args.finish
        if not outdir.is_dir():
            raise ValueError(f"{outdir} is not a directory")
        runs = sorted(outdir.glob('run_*'))
        if not runs:
            raise Exception(f"Did not find any runs in {outdir}")
        rundir = runs[-1]
        startrun = int(str(rundir).rpartition('_')[2])
        newtonsteps = int((rundir / 'newtonsteps').read_text())
        jstep = int((rundir / 'jstep').read_text())
        logfile, log = _open_log(outdir / f'log_finish_{startrun}')
        fa = list(infer_args) + ['--config', str(outdir / 'config.cfg'),
                                 '--couplings', str(rundir / 'J.npy'),
                                 '--outdir', str(outdir), '--newtonsteps', str(newtonsteps)]
        if (rundir / 'seqs').exists():
            fa += ['--seqs', str(rundir / 'seqs')]
        args = parser.parse_args(fa)
    else:
        args.outdir.mkdir(parents=True)
        logfile, log = _open_log(args.outdir / 'log')
        startrun, jstep = 0, 0
        parser.write_config_file(args, [str(args.outdir / 'config.cfg')])

    requireargs(args, 'bimarg alpha')
    print_node_startup(log, orig_args)
    if args.finish:
        log(f"Continuing from {rundir}")
        log("")
    log("Initialization")
    log("===============")
    p = attrdict({'outdir': args.outdir, 'round_stats': []})
    setup_seed(args, p, log)
    p.update(process_newton_args(args, log))
    p['L'], p['q'] = getLq(p.bimarg)
    unimarg = getUnimarg(p.bimarg)
    p['unimarg'] = unimarg
    p.update(process_potts_args(args, p.L, p.q, unimarg, log))
    p.update(process_sample_args(args, log))
    p.update(process_GPU_args(args, p.L, p.q, p.outdir, log))
    gpus = setup_GPUs(p, log)
    gpus.initMCMC(p.nsteps)
    gpus.initJstep()
    if p.distribute_jstep == 'head_node':
        raise Exception('"head_node" option only makes sense when using MPI')
    log("")

    gen_indep = args.seqs == 'independent' or args.init_model == 'independent'
    if gen_indep or args.reseed == 'independent':
        gpus.prepare_indep(unimarg)
    needed = None
    use_seed = p.reseed in ['single_best', 'single_random']
    if (p.preopt or p.reseed == 'none') and not gen_indep:
        needed = gpus.nseq['main']
    p.update(process_sequence_args(args, p.L, p.alpha, log, unimarg, nseqs=needed,
                                   needseed=use_seed))
    if p.reseed == 'msa':
        s = repeatseqs(loadSequenceFile(args.seedmsa, p.alpha, log), gpus.nseq['main'])
        p['seedmsa'] = np.split(s, gpus.ngpus)
    if p.preopt:
        if gen_indep:
            log("")
            log("Generating Indep seqs on GPUs")
            gpus.gen_indep('main')
        elif p.seqs is not None:
            log("")
            log("Initializing main seq buf with loaded seqs.")
            gpus.setSeqs('main', p.seqs, log)
        else:
            raise Exception("Need to specify initial seqs for preopt")
    elif p.reseed == 'none':
        if p.seqs is None:
            raise Exception("Need to provide seqs if not using seedseq")
        gpus.setSeqs('main', p.seqs, log)
    elif use_seed and p.seedseq is None:
        raise Exception("Must provide seedseq if using reseed=single_*")
    log("")

    log("Computation Overview")
    log("====================")
    log(f"Running {p.mcmcsteps} Newton-MCMC rounds")
    if p.equiltime == 'auto':
        log(f"In each round, running {p.nwalkers} MC walkers until equilibrated, with a "
            f"minimum of {p.min_equil} equilibration loops and maximum {p.max_equil}")
    else:
        log(f"In each round, running {p.nwalkers} MC walkers for {p.equiltime} equilibration "
            f"loops with {p.nsteps} MC steps per loop (Each walker equilibrated a total of "
            f"{p.nsteps * p.equiltime} MC steps, or {p.nsteps * p.equiltime / p.L:.1f} steps "
            f"per position).")
    f = p.bimarg
    expect_SSR = np.sum(f * (1 - f)) / p.nwalkers
    absexp = np.sqrt(2 / np.pi) * np.sqrt(f * (1 - f) / p.nwalkers) / f
    log("\nEstimated lowest achievable statistical error for this nwalkers and bimarg is:\n"
        f"MIN:    SSR = {expect_SSR:.4f}   rel% = {np.mean(absexp[f > 0.01]):.3f}")
    log("(Statistical error only. Modeling biases and perturbation procedure may cause "
        "additional error)")
    log("")
    log("")
    log("MCMC Run")
    log("========")
    mi3gpu.NewtonSteps.newtonMCMC(p, gpus, startrun, jstep, log)
    logfile.close()
    return p


def getEnergies(orig_args, args, log):
    """`energies`: Potts energies of a sequence file (Mi3.py:549-608)."""
    parser = Parser(progname + ' getEnergies', 'Compute Potts Energy of a set of sequences')
    parser.add_argument('out', default='output', help='Output File')
    parser.addopt('GPU Options', 'wgsize gpus profile')
    parser.addopt('Potts Model Options', 'alpha couplings')
    parser.addopt('Sequence Options', 'seqs')
    parser.addopt(None, 'outdir')
    args = parser.parse_args(args)
    requireargs(args, 'couplings alpha seqs')
    args.outdir.mkdir(parents=True)
    logfile, log = _open_log(args.outdir / 'log')
    print_node_startup(log, orig_args)
    args.init_model = None
    args.L = None
    p = attrdict({'outdir': args.outdir})
    p.update(process_potts_args(args, None, None, None, log))
    seqs = loadSequenceFile(args.seqs, p.alpha, log)
    n = seqs.shape[0]
    npad = ((n + 511) // 512) * 512
    seqs_p = np.zeros((npad, p.L), dtype='<u1')
    seqs_p[:n] = seqs
    args.nwalkers, args.nsteps, args.measurefperror = npad, 1, False
    p.update(process_GPU_args(args, p.L, p.q, p.outdir, log))
    p['rngseed'] = 0
    gpus = setup_GPUs(p, log)
    gpus.setSeqs('main', seqs_p, log)
    gpus.setBuf('J', p.couplings)
    gpus.calcEnergies('main')
    es = gpus.collect('E main')[:n]
    log(f"Saving results to file '{args.out}'")
    np.save(args.out, es)
    logfile.close()
    return es


def MCMCbenchmark(orig_args, args, log):
    """`benchmark`: time nloop MCMC kernel calls (Mi3.py:610-722)."""
    parser = Parser(progname + ' benchmark', 'Benchmark MCMC generation on the GPU')
    parser.add_argument('--nloop', type=np.uint32, required=True,
                        help="Number of kernel calls to benchmark")
    parser.addopt('GPU options', 'nwalkers nsteps wgsize gpus profile measurefperror')
    parser.addopt('Sequence Options', 'seedseq seqs')
    parser.addopt('Potts Model Options', 'alpha couplings L')
    parser.addopt(None, 'init_model outdir rngseed')
    args = parser.parse_args(args)
    if args.couplings is None:
        raise ValueError("--couplings is required")
    if args.nwalkers is None:
        raise ValueError("--nwalkers is required")
    args.outdir.mkdir(parents=True)
    logfile, log = _open_log(args.outdir / 'log')
    print_node_startup(log, orig_args)
    p = attrdict({'outdir': args.outdir})
    setup_seed(args, p, log)
    p.update(process_potts_args(args, args.L, None, None, log))
    p.update(process_GPU_args(args, p.L, p.q, p.outdir, log))
    gpus = setup_GPUs(p, log)
    gpus.initMCMC(p.nsteps)
    if args.seqs is not None:
        p.update(process_sequence_args(args, p.L, p.alpha, log, nseqs=gpus.nseq['main']))
        gpus.setSeqs('main', p.seqs, log)
    elif args.seedseq is not None:
        p.update(process_sequence_args(args, p.L, p.alpha, log, needseed=True))
        gpus.fillSeqs(p.seedseq)
    else:
        raise Exception("'seqs' or 'seedseq' option required")
    nloop = int(args.nloop)
    gpus.setBuf('J', p.couplings)

    def run():
        for _ in range(nloop):
            gpus.runMCMC()
        gpus.wait()
    log(f"Benchmarking MCMC for {nloop} loops, {p.nsteps} MC steps per loop")
    log("Warmup run...")
    run()
    log("Timed run...")
    t0 = time.perf_counter()
    run()
    dt = time.perf_counter() - t0
    tot = p.nwalkers * nloop * np.float64(p.nsteps)
    log("Elapsed time: ", dt)
    log(f"MC steps computed: {tot}")
    log(f"MC steps per second: {tot / dt:g}")
    logfile.close()
    return tot / dt


def equilibrate(orig_args, args, log):
    """`gen`: equilibrate walkers under given couplings and save them
    (Mi3.py:724-870)."""
    parser = Parser(progname + ' mcmc', 'Run a round of MCMC generation on the GPU')
    parser.addopt('GPU options', 'nwalkers nsteps wgsize gpus profile beta measurefperror')
    parser.addopt('Sequence Options', 'seedseq seqs indep_marg')
    parser.addopt('Sampling Options', 'equiltime min_equil max_equil trackequil tracked '
                                      'tempering nswaps_temp')
    parser.addopt('Potts Model Options', 'alpha couplings L')
    parser.addopt(None, 'init_model outdir rngseed')
    args = parser.parse_args(args)
    args.outdir.mkdir(parents=True)
    logfile, log = _open_log(args.outdir / 'log')
    print_node_startup(log, orig_args)
    p = attrdict({'outdir': args.outdir})
    setup_seed(args, p, log)
    p.update(process_potts_args(args, args.L, None, None, log))
    p.update(process_sample_args(args, log))
    p.update(process_GPU_args(args, p.L, p.q, p.outdir, log))
    gpus = setup_GPUs(p, log)
    gpus.initMCMC(p.nsteps)
    gen_indep = args.seqs == 'independent' or args.init_model == 'independent'
    if gen_indep:
        imarg = None
        if args.indep_marg is not None:
            imarg = np.load(args.indep_marg)
        elif args.init_model and Path(args.init_model, 'bimarg.npy').is_file():
            imarg = np.load(Path(args.init_model, 'bimarg.npy'))
        if imarg is None:
            raise ValueError("indep_marg must be supplied if generating independent-model "
                             "sequences")
        unimarg = imarg if imarg.shape == (p.L, p.q) else getUnimarg(imarg)
        gpus.prepare_indep(unimarg.astype('f4'))
    needseed = args.seedseq is not None
    nseqs = None if (needseed or gen_indep) else p.nwalkers
    p.update(process_sequence_args(args, p.L, p.alpha, log, nseqs=nseqs, needseed=needseed))
    if needseed:
        gpus.fillSeqs(p.seedseq)
    elif gen_indep:
        gpus.gen_indep('main')
    else:
        gpus.setSeqs('main', p.seqs, log)
    gpus.setBuf('J', p.couplings)
    log("Equilibrating")
    log("====================")
    bimarg, bicount, energies, e_rho, ptinfo, steps = mi3gpu.NewtonSteps.runMCMC(
        gpus, p.couplings, 'gen', p, log)
    seqs = gpus.collect('seq main')
    if bicount is not None:
        np.savetxt(p.outdir / 'bicounts', bicount, fmt='%d')
    np.save(p.outdir / 'bimarg', bimarg)
    np.save(p.outdir / 'energies', energies)
    writeSeqs(p.outdir / 'seqs', seqs, p.alpha)
    log("Mean energy:", np.mean(energies))
    log("Done!")
    logfile.close()
    return bimarg, energies


def _pad_rows(seqs, multiple):
    """`seqs` with zero rows appended up to a positive multiple of
    `multiple` rows (the walker count the device needs)."""
    n = seqs.shape[0]
    want = max(multiple, -(-n // multiple) * multiple)
    if want == n:
        return seqs
    return np.concatenate([seqs, np.zeros((want - n, seqs.shape[1]), dtype=seqs.dtype)])


def subseqFreq(orig_args, args, log):
    """`subseq`: log relative frequency of subsequences at fixed positions
    against a background sample (Mi3.py:872-976).  For every subsequence
    sequence n, the fixed positions of every background sequence are
    replaced by those of n and log f_n = logsumexp(E_orig - E_substituted)
    (or, with --iterbackgrounds, the roles of the two sets swap)."""
    from scipy.special import logsumexp
    parser = Parser(progname + ' subseqFreq',
                    'Compute relative frequency of subsequences at fixed positions')
    parser.add_argument('fixpos', help="comma separated list of fixed positions")
    parser.add_argument('--iterbackgrounds', action='store_true',
                        help='use if backgrounds is small')
    parser.addopt('GPU options', 'nsteps wgsize gpus profile beta')
    parser.addopt('Potts Model Options', 'alpha couplings L')
    parser.addopt(None, 'outdir')
    parser.add_argument('backgroundseqs', help="large sample of equilibrated sequences")
    parser.add_argument('subseqs', help="sequences from which to compute subseq freqs")
    args = parser.parse_args(args)
    args.measurefperror = False
    args.outdir.mkdir(parents=True)
    logfile, log = _open_log(args.outdir / 'log')
    print_node_startup(log, orig_args)
    log("Initialization")
    log("===============")
    log("")
    p = attrdict({'outdir': args.outdir})
    args.trackequil = 0
    args.init_model = None
    p.update(process_potts_args(args, getattr(args, 'L', None), None, None, log))
    L, q, alpha = p.L, p.q, p.alpha

    background = loadSequenceFile(args.backgroundseqs, alpha, log)
    subseqs = loadSequenceFile(args.subseqs, alpha, log)
    positions = np.array([int(x) for x in args.fixpos.split(',')])
    # One set stays in the large device buffer; the members of the other
    # (the walker buffer) are substituted into it one at a time.  Normally
    # the subsequences substitute their fixed positions into the background;
    # with --iterbackgrounds each background sequence substitutes everything
    # but those positions into the subsequence set.
    if args.iterbackgrounds:
        walkers, stored = background, subseqs
        substituted = np.setdiff1d(np.arange(L), positions)
    else:
        walkers, stored = subseqs, background
        substituted = positions
    nwalk = len(walkers)
    walkers = _pad_rows(walkers, 512)
    args.nwalkers = walkers.shape[0]
    p.update(process_GPU_args(args, L, q, p.outdir, log))
    p['rngseed'] = 0
    gpus = setup_GPUs(p, log, splitwalkers=False)
    gpus.initMCMC(p.nsteps)
    stored_parts = np.array_split(stored, gpus.ngpus)
    gpus.initLargeBufs(stored_parts[0].shape[0])
    gpus.initSubseq()
    log(substituted)
    marks = np.zeros(L, dtype='u1')
    marks[substituted] = 1

    gpus.setBuf('seq main', walkers)
    gpus.setBuf('seq large', stored_parts)
    gpus.markPos(marks)
    gpus.setBuf('J', p.couplings)
    log("")
    log("Subsequence Frequency Calculation")
    log("=================================")
    log("")

    def substituted_energies(n):
        """Energies of the stored set with walker n's marked residues."""
        gpus.copySubseq(n)
        gpus.calcEnergies('large')
        return gpus.collect('E large')

    gpus.calcEnergies('main')
    walkerE = gpus.collect('E main')[:nwalk]
    log("Getting substituted energies...")
    if args.iterbackgrounds:
        # log f of stored sequence k: logsumexp over background walkers n of
        # E_n - E(k with n's residues outside the fixed positions)
        logf = np.full(len(stored), -np.inf)
        for n in range(nwalk):
            logf = np.logaddexp(logf, walkerE[n] - substituted_energies(n))
    else:
        # log f of subsequence n: logsumexp over background sequences b of
        # E_b - E(b with n's fixed positions).  (The reference subtracts the
        # subsequence set's energies here, which only broadcasts when both
        # sets have the same size; the background energies are meant.)
        gpus.calcEnergies('large')
        storedE = gpus.collect('E large')
        logf = np.array([logsumexp(storedE - substituted_energies(n)) for n in range(nwalk)])
    log("Done! Saving result (log frequency) to file.")
    np.save(p.outdir / 'probs.npy', logf)
    logfile.close()
    return logf


def main(args):
    import argparse
    actions = {'infer': inverseIsing, 'energies': getEnergies, 'benchmark': MCMCbenchmark,
               'subseq': subseqFreq, 'gen': equilibrate}
    parser = argparse.ArgumentParser(description='Perform biophysical Potts Model calculations '
                                     'on the GPU', add_help=False)
    parser.add_argument('action', choices=actions.keys(), nargs='?', default=None)
    parser.add_argument('--clinfo', action='store_true', help="Display detected GPUs")
    parser.add_argument('--mpi', action='store_true', help="Enable MPI (not supported)")
    parser.add_argument('-h', '--help', action='store_true')
    known, rest = parser.parse_known_args(args)
    if known.clinfo:
        printGPUs(print)
        return
    if known.action is None:
        print(parser.format_help() if known.help else parser.format_usage())
        return
    if known.help:
        rest.append('-h')
    if known.mpi:
        raise Exception("MPI is not part of this build (one GPU per process)")
    return actions[known.action](args, rest, print)


if __name__ == '__main__':
    setup_exit_hook(print)
    main(sys.argv[1:])
