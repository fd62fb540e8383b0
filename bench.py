#!/usr/bin/env python3
"""Benchmark of the plain Mi3 sweep on one MI355X.

Workload: the production-size inference round's sampler -- a random Potts
model with L=232 sites and q=21 letters (fields h ~ N(0,1), couplings
J ~ N(0, 0.05) moved to the zero-sum gauge and folded into the fieldless
float32 form), 2^18 walkers.  One step = one MCMC kernel call of
8 sweeps (8*L single-site Metropolis updates per walker).  The value is
site-updates/s over the timed steps, with walker state and couplings resident
in HBM before timing starts.

Printed as one JSON line, with
  roofline:     algorithmic bytes (L+1 per site update: read the walker's L
                residues, write back one) per launch over the launch duration
                measured with HIP events on the engine's stream, against the
                8 TB/s HBM peak;
  cpu_baseline: the CPU port of the same sampler (oracle/oracle.c, float32
                as the reference CPU sampler) on this host's cores, on a
                bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / 'mi3-gpu_amd'))
sys.path.insert(0, str(REPO))

METRIC = "site-updates/s & energy-ESS/s at L=232 q=21; s per inference round at fixed SSR"
HBM_PEAK_GBS = 8000.0


def production_model(L=232, q=21, seed=1):
    """h ~ N(0,1), J ~ N(0,0.05), zero-sum gauge, folded to fieldless float32."""
    from mi3gpu.utils.changeGauge import zeroJGauge, fieldlessGaugeEven
    rng = np.random.default_rng(seed)
    h = rng.normal(0, 1.0, size=(L, q))
    J = rng.normal(0, 0.05, size=(L * (L - 1) // 2, q * q))
    h, J = zeroJGauge(h, J)
    return fieldlessGaugeEven(h, J)[1].astype('<f4')


def cpu_baseline(J, L, q, seed, target_s=12.0):
    """Time the CPU port of the sampler on this host's cores."""
    from oracle import cboracle
    nthr = int(min(16, os.cpu_count() or 1))
    walkers = 64 * nthr
    rng = np.random.default_rng(seed)
    s = rng.integers(0, q, size=(walkers, L)).astype('<u1')
    t0 = time.perf_counter()
    s = cboracle.metropolis_f32(s, J, seed, 0, L, nthreads=nthr)
    t1 = time.perf_counter() - t0
    sweeps = int(max(1, min(64, target_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    cboracle.metropolis_f32(s, J, seed, 2 * L, sweeps * L, nthreads=nthr)
    dt = time.perf_counter() - t0
    su = walkers * sweeps * L
    return {"value": su / dt, "unit": "site-updates/s", "cores": nthr, "kind": "port",
            "sample": f"{walkers} walkers x {sweeps} sweeps of the same L=232 q=21 model "
                      f"({su:.3g} site updates, {dt:.1f} s, oracle/oracle.c float32, "
                      f"{nthr} OpenMP threads)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--walkers', type=int, default=2 ** 18)
    ap.add_argument('--sweeps-per-step', type=int, default=8)
    ap.add_argument('--wgsize', type=int, default=512,
                    help="walkers per work-group of the sweep (128, 256, 512)")
    ap.add_argument('--no-cpu', action='store_true', help="skip the CPU baseline")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported by this build "
              "(the path shards by walker; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from mi3gpu.mcmcGPU import MCMCGPU
    L, q, N = 232, 21, a.walkers
    nsteps = a.sweeps_per_step * L
    J = production_model(L, q, seed=1)
    s0 = np.random.default_rng(2).integers(0, q, size=(N, L)).astype('<u1')
    with tempfile.TemporaryDirectory() as d:
        g = MCMCGPU((0, 0, None, None), L, q, N, a.wgsize, Path(d), 1024, 2, profile=True)
        g.initMCMC(nsteps, 0, 2 ** 63)
        g.setBuf('J', J)
        g.setBuf('seq main', s0)
        for _ in range(a.warmup):
            g.runMCMC()
        g.wait()
        g.profile_reset()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            g.runMCMC()
        g.wait()
        dt = time.perf_counter() - t0
        kms, kcnt = g.kernel_ms('mcmc')
        # energies of the final ensemble (sanity)
        g.calcEnergies('main')
        E = g.getBuf('E main').read()

    su_per_step = N * nsteps
    value = su_per_step * a.steps / dt
    alg_bytes = su_per_step * (L + 1)
    achieved = alg_bytes / (kms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain Mi3 Metropolis sweep, L=232 q=21, 2^18 walkers, random "
                               "Potts model h~N(0,1) J~N(0,0.05) zero-sum gauge folded "
                               "fieldless float32",
                   "L": L, "q": q, "walkers": N, "sweeps_per_step": a.sweeps_per_step,
                   "walkers_per_workgroup": a.wgsize,
                   "mean_energy_after": float(np.mean(E))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_sweep", "kernel_ms": kms, "launches": kcnt,
                     "alg_bytes_per_launch": alg_bytes},
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(J, L, q, seed=3)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
