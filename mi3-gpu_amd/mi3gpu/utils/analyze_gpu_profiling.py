#!/usr/bin/env python3
"""Summarise the `EVT name start end [bytes]` lines written by
MCMCGPU.logProfile into gpu-N.log (analyze_gpu_profiling.py of the
reference): total time and call count per kernel."""
import sys
from collections import defaultdict


def summarize(lines):
    tot, cnt, byt = defaultdict(float), defaultdict(int), defaultdict(int)
    for l in lines:
        p = l.split()
        if len(p) < 4 or p[0] != 'EVT':
            continue
        name, a, b = p[1], int(p[2]), int(p[3])
        tot[name] += (b - a) * 1e-9
        cnt[name] += 1
        if len(p) > 4:
            byt[name] += int(p[4])
    return {k: (tot[k], cnt[k], byt[k]) for k in tot}


def main(argv):
    with open(argv[1]) as f:
        res = summarize(f)
    for k, (t, n, b) in sorted(res.items(), key=lambda kv: -kv[1][0]):
        print(f"{k:24s} {t:10.4f} s  {n:8d} calls  {b / max(t, 1e-12) / 1e9:9.1f} GB/s")


if __name__ == '__main__':
    main(sys.argv)
