#!/usr/bin/env python3
"""Build libmi3hip.so (HIP, gfx950) into mi3gpu/ next to the Python host code.

    python mi3-gpu_amd/build.py [--verbose]

Translation units are compiled in parallel with hipcc and linked into one
shared library.  Packed FP32 instructions are disabled for device code
(`-Xclang -target-feature -Xclang -packed-fp32-ops`); the host pass warns that
it does not know the feature, which is harmless.
"""
import argparse
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent
CSRC = HERE / 'csrc'
OUT = HERE / 'mi3gpu' / 'libmi3hip.so'
BUILD = HERE / '_build'
ARCH = os.environ.get('MI3_OFFLOAD_ARCH', 'gfx950')

UNITS = ['runtime.hip', 'potts.hip', 'batch.hip', 'made.hip']

CFLAGS = [
    f'--offload-arch={ARCH}', '-O3', '-fPIC', '-std=c++17',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument',
    '-I', str(REPO / 'include'), '-I', str(CSRC),
]


def build(verbose=False):
    BUILD.mkdir(exist_ok=True)
    procs = []
    objs = []
    for u in UNITS:
        obj = BUILD / (Path(u).stem + '.o')
        objs.append(obj)
        cmd = ['hipcc', *CFLAGS, '-c', str(CSRC / u), '-o', str(obj)]
        if verbose:
            print(' '.join(cmd), flush=True)
        procs.append((u, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = False
    for u, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors='replace')
        # drop the host pass's note about the device-only feature flag
        text = '\n'.join(l for l in text.splitlines()
                         if 'packed-fp32-ops' not in l and 'ignoring feature' not in l)
        if p.returncode != 0:
            failed = True
            print(f'--- {u} failed ---\n{text}', file=sys.stderr)
        elif verbose and text.strip():
            print(text)
    if failed:
        raise RuntimeError('libmi3hip build failed')
    tmp = OUT.with_suffix('.so.tmp')
    cmd = ['hipcc', f'--offload-arch={ARCH}', '-shared', '-fPIC', *map(str, objs), '-o', str(tmp)]
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    ap = argparse.ArgumentParser()
    ap.add_argument('--verbose', action='store_true')
    a = ap.parse_args()
    print(build(a.verbose))
