"""MSA text format: one sequence per line, optional whitespace-padded id
column, optional leading '#' header lines, optional bzip2 compression
(seqload.py of the reference, loadSeqs/writeSeqs/mapSeqs/reduceSeqs).
"""
import bz2
import io
import json
from pathlib import Path

import numpy as np

from mi3gpu.utils.potts_common import alpha21

prot_alpha = alpha21.encode('ascii')
_BZ2_MAGIC = b"BZh"


def _open_read(fn, zipf=None):
    if hasattr(fn, 'read'):
        return fn, False
    with open(fn, 'rb') as f:
        start = f.read(3)
    if zipf is not False and start == _BZ2_MAGIC:
        return bz2.open(fn, 'rb'), True
    if zipf is True:
        raise Exception("bz2 file expected")
    return open(fn, 'rb'), True


def _alphabet(alpha, headers):
    if alpha is None:
        for h in headers:
            if h.startswith("PARAM"):
                alpha = json.loads(h[len('PARAM '):]).get('alpha', None)
        if alpha is None:
            raise Exception("Could not determine alphabet")
    return alpha.encode('ascii') if isinstance(alpha, str) else alpha


def _parse(data, alpha, count=None):
    """Parse the bytes of a whole file -> (seqs, ids, headers)."""
    lines = data.split(b'\n')
    headers = []
    k = 0
    while k < len(lines) and (lines[k].strip() == b'' or lines[k].startswith(b'#')):
        if lines[k].strip() != b'':
            headers.append(lines[k][1:].decode('utf-8'))
        k += 1
    body = lines[k:]
    # drop trailing empty lines and a Stockholm terminator
    while body and body[-1].strip() in (b'', b'//'):
        body.pop()
    if count is not None:
        body = body[:count]
    alpha = _alphabet(alpha, headers)
    if not body:
        return np.zeros((0, 0), dtype='u1'), None, headers
    first = body[0]
    lineL = len(first)
    idL = first.rfind(b' ') + 1
    for n, l in enumerate(body):
        if len(l) != lineL:
            raise ValueError("Sequence {} has wrong length (expected {})".format(n + 1, lineL))
    mat = np.frombuffer(b''.join(body), dtype=np.uint8).reshape(len(body), lineL)
    ids = None
    if idL:
        idcol = mat[:, :idL]
        bad = np.where(idcol[:, -1] != ord(' '))[0]
        if len(bad):
            raise ValueError("Sequence {} has wrong id field length. Expected id length "
                             "{}".format(bad[0] + 1, idL))
        ids = np.char.strip(np.ascontiguousarray(idcol).view('S{}'.format(idL)).ravel())
        mat = mat[:, idL:]
    table = np.full(256, 255, np.uint8)
    table[np.frombuffer(alpha, np.uint8)] = np.arange(len(alpha), dtype=np.uint8)
    seqs = table[mat]
    bl, bc = np.where(seqs == 255)
    if len(bl):
        raise ValueError("Invalid residue '{}' in sequence {} position {}".format(
            chr(mat[bl[0], bc[0]]), bl[0] + 1, bc[0]))
    return seqs, ids, headers


def loadSeqs(fn, alpha=prot_alpha, zipf=None, count=None):
    """(seqs uint8 (N, L), ids or None, header lines) (seqload.py:91-135)."""
    f, close = _open_read(fn, zipf)
    try:
        data = f.read()
    finally:
        if close:
            f.close()
    return _parse(data, alpha, count)


def mapSeqs(fn, mapper, alpha=prot_alpha, zipf=None):
    seqs, ids, headers = loadSeqs(fn, alpha, zipf)
    if mapper is None:
        return seqs, headers
    return mapper(seqs, (ids, headers)), headers


def reduceSeqs(fn, reduce_func, startval=None, alpha=prot_alpha):
    seqs, ids, headers = loadSeqs(fn, alpha)
    a = _alphabet(alpha, headers)
    return reduce_func(startval, seqs, ids, headers, a), headers


def bz2_compress_parallel(data, chunk=8 << 20, nthreads=None):
    """bzip2-compress `data` as concatenated bzip2 streams, one per ~8 MB
    piece split at line ends, compressed on a thread pool (the compressor
    releases the GIL).  Every bzip2 reader (bzip2 -d, Python's bz2.open)
    decompresses a multi-stream file to the concatenated text, so the file
    reads back exactly as a single-stream one."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    n = len(data)
    if n <= chunk:
        return bz2.compress(data)
    cuts = [0]
    while cuts[-1] + chunk < n:
        nl = data.find(b'\n', cuts[-1] + chunk)
        if nl < 0:
            break
        cuts.append(nl + 1)
    cuts.append(n)
    pieces = [data[a:b] for a, b in zip(cuts[:-1], cuts[1:]) if b > a]
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=nthreads) as ex:
        return b''.join(ex.map(bz2.compress, pieces))


def writeSeqs(fn, seqs, alpha=prot_alpha, ids=None, headers=None, write_param=False,
              zipf=None):
    """Write sequences as text lines, optionally with ids, a PARAM header and
    bzip2 compression (seqload.py:334-410)."""
    if isinstance(alpha, str):
        alpha = alpha.encode('ascii')
    seqs = np.asarray(seqs)
    table = np.frombuffer(alpha, dtype=np.uint8)
    text = table[seqs.astype(np.intp)] if seqs.size else np.zeros((seqs.shape[0], 0), 'u1')
    if ids is not None:
        ids = [x if isinstance(x, bytes) else str(x).encode('utf-8') for x in ids]
        idL = max(len(x) for x in ids) + 1
        idcol = np.frombuffer(b''.join(x.ljust(idL) for x in ids), dtype=np.uint8)
        text = np.concatenate([idcol.reshape(len(ids), idL), text], axis=1)
    out = np.empty((text.shape[0], text.shape[1] + 1), dtype=np.uint8)
    out[:, :-1] = text
    out[:, -1] = ord('\n')
    chunks = []
    if write_param:
        chunks.append('#PARAM {}\n'.format(json.dumps({'alpha': alpha.decode()})).encode())
    if headers:
        chunks.append("\n".join(h if h.startswith('#') else '#' + h for h in headers).encode()
                      + b"\n")
    chunks.append(out.tobytes())
    payload = b''.join(chunks)
    if hasattr(fn, 'write'):
        fn.write(payload)
        return
    if zipf:
        with open(fn, 'wb') as f:
            f.write(bz2_compress_parallel(payload))
    else:
        with open(fn, 'wb') as f:
            f.write(payload)
