# This file keeps the interface and behaviour of the file of the same name
# in Mi3-GPU, Copyright 2020 Allan Haldane, and is a derived work of it.
# Mi3-GPU is free software: you can redistribute it and/or modify it under
# the terms of the GNU General Public License as published by the Free
# Software Foundation, version 3 of the License; this file is distributed
# under the same licence, WITHOUT ANY WARRANTY; without even the implied
# warranty of MERCHANTABILITY or FITNESS FOR A PARTICULAR PURPOSE.  See the
# GNU General Public License <http://www.gnu.org/licenses/> for details.
"""Per-device host wrapper of the MI355X Mi3 engine.

Drop-in for the reference's mcmcGPU module (mi3gpu/mcmcGPU.py): same class
(`MCMCGPU`), method names, arguments, buffer names and dtypes, same
`FutureBuf` read semantics and the same module functions
(`setup_GPU_context`, `initGPU`, `wgsize_heuristic`, `histogram_heuristic`,
`printGPUs`, `unpackJ_CPU`).  Underneath, every call goes through the C ABI of
libmi3hip.so (include/mi3hip.h) to hand-written HIP kernels on one in-order
stream; there is no OpenCL and no CPU fallback.

Differences that callers can observe:
  * kernel wrappers accept `wait_for` for compatibility and ignore it: all
    work is stream ordered, which is what the reference's default
    `wait_for=None` chaining produces (mcmcGPU.py:211-222);
  * the device layout of "seq *" buffers is site-major (L rows of n bytes)
    instead of the (SWORDS, n) uint32 layout; setBuf/getBuf still take and
    return (n, L) uint8 arrays, so callers see no difference;
  * random streams are counter-based (Philox4x32-10 keyed by the seed), so
    the "rngstates" and "randpos" buffers do not exist;
  * `updateJ` followed by `reg_l1z`/`reg_l2z` runs as one fused kernel.
"""
import collections
import ctypes
import time
from pathlib import Path

import numpy as np

from mi3gpu import _lib
from mi3gpu._lib import lib, check, enc, as_ptr
from mi3gpu.utils.potts_common import printsome

rng_buf_mul = 1024


def printGPUs(log):
    """List the visible devices (mcmcGPU.py:58-63)."""
    n = _lib.device_count()
    log("Platform 0 'HIP':")
    for d in range(n):
        log(_lib.device_info(d))
    log("")


class FutureBuf:
    """Result of getBuf: `read()` blocks until the copy, enqueued in stream
    order when getBuf was called, has landed (mcmcGPU.py:95-108)."""

    def __init__(self, gpu, ticket, shape, dtype, postprocess=None):
        self._gpu = gpu
        self._ticket = ticket
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.postfunc = postprocess
        self._data = None

    def read(self):
        if self._data is None:
            buf = np.empty(self.shape, dtype=self.dtype)
            check(self._gpu._ctx, lib().mi3_ticket_wait(self._gpu._ctx, self._ticket,
                                                       as_ptr(buf), buf.nbytes))
            self._data = self.postfunc(buf) if self.postfunc is not None else buf
        return self._data


def _pending_flush(meth):
    """Run a deferred updateJ before any other device operation."""
    def wrapped(self, *a, **k):
        self._flush_update()
        return meth(self, *a, **k)
    wrapped.__name__ = meth.__name__
    wrapped.__doc__ = meth.__doc__
    return wrapped


class MCMCGPU:
    def __init__(self, gpuinfo, L, q, nseq, wgsize, outdir, vsize, seed, profile=False):
        if nseq % 512 != 0:
            raise ValueError("nwalkers/ngpus must be a multiple of 512")
        device, gpunum, _, _ = gpuinfo
        self.L, self.q = int(L), int(q)
        self.nPairs = self.L * (self.L - 1) // 2
        self.SWORDS = (self.L - 1) // 4 + 1
        self.SBYTES = 4 * self.SWORDS
        self.nseq = {'main': int(nseq)}
        self.nwalkers = int(nseq)
        self.gpunum = gpunum
        self.device = device
        self.wgsize = wgsize
        self.nhist, self.histws = histogram_heuristic(q)
        if nseq % wgsize != 0:
            raise Exception("nseq per GPU must be a multiple of wgsize")
        self.profile = profile
        self.logfn = Path(outdir) / 'gpu-{}.log'.format(gpunum)
        with open(self.logfn, "wt") as f:
            f.write(_lib.device_info(device) + "\n")

        ctx = ctypes.c_void_p()
        check(None, lib().mi3_create(device, self.L, self.q, self.nwalkers,
                                     int(seed) & 0xFFFFFFFFFFFFFFFF, 1 if profile else 0,
                                     ctypes.byref(ctx)))
        self._ctx = ctx
        check(self._ctx, lib().mi3_set_wgsize(self._ctx, int(wgsize)))
        self.seed = int(seed)
        self.initted = []
        self.buf_spec = {}
        self.largebufs = []
        self.nstoredseqs = 0
        self._pending = None

        nP, q2 = self.nPairs, self.q * self.q
        self._setupBuffer('J', '<f4', (nP, q2))
        self._setupBuffer('bi', '<f4', (nP, q2))
        self._setupBuffer('bicount', '<u4', (nP, q2))
        self._setupBuffer('seq main', '<u1', (self.nseq['main'], self.L))
        self._setupBuffer('cprob', '<f4', (self.L, self.q - 1))
        self._setupBuffer('E main', '<f4', (self.nseq['main'],))
        self._setupBuffer('minout', '<f4', (1,))
        self.log("Initialized context: L={} q={} nwalkers={}".format(L, q, nseq))

    def __del__(self):
        ctx = getattr(self, '_ctx', None)
        if ctx:
            try:
                lib().mi3_destroy(ctx)
            except Exception:
                pass
            self._ctx = None

    # -- logging ---------------------------------------------------------

    def log(self, msg):
        with open(self.logfn, "at") as f:
            print("{: 10.3f}".format(time.process_time()), msg, file=f)

    def logProfile(self):
        """Append EVT lines (name, start, end, bytes) for completed launches
        (mcmcGPU.py:224-237)."""
        if not self.profile:
            return
        lib().mi3_profile_flush(self._ctx, enc(str(self.logfn)))

    # -- buffers -----------------------------------------------------------

    def _setupBuffer(self, bufname, buftype, bufshape, pad=None, flags=None):
        nbytes = int(np.dtype(buftype).itemsize * int(np.prod(bufshape)))
        check(self._ctx, lib().mi3_buf_alloc(self._ctx, enc(bufname), nbytes))
        self.buf_spec[bufname] = (buftype, tuple(bufshape), flags)

    def require(self, *reqs):
        for r in reqs:
            if r not in self.initted:
                raise Exception("{} not initialized".format(r))

    def _initcomponent(self, cmp):
        if cmp in self.initted:
            raise Exception("Already initialized {}".format(cmp))
        self.initted.append(cmp)

    def initMCMC(self, nsteps, rng_offset, rng_span):
        self._initcomponent('MCMC')
        self.nsteps = int(nsteps)
        self._setupBuffer('Bs', '<f4', (self.nseq['main'],))
        self.setBuf('Bs', np.ones(self.nseq['main'], dtype='<f4'))
        check(self._ctx, lib().mi3_init_mcmc(self._ctx, self.nsteps, int(rng_offset),
                                             int(rng_span)))

    def initLargeBufs(self, nseq_large):
        self._initcomponent('Large')
        n = int(nseq_large)
        # rows of the site-major layout are padded to 64 bytes
        self.nseq['large'] = n
        self._large_ld = ((n + 63) // 64) * 64
        self._setupBuffer('seq large', '<u1', (self._large_ld, self.L))
        self.buf_spec['seq large'] = ('<u1', (n, self.L), None)
        self._setupBuffer('E large', '<f4', (self._large_ld,))
        self._setupBuffer('E tmp large', '<f4', (self._large_ld,))
        self._setupBuffer('weights large', '<f4', (self._large_ld,))
        for b in ('E large', 'E tmp large', 'weights large'):
            self.buf_spec[b] = ('<f4', (n,), None)
        self.largebufs.extend(['seq large', 'E large', 'weights large'])
        self.nstoredseqs = 0

    def initSubseq(self):
        self.require('Large')
        self._initcomponent('Subseq')
        self._setupBuffer('markpos', '<u1', (self.SBYTES,))

    def initJstep(self):
        self._initcomponent('Jstep')
        nP, q2 = self.nPairs, self.q * self.q
        self._setupBuffer('dJ', '<f4', (nP, q2))
        self._setupBuffer('bi target', '<f4', (nP, q2))
        self._setupBuffer('Creg', '<f4', (nP, q2))
        self._setupBuffer('Xlambdas', '<f4', (nP,))
        self._setupBuffer('weights', '<f4', (self.nseq['main'],))
        self._setupBuffer('weightstats', '<f4', (2,))
        self._setupBuffer('E tmp', '<f4', (self.nseq['main'],))

    # -- host-side sequence packing (reference layout helpers) --------------

    def packSeqs_4(self, seqs):
        """(n, L) uint8 -> (SWORDS, n) uint32, 4 residues per word, little
        endian (the reference's device layout, mcmcGPU.py:333-346)."""
        if seqs.dtype != np.dtype('<u1'):
            raise Exception("seqs must have u1 dtype")
        b = np.zeros((seqs.shape[0], self.SBYTES), dtype='<u1')
        b[:, :self.L] = seqs
        return np.ascontiguousarray(b.view('<u4').T)

    def unpackSeqs_4(self, mem):
        """Inverse of packSeqs_4 (mcmcGPU.py:348-354)."""
        b = np.ascontiguousarray(np.asarray(mem, dtype='<u4').T).view('<u1')
        return b[:, :self.L]

    def _site_major(self, seqs, ld):
        out = np.zeros((self.L, ld), dtype='<u1')
        out[:, :seqs.shape[0]] = seqs.T
        return out

    def _seq_ld(self, bufname):
        return self._large_ld if bufname == 'seq large' else self.nseq['main']

    # -- deferred coupling update ----------------------------------------------

    def _flush_update(self):
        if self._pending is not None:
            gamma, pc, Jbuf = self._pending
            self._pending = None
            check(self._ctx, lib().mi3_update_J(self._ctx, gamma, pc, enc(Jbuf)))

    # -- compute ---------------------------------------------------------------

    @_pending_flush
    def repackseqs_T(self, bufname, wait_for=None):
        """No-op: the device layout is already site-major (the transposed
        layout the reference builds here, mcmcGPU.py:356-373)."""
        return None

    @_pending_flush
    def prepare_indep(self, unimarg, wait_for=None):
        """Cumulative probabilities for the independent reseed
        (mcmcGPU.py:375-378)."""
        cprob = np.cumsum(np.asarray(unimarg, dtype=np.float64), axis=1)
        cprob = (cprob[:, :-1] / cprob[:, -1, None]).astype('<f4')
        return self.setBuf('cprob', cprob)

    @_pending_flush
    def gen_indep(self, bufname, wait_for=None):
        self.log("gen_indep")
        check(self._ctx, lib().mi3_gen_indep(self._ctx, enc('seq ' + bufname)))
        if bufname == 'large':
            self.nstoredseqs = self.nseq['large']

    def unpackJ(self, wait_for=None):
        """No-op: the library keeps the chunked copy of J in sync."""
        return wait_for

    @_pending_flush
    def runMCMC(self, wait_for=None):
        """nsteps Metropolis site updates of every walker (mcmcGPU.py:460-482)."""
        self.require('MCMC')
        self.log("runMCMC")
        beta = getattr(self, 'beta', None)
        check(self._ctx, lib().mi3_run_mcmc(self._ctx, 1.0 if beta is None else float(beta)))

    def measureFPerror(self, log, nloops=3):
        """Compare energies recomputed on device with a float64 host sum
        (the intent of mcmcGPU.py:484-499, which is broken there)."""
        from mi3gpu.utils.getSeqEnergies import E_potts
        for n in range(nloops):
            self.runMCMC()
            self.calcEnergies('main')
            e = self.getBuf('E main').read()
            seqs = self.getBuf('seq main').read()
            J = self.getBuf('J').read()
            ex = E_potts(seqs, J)
            log("Run", n, "Error:", np.mean((e - ex) ** 2))

    def _nseq_of(self, seqbufname):
        return self.nseq['main'] if seqbufname == 'main' else self.nstoredseqs

    @_pending_flush
    def calcBicounts(self, seqbufname, wait_for=None):
        self.log("calcBicounts " + seqbufname)
        check(self._ctx, lib().mi3_calc_bicounts(self._ctx, enc('seq ' + seqbufname),
                                                 self._nseq_of(seqbufname)))

    @_pending_flush
    def bicounts_to_bimarg(self, seqbufname='main', wait_for=None):
        self.log("bicounts_to_bimarg ")
        check(self._ctx, lib().mi3_bicounts_to_bimarg(self._ctx, self._nseq_of(seqbufname)))

    @_pending_flush
    def calcEnergies(self, seqbufname, Jbufname='J', wait_for=None):
        self.log("calcEnergies " + seqbufname)
        check(self._ctx, lib().mi3_calc_energies(self._ctx, enc('seq ' + seqbufname),
                                                 enc(Jbufname), enc('E ' + seqbufname),
                                                 self._nseq_of(seqbufname)))

    @_pending_flush
    def min_buf(self, buf, wait_for=None):
        self.require('Jstep')
        n = self._nseq_of(buf.split()[-1]) if buf.startswith(('E', 'weights')) else \
            int(np.prod(self.buf_spec[buf][1]))
        check(self._ctx, lib().mi3_min_buf(self._ctx, enc(buf), n))

    @_pending_flush
    def weight_statistics(self, buf='main', wait_for=None):
        self.require('Jstep')
        wname = 'weights' if buf == 'main' else 'weights large'
        check(self._ctx, lib().mi3_weight_stats(self._ctx, enc(wname), self._nseq_of(buf)))

    @_pending_flush
    def dE_to_weights(self, buf='main', offset=0., wait_for=None):
        self.require('Jstep')
        ename, wname = ('E main', 'weights') if buf == 'main' else ('E large', 'weights large')
        check(self._ctx, lib().mi3_dE_to_weights(self._ctx, enc(ename), enc(wname),
                                                 self._nseq_of(buf), float(offset)))

    @_pending_flush
    def fixed_beta_weights(self, ref_E, seqbufname='main', wait_for=None):
        self.require('Jstep')
        beta = getattr(self, 'beta', None)
        beta = 1.0 if beta is None else float(beta)
        wname = 'weights' if seqbufname == 'main' else 'weights large'
        check(self._ctx, lib().mi3_fixed_beta_weights(self._ctx, enc('E ' + seqbufname),
                                                      enc(wname), self._nseq_of(seqbufname),
                                                      float(ref_E), beta))

    @_pending_flush
    def weightedMarg(self, seqbufname='main', wait_for=None):
        self.require('Jstep')
        self.log("weightedMarg")
        wname = 'weights' if seqbufname == 'main' else 'weights large'
        check(self._ctx, lib().mi3_weighted_marg(self._ctx, enc('seq ' + seqbufname),
                                                 enc(wname), self._nseq_of(seqbufname)))

    @_pending_flush
    def renormalize_bimarg(self, wait_for=None):
        check(self._ctx, lib().mi3_renormalize_bimarg(self._ctx))

    @_pending_flush
    def addFloatBuf(self, dstname, srcname, wait_for=None):
        n = int(np.prod(self.buf_spec[dstname][1]))
        if n != int(np.prod(self.buf_spec[srcname][1])):
            raise Exception('Tried to add bufs of different sizes')
        check(self._ctx, lib().mi3_add_float_bufs(self._ctx, enc(dstname), enc(srcname), n))

    def addBiBuffer(self, bufname, otherbuf, wait_for=None):
        raise NotImplementedError("cross-device bimarg reduction: this build drives one GPU")

    @_pending_flush
    def updateJ(self, gamma, pc, Jbuf='dJ', wait_for=None):
        """Deferred: runs fused with a following reg_l1z/reg_l2z, or alone
        before the next device operation (updatedJ, mcmc.cl:829-843)."""
        self.require('Jstep')
        self.log("updateJ")
        self._pending = (float(gamma), float(pc), Jbuf)

    def _reg(self, kind, gamma, pc, lJ):
        self.require('Jstep')
        if self._pending is not None and self._pending[2] == 'dJ':
            g, p, _ = self._pending
            self._pending = None
            if (g, p) == (float(gamma), float(pc)):
                check(self._ctx, lib().mi3_update_J_reg(self._ctx, g, p, kind, float(lJ)))
                return
            check(self._ctx, lib().mi3_update_J(self._ctx, g, p, enc('dJ')))
        else:
            self._flush_update()
        fn = lib().mi3_reg_l1z if kind == 1 else lib().mi3_reg_l2z
        check(self._ctx, fn(self._ctx, float(gamma), float(pc), float(lJ)))

    def reg_l1z(self, gamma, pc, lJ, wait_for=None):
        self.log("reg_l1z")
        self._reg(1, gamma, pc, lJ)

    def reg_l2z(self, gamma, pc, lJ, wait_for=None):
        self.log("reg_l2z")
        self._reg(2, gamma, pc, lJ)

    @_pending_flush
    def reg_SCADJ(self, gamma, pc, r, a, wait_for=None):
        """SCAD penalty on the zero-gauge couplings (mcmcGPU.py:747-759)."""
        self.require('Jstep')
        self.log("reg_SCADJ")
        check(self._ctx, lib().mi3_reg_SCADJ(self._ctx, float(gamma), float(pc), float(r),
                                             float(a)))

    @_pending_flush
    def reg_X(self, gamma, pc, lX, wait_for=None):
        """lambda_X |X| penalty, X = sum J*C (mcmcGPU.py:775-787, kernel
        mcmc.cl:1004-1028)."""
        self.require('Jstep')
        self.log("reg X")
        check(self._ctx, lib().mi3_reg_X(self._ctx, float(gamma), float(pc), float(lX)))

    @_pending_flush
    def reg_Xij(self, gamma, pc, wait_for=None):
        """Per-pair lambda_X read from the 'Creg' buffer (mcmcGPU.py:761-773,
        kernel mcmc.cl:972-1002)."""
        self.require('Jstep')
        self.log("reg Xij")
        check(self._ctx, lib().mi3_reg_Xij(self._ctx, float(gamma), float(pc)))

    @_pending_flush
    def reg_SCADX(self, gamma, pc, s, r, a, wait_for=None):
        self.require('Jstep')
        self.log("reg_SCADX")
        check(self._ctx, lib().mi3_reg_SCADX(self._ctx, float(gamma), float(pc), float(s),
                                             float(r), float(a)))

    @_pending_flush
    def reg_expX(self, gamma, pc, lam, wait_for=None):
        self.require('Jstep')
        self.log("reg_expX")
        check(self._ctx, lib().mi3_reg_expX(self._ctx, float(gamma), float(pc), float(lam)))

    @_pending_flush
    def reg_ddE(self, gamma, pc, lam, wait_for=None):
        self.require('Jstep')
        self.log("reg ddE")
        check(self._ctx, lib().mi3_reg_ddE(self._ctx, float(gamma), float(pc), float(lam)))

    @_pending_flush
    def reg_SCADddE(self, gamma, pc, lam, r, wait_for=None):
        self.require('Jstep')
        self.log("reg SCADddE")
        check(self._ctx, lib().mi3_reg_SCADddE(self._ctx, float(gamma), float(pc), float(lam),
                                               float(r)))

    @_pending_flush
    def getBuf(self, bufname, truncateLarge=True, wait_for=None):
        """Snapshot of a buffer, taken in stream order (mcmcGPU.py:845-867)."""
        self.log("getBuf " + bufname)
        buftype, bufshape, _ = self.buf_spec[bufname]
        ticket = ctypes.c_int(-1)
        if bufname.split()[0] == 'seq':
            ld = self._seq_ld(bufname)
            n = bufshape[0]
            check(self._ctx, lib().mi3_buf_read_async(self._ctx, enc(bufname), self.L * ld, 0,
                                                      ctypes.byref(ticket)))
            nret = self.nstoredseqs if (bufname in self.largebufs and truncateLarge) else n

            def post(b, nret=nret):
                return np.ascontiguousarray(b[:, :nret].T)
            return FutureBuf(self, ticket.value, (self.L, ld), '<u1', post)
        nbytes = int(np.dtype(buftype).itemsize * int(np.prod(bufshape)))
        check(self._ctx, lib().mi3_buf_read_async(self._ctx, enc(bufname), nbytes, 0,
                                                  ctypes.byref(ticket)))
        if bufname in self.largebufs and truncateLarge:
            nret = self.nstoredseqs
            return FutureBuf(self, ticket.value, bufshape, buftype, lambda b: b[:nret])
        return FutureBuf(self, ticket.value, bufshape, buftype)

    @_pending_flush
    def setBuf(self, bufname, buf, wait_for=None):
        """Copy a host array (or another device buffer, by name) into a
        buffer, with the reference's dtype/shape checks (mcmcGPU.py:869-907)."""
        self.log("setBuf " + bufname)
        if isinstance(buf, str):
            buftype, bufshape, _ = self.buf_spec[bufname]
            nbytes = int(np.dtype(buftype).itemsize * int(np.prod(bufshape)))
            check(self._ctx, lib().mi3_buf_copy(self._ctx, enc(bufname), enc(buf), nbytes))
            return None
        buftype, bufshape, _ = self.buf_spec[bufname]
        if not isinstance(buf, np.ndarray):
            buf = np.array(buf, dtype=buftype)
        if np.dtype(buftype) != buf.dtype:
            raise ValueError("Buffer dtype mismatch.Expected {}, got {}".format(
                np.dtype(buftype), buf.dtype))
        if bufname.split()[0] == 'seq':
            if buf.ndim != 2 or buf.shape[1] != self.L or buf.shape[0] > bufshape[0] or \
                    (bufname == 'seq main' and buf.shape[0] != bufshape[0]):
                raise ValueError("Buffer size mismatch. Expected {}, got {}".format(
                    bufshape, buf.shape))
            if buf.size and int(buf.max()) >= self.q:
                raise ValueError("sequence residues must be < q = {}".format(self.q))
            dev = self._site_major(buf, self._seq_ld(bufname))
            check(self._ctx, lib().mi3_buf_write(self._ctx, enc(bufname), as_ptr(dev),
                                                 dev.nbytes, 0))
            if bufname == 'seq large':
                self.nstoredseqs = buf.shape[0]
            return None
        if bufshape != buf.shape and not (bufshape == (1,) or buf.size == 1):
            raise ValueError("Buffer size mismatch. Expected {}, got {}".format(
                bufshape, buf.shape))
        buf = np.ascontiguousarray(buf)
        check(self._ctx, lib().mi3_buf_write(self._ctx, enc(bufname), as_ptr(buf),
                                             buf.nbytes, 0))
        return None

    @_pending_flush
    def fillBuf(self, bufname, val, wait_for=None):
        self.log("fillBuf " + bufname)
        buftype = np.dtype(self.buf_spec[bufname][0])
        v = np.array([val], dtype=buftype)
        if buftype.itemsize == 4:
            pattern = int(v.view('<u4')[0])
        elif buftype.itemsize == 1:
            pattern = int(v.view('<u1')[0]) * 0x01010101
        else:
            raise ValueError("fillBuf supports 1- and 4-byte element types")
        check(self._ctx, lib().mi3_buf_fill_u32(self._ctx, enc(bufname), pattern))

    def markPos(self, marks, wait_for=None):
        self.require('Subseq')
        marks = np.asarray(marks).astype('<u1')
        if len(marks) == self.L:
            marks = np.resize(marks, self.SBYTES)
            marks[self.L:] = 0
        return self.setBuf('markpos', marks)

    @_pending_flush
    def fillSeqs(self, startseq, seqbufname='main', wait_for=None):
        self.log("fillSeqs " + seqbufname)
        s = np.ascontiguousarray(startseq, dtype='<u1')
        if s.shape != (self.L,):
            raise ValueError("seed sequence must have length L")
        check(self._ctx, lib().mi3_fill_seqs(self._ctx, enc('seq ' + seqbufname), as_ptr(s)))

    @_pending_flush
    def storeSeqs(self, seqs=None, wait_for=None):
        """Append sequences (or the whole main buffer) to the large buffer
        (mcmcGPU.py:934-978), as one device-side copy."""
        self.require('Large')
        offset = self.nstoredseqs
        self.log("storeSeqs " + str(offset))
        if seqs is None:
            nseq = self.nseq['main']
            if offset + nseq > self.nseq['large']:
                raise Exception("cannot store seqs past end of large buffer")
            src = 'seq main'
        else:
            seqs = np.asarray(seqs)
            nseq, L = seqs.shape
            if L != self.L:
                raise Exception("Sequences have wrong length: {} vs {}".format(L, self.L))
            if offset + nseq > self.nseq['large']:
                raise Exception("cannot store seqs past end of large buffer")
            if seqs.dtype != np.dtype('u1'):
                raise ValueError("sequences must be uint8")
            if seqs.size and int(seqs.max()) >= self.q:
                raise ValueError("sequence residues must be < q = {}".format(self.q))
            if not lib().mi3_buf_exists(self._ctx, b'seq stage'):
                check(self._ctx, lib().mi3_buf_alloc(self._ctx, b'seq stage',
                                                     self.L * self._large_ld))
            dev = self._site_major(seqs, self._large_ld)
            check(self._ctx, lib().mi3_buf_write(self._ctx, b'seq stage', as_ptr(dev),
                                                 dev.nbytes, 0))
            src = 'seq stage'
        check(self._ctx, lib().mi3_copy_seqs(self._ctx, enc(src), 0, b'seq large', offset,
                                             nseq))
        self.nstoredseqs += nseq

    def clearLargeSeqs(self):
        self.require('Large')
        self.nstoredseqs = 0

    @_pending_flush
    def restoreSeqs(self, wait_for=None):
        """Main buffer <- the first nseq stored large-buffer sequences (the
        reference's restoreSeqs reads an undefined offset; this one uses 0)."""
        self.require('Large')
        nseq = self.nseq['main']
        if self.nstoredseqs < nseq:
            raise Exception("not enough seqs stored in large buffer")
        check(self._ctx, lib().mi3_copy_seqs(self._ctx, b'seq large', 0, b'seq main', 0, nseq))

    @_pending_flush
    def copySubseq(self, seqind, wait_for=None):
        """Every large-buffer sequence takes main sequence `seqind`'s
        residues at the marked positions (mcmcGPU.py:1002-1014)."""
        self.require('Subseq')
        self.log("copySubseq " + str(seqind))
        if seqind >= self.nseq['main']:
            raise Exception("given index is past end of main seq buffer")
        check(self._ctx, lib().mi3_copy_subseq(self._ctx, b'seq main', int(seqind),
                                               b'seq large', b'markpos'))

    def wait(self):
        self._flush_update()
        check(self._ctx, lib().mi3_synchronize(self._ctx))

    # -- measurement helpers (not in the reference) -----------------------------

    def kernel_ms(self, name):
        """(average ms, count) of profiled launches named `name`."""
        avg = ctypes.c_double(0)
        cnt = ctypes.c_int(0)
        check(self._ctx, lib().mi3_profile_kernel_ms(self._ctx, enc(name), ctypes.byref(avg),
                                                     ctypes.byref(cnt)))
        return avg.value, cnt.value

    def profile_reset(self):
        check(self._ctx, lib().mi3_profile_reset(self._ctx))

    def mcmc_steps_done(self):
        s = ctypes.c_uint64(0)
        check(self._ctx, lib().mi3_mcmc_step_count(self._ctx, ctypes.byref(s)))
        return s.value


################################################################################

def unpackJ_CPU(couplings):
    """(L(L-1)/2, q*q) -> (L*L, q*q) with every ordered pair (mcmcGPU.py:1022-1033)."""
    nP, q2 = couplings.shape
    L = int(((1 + np.sqrt(1 + 8 * nP)) // 2) + 0.5)
    q = int(round(np.sqrt(q2)))
    full = np.zeros((L, L, q, q), dtype='<f4')
    i, j = np.triu_indices(L, k=1)
    blk = couplings.reshape(nP, q, q)
    full[i, j] = blk
    full[j, i] = blk.transpose(0, 2, 1)
    return full.reshape(L * L, q * q)


def _parse_gpuspec(gpuspec):
    if gpuspec is None:
        return list(range(_lib.device_count()))
    devs = []
    for a in str(gpuspec).split(','):
        try:
            devs.append(int(a.split(':')[-1]))
        except ValueError:
            raise Exception("Error: GPU specification must be comma separated "
                            " list of platform:gpus, eg '0:0,0:1'")
    return devs


def setup_GPU_context(scriptpath, scriptfile, param, log):
    """Select devices (mcmcGPU.py:1038-1100).  Nothing is compiled at run
    time: the kernels are prebuilt gfx950 code objects in libmi3hip.so and
    take L and q as arguments.  Returns (context info, device list,
    (build description, build log))."""
    devs = _parse_gpuspec(param.gpuspec)
    n = _lib.device_count()
    for d in devs:
        if d < 0 or d >= n:
            raise Exception("No GPU with specification {}".format(d))
    if len(devs) == 0:
        raise Exception("Error: No GPUs found")
    info = lib().mi3_version().decode()
    log("Using prebuilt kernels: " + info)
    return (None, None), devs, (info, "")


def initGPU(devnum, cldat, device, nwalkers, param, log):
    """Create the per-device wrapper (mcmcGPU.py:1102-1119)."""
    wgsize = param.wgsize
    if wgsize not in [1 << n for n in range(32)]:
        raise Exception("wgsize must be a power of two")
    gpu = MCMCGPU((device, devnum, None, None), param.L, param.q, nwalkers, wgsize,
                  param.outdir, 1024, param.rngseed, profile=param.profile)
    gpu.beta = param.beta
    return gpu


def wgsize_heuristic(q, wgsize='auto'):
    """Work-group size option of the reference (mcmcGPU.py:1121-1132); the
    HIP kernels size their own launches, the value is only validated."""
    if wgsize == 'auto':
        w = 256
        while w < q * q:
            w *= 2
        return w
    wgsize = int(wgsize)
    if wgsize < q * q:
        raise Exception("Must have wgsize >= q*q, but got {} < {}".format(wgsize, q * q))
    return wgsize


def histogram_heuristic(q):
    """(NHIST, HISTWS) of the reference's histogram kernels
    (mcmcGPU.py:1135-1163), kept for API compatibility."""
    nhist = 4096 // (q * q)
    if nhist == 0:
        raise Exception("alphabet size too large to make histogram on gpu")
    nhist = 2 ** int(np.log2(nhist))
    hist_ws = 512 if q <= 12 else (256 if q <= 16 else 128)
    return min(nhist, hist_ws), hist_ws
