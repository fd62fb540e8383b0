# This file keeps the interface and behaviour of the file of the same name
# in Mi3-GPU, Copyright 2020 Allan Haldane, and is a derived work of it.
# Mi3-GPU is free software: you can redistribute it and/or modify it under
# the terms of the GNU General Public License as published by the Free
# Software Foundation, version 3 of the License; this file is distributed
# under the same licence, WITHOUT ANY WARRANTY; without even the implied
# warranty of MERCHANTABILITY or FITNESS FOR A PARTICULAR PURPOSE.  See the
# GNU General Public License <http://www.gnu.org/licenses/> for details.
"""Device-group interface of the inference driver (node_manager.py of the
reference).  `GPU_node` keeps the reference's method surface, which is what
NewtonSteps and Mi3 call; on this build a node holds the devices of one host
and every method fans out to them in order.  With one device (the supported
configuration) merge_bimarg reduces to renormalisation
(node_manager.py:283-287)."""
import numpy as np


# How the per-device copies of a buffer become one host array, keyed by the
# first word of the buffer name (node_manager.py:168-196): per-walker
# buffers are stacked in device order, counts add up and bimarginals are
# averaged.  A name not listed here comes back as the list of device arrays.
def _stacked(parts):
    return np.concatenate(parts)


def _added(parts):
    out = np.array(parts[0], copy=True)
    for p in parts[1:]:
        out += p
    return out


def _averaged(parts):
    return _added(parts) / len(parts)


_COMBINE = {'seq': _stacked, 'E': _stacked, 'weights': _stacked, 'Bs': _stacked,
            'bicount': _added, 'bi': _averaged}


def combine(bufname, parts):
    """One host array from the per-device arrays `parts` of `bufname`."""
    fn = _COMBINE.get(bufname.split()[0])
    if fn is None:
        return parts
    return parts[0] if len(parts) == 1 else fn(parts)


class GPU_node:
    """The devices of one host behind one interface; every method runs on
    each device in order."""

    def __init__(self, gpus):
        self.gpus = list(gpus)
        self.head_node = self

    @property
    def head_gpu(self):
        """A node made of the first device alone."""
        return GPU_node(self.gpus[:1])

    @property
    def ngpus(self):
        return len(self.gpus)

    @property
    def nwalkers(self):
        return int(np.sum([g.nwalkers for g in self.gpus]))

    @property
    def nseq(self):
        """Sequences per buffer kind, over all devices."""
        kinds = ['main'] + (['large'] if 'large' in self.gpus[0].nseq else [])
        return {k: int(np.sum([g.nseq.get(k, 0) for g in self.gpus])) for k in kinds}

    @property
    def gpu_list(self):
        return [f"({g.gpunum}) device {g.device}" for g in self.gpus]

    def _each(self, meth, *args):
        for g in self.gpus:
            getattr(g, meth)(*args)

    def initMCMC(self, nsteps):
        """Disjoint walker-stream ranges per device (node_manager.py:73-85)."""
        span = np.uint64(2 ** 63) // np.uint64(self.ngpus)
        for n, g in enumerate(self.gpus):
            g.initMCMC(nsteps, int(np.uint64(n) * span), int(span))

    def initLargeBufs(self, nseq):
        if isinstance(nseq, list):
            for n, g in zip(nseq, self.gpus):
                g.initLargeBufs(n)
        else:
            self._each('initLargeBufs', nseq)

    def initJstep(self):
        self._each('initJstep')

    def initSubseq(self):
        self._each('initSubseq')

    def prepare_indep(self, unimarg):
        self._each('prepare_indep', unimarg)

    def gen_indep(self, bufname):
        self._each('gen_indep', bufname)

    def runMCMC(self):
        self._each('runMCMC')

    def calcEnergies(self, seqbufname='main', Jbufname='J'):
        self._each('calcEnergies', seqbufname, Jbufname)

    def calcBicounts(self, seqbufname):
        self._each('calcBicounts', seqbufname)

    def bicounts_to_bimarg(self, seqbufname='main'):
        self._each('bicounts_to_bimarg', seqbufname)

    def updateJ(self, gamma, pc, Jbuf='dJ'):
        self._each('updateJ', gamma, pc, Jbuf)

    def reg(self, name, param):
        self._each('reg_{}'.format(name), *param)

    def min_buf(self, buf):
        self._each('min_buf', buf)

    def weight_statistics(self, buf='main'):
        self._each('weight_statistics', buf)

    def dE_to_weights(self, buf='main', offset=0.):
        # positional order of NewtonSteps' call: (seqbuf, ref_dE)
        self._each('dE_to_weights', buf, offset)

    def fixed_beta_weights(self, ref_E, seqbufname='main'):
        self._each('fixed_beta_weights', ref_E, seqbufname)

    def weightedMarg(self, seqbufname):
        self._each('weightedMarg', seqbufname)

    def addFloatBuf(self, dstname, srcname):
        self._each('addFloatBuf', dstname, srcname)

    def renormalize_bimarg(self):
        self._each('renormalize_bimarg')

    def getBuf(self, bufname):
        return [g.getBuf(bufname) for g in self.gpus]

    def readBufs(self, bufnames):
        if isinstance(bufnames, str):
            return [b.read() for b in self.getBuf(bufnames)]
        futures = [self.getBuf(bn) for bn in bufnames]
        return [[f.read() for f in fl] for fl in futures]

    def collect(self, bufs):
        """Read buffers from every device and combine them per buffer type
        (node_manager.py:168-196)."""
        if isinstance(bufs, str):
            return combine(bufs, self.readBufs(bufs))
        return [combine(bn, arrs) for bn, arrs in zip(bufs, self.readBufs(bufs))]

    def setBuf(self, bufname, dat):
        if isinstance(dat, list):
            for g, b in zip(self.gpus, dat):
                g.setBuf(bufname, b)
        else:
            self._each('setBuf', bufname, dat)

    def fillBuf(self, bufname, val):
        self._each('fillBuf', bufname, val)

    def setSeqs(self, bufname, seqs, log=None):
        """Split sequences across devices (node_manager.py:210-232)."""
        if isinstance(seqs, np.ndarray) or len(seqs) == 1:
            if not isinstance(seqs, np.ndarray):
                seqs = seqs[0]
            nbuf = self.nseq[bufname]
            if bufname == 'main' and seqs.shape[0] != nbuf:
                raise Exception("Expected {} total sequences, got {}".format(nbuf, seqs.shape[0]))
            seqs = np.array_split(seqs, self.ngpus)
        elif len(seqs) != self.ngpus:
            raise Exception("Expected {} sequence bufs, got {}".format(self.ngpus, len(seqs)))
        if log:
            log("Transferring {} seqs to gpu's {} seq buffer...".format(
                str([len(s) for s in seqs]), bufname))
        for g, s in zip(self.gpus, seqs):
            g.setBuf('seq ' + bufname, s)

    def fillSeqs(self, seq):
        self._each('fillSeqs', seq)

    def markPos(self, marks):
        self._each('markPos', marks)

    def storeSeqs(self, seqs=None):
        self._each('storeSeqs', seqs)

    def clearLargeSeqs(self):
        self._each('clearLargeSeqs')

    def copySubseq(self, seqind):
        self._each('copySubseq', seqind)

    def merge_bimarg(self):
        """Combine the devices' weighted bimarginals and renormalise
        (node_manager.py:283-304).  Several devices of one host are reduced
        through the host (the reference's device-to-device adds need a shared
        OpenCL context)."""
        if len(self.gpus) == 1:
            self.gpus[0].renormalize_bimarg()
            return
        bi = _added([g.getBuf('bi').read() for g in self.gpus])
        bi = (bi / bi.sum(axis=1, keepdims=True)).astype('<f4')
        for g in self.gpus:
            g.setBuf('bi', bi)

    def wait(self):
        self._each('wait')

    def logProfile(self):
        self._each('logProfile')
