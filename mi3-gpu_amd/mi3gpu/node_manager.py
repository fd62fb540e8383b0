"""Device-group interface of the inference driver (node_manager.py of the
reference).  `GPU_node` keeps the reference's method surface, which is what
NewtonSteps and Mi3 call; on this build a node holds the devices of one host
and every method fans out to them in order.  With one device (the supported
configuration) merge_bimarg reduces to renormalisation
(node_manager.py:283-287)."""
import numpy as np


def sumarr(arrlist):
    tot = arrlist[0].copy()
    for a in arrlist[1:]:
        np.add(tot, a, tot)
    return tot


def meanarr(arrlist):
    return sumarr(arrlist) / len(arrlist)


def skip_single(f):
    def skipper(x):
        return x[0] if len(x) == 1 else f(x)
    return skipper


class GPU_node:
    groupfunc = {
        'weights': skip_single(np.concatenate),
        'E': skip_single(np.concatenate),
        'Bs': skip_single(np.concatenate),
        'bicount': skip_single(sumarr),
        'bi': skip_single(meanarr),
        'seq': skip_single(np.concatenate),
    }

    def __init__(self, gpus):
        self.gpus = gpus
        self.head_node = self

    @property
    def head_gpu(self):
        return type(self)([self.gpus[0]])

    @property
    def nwalkers(self):
        return sum(g.nwalkers for g in self.gpus)

    @property
    def nseq(self):
        ret = {'main': sum(g.nseq['main'] for g in self.gpus)}
        if 'large' in self.gpus[0].nseq:
            ret['large'] = sum(g.nseq.get('large', 0) for g in self.gpus)
        return ret

    @property
    def ngpus(self):
        return len(self.gpus)

    @property
    def gpu_list(self):
        return ["({}) device {}".format(g.gpunum, g.device) for g in self.gpus]

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
            f = self.groupfunc.get(bufs.split()[0], lambda x: x)
            return f(self.readBufs(bufs))
        return [self.groupfunc.get(bn.split()[0], lambda x: x)(arrs)
                for bn, arrs in zip(bufs, self.readBufs(bufs))]

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
        bi = sumarr([g.getBuf('bi').read() for g in self.gpus])
        bi = (bi / bi.sum(axis=1, keepdims=True)).astype('<f4')
        for g in self.gpus:
            g.setBuf('bi', bi)

    def wait(self):
        self._each('wait')

    def logProfile(self):
        self._each('logProfile')
