"""ctypes wrapper of the float32 OpenMP CPU baseline (``oracle/cpu_tracer.cpp``):
the reference's tracer, OSEM subset update and histogram loop on the host
cores.

Test infrastructure only: timed by ``bench.py``'s cpu_baseline leg and
checked by ``tests/``, never imported by the ``nipet_amd`` package.
"""
import ctypes as ct
import os
import platform

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
NTV, NTT = 1807, 10
NVOX = 320 * 320 * 127


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, 'libcputracer.so')
        if not os.path.exists(path):
            import subprocess
            subprocess.run(['make', '-s', '-C', _HERE], check=True)
        L = ct.CDLL(path)
        P, I, I64 = ct.c_void_p, ct.c_int, ct.c_int64
        geo = [P, P, P, I, P, P, P, I, I]   # rec tv subs nprj li2rng li2sn li2nos nli nsn
        L.cb_threads.restype = I
        L.cb_tx.argtypes = [P, P, I, P, P]
        L.cb_tx.restype = I
        L.cb_check.argtypes = geo
        L.cb_check.restype = I64
        L.cb_fprj.argtypes = [P, P] + geo + [I]
        L.cb_fprj.restype = I
        L.cb_bprj.argtypes = [P, P] + geo + [P]
        L.cb_bprj.restype = I
        L.cb_osem_subset.argtypes = [P, P, P, P, P] + geo + [P, P, P]
        L.cb_osem_subset.restype = I
        L.cb_time_pair.argtypes = [P] + geo + [P, P, P, I, P, P]
        L.cb_time_pair.restype = I
        L.cb_hst.argtypes = [P, I64, I, I, I, I] + [P] * 13
        L.cb_hst.restype = I64
        L.cb_time_hst.argtypes = [P, I64, I, I] + [P] * 13 + [I, P]
        L.cb_time_hst.restype = I64
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


class Tracer:
    """The reference's float32 transaxial tables (rec, step bytes) of all bins
    and the span-11 axial tables of ``params``."""

    def __init__(self, params):
        tx, ax, Cnt = params['txLUT'], params['axLUT'], params['Cnt']
        self.crs = np.ascontiguousarray(tx['crs'], np.float32)
        self.s2c = np.ascontiguousarray(tx['s2c'], np.int16)
        self.naw = len(self.s2c)
        self.li2rng = np.ascontiguousarray(ax['li2rng'], np.float32)
        self.li2sn = np.ascontiguousarray(ax['li2sn'], np.int16)
        self.li2nos = np.ascontiguousarray(ax['li2nos'], np.int8)
        self.nli = len(self.li2rng)
        self.nsn = Cnt['NSN11']
        self.rec = np.zeros((self.naw, NTT), np.float32)
        self.tv = np.zeros((self.naw, NTV), np.uint8)
        bad = lib().cb_tx(_p(self.crs), _p(self.s2c), self.naw, _p(self.rec), _p(self.tv))
        assert bad == 0, f'{bad} bins overflow the step table'
        self._work = None

    def _geo(self, subs):
        subs = np.ascontiguousarray(subs, np.int32)
        return subs, [_p(self.rec), _p(self.tv), _p(subs), len(subs), _p(self.li2rng),
                      _p(self.li2sn), _p(self.li2nos), self.nli, self.nsn]

    def work(self):
        if self._work is None:
            self._work = np.empty((lib().cb_threads(), NVOX), np.float32)
        return self._work

    def check(self, subs):
        """Out-of-image accesses the walks of bins ``subs`` would make."""
        subs, g = self._geo(subs)
        return int(lib().cb_check(*g))

    def fprj(self, img, subs, att=False):
        img = np.ascontiguousarray(img, np.float32)
        assert img.size == NVOX
        subs, g = self._geo(subs)
        out = np.empty((len(subs), self.nsn), np.float32)
        lib().cb_fprj(_p(out), _p(img), *g, int(att))
        return out

    def bprj(self, sino, subs):
        subs, g = self._geo(subs)
        sino = np.ascontiguousarray(sino, np.float32)
        assert sino.shape == (len(subs), self.nsn)
        out = np.empty((320, 320, 127), np.float32)
        lib().cb_bprj(_p(out), _p(sino), *g, _p(self.work()))
        return out

    def osem_subset(self, img, prompts, rho, sens, mask, subs):
        """One subset update of ``img`` (float32, updated in place)."""
        subs, g = self._geo(subs)
        prompts = np.ascontiguousarray(prompts, np.uint16)
        rho = np.ascontiguousarray(rho, np.float32)
        sens = np.ascontiguousarray(sens, np.float32)
        mask = np.ascontiguousarray(mask, np.uint8)
        assert img.dtype == np.float32 and img.flags.c_contiguous and img.size == NVOX
        assert prompts.shape == rho.shape == (self.naw, self.nsn)
        es = np.empty((len(subs), self.nsn), np.float32)
        bi = np.empty(NVOX, np.float32)
        lib().cb_osem_subset(_p(img), _p(prompts), _p(rho), _p(sens), _p(mask), *g, _p(es),
                             _p(bi), _p(self.work()))
        return img

    def time_pair(self, img, subs, runs=3):
        """Wall times [s] of `runs` FP + BP pairs over bins ``subs`` after one
        warm-up pair: (fp_s, bp_s)."""
        subs, g = self._geo(subs)
        img = np.ascontiguousarray(img, np.float32)
        sino = np.empty((len(subs), self.nsn), np.float32)
        bimg = np.empty(NVOX, np.float32)
        fp, bp = np.zeros(runs), np.zeros(runs)
        lib().cb_time_pair(_p(img), *g, _p(sino), _p(bimg), _p(self.work()), runs, _p(fp),
                           _p(bp))
        return fp, bp


def _hst_out(nitag):
    tn = (min(nitag, 5400) + 3) // 4
    return dict(psino=np.zeros(837 * 86688, np.uint32), ssrb=np.zeros(127 * 86688, np.uint32),
                rprmt=np.zeros(nitag, np.uint32), rdlyd=np.zeros(nitag, np.uint32),
                zR=np.zeros(nitag, np.uint64), zM=np.zeros(nitag, np.uint64),
                snview=np.zeros(tn * 127 * 344, np.uint32), fansums=np.zeros(64 * 504, np.uint32),
                bucks=np.zeros(2 * nitag * 224, np.uint32))


_HST_KEYS = ('psino', 'ssrb', 'rprmt', 'rdlyd', 'zR', 'zM', 'snview', 'fansums', 'bucks')


def _hst_luts(params):
    ax, tx = params['axLUT'], params['txLUT']
    return [np.ascontiguousarray(ax['sn1_sn11'], np.int16),
            np.ascontiguousarray(ax['sn1_ssrb'], np.int16),
            np.ascontiguousarray(ax['sn1_rno'], np.int16),
            np.ascontiguousarray(tx['s2cF'], np.int16)]


def hst(words, params, toff, nitag, tstart=0, tstop=None):
    """The reference's histogram loop (no bootstrap) over ``words``; returns
    the outputs (span-11 sinogram with delayeds in the upper 16 bits) and the
    number of counted events."""
    words = np.ascontiguousarray(words, np.int32)
    tstop = nitag if tstop is None else tstop
    out = _hst_out(nitag)
    luts = _hst_luts(params)
    n = lib().cb_hst(_p(words), len(words), toff, nitag, tstart, tstop,
                     *[_p(a) for a in luts], *[_p(out[k]) for k in _HST_KEYS])
    out['counted'] = int(n)
    return out


def time_hst(words, params, toff, nitag, runs=3):
    """Wall times [s] of `runs` histograms of ``words`` (outputs zeroed inside
    the timed region) after one warm-up; (secs, counted events)."""
    words = np.ascontiguousarray(words, np.int32)
    out = _hst_out(nitag)
    luts = _hst_luts(params)
    secs = np.zeros(runs)
    n = lib().cb_time_hst(_p(words), len(words), toff, nitag, *[_p(a) for a in luts],
                          *[_p(out[k]) for k in _HST_KEYS], runs, _p(secs))
    return secs, int(n)


def host_info():
    """Host cores and CPU model (nproc, the model name of /proc/cpuinfo, the
    OpenMP threads the baseline uses)."""
    model = platform.processor() or ''
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    return {'nproc': os.cpu_count(), 'affinity': len(os.sched_getaffinity(0)),
            'cpu_model': model, 'omp_threads': int(lib().cb_threads())}
