"""List-mode histogramming (``niftypet/nipet/lm/mmrhist.py``) on the GPU.

* ``lminfo``  -- ``mmr_lmproc.lminfo`` (``lm/src/lm_module.cu:68-171``): number of
  words, first/last time tag and their positions.
* ``mmrhist`` / ``hist`` -- ``lm/mmrhist.py:25-180``: histogram a frame of a
  list-mode file into span-11 prompt/delayed sinograms, SSRB sinogram, head
  curves, centre of mass, projection views, fan-sums and bucket singles, with
  optional bootstrap from a reproducible counter-based generator:
  non-parametric (``Cnt['BTP'] = 1``, each event redrawn from the events of
  its 8192-word patch) or parametric (``Cnt['BTP'] = 2``, Poisson weights of
  mean ``Cnt['BTPRT']``).
* ``LmHistogram`` -- the device session behind both: words are handed over in
  chunks (device tensors), the time-tag state carries across chunks.
* ``randoms`` / ``rand`` -- ``lm/mmrhist.py:188-221``: randoms sinogram and
  crystal singles map estimated from the delayed fan-sums
  (``Randoms``, the device session holding the crystal LUTs).

The file is streamed from a memory map through two pinned host buffers: the
copy of chunk k+1 runs on a copy stream while chunk k is histogrammed.
"""
import ctypes as ct
import logging
import os

import numpy as np

from .. import _lib

log = logging.getLogger(__name__)

__all__ = ['lminfo', 'mmrhist', 'hist', 'LmHistogram', 'hist_words', 'tag_range',
           'second_index', 'Randoms', 'randoms', 'rand', 'dynamic_timings']

CHUNK_WORDS = 1 << 26      # 256 MB of words per transfer


def _torch():
    import torch
    return torch


def _tag_positions(mm):
    """First / last time tag (ms) and their word positions in an int32 array."""
    step = 1 << 20
    n = len(mm)
    first = last = None
    for s in range(0, n, step):
        blk = np.asarray(mm[s:s + step])
        t = np.nonzero((blk >> 29) == -4)[0]
        if len(t):
            first = (int(blk[t[0]] & 0x1fffffff), s + int(t[0]))
            break
    for e in range(n, 0, -step):
        s = max(0, e - step)
        blk = np.asarray(mm[s:e])
        t = np.nonzero((blk >> 29) == -4)[0]
        if len(t):
            last = (int(blk[t[-1]] & 0x1fffffff), s + int(t[-1]))
            break
    if first is None:
        raise ValueError('no time tag in the list-mode data')
    return first, last


def lminfo(flm):
    """(number of words, (first tag ms, last tag ms), (first tag position,
    last tag position)) as ``mmr_lmproc.lminfo``; positions are 1-based from
    the start and 0-based from the end, like the reference."""
    mm = np.memmap(flm, dtype=np.int32, mode='r')
    (t0, a0), (t1, a1) = _tag_positions(mm)
    return len(mm), (t0, t1), (a0 + 1, a1)


class LmHistogram:
    """Device histogram of one frame of an mMR list-mode word stream."""

    def __init__(self, scanner_params, toff, nitag, tstart=0, tstop=None, span=None, btp=None,
                 btprt=None, seed=0, nrep=1, rep0=0):
        torch = _torch()
        Cnt, tx, ax = scanner_params['Cnt'], scanner_params['txLUT'], scanner_params['axLUT']
        self.span = Cnt['SPN'] if span is None else span
        self.btp = Cnt.get('BTP', 0) if btp is None else btp
        self.btprt = Cnt.get('BTPRT', 1.0) if btprt is None else btprt
        self.nitag, self.nrep = int(nitag), int(nrep)
        tstop = self.nitag if tstop is None else tstop
        if not 0 <= tstart <= tstop <= self.nitag:
            raise ValueError(f'frame [{tstart}, {tstop}) s outside the acquisition '
                             f'[0, {self.nitag}) s')
        if self.btp not in (0, 1, 2):
            raise ValueError(f'bootstrap type {self.btp} (0: none, 1: non-parametric, '
                             '2: parametric)')
        self.nsinos = {11: Cnt['NSN11'], 1: Cnt['NSN1'], 0: Cnt['NSEG0']}[self.span]
        self._keep = [np.ascontiguousarray(ax['sn1_sn11'], np.int16),
                      np.ascontiguousarray(ax['sn1_ssrb'], np.int16),
                      np.ascontiguousarray(ax['sn1_rno'], np.int16),
                      np.ascontiguousarray(tx['s2cF'], np.int16)]
        cfg = _lib.NipetLmConfig(int(toff), self.nitag, int(tstart), int(tstop), self.span,
                                 int(self.btp), float(self.btprt), int(seed), self.nrep,
                                 int(rep0), *[k.ctypes.data for k in self._keep])
        h = ct.c_void_p()
        _lib.check(_lib.lib().nipet_lm_create(ct.byref(cfg), ct.byref(h)), 'nipet_lm_create')
        self.handle = h
        z = dict(dtype=torch.int32, device='cuda')
        nb = self.nsinos * 86688
        self.tn = (min(self.nitag, 5400) + 3) // 4
        self.out = {
            'psino': torch.zeros((nb, self.nrep), **z), 'dsino': torch.zeros((nb, self.nrep), **z),
            'ssr': torch.zeros((127, 252, 344), **z),
            'phc': torch.zeros((self.nitag, self.nrep), **z),
            'dhc': torch.zeros((self.nitag, self.nrep), **z),
            'zR': torch.zeros(self.nitag, dtype=torch.int64, device='cuda'),
            'zM': torch.zeros(self.nitag, dtype=torch.int64, device='cuda'),
            'pvs': torch.zeros((self.tn, 127, 344), **z), 'fan': torch.zeros((64, 504), **z),
            'bck': torch.zeros((2, self.nitag, 224), **z)}
        self._o = _lib.NipetLmOut(*[self.out[k].data_ptr() for k in (
            'psino', 'dsino', 'ssr', 'phc', 'dhc', 'zR', 'zM', 'pvs', 'fan', 'bck')])

    def __del__(self):
        h = getattr(self, 'handle', None)
        # at interpreter exit the module globals may already be gone
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.nipet_lm_destroy(h)
            self.handle = None

    def add(self, words):
        """Histogram the next chunk (CUDA int32 tensor) on the current stream."""
        torch = _torch()
        if words.dtype != torch.int32 or not words.is_cuda:
            raise TypeError('words must be a CUDA int32 tensor')
        words = words.contiguous()
        _lib.check(_lib.lib().nipet_lm_chunk(self.handle, words.data_ptr(), words.numel(),
                                             ct.byref(self._o), _lib.stream_handle()),
                   'nipet_lm_chunk')

    def finish(self):
        """Derive the SSRB prompt sinogram; call after the last chunk."""
        _lib.check(_lib.lib().nipet_lm_finish(self.handle, ct.byref(self._o),
                                              _lib.stream_handle()), 'nipet_lm_finish')

    def status(self):
        n, bad, ovf = ct.c_int64(), ct.c_uint(), ct.c_uint()
        _lib.check(_lib.lib().nipet_lm_status(self.handle, ct.byref(n), ct.byref(bad),
                                              ct.byref(ovf)), 'nipet_lm_status')
        return n.value, bad.value, ovf.value


def tag_range(words):
    """(first, last) time-tag values [ms] of a word array (CUDA tensor: found
    on the device; else NumPy)."""
    torch = _torch()
    if isinstance(words, torch.Tensor) and words.is_cuda:
        w = words.contiguous()
        span = torch.empty(2, dtype=torch.int64, device='cuda')
        _lib.check(_lib.lib().nipet_lm_tag_span(_lib.ptr(w), w.numel(), _lib.ptr(span),
                                                _lib.stream_handle()), 'nipet_lm_tag_span')
        i0, i1 = (int(v) for v in span.cpu().numpy())
        if i1 == 0:
            raise ValueError('no time tags in the list-mode data')
        return int(w[i0]) & 0x1fffffff, int(w[i1 - 1]) & 0x1fffffff
    w = words.numpy() if isinstance(words, torch.Tensor) else np.asarray(words)
    (tf, _), (tl, _) = _tag_positions(w)
    return tf, tl


def second_index(words, toff, nsec):
    """Host int64 array [nsec + 1]: word index of the first time tag of each
    second s = 0 .. nsec after toff in a device word array (the number of
    words where none follows), found on the device."""
    torch = _torch()
    w = words.contiguous()
    first = torch.empty(nsec + 1, dtype=torch.int64, device='cuda')
    _lib.check(_lib.lib().nipet_lm_second_index(_lib.ptr(w), w.numel(), int(toff), int(nsec),
                                                _lib.ptr(first), _lib.stream_handle()),
               'nipet_lm_second_index')
    return first.cpu().numpy()


def hist_words(words, scanner_params, t0=0, t1=0, btp=None, seed=0, nrep=1, span=None, rep0=0):
    """Histogram a device (or host) int32 word array that holds a complete
    acquisition. Returns the ``LmHistogram`` (device outputs in ``.out``)."""
    torch = _torch()
    w = words if isinstance(words, torch.Tensor) else torch.from_numpy(np.asarray(words))
    tf, tl = tag_range(w)
    nitag = ((tl - tf) + 999) // 1000
    if t0 == t1:
        t0, t1 = 0, nitag
    h = LmHistogram(scanner_params, tf, nitag, t0, min(t1, nitag), span=span, btp=btp, seed=seed,
                    nrep=nrep, rep0=rep0)
    h.add(w.to('cuda', torch.int32))
    h.finish()
    return h


def _stream_file(h, flm, offset_words=0):
    """Copy the file through two pinned buffers into the device histogram."""
    torch = _torch()
    mm = np.memmap(flm, dtype=np.int32, mode='r')
    n = len(mm)
    size = min(CHUNK_WORDS, max(n - offset_words, 1))
    pin = [torch.empty(size, dtype=torch.int32).pin_memory() for _ in range(2)]
    dev = [torch.empty(size, dtype=torch.int32, device='cuda') for _ in range(2)]
    copy = torch.cuda.Stream()
    main = torch.cuda.current_stream()
    done = [torch.cuda.Event() for _ in range(2)]
    k = 0
    for s in range(offset_words, n, size):
        e = min(n, s + size)
        b = k % 2
        done[b].synchronize()            # the kernel that read dev[b] has finished
        pin[b][:e - s].numpy()[:] = mm[s:e]
        with torch.cuda.stream(copy):
            dev[b][:e - s].copy_(pin[b][:e - s], non_blocking=True)
        main.wait_stream(copy)
        h.add(dev[b][:e - s])
        done[b].record(main)
        k += 1
    h.finish()
    torch.cuda.synchronize()


def hist(datain, txLUT, axLUT, Cnt, t0=0, t1=0, cmass_sig=5, frms=None, use_stored=False,
         store=False, outpath=''):
    """Histogram frame [t0, t1) s of ``datain['lm_bf']`` (``lm/mmrhist.py:40-180``)."""
    import scipy.ndimage as ndi
    if use_stored and 'sinos' in datain and os.path.basename(datain['sinos']) == \
            f"sinos_s{Cnt['SPN']}_frm-{t0}-{t1}.npz":
        hstout = dict(np.load(datain['sinos'], allow_pickle=False))
        nitag = len(hstout['phc'])
    elif os.path.isfile(datain.get('lm_bf', '')):
        nele, ttags, _ = lminfo(datain['lm_bf'])
        nitag = int((ttags[1] - ttags[0] + 999) / 1000)
        if t1 > nitag:
            t1 = nitag
        if t0 >= nitag:
            raise ValueError('the time frame definition is not allowed '
                             '(outside acquisition time)')
        tstart, tstop = (0, nitag) if t0 == t1 else (t0, t1)
        h = LmHistogram({'Cnt': Cnt, 'txLUT': txLUT, 'axLUT': axLUT}, ttags[0], nitag,
                        tstart, tstop, span=Cnt['SPN'], btp=Cnt.get('BTP', 0),
                        btprt=Cnt.get('BTPRT', 1.0), seed=Cnt.get('SEED', 0))
        _stream_file(h, datain['lm_bf'], Cnt.get('LMOFF', 0) // 4)
        hstout = _host_outputs(h, Cnt)
        if store:
            fsino = os.path.dirname(datain['lm_bf']) if outpath == '' else \
                os.path.join(outpath, 'sino')
            os.makedirs(fsino, exist_ok=True)
            np.savez(os.path.join(fsino, f"sinos_s{Cnt['SPN']}_frm-{t0}-{t1}.npz"), **hstout)
    else:
        raise ValueError('input list-mode data is not defined')
    return _finish(hstout, Cnt, nitag, t0, t1, cmass_sig, ndi)


def _host_outputs(h, Cnt, rep=0):
    """The reference's raw output dictionary (``lm/mmrhist.py:93-110``) on the host."""
    o = {k: v.cpu().numpy() for k, v in h.out.items()}
    shp = (h.nsinos, Cnt['NSANGLES'], Cnt['NSBINS'])
    zR, zM = o['zR'].astype(np.float64), o['zM'].astype(np.float64)
    with np.errstate(divide='ignore', invalid='ignore'):
        mss = (zR / zM).astype(np.float32)
    return {
        'phc': o['phc'][:, rep].astype(np.uint32), 'dhc': o['dhc'][:, rep].astype(np.uint32),
        'mss': mss, 'pvs': o['pvs'].astype(np.uint32), 'bck': o['bck'].astype(np.uint32),
        'fan': o['fan'].astype(np.uint32),
        'psn': (o['psino'][:, rep].astype(np.uint32) & 0xffff).astype(np.uint16).reshape(shp),
        'dsn': (o['dsino'][:, rep].astype(np.uint32) & 0xffff).astype(np.uint16).reshape(shp),
        'ssr': o['ssr'].astype(np.uint32)}


def _finish(hstout, Cnt, nitag, t0, t1, cmass_sig, ndi):
    """Post-processing of ``lm/mmrhist.py:131-180``."""
    pvs_sgtl = np.right_shift(hstout['pvs'], 8).astype(np.float32)
    pvs_crnl = np.bitwise_and(hstout['pvs'], 255).astype(np.float32)
    cmass = Cnt['SO_VXZ'] * ndi.gaussian_filter(hstout['mss'], cmass_sig, mode='mirror')
    bck = hstout['bck'].copy()
    nsr = bck[1] >> 30
    bck[0, nsr > 0] = bck[0, nsr > 0] / nsr[nsr > 0]
    tmsk = np.sum(nsr, axis=1) > 0
    single_rate = np.copy(bck[0, tmsk, :])
    t = np.arange(nitag)[tmsk]
    buckets = np.int32(np.sum(single_rate, axis=0) / max(single_rate.shape[0], 1))
    if t0 == t1:
        t1 = t0 + nitag
    return {
        't0': t0, 't1': t1, 'dur': t1 - t0, 'phc': hstout['phc'], 'dhc': hstout['dhc'],
        'cmass': cmass, 'pvs_sgtl': pvs_sgtl, 'pvs_crnl': pvs_crnl, 'fansums': hstout['fan'],
        'sngl_rate': single_rate, 'tsngl': t, 'buckets': buckets,
        'psino': hstout['psn'].astype(np.uint16), 'dsino': hstout['dsn'].astype(np.uint16),
        'pssr': hstout['ssr']}


def mmrhist(datain, scanner_params, t0=0, t1=0, outpath='', frms=None, use_stored=False,
            store=False, cmass_sig=5):
    """Histogram, head curves and centre of mass of a list-mode frame
    (``lm/mmrhist.py:25-37``)."""
    Cnt, txLUT, axLUT = scanner_params['Cnt'], scanner_params['txLUT'], scanner_params['axLUT']
    return hist(datain, txLUT, axLUT, Cnt, frms=frms, use_stored=use_stored, store=store,
                outpath=outpath, t0=t0, t1=t1, cmass_sig=cmass_sig)


def dynamic_timings(flist, offset=0):
    """Start and end times of dynamic frames (``lm/mmrhist.py:492-556``).
    ``flist`` is a list of frame durations [15, 15, 30, ...] or ``['def',
    [repetitions, duration], ...]``; ``offset`` adds a first frame [0, offset]
    (durations list) or a zero-repetition entry (``'def'`` list), as the
    reference does. Returns {'total': end time, 'frames': uint16 durations,
    'timings': ['timings', [t0, t1], ...]}."""
    from collections.abc import Collection
    from numbers import Integral
    if not isinstance(flist, Collection) or isinstance(flist, str):
        raise TypeError('Wrong type of frame data input')
    if all(isinstance(t, Integral) for t in flist):
        tsum = offset
        t_frames = ['timings', [0, offset]] if offset > 0 else ['timings']
        for d in flist:
            t_frames.append([tsum, tsum + d])
            tsum += d
        frms = np.uint16(flist)
    elif flist[0] == 'def' and all(isinstance(t, Collection) and len(t) == 2
                                   for t in flist[1:]):
        defs = [list(t) for t in flist[1:]]
        if offset > 0:
            defs.insert(0, [0, offset])
        farray = np.asarray(defs, dtype=np.int64)
        frms = np.repeat(farray[:, 1], farray[:, 0]).astype(np.uint16)
        t_frames, tsum = ['timings'], 0
        for d in frms:
            t_frames.append([tsum, tsum + int(d)])
            tsum += int(d)
    else:
        raise TypeError('Unrecognised time frame definitions.')
    return {'total': tsum, 'frames': frms, 'timings': t_frames}


# ------------------------------------------------------------------------------
# randoms from the delayed fan-sums
# ------------------------------------------------------------------------------

RND_NITER = 10     # fixed-point steps of gpu_randoms (lm/src/rnd.cu:339)


class Randoms:
    """Device session of the randoms estimate from delayed fan-sums
    (``mmr_lmproc.rand``, ``lm/src/rnd.cu:180-420``). The crystal
    coincidence matrix and sinogram LUTs are uploaded once; ``__call__`` takes
    device or host fan-sums (64, 504) and returns device tensors:
    the randoms sinogram in the device layout (Naw, nsinos) and the crystal
    singles map (504, 64), zero in the gap crystals."""

    def __init__(self, scanner_params, span=None):
        Cnt, tx, ax = scanner_params['Cnt'], scanner_params['txLUT'], scanner_params['axLUT']
        self.span = Cnt['SPN'] if span is None else span
        if self.span not in (1, 11):
            raise ValueError(f'randoms: span {self.span} is not supported (1 or 11)')
        self.nsinos = Cnt['NSN11'] if self.span == 11 else Cnt['NSN1']
        self.naw = int(tx['Naw'])
        self._keep = [np.ascontiguousarray(tx['cij'], np.int8),
                      np.ascontiguousarray(tx['crsri'], np.int16),
                      np.ascontiguousarray(tx['s2cr'], np.int16),
                      np.ascontiguousarray(ax['sn1_rno'], np.int16),
                      np.ascontiguousarray(ax['sn1_sn11'], np.int16)]
        d = _lib.NipetRndDesc(self.naw, self.nsinos, Cnt['NSN1'], Cnt['MRD'],
                              *[k.ctypes.data for k in self._keep])
        h = ct.c_void_p()
        _lib.check(_lib.lib().nipet_rnd_create(ct.byref(d), ct.byref(h)), 'nipet_rnd_create')
        self.handle = h

    def __del__(self):
        h = getattr(self, 'handle', None)
        # at interpreter exit the module globals may already be gone
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.nipet_rnd_destroy(h)
            self.handle = None

    def __call__(self, fansums, niter=RND_NITER, rsino=None, cmap=None):
        torch = _torch()
        f = fansums if isinstance(fansums, torch.Tensor) else torch.from_numpy(
            np.ascontiguousarray(np.asarray(fansums).astype(np.uint32).view(np.int32)))
        if tuple(f.shape) != (64, 504) or f.element_size() != 4:
            raise ValueError('fan-sums must be a (64, 504) array of 32-bit counts')
        f = f.to('cuda').contiguous()
        if rsino is None:
            rsino = torch.empty((self.naw, self.nsinos), dtype=torch.float32, device='cuda')
        if cmap is None:
            cmap = torch.empty((504, 64), dtype=torch.float32, device='cuda')
        if (tuple(rsino.shape) != (self.naw, self.nsinos) or rsino.dtype != torch.float32
                or not rsino.is_contiguous() or tuple(cmap.shape) != (504, 64)
                or cmap.dtype != torch.float32 or not cmap.is_contiguous()):
            raise ValueError('randoms: bad output tensors')
        _lib.check(_lib.lib().nipet_rnd_estimate(self.handle, f.data_ptr(), int(niter),
                                                 rsino.data_ptr(), cmap.data_ptr(),
                                                 _lib.stream_handle()), 'nipet_rnd_estimate')
        return rsino, cmap


def rand(fansums, txLUT, axLUT, Cnt):
    """Randoms sinogram (nsinos, 252, 344) with zero gaps and crystal singles
    map (504, 64) as float32 host arrays (``lm/mmrhist.py:210-221``)."""
    from ..mmraux import putgaps
    r = Randoms({'Cnt': Cnt, 'txLUT': txLUT, 'axLUT': axLUT})
    rs, cmap = r(fansums)
    return putgaps(rs, txLUT, Cnt).cpu().numpy(), cmap.cpu().numpy()


def randoms(hst, scanner_params, gpu_dim=False):
    """Estimated randoms from the delayed fan-sums of a histogram
    (``lm/mmrhist.py:188-207``): (randoms sinogram, singles map); with
    ``gpu_dim`` the sinogram is in the device layout (Naw, nsinos)."""
    Cnt, txLUT, axLUT = scanner_params['Cnt'], scanner_params['txLUT'], scanner_params['axLUT']
    if gpu_dim:
        rs, cmap = Randoms(scanner_params)(hst['fansums'])
        return rs.cpu().numpy(), cmap.cpu().numpy()
    return rand(hst['fansums'], txLUT, axLUT, Cnt)
