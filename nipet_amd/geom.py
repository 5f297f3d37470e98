"""Device-resident scanner geometry.

One ``NipetGeom`` per (scanner parameters, span, device) is built on first use
and cached: the FP64 transaxial intersection lists of all 68,516 transaxial
bins are computed once on the GPU and stay in HBM (the reference recomputes
and re-uploads its tracing tables on every projector call,
prj/src/prj_module.cu:380-398).
"""
import ctypes as ct

import numpy as np

from . import _lib

_CACHE = {}


class DeviceGeom:
    def __init__(self, scanner_params, device):
        Cnt = scanner_params['Cnt']
        tx, ax = scanner_params['txLUT'], scanner_params['axLUT']
        span = Cnt['SPN']
        if span == 11:
            li2sn, nsn = ax['li2sn'], Cnt['NSN11']
            li2nos = ax['li2nos']
        elif span == 1:
            li2sn = ax['li2sn1']
            nsn = Cnt['NSN1'] if Cnt['RNG_END'] - Cnt['RNG_STRT'] == Cnt['NRNG'] else \
                (Cnt['RNG_END'] - Cnt['RNG_STRT'])**2
            li2nos = np.ones(len(ax['li2rng']), np.int8)
        else:
            raise ValueError(f'projection in span {span} is not supported (span 1 or 11)')
        self.span, self.nsn = span, nsn
        self.device = device
        self._params_ref = (tx, ax)  # keeps the cache key (ids) valid
        self._keep = [
            np.ascontiguousarray(tx['crs'], np.float32),
            np.ascontiguousarray(tx['s2c'], np.int16),
            np.ascontiguousarray(ax['li2rng'], np.float32),
            np.ascontiguousarray(li2sn, np.int16),
            np.ascontiguousarray(li2nos, np.int8)]
        crs, s2c, li2rng, li2sn, li2nos = self._keep
        d = _lib.NipetGeomDesc(
            crs.ctypes.data, crs.shape[0], s2c.ctypes.data, s2c.shape[0], li2rng.ctypes.data,
            li2sn.ctypes.data, li2nos.ctypes.data, li2rng.shape[0], nsn, float(Cnt['AXR']),
            float(Cnt['SZ_VOXY']), float(Cnt['SZ_VOXZ']), float(Cnt['TFOV2']))
        import torch
        h = ct.c_void_p()
        with torch.cuda.device(device):
            _lib.check(_lib.lib().nipet_geom_create(ct.byref(d), ct.byref(h)), 'nipet_geom_create')
        self.handle = h
        self.naw = s2c.shape[0]
        self.nli = li2rng.shape[0]
        n = ct.c_int64()
        _lib.check(_lib.lib().nipet_geom_info(h, None, ct.byref(n), None, None), 'geom_info')
        self.nseg = n.value

    def __del__(self):
        h = getattr(self, 'handle', None)
        # at interpreter exit the module globals may already be gone
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.nipet_geom_destroy(h)
            self.handle = None

    def seg_counts(self):
        """Segments per transaxial bin (int64 [naw]), cached host copy."""
        if getattr(self, '_segc', None) is None:
            off = np.zeros(self.naw + 1, np.int64)
            _lib.check(_lib.lib().nipet_tx_lists(self.handle, off.ctypes.data, None, None, None),
                       'nipet_tx_lists')
            self._segc = np.diff(off)
        return self._segc

    def tx_lists(self):
        """(offsets int64 [naw+1], col int32 (u + 320 v), t float32, len float32)."""
        off = np.zeros(self.naw + 1, np.int64)
        col = np.zeros(self.nseg, np.int32)
        t = np.zeros(self.nseg, np.float32)
        ln = np.zeros(self.nseg, np.float32)
        _lib.check(_lib.lib().nipet_tx_lists(self.handle, off.ctypes.data, col.ctypes.data,
                                             t.ctypes.data, ln.ctypes.data), 'nipet_tx_lists')
        return off, col, t, ln


def get_geom(scanner_params, device=None):
    """Cached ``DeviceGeom`` for these scanner parameters on ``device``."""
    import torch
    if not torch.cuda.is_available():
        raise _lib.NipetError('no HIP device available: the nipet_amd projectors run on the GPU '
                              'only')
    dev = torch.cuda.current_device() if device is None else torch.device(device).index
    if dev is None:
        dev = torch.cuda.current_device()
    Cnt = scanner_params['Cnt']
    key = (id(scanner_params['txLUT']), id(scanner_params['axLUT']), Cnt['SPN'],
           Cnt['RNG_STRT'], Cnt['RNG_END'], dev)
    g = _CACHE.get(key)
    if g is None:
        g = DeviceGeom(scanner_params, dev)
        _CACHE[key] = g
    return g


def count_terms(scanner_params, rays, tx=None):
    """Number I of (span-1 LOR, voxel) weight terms of the projector over the
    transaxial bins ``rays`` (z-plane splits count as separate terms): per
    bin, every LOR walks the bin's nseg segments, and an oblique LOR with ring
    difference d crosses floor(0.5 + 2d e/atn) - floor(0.5 + 2d s/atn) z-planes
    between the first segment start s and the last segment end e (the same for
    every r0 and for the mirror). ``tx``: optional (offsets, col, t, len) lists.
    Returns an int64 array per bin."""
    Cnt, txLUT, axLUT = scanner_params['Cnt'], scanner_params['txLUT'], scanner_params['axLUT']
    rays = np.asarray(rays, dtype=np.int64)
    if tx is None:
        tx = get_geom(scanner_params).tx_lists()
    off, _, t, ln = tx
    crs = txLUT['crs'].astype(np.float64)
    c = txLUT['s2c'][rays].astype(np.int64)
    p = 0.5 * (crs[c[:, 0], :2] + crs[c[:, 0], 2:])
    q = 0.5 * (crs[c[:, 1], :2] + crs[c[:, 1], 2:])
    atn = np.hypot(*(q - p).T)
    b, e = off[rays], off[rays + 1]
    nseg = (e - b).astype(np.int64)
    s0 = t[b].astype(np.float64)
    e1 = t[e - 1].astype(np.float64) + ln[e - 1].astype(np.float64)
    r = axLUT['li2rno'].astype(np.int64)
    dcount = np.bincount(r[:, 1] - r[:, 0])          # LOR pairs per ring difference
    d = np.arange(1, len(dcount))
    cross = (np.floor(0.5 + 2 * d[None, :] * e1[:, None] / atn[:, None]) -
             np.floor(0.5 + 2 * d[None, :] * s0[:, None] / atn[:, None])).astype(np.int64)
    nlor = dcount[0] + 2 * dcount[1:].sum()
    return nseg * nlor + 2 * (cross * dcount[1:][None, :]).sum(1)
