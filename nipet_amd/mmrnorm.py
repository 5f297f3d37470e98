"""Normalisation of the mMR (``niftypet/nipet/mmrnorm.py``).

* ``get_components`` -- ``mmrnorm.py:19-112``: the normalisation components
  from the scanner's component file (``datain['nrm_bf']``, float32 blocks) and
  the quantification factor from its Interfile header, found inside the
  header file (``datain['nrm_ima']`` or ``datain['nrm_dcm']``; the DICOM
  header embeds the Interfile text, which is located by its ``!INTERFILE``
  marker). The span-1 axial factors and axial scatter scale factors are the
  scanner calibration tables shipped with the reference package
  (``auxdata/``): pass their folder as ``auxdata`` (or ``datain['auxdata']``).
* ``get_norm_sino`` -- ``mmrnorm.py:155-195``: the normalisation sinogram
  from the components and the histogram's singles buckets, computed by the
  HIP kernels behind ``nipet_norm_sino`` (device layout (Naw, nsinos), or the
  Siemens layout with zero gaps).
* ``NormSino`` -- the device session (sinogram LUTs uploaded once).
"""
import ctypes as ct
import logging
import os
import re

import numpy as np

from . import _lib

log = logging.getLogger(__name__)

__all__ = ['get_components', 'get_norm_sino', 'get_sinog', 'get_sino', 'NormSino']

QF_LOC = 0.205      # local quantification correction factor (mmrnorm.py:103)


def _torch():
    import torch
    return torch


def _interfile_header(fname):
    """The Interfile text of a normalisation header file (a plain Interfile
    header or a DICOM file carrying one)."""
    raw = open(fname, 'rb').read()
    i = raw.find(b'!INTERFILE')
    if i < 0:
        raise ValueError(f'{fname}: no Interfile header found')
    j = raw.find(b'\x00', i)
    txt = raw[i:j if j > 0 else len(raw)].decode('latin-1')
    if 'scanner quantification factor' not in txt:
        raise ValueError(f'{fname}: the Interfile header has no scanner quantification factor')
    return txt


def get_components(datain, Cnt, auxdata=None):
    """Normalisation components and the Interfile header text
    (``mmrnorm.py:19-112``)."""
    if os.path.isfile(datain.get('nrm_ima', '')) and os.path.isfile(datain.get('nrm_bf', '')):
        fdat, fhdr = datain['nrm_bf'], datain['nrm_ima']
    elif os.path.isfile(datain.get('nrm_dcm', '')) and os.path.isfile(datain.get('nrm_bf', '')):
        fdat, fhdr = datain['nrm_bf'], datain['nrm_dcm']
    else:
        raise NameError('norm file does not exist or it is incomplete')
    NSB, NS0, NCRS, NRNG, NSN11 = (Cnt['NSBINS'], Cnt['NSEG0'], Cnt['NCRS'], Cnt['NRNG'],
                                   Cnt['NSN11'])
    with open(fdat, 'rb') as f:
        geo = np.fromfile(f, np.float32, NSB * NS0).reshape(NS0, NSB)
        cinf = np.fromfile(f, np.float32, 9 * NSB).reshape(NSB, 9)
        ceff = np.fromfile(f, np.float32, NCRS * NRNG).reshape(NRNG, NCRS)
        axe1 = np.fromfile(f, np.float32, NSN11)
        dtp = np.fromfile(f, np.float32, NRNG)
        dtnp = np.fromfile(f, np.float32, NRNG)
        dtc = np.fromfile(f, np.float32, 9)
        axe2 = np.fromfile(f, np.float32, NSN11)
    if axe2.size != NSN11:
        raise ValueError(f'{fdat}: the component file is truncated')
    aux = auxdata or datain.get('auxdata')
    if aux is None or not os.path.isdir(aux):
        raise ValueError('the scanner calibration tables (AxialFactorForSpan1.npy, '
                         'RelativeScaleFactors_scatter_axial_ssrTOspan{1,11}.f32) are needed: '
                         'pass their folder as auxdata= or datain["auxdata"]')
    axf1 = np.load(os.path.join(aux, 'AxialFactorForSpan1.npy'), allow_pickle=False)
    sax_f11 = np.fromfile(os.path.join(aux, 'RelativeScaleFactors_scatter_axial_ssrTOspan11.f32'),
                          np.float32, NSN11)
    sax_f1 = np.fromfile(os.path.join(aux, 'RelativeScaleFactors_scatter_axial_ssrTOspan1.f32'),
                         np.float32, Cnt['NSN1'])
    nhdr = _interfile_header(fhdr)
    f0 = nhdr.find('scanner quantification factor')
    f1 = f0 + nhdr[f0:].find('\n')
    p = re.compile(r'(?<=:=)\s*\d{1,5}[.]\d{3,10}[e][+-]\d{1,4}')
    found = p.findall(nhdr[f0:f1] if f1 > f0 else nhdr[f0:])
    if not found:
        raise ValueError(f'{fhdr}: unreadable scanner quantification factor')
    qf = float(found[0])
    return {'qf': qf, 'qf_loc': QF_LOC, 'geo': geo, 'cinf': cinf, 'ceff': ceff, 'axe1': axe1,
            'dtp': dtp, 'dtnp': dtnp, 'dtc': dtc, 'axe2': axe2, 'axf1': axf1,
            'sax_f11': sax_f11, 'sax_f1': sax_f1}, nhdr


class NormSino:
    """Device session of the normalisation sinogram from components
    (``mmr_auxe.norm``, ``src/norm.cu``)."""

    def __init__(self, scanner_params, span=None):
        Cnt, tx, ax = scanner_params['Cnt'], scanner_params['txLUT'], scanner_params['axLUT']
        self.span = Cnt['SPN'] if span is None else span
        if self.span not in (1, 11):
            raise ValueError(f'unrecognised span {self.span}')
        if 'rNSN1' in Cnt and 'rLUT' in ax:
            raise ValueError('Full FOV has to be used for normalisation - switch off reduced '
                             'rings mode.')
        self.nsinos = Cnt['NSN1'] if self.span == 1 else Cnt['NSN11']
        self.naw = int(tx['Naw'])
        self._keep = [np.ascontiguousarray(tx['aw2ali'], np.int32),
                      np.ascontiguousarray(ax['sn1_rno'], np.int16),
                      np.ascontiguousarray(ax['sn1_sn11'], np.int16),
                      np.ascontiguousarray(ax['sn1_sn11no'], np.int8)]
        d = _lib.NipetNormDesc(self.naw, self.span, *[k.ctypes.data for k in self._keep])
        h = ct.c_void_p()
        _lib.check(_lib.lib().nipet_norm_create(ct.byref(d), ct.byref(h)), 'nipet_norm_create')
        self.handle = h

    def __del__(self):
        h = getattr(self, 'handle', None)
        # at interpreter exit the module globals may already be gone
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.nipet_norm_destroy(h)
            self.handle = None

    def __call__(self, normcomp, buckets, out=None):
        """Device normalisation sinogram (Naw, nsinos) float32."""
        torch = _torch()
        shapes = {'geo': (None, 344), 'cinf': (344, 9), 'ceff': (64, 504), 'axe1': (837,),
                  'dtp': (64,), 'dtnp': (64,), 'axf1': (4084,)}
        arrs = {}
        for k, shp in shapes.items():
            if k == 'axf1' and self.span == 11 and k not in normcomp:
                arrs[k] = None
                continue
            a = np.ascontiguousarray(np.asarray(normcomp[k], np.float32))
            if a.ndim != len(shp) or any(s is not None and s != n for s, n in zip(shp, a.shape)):
                raise ValueError(f'normalisation component {k} has shape {a.shape}')
            arrs[k] = a
        arrs['geo'] = np.ascontiguousarray(arrs['geo'][0])
        b = np.ascontiguousarray(np.asarray(buckets).astype(np.int32)).ravel()
        if b.size != 224:
            raise ValueError('buckets must hold the 224 singles rates')
        comp = _lib.NipetNormComp(*[None if arrs[k] is None else arrs[k].ctypes.data
                                    for k in ('geo', 'cinf', 'ceff', 'axe1', 'axf1', 'dtp',
                                              'dtnp')])
        if out is None:
            out = torch.empty((self.naw, self.nsinos), dtype=torch.float32, device='cuda')
        if (tuple(out.shape) != (self.naw, self.nsinos) or out.dtype != torch.float32
                or not out.is_cuda or not out.is_contiguous()):
            raise ValueError('norm: bad output tensor')
        _lib.check(_lib.lib().nipet_norm_sino(self.handle, ct.byref(comp), b.ctypes.data,
                                              out.data_ptr(), _lib.stream_handle()),
                   'nipet_norm_sino')
        return out


def get_norm_sino(datain, scanner_params, hst, normcomp=None, gpu_dim=False):
    """Normalisation sinogram (``mmrnorm.py:155-195``): float32 host array in
    the device layout (Naw, nsinos) with ``gpu_dim``, else the Siemens layout
    (nsinos, 252, 344) with zero gaps."""
    from .mmraux import putgaps
    Cnt, txLUT = scanner_params['Cnt'], scanner_params['txLUT']
    if normcomp is None:
        normcomp, _ = get_components(datain, Cnt)
    s = NormSino(scanner_params)(normcomp, hst['buckets'])
    if gpu_dim:
        return s.cpu().numpy()
    return putgaps(s, txLUT, Cnt).cpu().numpy()


def get_sinog(datain, hst, axLUT, txLUT, Cnt, normcomp=None):
    """Device-layout normalisation sinogram (``mmrnorm.py:115-134``)."""
    return get_norm_sino(datain, {'Cnt': Cnt, 'txLUT': txLUT, 'axLUT': axLUT}, hst,
                         normcomp=normcomp, gpu_dim=True)


def get_sino(datain, hst, axLUT, txLUT, Cnt):
    """Siemens-layout normalisation sinogram (``mmrnorm.py:137-152``)."""
    return get_norm_sino(datain, {'Cnt': Cnt, 'txLUT': txLUT, 'axLUT': axLUT}, hst)
