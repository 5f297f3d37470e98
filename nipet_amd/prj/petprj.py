"""The ``petprj`` extension interface (``prj/src/prj_module.cu:37-43``) over the
C ABI of ``libnipet_hip.so``.

The reference's Python layer calls ``petprj.fprj``, ``petprj.bprj`` and
``petprj.osem`` with NumPy / CuVec arrays, the LUT dictionaries and the
constants dictionary, and the extension writes its output array in place.
These functions take the same arguments in the same order and write their
output in place the same way, so code written against the extension module
can import this one instead:

* ``fprj(sino, im, txLUT, axLUT, subs, cnst, att, sync=True)``
  (``prj/src/prj_module.cu:232-425``);
* ``bprj(bimg, sino, txLUT, axLUT, subs, cnst, sync=True)``
  (``prj/src/prj_module.cu:430-608``);
* ``osem(imgout, psng, rsng, ssng, nsng, asng, subs, imgsens, rcnmsk, krnl,
  txLUT, axLUT, cnst)`` (``prj/src/prj_module.cu:613-839``): one iteration
  over all subsets, ``nipet_osem_iter``.

Arrays may be NumPy (copied to the device and back, like the reference's
host arrays) or CUDA ``torch`` tensors (used in place). Errors raise
``NipetError`` instead of exiting the process. ``subs`` for ``fprj`` / ``bprj``
is a bin list (``[-1]`` = all bins); for ``osem`` it is the (Nsub, Nprj)
table whose row i holds the subset's bin count and then its bins
(``prj/mmrrec.py:286-294``).
"""
import numpy as np

from .. import _lib
from ..geom import get_geom

__all__ = ['fprj', 'bprj', 'osem', 'subset_table']


def _torch():
    import torch
    return torch


def _params(txLUT, axLUT, cnst):
    return {'Cnt': cnst, 'txLUT': txLUT, 'axLUT': axLUT}


def _dev(x, dtype):
    """(CUDA tensor, host array to write back or None)."""
    torch = _torch()
    if isinstance(x, torch.Tensor):
        if not x.is_cuda or x.dtype != dtype or not x.is_contiguous():
            raise TypeError('torch arguments must be contiguous CUDA tensors of dtype '
                            f'{dtype}')
        return x, None
    a = np.asarray(x)
    return torch.from_numpy(np.ascontiguousarray(a)).to('cuda', dtype=dtype), a


def _done(t, host, sync=True):
    torch = _torch()
    if host is not None:
        host[...] = t.cpu().numpy().reshape(host.shape)
    elif sync:
        torch.cuda.synchronize()


def _rows(subs, naw):
    torch = _torch()
    s = np.asarray(subs.cpu() if hasattr(subs, 'cpu') else subs, np.int32).ravel()
    if s.size == 0:
        raise ValueError('subs is empty')
    if s[0] < 0:
        return None, naw
    return torch.from_numpy(s).cuda(), len(s)


def fprj(sino, im, txLUT, axLUT, subs, cnst, att, sync=True):
    """Forward projection of device image ``im`` (320, 320, 127) into ``sino``
    (nprj, nsn): emission (att = 0) or attenuation factors (att = 1)."""
    torch = _torch()
    g = get_geom(_params(txLUT, axLUT, cnst))
    rows, nprj = _rows(subs, g.naw)
    out, host = _dev(sino, torch.float32)
    img, _ = _dev(im, torch.float32)
    if out.numel() != nprj * g.nsn or img.numel() != 320 * 320 * 127:
        raise ValueError(f'fprj: sino must hold ({nprj}, {g.nsn}) values and im 320*320*127')
    _lib.check(_lib.lib().nipet_fprj(g.handle, _lib.ptr(out), _lib.ptr(img), _lib.ptr(rows),
                                     nprj, int(att), _lib.stream_handle()), 'nipet_fprj')
    _done(out, host, sync)


def bprj(bimg, sino, txLUT, axLUT, subs, cnst, sync=True):
    """Back projection of ``sino`` (nprj, nsn) into ``bimg`` (320, 320, 127),
    which is overwritten."""
    torch = _torch()
    g = get_geom(_params(txLUT, axLUT, cnst))
    rows, nprj = _rows(subs, g.naw)
    out, host = _dev(bimg, torch.float32)
    y, _ = _dev(sino, torch.float32)
    if y.numel() != nprj * g.nsn or out.numel() != 320 * 320 * 127:
        raise ValueError(f'bprj: sino must hold ({nprj}, {g.nsn}) values and bimg 320*320*127')
    _lib.check(_lib.lib().nipet_bprj(g.handle, _lib.ptr(out), _lib.ptr(y), _lib.ptr(rows), nprj,
                                     _lib.stream_handle()), 'nipet_bprj')
    _done(out, host, sync)


def subset_table(subsets):
    """The (Nsub, Nprj) int32 subset table of ``petprj.osem``: row i = (number
    of bins, bins of subset i, zero padding) (``prj/mmrrec.py:286-294``)."""
    n = max(len(s) for s in subsets)
    t = np.zeros((len(subsets), n + 1), np.int32)
    for i, s in enumerate(subsets):
        t[i, 0] = len(s)
        t[i, 1:1 + len(s)] = s
    return t


def osem(imgout, psng, rsng, ssng, nsng, asng, subs, imgsens, rcnmsk, krnl, txLUT, axLUT,
         cnst):
    """One OSEM iteration over the subsets of ``subs`` (``petprj.osem``):
    ``imgout`` (320, 320, 127) float32 is updated in place; ``psng`` prompts
    (Naw, nsn) (converted to uint16 like the extension's NPY_UINT16 input,
    saturating), ``rsng``/``ssng``/``nsng``/``asng`` randoms, scatter, norm and
    attenuation (Naw, nsn) float32; ``imgsens`` (Nsub, 320, 320, 127)
    sensitivity images BP(a n); ``rcnmsk`` the reconstruction mask; ``krnl``
    (3, 17) PSF kernel, resolution modelling when ``krnl[0, 0] >= 0``."""
    torch = _torch()
    from .osem import prompts_u16
    g = get_geom(_params(txLUT, axLUT, cnst))
    table = np.ascontiguousarray(np.asarray(subs.cpu() if hasattr(subs, 'cpu') else subs,
                                            np.int32))
    if table.ndim != 2:
        raise ValueError('subs must be the (Nsub, Nprj) subset table')
    nsub, stride = table.shape
    img, host = _dev(imgout, torch.float32)
    if isinstance(psng, torch.Tensor):
        p = prompts_u16(psng.cuda()).contiguous()
    else:
        p = torch.from_numpy(prompts_u16(psng)).cuda()
    r, _ = _dev(rsng, torch.float32)
    s, _ = _dev(ssng, torch.float32)
    n, _ = _dev(nsng, torch.float32)
    a, _ = _dev(asng, torch.float32)
    sens, _ = _dev(imgsens, torch.float32)
    m = rcnmsk if isinstance(rcnmsk, torch.Tensor) else torch.from_numpy(
        np.ascontiguousarray(np.asarray(rcnmsk)).astype(np.uint8))
    m = m.to('cuda', torch.uint8).contiguous()
    nel = g.naw * g.nsn
    if any(x.numel() != nel for x in (p, r, s, n, a)):
        raise ValueError(f'osem: sinograms must hold ({g.naw}, {g.nsn}) values')
    if img.numel() != 320 * 320 * 127 or m.numel() != img.numel() or \
            sens.numel() != nsub * img.numel():
        raise ValueError('osem: image, mask and (Nsub) sensitivity images must be 320x320x127')
    k = None
    if krnl is not None:
        k = np.ascontiguousarray(np.asarray(krnl, np.float32))
        if k.shape != (3, 17):
            raise ValueError('krnl must be (3, 17)')
    _lib.check(_lib.lib().nipet_osem_iter(
        g.handle, _lib.ptr(img), _lib.ptr(p), _lib.ptr(r), _lib.ptr(s), _lib.ptr(n), _lib.ptr(a),
        table.ctypes.data, nsub, stride, _lib.ptr(sens), _lib.ptr(m),
        None if k is None else k.ctypes.data, _lib.stream_handle()), 'nipet_osem_iter')
    _done(img, host)
