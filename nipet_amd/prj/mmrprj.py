"""Forward and back projector operator API (``niftypet/nipet/prj/mmrprj.py``).

``frwd_prj`` / ``back_prj`` keep the reference's signature, defaults, accepted
shapes and output layouts (``prj/mmrprj.py:46-139, 147-221``). The projection
itself always runs on the GPU through ``libnipet_hip.so``:

* NumPy inputs are copied to the device and the result comes back as NumPy
  (the reference's CuVec outputs are host-visible arrays);
* CUDA ``torch.Tensor`` inputs stay on the device and the result is a CUDA
  tensor on the current stream (the ``dev_out=True, sync=False`` usage of the
  reference's notebook, ``examples/demo.ipynb``).
"""
import logging

import numpy as np

from .. import _lib
from ..geom import get_geom

log = logging.getLogger(__name__)

ISUB_DEFAULT = np.array([-1], dtype=np.int32)

__all__ = ['frwd_prj', 'back_prj', 'trnx_prj', 'ISUB_DEFAULT']


def _torch():
    import torch
    return torch


def _nsinos(Cnt):
    if Cnt['SPN'] == 1:
        nrng = Cnt['RNG_END'] - Cnt['RNG_STRT']
        return nrng**2 - (12 if nrng == 64 else 0)
    if Cnt['SPN'] == 11:
        return Cnt['NSN11']
    raise ValueError(f"projection in span {Cnt['SPN']} is not supported (span 1 or 11)")


def _dev_tensor(x, dtype):
    """(cuda tensor, came_from_host)"""
    torch = _torch()
    if isinstance(x, torch.Tensor):
        if not x.is_cuda:
            return x.to('cuda', dtype=dtype).contiguous(), True
        return x.to(dtype=dtype).contiguous(), False
    return torch.from_numpy(np.ascontiguousarray(x)).to('cuda', dtype=dtype), True


def _subs(isub, geom):
    torch = _torch()
    isub = np.asarray(isub, dtype=np.int32).ravel()
    if isub.size == 0:
        raise ValueError('isub is empty')
    if isub[0] < 0:
        return None, geom.naw
    if isub.min() < 0 or isub.max() >= geom.naw:
        raise ValueError(f'isub entries must be in [0, {geom.naw})')
    return torch.from_numpy(isub).to('cuda'), len(isub)


def _image_dev(im, Cnt):
    """Device-layout float32 CUDA image (320,320,127) from an E7 or device image."""
    torch = _torch()
    shp = tuple(im.shape)
    if shp == (Cnt['SO_IMZ'], Cnt['SO_IMY'], Cnt['SO_IMX']):
        from ..img.mmrimg import convert2dev
        if isinstance(im, torch.Tensor):
            m = (Cnt['SO_IMX'] - Cnt['SZ_IMX']) // 2
            t = im[:, m:m + Cnt['SZ_IMY'], m:m + Cnt['SZ_IMX']].permute(1, 2, 0)
            return _dev_tensor(t, torch.float32)
        return _dev_tensor(convert2dev(im, Cnt), torch.float32)
    if shp == (Cnt['SZ_IMY'], Cnt['SZ_IMX'], Cnt['SZ_IMZ']):
        return _dev_tensor(im, torch.float32)
    raise ValueError('wrong image size; it has to be one of these: (z,y,x) = (127,344,344)'
                     ' or (y,x,z) = (320,320,127)')


def _sync(sync):
    if sync:
        _torch().cuda.current_stream().synchronize()


def frwd_prj(im, scanner_params, isub=ISUB_DEFAULT, dev_out=False, attenuation=False,
             fullsino_out=True, output=None, sync=True):
    """Forward projection of an emission image or a mu-map (``attenuation=True``:
    exp(-line integral), span-11 bins average their span-1 LORs).

    Returns (Naw | len(isub), nsinos) when ``dev_out`` (rows in ``isub`` order
    unless ``fullsino_out`` scatters them into a full (Naw, nsinos) sinogram),
    else the Siemens sinogram (nsinos, 252, 344) with gaps.
    """
    torch = _torch()
    Cnt, txLUT = scanner_params['Cnt'], scanner_params['txLUT']
    nsinos = _nsinos(Cnt)
    img, host = _image_dev(im, Cnt)
    geom = get_geom(scanner_params)
    subs, nprj = _subs(isub, geom)
    if output is None:
        sinog = torch.empty((nprj, nsinos), dtype=torch.float32, device='cuda')
    else:
        sinog, _ = _dev_tensor(output, torch.float32)
        if tuple(sinog.shape) != (nprj, nsinos):
            raise ValueError(f'output must have shape {(nprj, nsinos)}')
    _lib.check(_lib.lib().nipet_fprj(geom.handle, _lib.ptr(sinog), _lib.ptr(img), _lib.ptr(subs),
                                     nprj, int(bool(attenuation)), _lib.stream_handle()),
               'nipet_fprj')
    sino = sinog
    if fullsino_out and subs is not None:
        sino = torch.zeros((geom.naw, nsinos), dtype=torch.float32, device='cuda')
        sino[subs.long()] = sinog
    if not dev_out and fullsino_out:
        sino = _gaps_dev(sino, txLUT, Cnt, nsinos, put=True)
    if output is not None and sino is sinog and isinstance(output, np.ndarray):
        _sync(True)
        output[...] = sinog.cpu().numpy()
        return output
    if host:
        _sync(True)
        return sino.cpu().numpy()
    _sync(sync)
    return sino


def back_prj(sino, scanner_params, isub=ISUB_DEFAULT, dev_out=False, output=None, sync=True):
    """Back projection (adjoint of the emission projector) of a Siemens
    sinogram (nsinos, 252, 344) or a device sinogram (Naw | len(isub), nsinos).
    Returns the device image (320,320,127) when ``dev_out``, else the E7 image
    (127,344,344)."""
    torch = _torch()
    Cnt, txLUT = scanner_params['Cnt'], scanner_params['txLUT']
    nsinos = _nsinos(Cnt)
    geom = get_geom(scanner_params)
    subs, nprj = _subs(isub, geom)
    shp = tuple(sino.shape)
    if len(shp) == 3:
        if shp != (nsinos, Cnt['NSANGLES'], Cnt['NSBINS']):
            raise ValueError('Unexpected sinogram array dimensions/shape for Siemens defaults.')
        s, host = _dev_tensor(sino, torch.float32)
        sinog = _gaps_dev(s, txLUT, Cnt, nsinos, put=False)
        if subs is not None:
            sinog = sinog[subs.long()].contiguous()
    elif len(shp) == 2:
        if subs is None and shp[0] != geom.naw:
            raise ValueError('Unexpected number of transaxial elements in the full sinogram.')
        if subs is not None and shp[0] != nprj:
            raise ValueError('Unexpected number of transaxial elements in the subset sinogram.')
        if shp[1] != nsinos:
            raise ValueError('Inconsistent number of sinograms in the array.')
        sinog, host = _dev_tensor(sino, torch.float32)
    else:
        raise ValueError('Unexpected shape of the input sinogram.')
    shape = (Cnt['SZ_IMY'], Cnt['SZ_IMX'], Cnt['SZ_IMZ'])
    if output is None:
        bimg = torch.empty(shape, dtype=torch.float32, device='cuda')
    else:
        bimg, _ = _dev_tensor(output, torch.float32)
        if tuple(bimg.shape) != shape:
            raise ValueError(f'output must have shape {shape}')
    _lib.check(_lib.lib().nipet_bprj(geom.handle, _lib.ptr(bimg), _lib.ptr(sinog), _lib.ptr(subs),
                                     nprj, _lib.stream_handle()), 'nipet_bprj')
    if not dev_out:
        from ..img.mmrimg import convert2e7
        _sync(True)
        return convert2e7(bimg.cpu().numpy(), Cnt)
    if host:
        _sync(True)
        out = bimg.cpu().numpy()
        if isinstance(output, np.ndarray):
            output[...] = out
            return output
        return out
    _sync(sync)
    return bimg


def trnx_prj(scanner_params, sino=None, im=None):
    """Transaxial intersection lists (the reference's ``tv``/``tt`` tables,
    ``prj/mmrprj.py:19-38``), here as CSR arrays:
    {'offsets', 'col' (u + 320 v), 't' (segment start, cm), 'len' (cm)}."""
    if sino is not None and im is not None:
        raise ValueError('Only one input should be given: sinogram or image.')
    off, col, t, ln = get_geom(scanner_params).tx_lists()
    return {'offsets': off, 'col': col, 't': t, 'len': ln}


_AW2ALI = {}


def _aw2ali_dev(txLUT):
    torch = _torch()
    key = (id(txLUT), torch.cuda.current_device())
    t = _AW2ALI.get(key)
    if t is None:
        t = (torch.from_numpy(np.ascontiguousarray(txLUT['aw2ali'], np.int32)).to('cuda'), txLUT)
        _AW2ALI[key] = t
    return t[0]


def _gaps_dev(s, txLUT, Cnt, nsinos, put, output=None):
    """GPU gap insertion/removal on CUDA tensors (kernels nipet_putgaps /
    nipet_remgaps)."""
    torch = _torch()
    aw2ali = _aw2ali_dev(txLUT)
    naw, nba = txLUT['Naw'], Cnt['NSANGLES'] * Cnt['NSBINS']
    st = _lib.stream_handle()
    if put:
        s = s.to(torch.float32).contiguous()
        if tuple(s.shape) != (naw, nsinos):
            raise ValueError(f'expected device sinogram of shape {(naw, nsinos)}')
        out = torch.empty((nsinos, Cnt['NSANGLES'], Cnt['NSBINS']), dtype=torch.float32,
                          device='cuda')
        _lib.check(_lib.lib().nipet_putgaps(_lib.ptr(out), _lib.ptr(s), _lib.ptr(aw2ali), naw,
                                            nsinos, nba, st), 'nipet_putgaps')
        return out
    if s.dtype == torch.uint16:
        dt = 1
    else:
        s = s.to(torch.float32)
        dt = 0
    s = s.contiguous()
    if s.dim() != 3 or tuple(s.shape[1:]) != (Cnt['NSANGLES'], Cnt['NSBINS']):
        raise ValueError('expected a Siemens sinogram (n, 252, 344)')
    out = output if output is not None else torch.empty((naw, s.shape[0]), dtype=torch.float32,
                                                        device='cuda')
    _lib.check(_lib.lib().nipet_remgaps(_lib.ptr(out), _lib.ptr(s), dt, _lib.ptr(aw2ali), naw,
                                        s.shape[0], nba, st), 'nipet_remgaps')
    return out
