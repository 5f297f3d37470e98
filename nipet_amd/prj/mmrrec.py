"""OSEM reconstruction driver (``niftypet/nipet/prj/mmrrec.py``).

* ``get_subsets14`` -- ``prj/mmrrec.py:43-93``: 14 balanced angular subsets of
  18 angles each (14 angles taken one per 18-angle block in a rotating pattern
  plus 4 of the remaining angles), returned as transaxial-bin (aw) indices in
  (radial bin, angle) order.
* ``psf_config``    -- ``prj/mmrrec.py:96-155``.
* ``osemone``       -- ``prj/mmrrec.py:158-470`` (corrections are arrays; the
  randoms come from the delayed fan-sums when no randoms sinogram is given;
  this path has no scanner-file readers or scatter model).
"""
import logging
import os
import time
from collections.abc import Iterable
from numbers import Real

import numpy as np

log = logging.getLogger(__name__)

__all__ = ['get_subsets14', 'psf_config', 'fwhm2sig', 'osemone', 'osem_prepare', 'osem_run',
           'RI_LUT', 'quant_factor', 'recon_inputs', 'rec_output']

# Radioisotopes: half-life [s] and positron branching fraction. The reference
# reads them from ``resources.riLUT`` of its un-vendored ``niftypet.ninst``
# dependency (>= 0.12.0, ``setup.cfg:45``) at ``prj/mmrrec.py:309, 321, 327``;
# the values here are the published decay data that table holds. The table
# itself cannot be read here, so these constants are not pinned against it.
RI_LUT = {
    'Ge68': {'BF': 0.891, 'thalf': 270.9516 * 24 * 60 * 60},
    'Ga68': {'BF': 0.891, 'thalf': 67.71 * 60},
    'F18': {'BF': 0.967, 'thalf': 109.77120 * 60},
    'O15': {'BF': 0.999, 'thalf': 122.2416},
    'C11': {'BF': 0.998, 'thalf': 20.38 * 60},
}


def quant_factor(Cnt, hst, ncmp, decay_ref_time=None):
    """Factor dcycrr * qf * qf_loc that ``osemone`` applies to the image
    (``prj/mmrrec.py:308-331, 384-388``): with the normalisation components
    ``ncmp`` and a histogram carrying ``t0`` and ``dur``, qf = ncmp['qf'] /
    BF / dur, qf_loc = ncmp['qf_loc'], and with ``Cnt['DCYCRR']`` the decay
    correction to ``decay_ref_time`` (default the frame start t0),
    dcycrr = exp(l tref) l dur / (1 - exp(-l dur)), l = ln 2 / T_half.
    Otherwise 1."""
    if ncmp is None or 't0' not in hst or 'dur' not in hst:
        return 1.0
    iso = RI_LUT[Cnt['ISOTOPE']]
    lmbd = np.log(2) / iso['thalf']
    dur = float(hst['dur'])
    if Cnt.get('DCYCRR'):
        tref = hst['t0'] if decay_ref_time is None else decay_ref_time
        dcycrr = np.exp(lmbd * tref) * lmbd * dur / (1 - np.exp(-lmbd * dur))
    else:
        dcycrr = 1.
    qf = ncmp['qf'] / iso['BF'] / dur
    return float(dcycrr * qf * ncmp['qf_loc'])


def fwhm2sig(fwhm, voxsize=1.):
    return (fwhm/voxsize) / (2 * (2 * np.log(2))**.5)


def get_subsets14(n, params):
    """The n-th of 14 transaxial subsets: (aw indices, angle table S (14,18))."""
    Cnt, txLUT = params['Cnt'], params['txLUT']
    N = 14
    NSA = Cnt['NSANGLES']
    P = NSA // N                    # 18 angles per subset / per block
    fs = N / float(P - N)           # spread of the 4 remaining angles per block
    blocks = np.array([np.arange(i, NSA, P) for i in range(N)])   # (14 offsets, 14 blocks)
    S = np.zeros((N, P), dtype=np.int16)
    for s in range(N):
        angles = [blocks[(s + b) % N, b] for b in range(N)]
        extra = np.int16(np.floor(np.arange(s, 2 * N, fs)[:P - N] % N))
        for i in range(P - N):
            angles.append(blocks[-1, extra[i]] + i + 1)
        S[s] = angles
    live = txLUT['msino'] > 0
    awmap = -np.ones(live.shape, dtype=np.int32)
    awmap[live] = np.arange(Cnt['Naw'], dtype=np.int32)
    iprj = awmap[:, S[n]]
    iprj = iprj[iprj >= 0]
    return iprj, S


def psf_config(psf, Cnt):
    """Separable PSF kernel (3, 17): rows act on z, x and y of the device image
    (the reference's order, built from FWHM [x, y, z] in mm); ``[0,0] < 0``
    switches resolution modelling off (``prj/mmrrec.py:96-155``)."""
    R = Cnt['RSZ_PSF_KRNL']

    def _config(fwhm3, check_len=True):
        if check_len and (len(fwhm3) != 3 or any(f < 0 for f in fwhm3)):
            raise ValueError('Incorrect separable kernel FWHM definition')
        kernel = np.empty((3, 2*R + 1), dtype=np.float32)
        for i, f in enumerate(fwhm3):
            vox = Cnt['SZ_VOXZ'] * 10 if i == 2 else Cnt['SZ_VOXY'] * 10
            sig = fwhm2sig(f, voxsize=vox)
            x = np.arange(-R, R + 1)
            kernel[i] = np.exp(-0.5 * (x**2 / sig**2))
            kernel[i] /= np.sum(kernel[i])
        out = np.empty((3, 2*R + 1), dtype=np.float32)
        out[0], out[1], out[2] = kernel[2], kernel[0], kernel[1]
        return out

    if psf is None:
        k = _config([], False)
        k[0, 0] = -1
        return k
    if isinstance(psf, str):
        raise ValueError("measured PSF kernels need the scanner's calibration (not available)")
    if isinstance(psf, Real):
        return _config([psf] * 3)
    if isinstance(psf, Iterable):
        psf = np.asanyarray(psf)
        if psf.shape == (3, 2*R + 1):
            k = _config([], False)
            k[0], k[1], k[2] = psf[2], psf[0], psf[1]
            return k
        if len(psf) == 3:
            return _config(psf)
        raise ValueError(f"invalid PSF dimensions ({psf.shape})")
    raise ValueError(f"unrecognised PSF definition ({psf})")


def osem_prepare(*args, **kwargs):
    from .osem import osem_prepare as f
    return f(*args, **kwargs)


def osem_run(*args, **kwargs):
    from .osem import osem_run as f
    return f(*args, **kwargs)


def recon_inputs(datain, mumaps, hst, scanner_params, recmod=3, attnsino=None, sctsino=None,
                 randsino=None, normcomp=None, normsino=None):
    """The sinograms ``osemone`` reconstructs from (``prj/mmrrec.py:196-265``),
    in the device layout (Naw, nsn): dict with ``psng`` (prompts of ``hst``),
    ``asng`` (attenuation factors: ones for recmod 0, ``attnsino``, or the
    attenuation projection of the mu-maps, kept on the device), ``nsng``
    (``normsino``, else from ``normcomp`` or the component files of
    ``datain`` with ``hst['buckets']``, else ones), ``rsng`` (``randsino``,
    else from ``hst['fansums']``, else None), ``ssng`` (``sctsino`` for
    recmod >= 2, else None) and ``ncmp`` (the normalisation components used)."""
    from ..img import mmrimg
    from . import mmrprj
    Cnt, txLUT = scanner_params['Cnt'], scanner_params['txLUT']

    def dev(s):
        if s is None:
            return None
        if hasattr(s, 'is_cuda'):
            if s.dim() == 3:
                from ..mmraux import remgaps
                return remgaps(s, txLUT, Cnt)
            return s
        s = np.asarray(s)
        if s.ndim == 3:
            from ..mmraux import remgaps
            return remgaps(s, txLUT, Cnt)
        return s

    psng = dev(hst['psino'])
    if recmod == 0:
        asng = np.ones(tuple(psng.shape), dtype=np.float32)
    elif attnsino is not None:
        asng = dev(attnsino)
    else:
        import torch
        muh, muo = mumaps
        mus = torch.from_numpy(np.ascontiguousarray(
            mmrimg.convert2dev(np.asarray(muo) + np.asarray(muh), Cnt), np.float32)).cuda()
        asng = mmrprj.frwd_prj(mus, scanner_params, dev_out=True, attenuation=True)
    ncmp = normcomp
    if ncmp is None and datain.get('nrm_bf') and (datain.get('nrm_ima') or datain.get('nrm_dcm')):
        from ..mmrnorm import get_components
        ncmp = get_components(datain, Cnt)[0]
    if normsino is not None:
        nsng = dev(normsino)
    elif ncmp is not None:
        # prj/mmrrec.py:209-215: the normalisation sinogram from its components
        from ..mmrnorm import NormSino
        nsng = NormSino(scanner_params)(ncmp, hst['buckets'])
    else:
        nsng = np.ones(tuple(psng.shape), np.float32)
    if randsino is not None:
        rsng = dev(randsino)
    elif hst.get('fansums') is not None:
        from ..lm.mmrhist import Randoms
        rsng, _ = Randoms(scanner_params)(hst['fansums'])
    else:
        rsng = None
    ssng = None if (sctsino is None or recmod < 2) else dev(sctsino)
    return {'psng': psng, 'asng': asng, 'nsng': nsng, 'rsng': rsng, 'ssng': ssng, 'ncmp': ncmp}


def rec_output(img, iterations, datain, hst, scanner_params, ncmp=None, decay_ref_time=None,
               fwhm=0., frmno='', fcomment='', outpath=None, fout=None):
    """``osemone``'s result from the device image ``img`` (320, 320, 127) after
    ``iterations`` iterations (``prj/mmrrec.py:384-470``): quantification and
    decay correction, E7 image, the reference's output file name and, with
    ``fwhm > 0``, the smoothed image. Returns ``RecOut(im, fpet, imsmo, fsmo,
    affine)``."""
    from collections import namedtuple

    from ..img import mmrimg
    Cnt = scanner_params['Cnt']
    qscale = quant_factor(Cnt, hst, ncmp, decay_ref_time)
    if qscale != 1.0:
        log.info('applying decay correction and quantification factor %r', qscale)
        img = img * qscale
    im = mmrimg.convert2e7(img, Cnt)
    # output file name of the reference (prj/mmrrec.py:175-186, 406-413)
    opth = outpath if outpath is not None else (
        os.path.join(datain['corepath'], 'reconstructed') if datain.get('corepath') else '')
    fpet = None
    if fout is not None:
        fpet = os.path.join(opth, os.path.basename(fout).split('.')[0] +
                            f'_itr{iterations}{fcomment}.nii.gz')
    elif datain.get('lm_bf') and 't0' in hst and 't1' in hst:
        fpet = os.path.join(opth, os.path.basename(datain['lm_bf']).split('.')[0] +
                            f"{frmno}_t{hst['t0']}-{hst['t1']}sec_itr{iterations}"
                            f"{fcomment}.nii.gz")
    im_smo = None
    if fwhm > 0:
        # prj/mmrrec.py:420-421
        from scipy import ndimage as ndi
        im_smo = ndi.gaussian_filter(im, fwhm2sig(fwhm, voxsize=Cnt['SZ_VOXY'] * 10),
                                     mode='mirror')
    RecOut = namedtuple('RecOut', 'im, fpet, imsmo, fsmo, affine')
    return RecOut(im, fpet, im_smo, None, None)


def osemone(datain, mumaps, hst, scanner_params, recmod=3, itr=4, fwhm=0., psf=None,
            mask_radius=29., decay_ref_time=None, attnsino=None, sctsino=None, randsino=None,
            normcomp=None, emmskS=False, frmno='', fcomment='', outpath=None, fout=None,
            store_img=False, store_itr=None, ret_sinos=False, normsino=None):
    """OSEM reconstruction (``prj/mmrrec.py:158-470``) on the GPU.

    ``hst['psino']`` holds the prompts (Siemens (837,252,344) or device
    (68516,837)). Corrections are arrays in either shape: ``attnsino``
    (attenuation factors; computed from ``mumaps`` by the attenuation forward
    projector when None and recmod>0), ``normsino`` (normalisation factors;
    when None, computed from ``normcomp`` or the component files named in
    ``datain`` with ``hst['buckets']``, as ``prj/mmrrec.py:209-215``, else
    ones), ``randsino`` (additive; estimated from ``hst['fansums']``
    when None, as ``prj/mmrrec.py:244-250``, zero without fan-sums) and
    ``sctsino`` (additive; zero when None). ``emmskS`` only shapes the
    reference's scatter estimate (``prj/mmrrec.py:365``), which is not modelled
    here, so it has no effect.
    As the reference, the image is scaled by ``quant_factor`` (decay correction
    and quantification, when the normalisation components are known and ``hst``
    carries ``t0`` / ``dur``) and, with ``fwhm > 0``, also returned smoothed by
    a Gaussian of that FWHM [mm]. Returns a namedtuple ``(im, fpet, imsmo,
    fsmo, affine)``: the E7 image, the reference's output file name (when
    ``datain`` names the list-mode file, or ``fout`` is given; nothing is
    written), the smoothed image, and None for the smoothed file name and the
    affine (which needs the bed position from the DICOM headers). Writing
    NIfTI files (``store_img``, ``store_itr``) is not part of this package.
    """
    from . import osem as _osem
    if store_img or store_itr:
        raise NotImplementedError('osemone: writing NIfTI images (store_img / store_itr) is not '
                                  'part of nipet_amd; save RecOut.im with a NIfTI library')
    if ret_sinos:
        raise NotImplementedError('osemone: ret_sinos returns the scatter model\'s sinograms; '
                                  'scatter is not modelled here (pass it as sctsino)')
    d = recon_inputs(datain, mumaps, hst, scanner_params, recmod, attnsino, sctsino, randsino,
                     normcomp, normsino)
    psfkernel = psf_config(psf, scanner_params['Cnt'])
    stime = time.time()
    state = _osem.osem_prepare(scanner_params, d['psng'], d['asng'], d['nsng'], d['rsng'],
                               d['ssng'], mask_radius=mask_radius, psf=psfkernel)
    img = _osem.osem_run(state, itr)
    log.info('recon time: %.3g', time.time() - stime)
    return rec_output(img, state.iterations, datain, hst, scanner_params, d['ncmp'],
                      decay_ref_time, fwhm, frmno, fcomment, outpath, fout)
