"""Frame-by-frame reconstruction chain (``niftypet/nipet/img/pipe.py``).

``mmrchain`` keeps the reference's signature and output dictionary
(``img/pipe.py:20-373``): frames given as ``['fluid', [t0, t1], ...]``
(or ``'timings'``), ``['def', [n, dur], ...]`` or a list of durations
(``dynamic_timings``); each frame is histogrammed (``mmrhist``), frames whose
delayeds exceed 0.99 of the prompts are skipped (``img/pipe.py:317-324``),
and the others are reconstructed with the settings of ``osemone``; the
dynamic image stack is ``output['im']``.

Frames that share attenuation and normalisation (hence the 14 sensitivity
images) are reconstructed together as one batch: their prompts and randoms
(+ scatter) are stacked frame-innermost and the ``itr`` x 14 subset updates of
all frames run as one HIP graph launch (``prj/batch.py``). Each frame's image
is the one ``osemone`` gives for it alone (same arithmetic per frame, only
the order of the back projector's float atomics differs). ``batch=False``
reconstructs frame by frame with ``osemone``.

Not part of this path (the reference's NiftyReg / NIfTI / PVC tooling):
per-frame mu-map registration (``tAffine``), trimming / up-sampling
(``trim``), partial-volume correction (``pvcroi``) and NIfTI output: the file
names the reference would write are returned, nothing is written.
"""
import logging
import os
from numbers import Integral

import numpy as np

log = logging.getLogger(__name__)

__all__ = ['mmrchain', 'frame_timings']


def frame_timings(frames):
    """[[t0, t1], ...] of a frame definition in any of the reference's three
    forms (``img/pipe.py:95-125``)."""
    from ..lm.mmrhist import dynamic_timings
    if not isinstance(frames, list):
        raise TypeError('Wrong data type for dynamic frames')
    if (frames and isinstance(frames[0], str) and frames[0] in ('fluid', 'timings')
            and all(isinstance(t, list) and len(t) == 2 for t in frames[1:])):
        return [list(t) for t in frames[1:]]
    if (frames and isinstance(frames[0], str) and frames[0] == 'def'
            and all(isinstance(t, list) and len(t) == 2 for t in frames[1:])):
        return dynamic_timings(frames)['timings'][1:]
    if frames and all(isinstance(t, Integral) for t in frames):
        return dynamic_timings(frames)['timings'][1:]
    raise TypeError('frames must be a 1D list of durations [15, 15, 30, ...], '
                    "['def', [n, dur], ...] or ['fluid', [t0, t1], ...]")


def _mu(m, Cnt, what):
    """(exists, E7 image) of a mu-map given as a dict with 'im', an array or
    nothing (the reference's obtain_image, without its NIfTI reader)."""
    if m is None or (isinstance(m, (list, tuple)) and len(m) == 0):
        return False, np.zeros((Cnt['SO_IMZ'], Cnt['SO_IMY'], Cnt['SO_IMX']), np.float32)
    if isinstance(m, dict):
        if 'im' not in m:
            raise ValueError(f'{what}: a dictionary must hold the image under "im"')
        m = m['im']
    if isinstance(m, str):
        raise NotImplementedError(f'{what}: reading NIfTI files is not part of nipet_amd; '
                                  'pass the image array')
    im = np.asarray(m, np.float32)
    if im.shape != (Cnt['SO_IMZ'], Cnt['SO_IMY'], Cnt['SO_IMX']):
        raise ValueError(f'{what}: E7 image shape (127, 344, 344) expected, got {im.shape}')
    return True, im


def mmrchain(datain, scanner_params, outpath=None, fout=None, frames=None, mu_h=None, mu_o=None,
             tAffine=None, itr=4, fwhm=0., psf=None, recmod=-1, histo=None, decay_ref_time=None,
             trim=False, trim_scale=2, trim_interp=0, trim_memlim=True, trim_store=True,
             pvcroi=None, pvcreg_tool='niftyreg', store_rois=False, pvcpsf=None, pvcitr=5,
             fcomment='', ret_sinos=False, ret_histo=False, store_img=True,
             store_img_intrmd=False, store_itr=None, del_img_intrmd=False, batch=True,
             mask_radius=29., normcomp=None):
    """Reconstruct the frames of a list-mode acquisition (``img/pipe.py:20-373``).

    Returns the reference's dictionary: 'recmod', 'frames' ([t0, t1] list),
    '#frames', 'im' (the frames' E7 images, squeezed; the smoothed ones when
    ``fwhm > 0``; skipped frames stay zero), 'tuple' (the ``osemone`` result
    of a single frame), 'hst' (histograms by 't0-t1' with ``ret_histo``) and
    'fpet' (the file name the reference writes with ``store_img``; nothing is
    written). Beyond the reference: ``batch`` (reconstruct frames that share
    attenuation and normalisation as one graph launch), ``mask_radius`` and
    ``normcomp`` (normalisation components, else read from ``datain``)."""
    from ..lm.mmrhist import mmrhist
    from ..prj import mmrrec
    Cnt = scanner_params['Cnt']
    if tAffine is not None:
        raise NotImplementedError('mmrchain: per-frame mu-map registration (tAffine) needs '
                                  'NiftyReg, which is not part of nipet_amd')
    if trim or pvcroi:
        raise NotImplementedError('mmrchain: trimming and partial-volume correction are not '
                                  'part of nipet_amd')
    if ret_sinos:
        raise NotImplementedError('mmrchain: ret_sinos returns the scatter model\'s sinograms; '
                                  'scatter is not modelled here')
    if store_itr or store_img_intrmd:
        raise NotImplementedError('mmrchain: writing NIfTI images is not part of nipet_amd')
    if frames is None:
        frames = ['fluid', [0, 0]]
    # a given histogram fixes the frame (img/pipe.py:84-91)
    if histo is not None and 'psino' in histo:
        frames = ['fluid', [histo['t0'], histo['t1']]]
    else:
        if histo is not None:
            log.warning('the given histogram does not contain a prompt sinogram--will generate '
                        'a histogram.')
        histo = None
    t_frms = frame_timings(frames)
    nfrm = len(t_frms)
    if nfrm < 1:
        raise TypeError('Unrecognised/confusing time frames!')

    muo_ok, muo = _mu(mu_o, Cnt, 'object mu-map')
    muh_ok, muh = _mu(mu_h, Cnt, 'hardware mu-map')
    if muo_ok and muh_ok and recmod == -1:
        recmod = 3
    elif (muo_ok or muh_ok) and recmod == -1:
        recmod = 1
        log.warning('partial mu-map:  scatter correction is switched off.')
    elif recmod == -1:
        recmod = 0
        log.warning('no mu-map provided: scatter and attenuation corrections are switched off.')

    output = {'recmod': recmod, 'frames': t_frms, '#frames': nfrm}
    petdir = os.path.join(outpath, 'PET') if outpath is not None else (
        os.path.join(datain['corepath'], 'reconstructed') if datain.get('corepath') else '')
    petimg = os.path.join(petdir, 'multiple-frames' if nfrm > 1 else 'single-frame')
    dynim = np.zeros((nfrm, Cnt['SO_IMZ'], Cnt['SO_IMY'], Cnt['SO_IMX']), np.float32)
    hsts = {}
    # per-frame histograms and reconstruction inputs
    todo = []
    attn = None
    for ifrm, (t0, t1) in enumerate(t_frms):
        t0, t1 = int(t0), int(t1)
        log.info('dynamic frame times t0=%d, t1=%d:', t0, t1)
        hst = mmrhist(datain, scanner_params, t0=t0, t1=t1) if histo is None else histo
        if ret_histo:
            hsts[f'{t0}-{t1}'] = hst
        if np.sum(hst['dhc']) > 0.99 * np.sum(hst['phc']):
            log.warning('amount of randoms is the greater part of prompts => omitting '
                        'reconstruction of frame %d', ifrm)
            continue
        d = mmrrec.recon_inputs(datain, [muh, muo], hst, scanner_params, recmod,
                                attnsino=attn, normcomp=normcomp)
        if recmod > 0:
            attn = d['asng']          # the mu-maps are the same for every frame
        todo.append((ifrm, hst, d))

    frmno = (lambda i: f'_frm{i}') if nfrm > 1 else (lambda i: '')
    recs = {}
    groups = _batch_groups(todo, scanner_params) if (batch and len(todo) > 1) else \
        [[k] for k in range(len(todo))]
    for grp in groups:
        if len(grp) == 1:
            ifrm, hst, d = todo[grp[0]]
            recs[ifrm] = _single(datain, hst, d, scanner_params, itr, fwhm, psf, mask_radius,
                                 decay_ref_time, frmno(ifrm), fcomment, petimg, fout)
        else:
            recs.update(_batched([todo[k] for k in grp], datain, scanner_params, itr, fwhm, psf,
                                 mask_radius, decay_ref_time, frmno, fcomment, petimg, fout))
    for ifrm, rec in recs.items():
        dynim[ifrm] = rec.imsmo if fwhm > 0 else rec.im
        if nfrm == 1:
            output['tuple'] = rec
    output['im'] = np.squeeze(dynim)
    if ret_histo:
        output['hst'] = hsts
    if store_img and recs:
        last = recs[max(recs)]
        base = os.path.basename(last.fpet)[:8] if last.fpet else 'nipet'
        if nfrm == 1:
            t0, t1 = t_frms[0]
            name = (os.path.basename(fout) + '.nii.gz' if fout is not None else
                    f'{base}_t-{t0}-{t1}sec_itr-{itr}{fcomment}.nii.gz')
        else:
            name = (os.path.basename(fout).split('.')[0] + f'_nfrm-{nfrm}.nii.gz'
                    if fout is not None else f'{base}_nfrm-{nfrm}_itr-{itr}{fcomment}.nii.gz')
        output['fpet'] = os.path.join(petimg, name)
    return output


def _single(datain, hst, d, scanner_params, itr, fwhm, psf, mask_radius, decay_ref_time, frmno,
            fcomment, outpath, fout):
    from ..prj import mmrrec
    from ..prj import osem as _osem
    st = _osem.osem_prepare(scanner_params, d['psng'], d['asng'], d['nsng'], d['rsng'],
                            d['ssng'], mask_radius=mask_radius,
                            psf=mmrrec.psf_config(psf, scanner_params['Cnt']))
    img = _osem.osem_run(st, itr)
    return mmrrec.rec_output(img, st.iterations, datain, hst, scanner_params, d['ncmp'],
                             decay_ref_time, fwhm, frmno, fcomment, outpath, fout)


def _as_dev(x, dtype=None):
    import torch
    t = x if isinstance(x, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(x))
    t = t.to('cuda')
    return t if dtype is None else t.to(dtype)


def _batch_groups(todo, scanner_params):
    """Indices of frames that can share one batch: span 11 and equal
    attenuation x normalisation (the same sensitivity images)."""
    import torch
    if scanner_params['Cnt']['SPN'] != 11:
        return [[k] for k in range(len(todo))]
    groups, keys = [], []
    for k, (_, _, d) in enumerate(todo):
        an = _as_dev(d['asng'], torch.float32) * _as_dev(d['nsng'], torch.float32)
        for g, key in zip(groups, keys):
            if torch.equal(key, an):
                g.append(k)
                break
        else:
            groups.append([k])
            keys.append(an)
    return groups


def _batched(items, datain, scanner_params, itr, fwhm, psf, mask_radius, decay_ref_time, frmno,
             fcomment, outpath, fout):
    """Frames sharing attenuation and normalisation as one batched graph
    launch; frames whose image sum falls below 0.1 (the reference's early stop,
    prj/mmrrec.py:357-361) are reconstructed again alone."""
    import torch

    from ..prj import mmrrec
    from ..prj.batch import GROUP, OsemGraph, autotune, prepare_groups
    from ..prj.osem import prompts_u16
    d0 = items[0][2]
    nf = len(items)
    prompts = torch.stack([prompts_u16(_as_dev(d['psng'])) for _, _, d in items], dim=2)
    naw, nsn = prompts.shape[:2]
    rs = torch.zeros((naw, nsn, nf), dtype=torch.float32, device='cuda')
    for f, (_, _, d) in enumerate(items):
        for x in (d['rsng'], d['ssng']):
            if x is not None:
                rs[..., f] += _as_dev(x, torch.float32)
    groups = [prompts[..., f0:f0 + GROUP].contiguous() for f0 in range(0, nf, GROUP)]
    krnl = mmrrec.psf_config(psf, scanner_params['Cnt'])
    states = prepare_groups(scanner_params, groups, d0['asng'], d0['nsng'], rs, None, psf=krnl,
                            mask_radius=mask_radius)
    autotune(states[0])
    imgs = OsemGraph(states, itr).run()
    out = {}
    for f, (ifrm, hst, d) in enumerate(items):
        img = imgs[..., f].contiguous()
        if float(torch.nansum(img)) < 0.1:
            out[ifrm] = _single(datain, hst, d, scanner_params, itr, fwhm, psf, mask_radius,
                                decay_ref_time, frmno(ifrm), fcomment, outpath, fout)
            continue
        out[ifrm] = mmrrec.rec_output(img, itr, datain, hst, scanner_params, d['ncmp'],
                                      decay_ref_time, fwhm, frmno(ifrm), fcomment, outpath, fout)
    return out
