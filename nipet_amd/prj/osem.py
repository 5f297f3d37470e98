"""Device-resident OSEM (the ``petprj.osem`` loop, ``prj/src/recon.cu:318-521``,
driven by ``osemone``, ``prj/mmrrec.py:279-389``).

Differences in structure (not in arithmetic) from the reference:
* prompts, corrections, sensitivity images, mask and image are uploaded once
  and stay in HBM for all iterations (the reference re-uploads about 1.8 GB per
  iteration, ``prj/src/recon.cu:333-419``);
* the transaxial intersection lists are built once per geometry;
* per subset, three launches: FP fused with ``y / (FP + rho)``, the gather BP,
  and the fused ``img * BP / sens`` masked update (``nipet_osem_subset``);
* without resolution modelling (span 11), a subset's sensitivity image is
  back-projected in its first update, in the same launch as the ratio
  (``nipet_osem_subset_sens``: two images per traversal of the rays), not in a
  separate pass before the loop (``prj/mmrrec.py:291-298``). The images are
  the same; reading ``OsemState.sens`` completes any missing ones.
"""
import logging

import numpy as np

from .. import _lib
from ..geom import get_geom

log = logging.getLogger(__name__)

__all__ = ['OsemState', 'osem_prepare', 'osem_run', 'sensitivity', 'cylinder_mask', 'mask_window',
           'prompts_u16']


NVOX = 320 * 320 * 127


def _torch():
    import torch
    return torch


def _longest_first(geom, subs):
    """A subset's bins ordered by decreasing segment count (device int32).
    Workgroups are dispatched in bin order, so the long rays start first and
    the last wave of workgroups holds short ones (a shorter tail). FP and BP
    walk the same list, so the OSEM result does not depend on the order.
    Device subsets are ordered on the device (no host copy, which would wait
    for all queued work); host ones on the host."""
    torch = _torch()
    if isinstance(subs, torch.Tensor) and subs.is_cuda:
        src = subs.to(torch.int32).contiguous()
        out = torch.empty_like(src)
        _lib.check(_lib.lib().nipet_bins_longest_first(geom.handle, _lib.ptr(src), src.numel(),
                                                       _lib.ptr(out), _lib.stream_handle()),
                   'nipet_bins_longest_first')
        return out
    rows = subs.cpu().numpy() if isinstance(subs, torch.Tensor) else np.asarray(subs)
    segc = geom.seg_counts()
    valid = (rows >= 0) & (rows < len(segc))
    key = np.where(valid, segc[np.clip(rows, 0, len(segc) - 1)], -1)
    order = np.argsort(-key, kind='stable')
    return _dev(np.ascontiguousarray(rows[order], np.int32), torch.int32)


_MASKS = {}


def cylinder_mask(Cnt, radius=29.):
    """Device uint8 mask (320, 320, nz) equal to ``get_cylinder(Cnt, rad=radius,
    gpu_dim=True) > 0.9`` (the default OSEM mask, ``prj/mmrrec.py:303``),
    made from the one transaxial disc broadcast over the planes on the device and
    cached per device. Callers must not modify it."""
    torch = _torch()
    from ..img.mmrimg import cylinder_disc
    nvz = Cnt.get('rSO_IMZ', Cnt['SO_IMZ'])
    key = (float(radius), Cnt['SO_IMX'], Cnt['SZ_IMX'], float(Cnt['SO_VXY']), nvz,
           torch.cuda.current_device())
    m = _MASKS.get(key)
    if m is None:
        margin = (Cnt['SO_IMX'] - Cnt['SZ_IMX']) // 2
        disc = cylinder_disc(Cnt, rad=radius)[margin:margin + Cnt['SZ_IMY'],
                                              margin:margin + Cnt['SZ_IMX']] > 0.9
        d = torch.from_numpy(np.ascontiguousarray(disc, np.uint8)).to('cuda')
        m = d[:, :, None].expand(d.shape[0], d.shape[1], nvz).contiguous()
        _MASKS[key] = m
    return m


def mask_window(geom, mask, psf=False):
    """Device int32 (Naw, 2) segment windows of a uint8 device mask
    (``nipet_osem_window``): the first segment and the count of each bin's
    transaxial list whose columns hold a mask voxel; with resolution modelling
    the columns are dilated by the kernel's reach (8 voxels)."""
    torch = _torch()
    win = torch.empty((geom.naw, 2), dtype=torch.int32, device='cuda')
    _lib.check(_lib.lib().nipet_osem_window(geom.handle, _lib.ptr(mask), 8 if psf else 0,
                                            _lib.ptr(win), _lib.stream_handle()),
               'nipet_osem_window')
    return win


def prompts_u16(psng):
    """Prompt counts as uint16, saturated to 0 .. 65535 (fractional values
    truncated), for NumPy arrays and torch tensors alike (the result stays in
    the input's library and device). ``petprj.osem`` takes uint16 prompts
    (``prj/src/prj_module.cu:649-651``); the reference's histogrammer packs
    16-bit counts, which wrap at 65536 (``lm/src/hst.cu:173``), whereas the
    histogram here counts in 32 bits and saturates on the way to uint16."""
    torch = _torch()
    if isinstance(psng, torch.Tensor):
        if psng.dtype == torch.uint16:
            return psng
        x = psng.trunc() if psng.is_floating_point() else psng.to(torch.int64)
        return x.clamp(0, 65535).to(torch.int32).to(torch.uint16)
    x = np.asarray(psng)
    if x.dtype == np.uint16:
        return x
    if x.dtype.kind == 'f':
        x = np.trunc(x)
    return np.clip(x, 0, 65535).astype(np.uint16)


def _dev(x, dtype):
    torch = _torch()
    if isinstance(x, torch.Tensor):
        return x.to('cuda', dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(x)).to('cuda', dtype=dtype)


class OsemState:
    """Everything one OSEM reconstruction keeps on the device."""

    def __init__(self, geom, subsets, prompts, rho, an, sens, mask, krnl=None, pending=None,
                 win=None):
        torch = _torch()
        self.geom = geom
        self.subsets = subsets
        # the mask's segment windows (nipet_osem_window), used once the image
        # is zero outside the mask
        self.win = win
        self.nprj = max(len(s) for s in subsets)
        self.prompts, self.rho, self.an, self._sens, self.mask = prompts, rho, an, sens, mask
        # subsets whose sensitivity image is still to be back-projected (in
        # their first update)
        self.pending = set() if pending is None else set(pending)
        self.work_sino = torch.empty((self.nprj, geom.nsn), dtype=torch.float32, device='cuda')
        self.work_img = torch.empty((320, 320, 127), dtype=torch.float32, device='cuda')
        # resolution modelling: host kernel (3, 17) and a second work image
        self.krnl = krnl
        self.work_img2 = None if krnl is None else torch.empty_like(self.work_img)
        self._pair = None

    @property
    def sens(self):
        """Sensitivity images [nsub, 320, 320, 127], completing any not yet
        back-projected. (An image back-projected inside an update walked the
        mask's segment windows: it is exact on the mask, which is all the
        update reads.)"""
        for k in sorted(self.pending):
            sensitivity(self.geom, self.an, self.subsets[k], out=self._sens[k])
        self.pending.clear()
        return self._sens

    def pair_work(self):
        """Work buffers of the two-image first update (allocated while subsets
        are pending)."""
        torch = _torch()
        if self._pair is None:
            self._pair = (torch.empty((self.nprj, self.geom.nsn, 2), dtype=torch.float32,
                                      device='cuda'),
                          torch.empty((320, 320, 127, 2), dtype=torch.float32, device='cuda'))
        return self._pair


def sensitivity(geom, an, subs, out=None):
    """BP_subset(a*n) (``prj/mmrrec.py:291-298``) for one subset (device tensors:
    ``an`` the full (Naw, nsn) a*n sinogram, read at the subset's bins)."""
    torch = _torch()
    if out is None:
        out = torch.empty((320, 320, 127), dtype=torch.float32, device='cuda')
    _lib.check(_lib.lib().nipet_osem_sens(geom.handle, _lib.ptr(out), _lib.ptr(an),
                                          _lib.ptr(subs), len(subs), _lib.stream_handle()),
               'nipet_osem_sens')
    return out


def _psf_kernel(psf):
    """Host float32 (3, 17) kernel, or None when resolution modelling is off
    (``psf`` None or ``psf[0, 0] < 0``, the reference's switch)."""
    if psf is None:
        return None
    k = np.ascontiguousarray(np.asarray(psf, dtype=np.float32))
    if k.shape != (3, 17):
        raise ValueError('psf must be a (3, 17) separable kernel (psf_config)')
    return None if k[0, 0] < 0 else k


def psf_conv(img, krnl, out=None, tmp=None):
    """Resolution-model convolution of device image(s) [320, 320, 127(, nb)]
    with a (3, 17) kernel (``nipet_psf_conv``); returns ``out``."""
    torch = _torch()
    k = _psf_kernel(krnl)
    if k is None:
        raise ValueError('psf kernel switched off')
    nb = 1 if img.dim() == 3 else img.shape[3]
    out = torch.empty_like(img) if out is None else out
    tmp = torch.empty_like(img) if tmp is None else tmp
    _lib.check(_lib.lib().nipet_psf_conv(_lib.ptr(out), _lib.ptr(img), _lib.ptr(tmp),
                                         k.ctypes.data, nb, _lib.stream_handle()),
               'nipet_psf_conv')
    return out


def osem_prepare(scanner_params, psng, asng, nsng, rsng=None, ssng=None, mask_radius=29.,
                 subsets=None, mask=None, psf=None):
    """Upload the data and compute rho = (r+s)/(a n), a n and the per-subset
    sensitivity images. Sinograms are device-shaped (Naw, nsn); prompts are
    converted to uint16 as in ``petprj.osem`` (``prj/src/prj_module.cu:649-651``).
    ``psf``: (3, 17) kernel from ``psf_config`` (resolution modelling: the
    sensitivity images are convolved once, ``prj/src/recon.cu:441-445``)."""
    torch = _torch()
    from .mmrrec import get_subsets14
    geom = get_geom(scanner_params)
    Cnt = scanner_params['Cnt']
    shape = (geom.naw, geom.nsn)
    if isinstance(psng, torch.Tensor):
        prompts = prompts_u16(psng.to('cuda')).contiguous()
    else:
        prompts = _dev(prompts_u16(psng), torch.uint16)
    if tuple(prompts.shape) != shape:
        raise ValueError(f'prompts must be device-shaped {shape}')
    a = _dev(asng, torch.float32)
    n = _dev(nsng, torch.float32)
    r = None if rsng is None else _dev(rsng, torch.float32)
    s = None if ssng is None else _dev(ssng, torch.float32)
    for x in (a, n, r, s):
        if x is not None and tuple(x.shape) != shape:
            raise ValueError(f'corrections must be device-shaped {shape}')
    rho = torch.empty(shape, dtype=torch.float32, device='cuda')
    an = torch.empty(shape, dtype=torch.float32, device='cuda')
    _lib.check(_lib.lib().nipet_osem_corr(_lib.ptr(rho), _lib.ptr(an), _lib.ptr(a), _lib.ptr(n),
                                          _lib.ptr(r), _lib.ptr(s), rho.numel(),
                                          _lib.stream_handle()), 'nipet_osem_corr')
    if subsets is None:
        subsets = [get_subsets14(k, scanner_params)[0] for k in range(14)]
    subs = [_longest_first(geom, x) for x in subsets]
    krnl = _psf_kernel(psf)
    if krnl is None and geom.nsn <= 837:
        # back-projected with each subset's first update
        sens = torch.empty((len(subs), 320, 320, 127), dtype=torch.float32, device='cuda')
        pending = range(len(subs))
    else:
        sens = torch.empty((len(subs), 320, 320, 127), dtype=torch.float32, device='cuda')
        for k, x in enumerate(subs):
            sensitivity(geom, an, x, out=sens[k])
        pending = None
    if krnl is not None:
        tmp = torch.empty((320, 320, 127), dtype=torch.float32, device='cuda')
        for k in range(len(subs)):
            tmp.copy_(sens[k])
            psf_conv(tmp, krnl, out=sens[k], tmp=tmp)
    if mask is None:
        mask = cylinder_mask(Cnt, mask_radius)
    msk = _dev(mask if isinstance(mask, torch.Tensor) else np.asarray(mask, np.uint8),
               torch.uint8)
    return OsemState(geom, subs, prompts, rho, an, sens, msk, krnl, pending,
                     mask_window(geom, msk, krnl is not None))


def osem_iteration(state, img, first_full=True):
    """One OSEM iteration (all subsets) in place on the device image ``img``.
    Every update after the first walks the mask's segment windows (the image
    is then zero outside the mask); the first does too when ``first_full`` is
    False (an image already masked by an earlier update)."""
    st = _lib.stream_handle()
    L = _lib.lib()
    for k, subs in enumerate(state.subsets):
        win = None if (k == 0 and first_full) else _lib.ptr(state.win)
        if state.krnl is not None:
            _lib.check(L.nipet_osem_subset_rm_b(
                state.geom.handle, _lib.ptr(img), _lib.ptr(state.prompts), _lib.ptr(state.rho), 0,
                _lib.ptr(state.sens[k]), 0, _lib.ptr(state.mask), win, _lib.ptr(subs), len(subs), 1,
                state.krnl.ctypes.data, _lib.ptr(state.work_sino), _lib.ptr(state.work_img),
                _lib.ptr(state.work_img2), st), 'nipet_osem_subset_rm_b')
            continue
        if k in state.pending:
            ws2, wi2 = state.pair_work()
            _lib.check(L.nipet_osem_subset_sens(
                state.geom.handle, _lib.ptr(img), _lib.ptr(state.prompts), _lib.ptr(state.rho),
                _lib.ptr(state.an), _lib.ptr(state._sens[k]), _lib.ptr(state.mask), win,
                _lib.ptr(subs), len(subs), _lib.ptr(state.work_sino), _lib.ptr(ws2),
                _lib.ptr(wi2), st), 'nipet_osem_subset_sens')
            state.pending.discard(k)
            if not state.pending:
                state._pair = None
            continue
        _lib.check(L.nipet_osem_subset(state.geom.handle, _lib.ptr(img), _lib.ptr(state.prompts),
                                       _lib.ptr(state.rho), _lib.ptr(state._sens[k]),
                                       _lib.ptr(state.mask), win, _lib.ptr(subs), len(subs),
                                       _lib.ptr(state.work_sino), _lib.ptr(state.work_img), st),
                   'nipet_osem_subset')


def osem_run(state, niter, img0=None, check_counts=True):
    """``niter`` OSEM iterations from ``img0`` (ones by default, ``prj/mmrrec.py:306``).
    Stops early, like ``prj/mmrrec.py:357-361``, when the image sum drops below 0.1."""
    torch = _torch()
    img = torch.ones((320, 320, 127), dtype=torch.float32, device='cuda') if img0 is None \
        else _dev(img0, torch.float32).clone()
    state.iterations = 0
    for k in range(niter):
        osem_iteration(state, img, first_full=k == 0)
        state.iterations = k + 1
        if check_counts and float(torch.nansum(img)) < 0.1:
            log.warning('it seems there is not enough true data to render reasonable image')
            state.iterations = k      # as prj/mmrrec.py:357-361
            break
    return img


class OsemBatch:
    """Device state of nb reconstructions that share geometry, attenuation x
    norm (hence sensitivity images) and mask: bootstrap replicates of one frame
    (shared rho) or frames of one scan (rho per frame). Images are
    [320, 320, 127, nb] (replicate innermost)."""

    def __init__(self, geom, subsets, prompts, rho, rho_per_rep, sens, mask, nb, work=None,
                 krnl=None, win=None):
        torch = _torch()
        self.geom, self.subsets, self.nb = geom, subsets, nb
        self.krnl = krnl
        self.win = mask_window(geom, mask, krnl is not None) if win is None else win
        self.prompts, self.rho, self.rho_per_rep, self.sens, self.mask = (
            prompts, rho, rho_per_rep, sens, mask)
        self.nprj = max(len(s) for s in subsets)
        if work is not None:  # (work_sino, work_img[, work_img2]) of another batch
            ws, wi = work[0], work[1]
            if ws.numel() < self.nprj * geom.nsn * nb or wi.numel() < NVOX * nb:
                raise ValueError('shared work buffers are too small for this batch')
            self.work_sino, self.work_img = ws, wi
        else:
            self.work_sino = torch.empty((self.nprj, geom.nsn, nb), dtype=torch.float32,
                                         device='cuda')
            self.work_img = torch.empty((320, 320, 127, nb), dtype=torch.float32,
                                        device='cuda')
        self.work_img2 = None
        if krnl is not None:  # resolution modelling: a second work image
            self.work_img2 = (work[2] if work is not None and len(work) > 2 and
                              work[2] is not None and work[2].numel() >= NVOX * nb else
                              torch.empty((320, 320, 127, nb), dtype=torch.float32,
                                          device='cuda'))


def osem_prepare_batch(scanner_params, prompts, asng, nsng, rsng=None, ssng=None,
                       mask_radius=29., subsets=None, mask=None, sens=None, work=None,
                       psf=None):
    """Batch state: ``prompts`` (Naw, nsn, nb) (uint16 or counts), shared
    attenuation / norm (Naw, nsn); randoms / scatter either shared (Naw, nsn) or
    per replicate (Naw, nsn, nb). ``sens`` may be passed from an earlier batch
    with the same a n (a (14, 320, 320, 127) tensor)."""
    torch = _torch()
    from .mmrrec import get_subsets14
    geom = get_geom(scanner_params)
    Cnt = scanner_params['Cnt']
    p = prompts if isinstance(prompts, torch.Tensor) else torch.from_numpy(prompts_u16(prompts))
    p = prompts_u16(p.to('cuda')).contiguous()
    if p.dim() != 3 or tuple(p.shape[:2]) != (geom.naw, geom.nsn):
        raise ValueError(f'prompts must be (Naw, nsn, nb) = ({geom.naw}, {geom.nsn}, nb)')
    nb = p.shape[2]
    a = _dev(asng, torch.float32)
    n = _dev(nsng, torch.float32)
    per_rep = any(x is not None and tuple(x.shape) == (geom.naw, geom.nsn, nb)
                  for x in (rsng, ssng))
    shape = (geom.naw, geom.nsn, nb) if per_rep else (geom.naw, geom.nsn)

    def add(x):
        if x is None:
            return None
        x = _dev(x, torch.float32)
        return x if x.dim() == len(shape) else x[..., None].expand(shape).contiguous()

    r, s = add(rsng), add(ssng)
    rs = None if (r is None and s is None) else ((r if r is not None else 0) +
                                                   (s if s is not None else 0)).contiguous()
    an = (a * n).contiguous()
    anb = an if not per_rep else an[..., None].expand(shape).contiguous()
    rho = torch.empty(shape, dtype=torch.float32, device='cuda')
    tmp = torch.empty(shape, dtype=torch.float32, device='cuda')
    one = torch.ones(shape, dtype=torch.float32, device='cuda')
    _lib.check(_lib.lib().nipet_osem_corr(_lib.ptr(rho), _lib.ptr(tmp), _lib.ptr(anb),
                                          _lib.ptr(one), _lib.ptr(rs), None, rho.numel(),
                                          _lib.stream_handle()), 'nipet_osem_corr')
    del tmp, one
    if subsets is None:
        subsets = [get_subsets14(k, scanner_params)[0] for k in range(14)]
    subs = [_longest_first(geom, x) for x in subsets]
    krnl = _psf_kernel(psf)
    if sens is None:
        sens = torch.empty((len(subs), 320, 320, 127), dtype=torch.float32, device='cuda')
        for k, x in enumerate(subs):
            sensitivity(geom, an, x, out=sens[k])
        if krnl is not None:
            tmp = torch.empty((320, 320, 127), dtype=torch.float32, device='cuda')
            for k in range(len(subs)):
                tmp.copy_(sens[k])
                psf_conv(tmp, krnl, out=sens[k], tmp=tmp)
    if mask is None:
        mask = cylinder_mask(Cnt, mask_radius)
    msk = _dev(mask if isinstance(mask, torch.Tensor) else np.asarray(mask, np.uint8),
               torch.uint8)
    return OsemBatch(geom, subs, p, rho, per_rep, sens, msk, nb, work=work, krnl=krnl)


def osem_iteration_batch(state, img, first_full=True):
    """One OSEM iteration of all nb images; windows as ``osem_iteration``."""
    st = _lib.stream_handle()
    L = _lib.lib()
    for k, subs in enumerate(state.subsets):
        win = None if (k == 0 and first_full) else _lib.ptr(state.win)
        if state.krnl is not None:
            _lib.check(L.nipet_osem_subset_rm_b(
                state.geom.handle, _lib.ptr(img), _lib.ptr(state.prompts), _lib.ptr(state.rho),
                int(state.rho_per_rep), _lib.ptr(state.sens[k]), 0, _lib.ptr(state.mask), win,
                _lib.ptr(subs), len(subs), state.nb, state.krnl.ctypes.data,
                _lib.ptr(state.work_sino), _lib.ptr(state.work_img), _lib.ptr(state.work_img2),
                st), 'nipet_osem_subset_rm_b')
            continue
        _lib.check(L.nipet_osem_subset_b(
            state.geom.handle, _lib.ptr(img), _lib.ptr(state.prompts), _lib.ptr(state.rho),
            int(state.rho_per_rep), _lib.ptr(state.sens[k]), 0, _lib.ptr(state.mask), win,
            _lib.ptr(subs), len(subs), state.nb, _lib.ptr(state.work_sino),
            _lib.ptr(state.work_img), st), 'nipet_osem_subset_b')


def osem_run_batch(state, niter, img0=None):
    """``niter`` OSEM iterations of all nb images (ones by default); returns
    [320, 320, 127, nb]."""
    torch = _torch()
    img = torch.ones((320, 320, 127, state.nb), dtype=torch.float32, device='cuda') \
        if img0 is None else _dev(img0, torch.float32).clone()
    for k in range(niter):
        osem_iteration_batch(state, img, first_full=k == 0)
    return img
