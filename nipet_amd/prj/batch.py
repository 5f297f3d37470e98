"""Batched reconstruction of bootstrap replicates and dynamic frames.

The reference reconstructs frames and bootstrap realisations one after the
other, each with its own list-mode pass and ``osemone`` call (``mmrchain``,
img/pipe.py:295-373; bootstrap weights in the histogrammer,
lm/src/hst.cu:110-139). Here nb reconstructions that share attenuation and
normalisation (hence the 14 sensitivity images) run as one batch:

* one list-mode pass histograms all bootstrap replicates (replicate-innermost
  counts, Philox counter = (word index / 4, replicate), so replicate r is the
  same whichever session produces it), and ``nipet_remgaps_u16b`` turns them into
  the batched uint16 prompts (Naw, 837, nb);
* dynamic frames take one histogram session per time window and carry their
  own randoms + scatter (rho per frame);
* the whole ``itr`` x 14 OSEM of all nb images is captured once as a HIP
  graph (``torch.cuda.CUDAGraph`` over the library's stream launches: no host
  synchronisation, allocation or memset inside) and runs as one graph launch;
* replicates are stored in contiguous groups of ``GROUP`` ([group][voxel][4],
  prompts [group][bin][837][4]): a kernel launch covers one group, so every
  cache line it touches holds only that group's replicates;
* the replicate sub-batch width of the FP and BP kernels (4, 2 or 1 images per
  launch), and the BP's run length at each width, are picked by timing the
  kernels on the problem's own geometry and group size (``autotune``).
"""
import numpy as np

from .. import _lib
from .osem import OsemBatch, osem_iteration_batch, osem_prepare_batch

__all__ = ['GROUP', 'autotune', 'prepare_groups', 'OsemGraph', 'bootstrap_prompts',
           'frame_prompts', 'bootstrap_osem', 'frames_osem', 'FRAMES24', 'replicates_and_frames']

# 24 dynamic frames of uneven length (14 x 10 s, 4 x 15 s, 5 x 20 s, 1 x 300 s)
# covering a 600 s acquisition
FRAMES24 = ([(10 * i, 10 * i + 10) for i in range(14)] +
            [(140 + 15 * i, 155 + 15 * i) for i in range(4)] +
            [(200 + 20 * i, 220 + 20 * i) for i in range(5)] +
            [(300, 600)])

GROUP = 4  # replicates per contiguous group (the widest kernel instance)


def _torch():
    import torch
    return torch


_TUNED = {}


def autotune(state, widths=(4, 2, 1), repeats=2, probe=4):
    """Time the emission FP and the BP kernels on the first subset of
    ``state`` (an ``OsemBatch``) over a probe batch of min(nb, ``probe``)
    replicates, and keep the fastest shape per kernel: the FP sub-batch width,
    and the BP sub-batch width together with its run length (segments per
    wave run: each width has two compiled run lengths). Results are cached
    per (geometry, subset size) for the process. Returns {'fprj': w, 'bprj': w,
    'bprj_run': {w: segments}, 'ms': {(kernel, w[, run]): ms per image}}."""
    import ctypes as ct
    torch = _torch()
    L = _lib.lib()
    key = (id(state.geom), len(state.subsets[0]))
    if key in _TUNED:
        res = _TUNED[key]
        _lib.check(L.nipet_tune_set(_lib.PROF_FPRJ, res['fprj']), 'nipet_tune_set')
        _lib.check(L.nipet_tune_set(_lib.PROF_BPRJ, res['bprj']), 'nipet_tune_set')
        for w, run in res['bprj_run'].items():
            _lib.check(L.nipet_tune_set_run(w, run), 'nipet_tune_set_run')
        return res
    st = _lib.stream_handle()
    subs = state.subsets[0]
    nprj = len(subs)
    nb = min(state.nb, probe)
    img = torch.ones((320, 320, 127, nb), dtype=torch.float32, device='cuda')
    sino = torch.empty((nprj, state.geom.nsn, nb), dtype=torch.float32, device='cuda')
    bimg = torch.empty_like(img)
    res = {'ms': {}, 'bprj_run': {}}

    def timed(launch, n):
        launch()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(repeats):
            launch()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / repeats / n

    def fprj(n):
        return lambda: _lib.check(L.nipet_fprj_b(state.geom.handle, _lib.ptr(sino), _lib.ptr(img),
                                                 _lib.ptr(subs), nprj, n, 0, st), 'nipet_fprj_b')

    def bprj(n):
        return lambda: _lib.check(L.nipet_bprj_b(state.geom.handle, _lib.ptr(bimg),
                                                 _lib.ptr(sino), _lib.ptr(subs), nprj, n, st),
                                  'nipet_bprj_b')
    best = None
    for w in widths:
        n = min(w, nb)
        _lib.check(L.nipet_tune_set(_lib.PROF_FPRJ, int(w)), 'nipet_tune_set')
        ms = timed(fprj(n), n)
        res['ms'][('fprj', w)] = ms
        if best is None or ms < best[1]:
            best = (w, ms)
    _lib.check(L.nipet_tune_set(_lib.PROF_FPRJ, best[0]), 'nipet_tune_set')
    res['fprj'] = best[0]
    best = None
    for w in widths:
        n = min(w, nb)
        _lib.check(L.nipet_tune_set(_lib.PROF_BPRJ, int(w)), 'nipet_tune_set')
        cur, alt = ct.c_int(), ct.c_int()
        _lib.check(L.nipet_tune_get_run(int(w), ct.byref(cur), ct.byref(alt)), 'tune_get_run')
        wbest = None
        for run in (cur.value, alt.value):
            _lib.check(L.nipet_tune_set_run(int(w), run), 'nipet_tune_set_run')
            ms = timed(bprj(n), n)
            res['ms'][('bprj', w, run)] = ms
            if wbest is None or ms < wbest[1]:
                wbest = (run, ms)
        _lib.check(L.nipet_tune_set_run(int(w), wbest[0]), 'nipet_tune_set_run')
        res['bprj_run'][int(w)] = wbest[0]
        if best is None or wbest[1] < best[1]:
            best = (w, wbest[1])
    _lib.check(L.nipet_tune_set(_lib.PROF_BPRJ, best[0]), 'nipet_tune_set')
    res['bprj'] = best[0]
    _TUNED[key] = res
    return res


def _per_rep(x):
    return x is not None and hasattr(x, 'dim') and x.dim() == 3


def prepare_groups(scanner_params, prompt_groups, asng, nsng, rsng=None, ssng=None, sens=None,
                   psf=None, mask_radius=29.):
    """One ``OsemBatch`` per replicate group (``prompt_groups[i]``: CUDA uint16
    (Naw, 837, g_i), the first group the largest). Attenuation and norm are
    shared; randoms / scatter are shared (Naw, 837) or per replicate
    (Naw, 837, nb) CUDA tensors, sliced by group. Sensitivity images, mask,
    shared rho and the work buffers are common to all groups. ``psf``: (3, 17)
    resolution-modelling kernel (psf_config) or None."""
    states, off, first = [], 0, None
    per = _per_rep(rsng) or _per_rep(ssng)
    for p in prompt_groups:
        g = p.shape[2]
        work = None if first is None else (first.work_sino, first.work_img, first.work_img2)
        if first is not None and not per:
            st = OsemBatch(first.geom, first.subsets, p, first.rho, False, first.sens,
                           first.mask, g, work=work, krnl=first.krnl, win=first.win)
        else:
            def sl(x):
                return x[..., off:off + g].contiguous() if _per_rep(x) else x
            st = osem_prepare_batch(
                scanner_params, p, asng, nsng, sl(rsng), sl(ssng), mask_radius=mask_radius,
                sens=sens if first is None else first.sens, work=work, psf=psf)
        if first is None:
            first = st
        states.append(st)
        off += g
    return states


class OsemGraph:
    """``niter`` OSEM iterations of every replicate group (``OsemBatch`` or a
    list of them) captured as one HIP graph on persistent image buffers
    ([320, 320, 127, g] per group); ``run`` resets the images (ones, or
    ``img0`` [320, 320, 127, nb]) and launches the graph."""

    def __init__(self, states, niter):
        torch = _torch()
        states = [states] if isinstance(states, OsemBatch) else list(states)
        if not states or not all(isinstance(st, OsemBatch) for st in states):
            raise TypeError('states must be OsemBatch objects')
        self.states, self.niter = states, int(niter)
        self.imgs = [torch.ones((320, 320, 127, st.nb), dtype=torch.float32, device='cuda')
                     for st in states]
        _lib.check(_lib.lib().nipet_prof_enable(0), 'nipet_prof_enable')
        # load every kernel instance before capture: one subset update over a
        # single transaxial bin per group (the images are reset before each run)
        for st, img in zip(states, self.imgs):
            first = st.subsets[0][:1].contiguous()
            _lib.check(_lib.lib().nipet_osem_subset_b(
                st.geom.handle, _lib.ptr(img), _lib.ptr(st.prompts), _lib.ptr(st.rho),
                int(st.rho_per_rep), _lib.ptr(st.sens[0]), 0, _lib.ptr(st.mask), None,
                _lib.ptr(first), 1, st.nb, _lib.ptr(st.work_sino), _lib.ptr(st.work_img),
                _lib.stream_handle()), 'nipet_osem_subset_b')
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            for it in range(self.niter):
                for st, img in zip(states, self.imgs):
                    osem_iteration_batch(st, img, first_full=it == 0)
        torch.cuda.synchronize()

    def run_groups(self, img0=None):
        """Reset the images (ones, or ``img0`` [320, 320, 127, nb]), launch the
        graph and return the per-group images [320, 320, 127, g] (the graph's
        own buffers: valid until the next run)."""
        off = 0
        for img in self.imgs:
            if img0 is None:
                img.fill_(1.0)
            else:
                img.copy_(img0[..., off:off + img.shape[3]])
            off += img.shape[3]
        self.graph.replay()
        return self.imgs

    def run(self, img0=None):
        """``run_groups`` joined into one new [320, 320, 127, nb] tensor."""
        return _join(self.run_groups(img0))


def _join(imgs):
    """Per-group images [320, 320, 127, g] copied into one new
    [320, 320, 127, sum g] tensor."""
    torch = _torch()
    nb = sum(x.shape[3] for x in imgs)
    out = torch.empty((320, 320, 127, nb), dtype=torch.float32, device='cuda')
    off = 0
    for x in imgs:
        out[..., off:off + x.shape[3]].copy_(x)
        off += x.shape[3]
    return out


def bootstrap_prompts(words, scanner_params, nrep, seed, rep0=0, t0=0, t1=0, hist=None,
                      group=None):
    """Batched uint16 prompts of the bootstrap replicates rep0 .. rep0+nrep-1
    of the frame [t0, t1) of a device word array (one list-mode pass): one
    (Naw, 837, nrep) tensor, or with ``group`` a list of (Naw, 837, g) tensors
    of consecutive replicate groups. Returns (prompts, LmHistogram)."""
    from ..lm.mmrhist import hist_words
    h = hist if hist is not None else hist_words(words, scanner_params, t0, t1, btp=2, seed=seed,
                                                 nrep=nrep, rep0=rep0)
    if group is None:
        return _remgaps_b(h, scanner_params, h.nrep, 0, h.nrep), h
    return [_remgaps_b(h, scanner_params, h.nrep, r0, min(group, h.nrep - r0))
            for r0 in range(0, h.nrep, group)], h


def _remgaps_b(h, scanner_params, nrep, r0, nb, out=None, out_b0=0):
    """uint16 prompts (Naw, nsn, nb) of replicates r0 .. r0+nb-1 of a histogram
    session, or written into columns out_b0 .. out_b0+nb-1 of ``out``."""
    torch = _torch()
    ali = torch.from_numpy(np.asarray(scanner_params['txLUT']['aw2ali'], np.int32)).cuda()
    naw = ali.numel()
    nsn = h.nsinos
    if out is None:
        out = torch.empty((naw, nsn, nb), dtype=torch.uint16, device='cuda')
    if out.shape[:2] != (naw, nsn) or not out.is_contiguous():
        raise ValueError(f'prompts output must be a contiguous ({naw}, {nsn}, n) tensor')
    _lib.check(_lib.lib().nipet_remgaps_u16b(_lib.ptr(out), _lib.ptr(h.out['psino']),
                                             _lib.ptr(ali), naw, nsn, 86688, nrep, r0, nb,
                                             out.shape[2], out_b0, _lib.stream_handle()),
               'nipet_remgaps_u16b')
    return out


def frame_prompts(words, scanner_params, frames):
    """Batched uint16 prompts (Naw, 837, nf) of the time frames [(t0, t1), ...]
    (seconds from the first tag) of a device word array, one histogram session
    per frame over the words of its window. Returns (prompts, [delayed totals])."""
    torch = _torch()
    from ..lm.mmrhist import LmHistogram, second_index, tag_range
    tf, tl = tag_range(words)
    nitag = ((tl - tf) + 999) // 1000
    for t0, t1 in frames:
        if not (0 <= t0 < t1 <= nitag):
            raise ValueError(f'frame ({t0}, {t1}) outside 0..{nitag} s')
    # word index of the first tag of every second (lm/src/lmaux.cu:17-200)
    first = second_index(words, tf, nitag)
    naw = len(scanner_params['txLUT']['aw2ali'])
    out = torch.empty((naw, scanner_params['Cnt']['NSN11'], len(frames)), dtype=torch.uint16,
                      device='cuda')
    dtot = []
    for f, (t0, t1) in enumerate(frames):
        # words from the first tag of second t0 to the first tag of second t1
        i0, i1 = int(first[t0]), int(first[t1])
        h = LmHistogram(scanner_params, tf, nitag, t0, t1, btp=0)
        h.add(words[i0:i1])
        _remgaps_b(h, scanner_params, 1, 0, 1, out=out, out_b0=f)
        dtot.append(int(h.out['dhc'].cpu().numpy().astype(np.int64).sum()))
        del h
    return out, dtot


def bootstrap_osem(words, scanner_params, asng, nsng, rsng=None, ssng=None, nrep=64, seed=1,
                   rep0=0, itr=4, t0=0, t1=0, tune=True, graph=True, sens=None, psf=None):
    """OSEM (``itr`` x 14) of bootstrap replicates rep0 .. rep0+nrep-1 of a
    device list-mode word array, batched; corrections (Naw, 837) shared by all
    replicates. Returns (images [320, 320, 127, nrep], info)."""
    import time
    torch = _torch()
    tm = {}

    def mark(name, t0=[None]):
        torch.cuda.synchronize()
        now = time.perf_counter()
        if t0[0] is not None:
            tm[name] = round(now - t0[0], 4)
        t0[0] = now

    mark(None)
    groups, h = bootstrap_prompts(words, scanner_params, nrep, seed, rep0, t0, t1,
                                  group=GROUP)
    del h
    mark('histogram_s')
    states = prepare_groups(scanner_params, groups, asng, nsng, rsng, ssng, sens=sens, psf=psf)
    st = states[0]
    mark('prepare_s')
    info = {'tune': autotune(st) if tune else None}
    mark('tune_s')
    if graph:
        g = OsemGraph(states, itr)
        mark('capture_s')
        img = g.run()
    else:
        imgs = [torch.ones((320, 320, 127, x.nb), dtype=torch.float32, device='cuda')
                for x in states]
        for it in range(itr):
            for x, im in zip(states, imgs):
                osem_iteration_batch(x, im, first_full=it == 0)
        img = _join(imgs)
    mark('osem_s')
    info['sens'] = st.sens
    info['timing'] = tm
    return img, info


def frames_osem(words, scanner_params, frames, asng, nsng, rho_frames, itr=4, tune=True,
                graph=True, sens=None, psf=None):
    """OSEM of dynamic frames [(t0, t1), ...] of a device word array, batched;
    attenuation / norm (Naw, 837) shared, ``rho_frames`` = randoms + scatter
    per frame (Naw, 837, nf). Returns (images [320, 320, 127, nf], info)."""
    torch = _torch()
    prompts, dtot = frame_prompts(words, scanner_params, frames)
    nf = prompts.shape[2]
    groups = [prompts[..., f0:f0 + GROUP].contiguous() for f0 in range(0, nf, GROUP)]
    states = prepare_groups(scanner_params, groups, asng, nsng, rho_frames, None, sens=sens,
                            psf=psf)
    info = {'tune': autotune(states[0]) if tune else None, 'delayed': dtot}
    if graph:
        img = OsemGraph(states, itr).run()
    else:
        imgs = [torch.ones((320, 320, 127, x.nb), dtype=torch.float32, device='cuda')
                for x in states]
        for it in range(itr):
            for x, im in zip(states, imgs):
                osem_iteration_batch(x, im, first_full=it == 0)
        img = _join(imgs)
    return img, info


def replicates_and_frames(words, scanner_params, asng, nsng, rs, nrep=64, seed=1,
                          frames=FRAMES24, itr=4, tune=True, psf=None):
    """Bootstrap replicates 0 .. nrep-1 of a whole acquisition and its dynamic
    frames [(t0, t1), ...], reconstructed ``itr`` x 14 as ONE batch: one
    list-mode pass for all replicates, one session per frame, shared
    attenuation / norm (Naw, 837) and sensitivity images, randoms + scatter
    ``rs`` (Naw, 837) of the whole acquisition for every replicate and
    rs * (t1 - t0) / duration per frame, and every subset update of every
    replicate and frame group captured in one HIP graph. Returns (list of
    nrep replicate images, list of frame images) as [320, 320, 127] views of
    the graph's buffers, and an info dict (autotune, timing)."""
    import time
    torch = _torch()
    from ..lm.mmrhist import tag_range
    tm = {}

    def mark(name, t0=[None]):
        torch.cuda.synchronize()
        now = time.perf_counter()
        if t0[0] is not None:
            tm[name] = round(now - t0[0], 4)
        t0[0] = now

    mark(None)
    rep_groups, h = bootstrap_prompts(words, scanner_params, nrep, seed, group=GROUP)
    del h
    mark('histogram_s')
    tf, tl = tag_range(words)
    dur = ((tl - tf) + 999) // 1000
    fprompts, dtot = frame_prompts(words, scanner_params, frames)
    nf = fprompts.shape[2]
    frho = torch.stack([rs * ((t1 - t0) / dur) for t0, t1 in frames], dim=2).contiguous()
    mark('frame_histograms_s')
    reps = prepare_groups(scanner_params, rep_groups, asng, nsng, rs, None, psf=psf)
    fgroups = [fprompts[..., f0:f0 + GROUP].contiguous() for f0 in range(0, nf, GROUP)]
    del fprompts
    first = reps[0]
    frm = []
    off = 0
    for p in fgroups:
        g = p.shape[2]
        frm.append(osem_prepare_batch(scanner_params, p, asng, nsng,
                                      frho[..., off:off + g].contiguous(), None,
                                      sens=first.sens, psf=psf,
                                      work=(first.work_sino, first.work_img, first.work_img2)))
        off += g
    del frho
    mark('prepare_s')
    info = {'tune': autotune(first) if tune else None, 'delayed': dtot}
    mark('tune_s')
    graph = OsemGraph(reps + frm, itr)
    mark('capture_s')
    imgs = graph.run_groups()
    mark('osem_s')
    out = [img[..., b] for img in imgs for b in range(img.shape[3])]
    info['timing'] = tm
    info['graph'] = graph
    return out[:nrep], out[nrep:], info
