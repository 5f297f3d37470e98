#!/usr/bin/env python
"""Benchmark: FP+BP pairs/s of a full mMR span-11 OSEM reconstruction on one
MI355X (the workload of the reference's osemone -> petprj.osem path).

One step = one complete reconstruction from data already resident in HBM:
corrections rho = (r+s)/(a n) and a n, the 14 subset sensitivity images
BP_s(a n), and 4 iterations x 14 subsets of FP (fused with y/(FP+rho)), BP and
the masked multiplicative update. value = 56 FP+BP pairs per step / step time
(the 14 sensitivity back projections are inside the timed step but not
counted as pairs).

Data (synthetic, seeded): device-layout phantom = water ellipsoid (activity 1)
plus 20 spheres r U(0.5, 2.5) cm activity U(0.5, 4); mu-map water 0.096/cm
with a bone shell 0.15/cm; prompts = Poisson(scale (a n FP(phantom)) + randoms
+ scatter) at 3e8 expected counts; norm U(0.8, 1.2); randoms flat 12 % and
scatter smooth 25 % of the trues.

Roofline (field "roofline"): the dominant kernel's HBM throughput from its
rocprofv3-measured DRAM bytes per launch (profiles/pmc_<kernel>_osem4x14.json)
over its live mean launch time (HIP events on the library's stream), against
8 TB/s; the SQ instruction counters of the same kernel
(profiles/sq_<kernel>_osem4x14.json) over the same time give the issue rates
that actually bind it; the per-term byte model is reported as model_GBs.

Batched bootstrap + dynamic frames (extra field "batched", skipped by
--no-batch): a synthetic 600 s list-mode stream of 6e8 words (1 ms tags, 85 %
prompts drawn from the expected prompts a n FP(phantom) + randoms + scatter,
15 % delayeds) is histogrammed once into 64 Poisson(1) bootstrap replicates
and cut into 24 dynamic frames of 10-300 s; all 88 are reconstructed 4x14 as
one batch (one HIP graph launch, autotuned sub-batch widths) and compared
with plain reconstructions of single replicates and frames (one histogram
session, osem_prepare, osem_run each), whose images they must equal. The
"histogram" field times the histogramming of the stream alone (plain and with
64 bootstrap replicates).

CPU baseline (field "cpu_baseline", skipped by --no-cpu): the float32 OpenMP
transcription of the reference's tracer on the host cores, one full subset
FP + BP pair (median of 3 after a warm-up), plus the reference's histogram
loop over the first 3e7 words of the stream.

Autograd (extra field "autograd", --map-steps 0 skips): 200 gradient MAP steps
of the Poisson likelihood + quadratic prior through the differentiable
projectors, one checkpoint per subset.

Corrections (extra field "corrections", skipped by --no-corr): randoms from
fan-sums, the normalisation sinogram from components and one PSF convolution
of the phantom, each timed with HIP events.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batch]
                       [--no-corr] [--map-steps N]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP+BP pairs/s & % HBM peak (mMR span-11); batched bootstrap recons/s vs plain"
HBM_PEAK_GBS = 8000.0
CLOCK_HZ, N_CU = 2.4e9, 256   # MI355X peak engine clock, compute units
NITER, NSUB = 4, 14


_T0 = time.time()


def log(msg):
    print(f'[bench {time.time() - _T0:7.1f}s] {msg}', file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--no-batch', action='store_true',
                    help='skip the list-mode legs (histogram and batched reconstruction)')
    ap.add_argument('--no-recon', action='store_true',
                    help='list-mode legs: histogram only, no batched reconstruction')
    ap.add_argument('--map-steps', type=int, default=200,
                    help='gradient MAP steps of the autograd leg (0: skip)')
    ap.add_argument('--no-corr', action='store_true',
                    help='skip the randoms / normalisation / PSF timing leg')
    ap.add_argument('--nrep', type=int, default=64, help='bootstrap replicates of the batch')
    ap.add_argument('--lm-words', type=float, default=6e8, help='list-mode words (600 s)')
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='BYTES',
                    help='only run a plain device copy of BYTES (rocprof counter scaling)')
    return ap.parse_args()


def traffic_per_launch(kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), read from
    profiles/pmc_<kernel>_osem4x14.json when that summary exists."""
    path = os.path.join(ROOT, 'profiles', f'pmc_{kernel}_osem4x14.json')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get('hbm_bytes_per_launch')


def cpu_baseline(params, words=None, runs=3):
    """The float32 OpenMP transcription of the reference's tracer
    (oracle/cpu_tracer.cpp: sddn_tx tables, fprj/bprj_drct+oblq with
    per-thread BP images) timed on the host cores: one FP + BP pair over all
    4,894 bins of subset 0 of a random image, one warm-up pair then the median
    of `runs`, as pairs/s. Beside it, the reference's histogram loop (hst)
    over a bounded sample of the list-mode stream (words/s), and the FP64
    oracle over a few subset-0 rays (information only)."""
    import numpy as np

    import nipet_amd as nip
    from oracle import cpu_tracer as C
    from oracle import oracle as O
    sub = nip.get_subsets14(0, params)[0]
    rng = np.random.default_rng(0)
    img = rng.random((320, 320, 127), dtype=np.float32)
    T = C.Tracer(params)
    fp, bp = T.time_pair(img, sub, runs=runs)
    t_pair = float(np.median(fp + bp))
    info = C.host_info()
    out = {'value': round(1.0 / t_pair, 4), 'unit': 'pairs/s', 'cores': info['omp_threads'],
           'kind': 'port',
           'sample': f'float32 OpenMP transcription of the reference tracer: FP + BP over all '
                     f'{len(sub)} bins x 837 sinograms of subset 0 (one pair), warm-up + median '
                     f'of {runs}; FP {np.median(fp):.3f} s, BP {np.median(bp):.3f} s',
           'nproc': info['nproc'], 'cpu_model': info['cpu_model']}
    if words is not None:
        from nipet_amd.lm.mmrhist import tag_range
        nw = min(int(words.numel()), 30_000_000)
        w = words[:nw].cpu().numpy()
        tf, tl = tag_range(w)
        nitag = ((tl - tf) + 999) // 1000 + 1
        secs, counted = C.time_hst(w, params, tf, nitag, runs=runs)
        out['histogram'] = {'words_per_s': round(nw / float(np.median(secs)), 1),
                            'sample': f'hst event loop over the first {nw} words '
                                      f'({counted} events counted), warm-up + median of {runs}'}
    rows = np.sort(rng.choice(sub, 256, replace=False)).astype(np.int32)
    t0 = time.perf_counter()
    O.bprj(params, O.fprj(params, img, rows), rows)
    dt = time.perf_counter() - t0
    out['oracle_fp64_info'] = {'pairs_per_s': round(len(rows) / (dt * len(sub)), 4),
                               'sample': f'FP64 oracle over {len(rows)} subset-0 rays (not the '
                                         'baseline)'}
    return out


def sq_counters(kernel):
    """Per-launch SQ counters of `kernel` over the full OSEM (rocprofv3 --pmc),
    read from profiles/sq_<kernel>_osem4x14.json when that summary exists."""
    path = os.path.join(ROOT, 'profiles', f'sq_{kernel}_osem4x14.json')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def copy_calibration(nbytes):
    """Plain device copy of nbytes (torch copy_), for scaling the rocprofv3
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = nbytes // 4
    src = torch.rand(n, device='cuda')
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({'copy_bytes_read': 4 * n, 'copy_bytes_written': 4 * n, 'launches': 3}),
          flush=True)


LM_SECONDS = 600


def lm_stream(params, d, nwords):
    """Synthetic 600 s list-mode stream on the device whose prompt bins follow
    the expected prompts of the OSEM data set (a n FP(phantom) + randoms +
    scatter); returns (words, randoms + scatter at the stream's count level)."""
    import torch

    from nipet_amd.prj.mmrsim import lm_workload
    per_ms = max(1, int(round(nwords / (LM_SECONDS * 1000))))
    log(f'synthesising {LM_SECONDS} s list-mode stream ({per_ms} events/ms)')
    words, rs = lm_workload(params, d, seconds=LM_SECONDS, per_ms=per_ms, seed=5)
    torch.cuda.synchronize()
    return words, rs


def histogram(params, words, nrep=64):
    """List-mode histogramming of the 600 s stream (6e8 words): plain (span-11
    prompts / delayeds, head curves, fan-sums, buckets, SSRB) and with nrep
    Poisson(1) bootstrap replicates, device-resident words, wall clock."""
    import torch

    from nipet_amd.lm.mmrhist import hist_words
    out = {}
    for n, btp in ((1, 0), (nrep, 2)):
        hist_words(words[:1 << 20], params, btp=btp, nrep=n)  # load the kernels
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        h = hist_words(words, params, btp=btp, seed=7, nrep=n)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        key = 'plain' if n == 1 else f'bootstrap_{n}'
        out[key] = {'s': round(dt, 4), 'words_per_s': round(words.numel() / dt, 1),
                    'prompts': int(h.out['phc'][:, 0].sum(dtype=torch.int64))}
        del h
    out['workload'] = (f'{words.numel()} words over {LM_SECONDS} s (1 ms tags, 85 % prompts), '
                       'span-11; wall clock incl. session set-up and output zeroing')
    return out


def map_leg(params, d, steps):
    """Gradient MAP through the differentiable projectors (torch.autograd):
    `steps` preconditioned gradient steps of the Poisson NLL over all 14
    subsets + quadratic prior (beta 1e-3), one checkpoint per subset."""
    import torch

    import nipet_amd as nip
    from nipet_amd.prj.autograd import map_recon
    subs = [nip.get_subsets14(k, params)[0] for k in range(NSUB)]
    dev = lambda a: a if isinstance(a, torch.Tensor) else torch.from_numpy(a).cuda()  # noqa: E731
    prompts = dev(d['prompts']).to(torch.int32).to(torch.float32)
    mu = dev(d['mu'])
    rs = dev(d['randoms']) + dev(d['scatter'])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, hist = map_recon(params, prompts, dev(d['norm']), rs, mu, subs, beta=1e-3, steps=steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    log(f'MAP: {steps} steps in {dt:.2f} s')
    return {'workload': f'{steps} gradient MAP steps (Poisson NLL over 14 subsets + quadratic '
                        'prior beta 1e-3, checkpoint per subset), 320x320x127 image',
            'steps': steps, 's': round(dt, 3), 'ms_per_step': round(1e3 * dt / steps, 2),
            'objective_first_last': [float(hist[0]), float(hist[-1])],
            'timing': 'wall clock incl. attenuation FP and sensitivity BP of the set-up'}


def _rel_l2(torch, a, b):
    return float(torch.linalg.vector_norm((a - b).double()) / torch.linalg.vector_norm(b.double()))


def batched(params, d, words, rs, nrep):
    """64 Poisson(1) bootstrap replicates of the 600 s stream and its 24
    dynamic frames (10-300 s) reconstructed 4x14 as ONE batch (one list-mode
    pass for the replicates, one session per frame, one HIP graph launch for
    every subset update), against plain reconstructions of single replicates
    (one histogram session with that replicate, osem_prepare, osem_run) and
    of single frames (one frame session, osem_prepare, osem_run), whose images
    the batch must equal. Randoms + scatter: the stream's own (12 % + 25 % of
    the trues), scaled by duration per frame."""
    import numpy as np
    import torch

    from nipet_amd.lm.mmrhist import hist_words
    from nipet_amd.prj.batch import FRAMES24, _remgaps_b, frame_prompts, replicates_and_frames
    from nipet_amd.prj.osem import osem_prepare, osem_run
    seed = 2024
    torch.cuda.synchronize()
    log(f'batch: {nrep} replicates + {len(FRAMES24)} frames, one graph')
    t0 = time.perf_counter()
    reps, frms, info = replicates_and_frames(words, params, d['attn'], d['norm'], rs, nrep=nrep,
                                             seed=seed, itr=NITER)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    log(f'batch: {t_batch:.2f} s; {info["timing"]}')
    t_rep, t_frm, rels = [], [], {}
    for rep in (0, nrep - 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        h = hist_words(words, params, btp=2, seed=seed, nrep=1, rep0=rep)
        p = _remgaps_b(h, params, 1, 0, 1)[..., 0].contiguous()
        del h
        st = osem_prepare(params, p, d['attn'], d['norm'], rs)
        img = osem_run(st, NITER, check_counts=False)
        torch.cuda.synchronize()
        t_rep.append(time.perf_counter() - t0)
        rels[f'replicate_{rep}'] = _rel_l2(torch, reps[rep], img)
        del st, img, p
    lengths = [t1 - t0 for t0, t1 in FRAMES24]
    dur = LM_SECONDS
    for f in (int(np.argmin(lengths)), int(np.argmax(lengths))):
        fa, fb = FRAMES24[f]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        p, _ = frame_prompts(words, params, [FRAMES24[f]])
        st = osem_prepare(params, p[..., 0].contiguous(), d['attn'], d['norm'],
                          rs * ((fb - fa) / dur))
        img = osem_run(st, NITER, check_counts=False)
        torch.cuda.synchronize()
        t_frm.append(time.perf_counter() - t0)
        rels[f'frame_{f}'] = _rel_l2(torch, frms[f], img)
        del st, img, p
    tr, tf = sum(t_rep) / len(t_rep), sum(t_frm) / len(t_frm)
    n = nrep + len(FRAMES24)
    t_seq = nrep * tr + len(FRAMES24) * tf
    log(f'plain: {tr:.3f} s per replicate, {tf:.3f} s per frame; rel L2 {rels}')
    del info['graph']
    return {'workload': f'{nrep} Poisson(1) bootstrap replicates of a {dur} s list-mode stream '
                        f'of {words.numel()} words + its {len(FRAMES24)} dynamic frames of '
                        f'10-300 s, OSEM {NITER}x{NSUB} each, span-11, one graph launch',
            'recons_per_s': round(n / t_batch, 4), 'plain_recons_per_s': round(n / t_seq, 4),
            'speedup_vs_plain': round(t_seq / t_batch, 3), 'batch_s': round(t_batch, 3),
            'plain_s_per_replicate': round(tr, 4), 'plain_s_per_frame': round(tf, 4),
            'plain_checked': len(rels), 'max_rel_l2_vs_plain': max(rels.values()),
            'rel_l2_vs_plain': rels,
            'autotuned': {'fprj_width': info['tune']['fprj'], 'bprj_width': info['tune']['bprj'],
                          'bprj_run_by_width': info['tune']['bprj_run'],
                          'ms_per_image': {'/'.join(str(x) for x in k): round(v, 4)
                                           for k, v in info['tune']['ms'].items()}},
            'batch_breakdown_s': info['timing'],
            'timing': 'wall clock incl. histogramming, gap removal, corrections, sensitivity, '
                      'autotune, graph capture; plain = sequential estimate from the timed '
                      'plain runs (replicates and frames separately)'}


def _event_ms(fn, reps):
    """Mean milliseconds of fn() over reps calls, HIP events on the current
    stream (the stream the library launches on)."""
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def corrections(params, d):
    """Randoms from fan-sums, normalisation from components and the PSF
    convolution at full size (the corrections that feed one reconstruction)."""
    import numpy as np
    import torch

    from nipet_amd.lm.mmrhist import Randoms
    from nipet_amd.mmrnorm import NormSino
    from nipet_amd.prj.mmrrec import psf_config
    from nipet_amd.prj.osem import psf_conv
    rng = np.random.default_rng(7)
    fan = torch.from_numpy(rng.integers(10**6, 2 * 10**6, (64, 504)).astype(np.int32)).cuda()
    R = Randoms(params)
    rs = torch.empty((params['txLUT']['Naw'], params['Cnt']['NSN11']), device='cuda')
    cm = torch.empty((504, 64), device='cuda')
    t_rnd = _event_ms(lambda: R(fan, rsino=rs, cmap=cm), 10)
    comp = {'geo': rng.uniform(0.8, 1.2, (127, 344)), 'cinf': rng.uniform(0.9, 1.1, (344, 9)),
            'ceff': rng.uniform(0.7, 1.3, (64, 504)), 'axe1': rng.uniform(0.9, 1.1, 837),
            'axf1': rng.uniform(0.75, 1.17, 4084), 'dtp': rng.uniform(1e-7, 1e-6, 64),
            'dtnp': rng.uniform(1e-7, 1e-6, 64)}
    bk = rng.integers(10_000, 100_000, 224).astype(np.int32)
    N = NormSino(params)
    ns = torch.empty_like(rs)
    t_nrm = _event_ms(lambda: N(comp, bk, out=ns), 10)
    img = torch.from_numpy(d['img']).cuda()
    k = psf_config(4.0, params['Cnt'])
    po, pt = torch.empty_like(img), torch.empty_like(img)
    t_psf = _event_ms(lambda: psf_conv(img, k, out=po, tmp=pt), 20)
    nvox = img.numel()
    out_bytes = rs.numel() * 4
    return {'randoms_ms': round(t_rnd, 4), 'norm_ms': round(t_nrm, 4),
            'psf_conv_ms': round(t_psf, 4),
            'randoms_sino_GBs': round(out_bytes / (t_rnd * 1e-3) / 1e9, 1),
            'norm_sino_GBs': round(out_bytes / (t_nrm * 1e-3) / 1e9, 1),
            'psf_GBs': round(3 * 8 * nvox / (t_psf * 1e-3) / 1e9, 1),
            'note': 'randoms: 10 fixed-point steps + span-11 sinogram (68516 x 837); norm: '
                    'host components copied in, span-11 sinogram; psf: 3 passes over one '
                    '320x320x127 image, 8 B/voxel/pass'}


def main():
    args = parse()
    if args.gpus != 1:
        print(f'bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported '
              '(the multi-GPU sharding is a design note, not an implementation)',
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    import numpy as np
    import torch

    import nipet_amd as nip
    from nipet_amd import _lib
    from nipet_amd.geom import count_terms, get_geom
    from nipet_amd.prj import mmrsim
    from nipet_amd.prj.osem import osem_prepare, osem_run
    assert torch.cuda.is_available(), 'bench.py needs a HIP device'
    params = nip.get_mmrparams()
    Cnt = params['Cnt']
    log('building geometry')
    geom = get_geom(params)
    subsets = [nip.get_subsets14(k, params)[0] for k in range(NSUB)]
    subsets_dev = [torch.from_numpy(s).cuda() for s in subsets]
    mask = nip.get_cylinder(Cnt, rad=29., unival=1, gpu_dim=True) > 0.9
    mask_dev = torch.from_numpy(mask.astype(np.uint8)).cuda()
    log('generating synthetic data')
    d = mmrsim.osem_inputs(params, counts=3e8)
    torch.cuda.synchronize()
    log('data ready')

    def step():
        st = osem_prepare(params, d['prompts'], d['attn'], d['norm'], d['randoms'], d['scatter'],
                          subsets=subsets_dev, mask=mask_dev)
        return osem_run(st, NITER, check_counts=False)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log('warmup done')
    with _lib.KernelTimer() as kt:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            img = step()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    ms_step = 1e3 * dt / args.steps
    log(f'timed {args.steps} steps: {ms_step:.1f} ms/step')
    pairs = NITER * NSUB
    value = pairs * args.steps / dt

    # byte model of the dominant kernel (BP): B = 8 I + 4 M per launch
    tx = geom.tx_lists()
    terms = np.array([count_terms(params, s, tx).sum() for s in subsets], dtype=np.float64)
    M = np.array([len(s) * geom.nsn for s in subsets], dtype=np.float64)
    b_bp = float(np.mean(8 * terms + 4 * M))
    b_fp = float(np.mean(4 * terms + 4 * M))
    t_bp = kt.mean_ms(_lib.PROF_BPRJ) * 1e-3
    t_fp = kt.mean_ms(_lib.PROF_FPRJ) * 1e-3
    dominant, b_dom, t_dom = ('k_bprj', b_bp, t_bp) if t_bp >= t_fp else ('k_fprj', b_fp, t_fp)
    # HBM roofline of the dominant kernel from its measured DRAM bytes per
    # launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated); the
    # per-term byte model (every (LOR, voxel) term an HBM access) is reported
    # beside it: it counts LDS and L2 hits and exceeds the peak
    traffic = traffic_per_launch(dominant)
    achieved = traffic / t_dom / 1e9 if traffic else None
    sq = sq_counters(dominant)
    issue = None
    if sq:
        cyc = t_dom * CLOCK_HZ * N_CU   # CU-cycles of one launch
        issue = {'clock_GHz': CLOCK_HZ / 1e9, 'cus': N_CU,
                 'valu_per_cu_cycle': round(sq['SQ_INSTS_VALU'] / cyc, 3),
                 'salu_per_cu_cycle': round(sq['SQ_INSTS_SALU'] / cyc, 3),
                 'lds_idx_active_per_cu_cycle': round(sq['SQ_LDS_IDX_ACTIVE'] / cyc, 3),
                 'lds_bank_conflict_frac': round(sq['SQ_LDS_BANK_CONFLICT'] /
                                                 sq['SQ_LDS_IDX_ACTIVE'], 3),
                 'note': 'SQ counters per launch from the committed rocprofv3 --pmc summary '
                         'over this live launch time; one VALU (wave64) and one SALU '
                         'instruction per CU-cycle is the issue limit'}
    roofline = {'bound': 'hbm', 'achieved': round(achieved, 1) if achieved else None,
                'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                'frac': round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                'traffic': traffic, 'kernel': dominant, 'binding': 'instruction issue',
                'issue': issue,
                'model_GBs': round(b_dom / t_dom / 1e9, 1),
                'bytes_model': 'BP 8I+4M, FP 4I+4M per subset launch (I = (LOR, voxel) '
                               'terms, M = bins x sinograms): every term counted as HBM',
                'I_per_subset': float(terms.mean()),
                'fp_ms': round(t_fp * 1e3, 4), 'bp_ms': round(t_bp * 1e3, 4),
                'fp_model_GBs': round(b_fp / t_fp / 1e9, 1),
                'bp_model_GBs': round(b_bp / t_bp / 1e9, 1),
                'update_ms': round(kt.mean_ms(_lib.PROF_OSEM_UPDATE), 4),
                # first iteration: sensitivity and ratio back-projected together
                'bp_pair_ms': round(kt.mean_ms(_lib.PROF_BPRJ_PAIR), 4)}
    out = {
        'metric': METRIC, 'value': round(value, 3), 'unit': 'pairs/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': round(ms_step, 3),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32',
        'data': 'synthetic (seeded phantom, mu-map, Poisson prompts at 3e8 counts)',
        'config': {'workload': 'full OSEM 4 iterations x 14 subsets, mMR span-11, '
                               '4894 bins x 837 sinograms per subset, image 320x320x127',
                   'iterations': NITER, 'subsets': NSUB, 'counts': 3e8,
                   'recons_per_s': round(1e3 / ms_step, 4),
                   'image_sum': float(img.sum(dtype=torch.float64))},
        'roofline': roofline,
    }
    log('byte model done')
    if not args.no_corr:
        log('corrections')
        out['corrections'] = corrections(params, d)
    words = None
    if not args.no_batch:
        words, rs = lm_stream(params, d, args.lm_words)
        out['histogram'] = histogram(params, words, args.nrep)
        if not args.no_recon:
            out['batched'] = batched(params, d, words, rs, args.nrep)
        del rs
    if args.map_steps > 0:
        out['autograd'] = map_leg(params, d, args.map_steps)
    if not args.no_cpu:
        log('CPU baseline')
        out['cpu_baseline'] = cpu_baseline(params, words)
    del words
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
