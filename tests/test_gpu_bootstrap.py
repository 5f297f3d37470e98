"""Batched bootstrap / dynamic-frame reconstruction (nipet_amd.prj.batch) on
the GPU: every batched image equals its sequential plain reconstruction (one
histogram session with nrep = 1, rep0 = r, then osem_prepare / osem_run)
within 1e-5 relative L2; the captured HIP graph equals eager launches and
replays consistently; the autotuner's widths do not change the images."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ITR = 2


@pytest.fixture(scope='module')
def torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


def _rel(torch, a, b):
    return float(torch.linalg.vector_norm((a - b).double()) / torch.linalg.vector_norm(b.double()))


@pytest.fixture(scope='module')
def data(torch, params):
    """20 s list-mode stream whose prompts follow the phantom's projection,
    plus corrections (attenuation of the mu-map phantom, U(0.8, 1.2) norm,
    flat randoms + scatter)."""
    from nipet_amd.lm.lmsim import synth_lm
    from nipet_amd.prj import mmrsim
    from nipet_amd.prj.mmrprj import frwd_prj
    img = torch.from_numpy(mmrsim.ellipsoid_phantom(1)).cuda()
    mu = torch.from_numpy(mmrsim.mumap_phantom(2)).cuda()
    attn = frwd_prj(mu, params, dev_out=True, attenuation=True)
    emis = frwd_prj(img, params, dev_out=True)
    words = synth_lm(params, 20, 1500, weights=attn * emis, seed=11)
    g = torch.Generator(device='cuda').manual_seed(3)
    norm = 0.8 + 0.4 * torch.rand(attn.shape, generator=g, device='cuda')
    rs = torch.full_like(attn, 0.02)
    return dict(words=words, attn=attn, norm=norm, rs=rs)


def _plain(torch, params, d, prompts):
    from nipet_amd.prj.osem import osem_prepare, osem_run
    st = osem_prepare(params, prompts, d['attn'], d['norm'], d['rs'])
    return osem_run(st, ITR, check_counts=False)


def test_bootstrap_batch_equals_sequential(torch, params, data):
    from nipet_amd.lm.mmrhist import hist_words
    from nipet_amd.prj.batch import _remgaps_b, bootstrap_osem
    nrep, seed = 5, 77
    imgs, info = bootstrap_osem(data['words'], params, data['attn'], data['norm'], data['rs'],
                                nrep=nrep, seed=seed, itr=ITR)
    assert info['tune']['fprj'] in (1, 2, 4) and info['tune']['bprj'] in (1, 2, 4)
    for r in (0, 3, 4):
        h = hist_words(data['words'], params, btp=2, seed=seed, nrep=1, rep0=r)
        p = _remgaps_b(h, params, 1, 0, 1)[..., 0].contiguous()
        ref = _plain(torch, params, data, p)
        rel = _rel(torch, imgs[..., r], ref)
        print(f'replicate {r}: rel L2 vs sequential {rel:.3g}')
        assert rel <= 1e-5
    # replicates differ from each other
    assert _rel(torch, imgs[..., 0], imgs[..., 1]) > 1e-3


def test_graph_matches_eager_and_replays(torch, params, data):
    from nipet_amd.prj.batch import OsemGraph, bootstrap_prompts
    from nipet_amd.prj.osem import osem_iteration_batch, osem_prepare_batch
    prompts, _ = bootstrap_prompts(data['words'], params, 3, seed=5)
    st = osem_prepare_batch(params, prompts, data['attn'], data['norm'], data['rs'])
    eager = torch.ones((320, 320, 127, 3), device='cuda')
    for it in range(ITR):
        osem_iteration_batch(st, eager, first_full=it == 0)
    g = OsemGraph(st, ITR)
    a = g.run().clone()
    torch.cuda.synchronize()
    _ = torch.ones(10, device='cuda').sum().item()      # eager work between replays
    b = g.run().clone()
    assert _rel(torch, a, eager) <= 1e-6 and _rel(torch, b, a) <= 1e-6


def test_widths_do_not_change_images(torch, params, data):
    from nipet_amd import _lib
    from nipet_amd.prj.batch import bootstrap_prompts
    from nipet_amd.prj.osem import osem_prepare_batch, osem_run_batch
    prompts, _ = bootstrap_prompts(data['words'], params, 4, seed=9)
    st = osem_prepare_batch(params, prompts, data['attn'], data['norm'], data['rs'])
    out = []
    for w in (4, 1):
        for k in (_lib.PROF_FPRJ, _lib.PROF_BPRJ):
            _lib.check(_lib.lib().nipet_tune_set(k, w), 'tune')
        out.append(osem_run_batch(st, 1))
    # the other compiled BP run length of each width
    import ctypes as ct
    runs = {}
    for w in (4, 2, 1):
        cur, alt = ct.c_int(), ct.c_int()
        _lib.check(_lib.lib().nipet_tune_get_run(w, ct.byref(cur), ct.byref(alt)), 'get_run')
        runs[w] = (cur.value, alt.value)
        _lib.check(_lib.lib().nipet_tune_set_run(w, alt.value), 'set_run')
    for k in (_lib.PROF_FPRJ, _lib.PROF_BPRJ):
        _lib.check(_lib.lib().nipet_tune_set(k, 4), 'tune')
    out.append(osem_run_batch(st, 1))
    for w, (cur, alt) in runs.items():
        _lib.check(_lib.lib().nipet_tune_set_run(w, cur), 'set_run')
    assert _rel(torch, out[0], out[1]) <= 1e-6
    assert _rel(torch, out[0], out[2]) <= 1e-6
    assert _lib.lib().nipet_tune_set(_lib.PROF_FPRJ, 3) != 0
    assert _lib.lib().nipet_tune_set_run(4, 5) != 0


def test_frames_batch_equals_sequential(torch, params, data):
    from nipet_amd.prj.batch import frame_prompts, frames_osem
    frames = [(0, 4), (4, 10), (10, 20)]
    nf = len(frames)
    rho = torch.stack([data['rs'] * (t1 - t0) / 10.0 for t0, t1 in frames], dim=2).contiguous()
    imgs, info = frames_osem(data['words'], params, frames, data['attn'], data['norm'], rho,
                             itr=ITR)
    prompts, _ = frame_prompts(data['words'], params, frames)
    # frame windows partition the events of [0, 20) s
    assert int(prompts.to(torch.int64).sum()) > 0
    from nipet_amd.prj.osem import osem_prepare, osem_run
    for f in range(nf):
        st = osem_prepare(params, prompts[..., f].contiguous(), data['attn'], data['norm'],
                          rho[..., f].contiguous())
        ref = osem_run(st, ITR, check_counts=False)
        rel = _rel(torch, imgs[..., f], ref)
        print(f'frame {frames[f]}: rel L2 vs sequential {rel:.3g}')
        assert rel <= 1e-5


def test_bootstrap_psf_batch_equals_sequential(torch, params, data):
    """Resolution modelling in the grouped batch equals sequential plain runs."""
    from nipet_amd.lm.mmrhist import hist_words
    from nipet_amd.prj.batch import _remgaps_b, bootstrap_osem
    from nipet_amd.prj.mmrrec import psf_config
    from nipet_amd.prj.osem import osem_prepare, osem_run
    k = psf_config(4.0, params['Cnt'])
    imgs, _ = bootstrap_osem(data['words'], params, data['attn'], data['norm'], data['rs'],
                             nrep=5, seed=31, itr=ITR, psf=k)
    for r in (1, 4):
        h = hist_words(data['words'], params, btp=2, seed=31, nrep=1, rep0=r)
        p = _remgaps_b(h, params, 1, 0, 1)[..., 0].contiguous()
        st = osem_prepare(params, p, data['attn'], data['norm'], data['rs'], psf=k)
        ref = osem_run(st, ITR, check_counts=False)
        assert _rel(torch, imgs[..., r], ref) <= 1e-5
