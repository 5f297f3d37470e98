"""Batched projection / reconstruction (replicate-innermost layouts) against
the single-image path: FP and BP of a batch equal the per-image results
(1e-6 relative), batched OSEM images equal sequential plain runs within 1e-5."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


def _rel(a, b):
    a, b = a.double(), b.double()
    return float((a - b).abs().max() / b.abs().max())


@pytest.mark.parametrize('nb', [2, 4, 7])
def test_fp_bp_batch_equals_single(torch, params, subsets, nb):
    from nipet_amd import _lib
    from nipet_amd.geom import get_geom
    g = get_geom(params)
    sub = torch.from_numpy(subsets[1][::3].copy()).cuda()
    n = len(sub)
    gen = torch.Generator(device='cuda').manual_seed(nb)
    imgs = torch.rand((320, 320, 127, nb), generator=gen, device='cuda')
    sino = torch.empty((n, 837, nb), device='cuda')
    L, st = _lib.lib(), _lib.stream_handle()
    for att in (0, 1):
        _lib.check(L.nipet_fprj_b(g.handle, sino.data_ptr(), imgs.data_ptr(), sub.data_ptr(), n,
                                  nb, att, st))
        for b in range(nb):
            one = torch.empty((n, 837), device='cuda')
            x = imgs[..., b].contiguous()
            _lib.check(L.nipet_fprj(g.handle, one.data_ptr(), x.data_ptr(), sub.data_ptr(), n,
                                    att, st))
            assert _rel(sino[..., b], one) < 1e-6
    y = torch.rand((n, 837, nb), generator=gen, device='cuda')
    bimg = torch.empty((320, 320, 127, nb), device='cuda')
    _lib.check(L.nipet_bprj_b(g.handle, bimg.data_ptr(), y.data_ptr(), sub.data_ptr(), n, nb,
                              st))
    for b in range(nb):
        one = torch.empty((320, 320, 127), device='cuda')
        yb = y[..., b].contiguous()
        _lib.check(L.nipet_bprj(g.handle, one.data_ptr(), yb.data_ptr(), sub.data_ptr(), n, st))
        assert _rel(bimg[..., b], one) < 1e-6


def test_osem_batch_equals_sequential(torch, params, subsets):
    from nipet_amd.prj import mmrsim
    from nipet_amd.prj.osem import (osem_prepare, osem_prepare_batch, osem_run,
                                    osem_run_batch)
    d = mmrsim.osem_inputs(params, counts=2e7, seed=3)
    nb = 5
    gen = torch.Generator(device='cuda').manual_seed(11)
    # bootstrap-like replicates: Poisson(1) resampling of the prompts
    w = torch.poisson(torch.ones(d['prompts'].shape + (nb,), device='cuda'), generator=gen)
    pr = (d['prompts'].to(torch.float32)[..., None] * w).clamp(max=65535).to(torch.int32)
    pr = pr.to(torch.uint16)
    mini = [s[::4] for s in subsets[:6]]
    bs = osem_prepare_batch(params, pr, d['attn'], d['norm'], d['randoms'], d['scatter'],
                            subsets=mini)
    imgs = osem_run_batch(bs, 2)
    for b in range(nb):
        st = osem_prepare(params, pr[..., b].contiguous(), d['attn'], d['norm'], d['randoms'],
                          d['scatter'], subsets=mini)
        ref = osem_run(st, 2)
        diff = (imgs[..., b].double() - ref.double()).norm() / ref.double().norm()
        assert float(diff) < 1e-5, (b, float(diff))
