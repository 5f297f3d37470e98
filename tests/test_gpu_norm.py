"""Normalisation sinogram from components (nipet_norm_* through
nipet_amd.mmrnorm.NormSino) and the online variance (nipet_varon through
nipet_amd.mmraux.varon) on the GPU against the FP64 oracle (dev_norm
src/norm.cu:13-51, var src/auxmath.cu:15-22). float32 on the GPU: within
1e-5 relative L2 and 1e-5 relative per element."""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


def _comp(seed=1):
    rng = np.random.default_rng(seed)
    c = {'geo': rng.uniform(0.8, 1.2, (127, 344)), 'cinf': rng.uniform(0.9, 1.1, (344, 9)),
         'ceff': rng.uniform(0.7, 1.3, (64, 504)), 'axe1': rng.uniform(0.9, 1.1, 837),
         'axf1': rng.uniform(0.75, 1.17, 4084), 'dtp': rng.uniform(1e-7, 1e-6, 64),
         'dtnp': rng.uniform(1e-7, 1e-6, 64)}
    c = {k: v.astype(np.float32) for k, v in c.items()}
    bk = rng.integers(10_000, 100_000, 224).astype(np.int32)
    return c, bk


def _cmp(a, b, what, tol=1e-5):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    assert np.isfinite(a).all()
    l2 = np.linalg.norm(a - b) / np.linalg.norm(b)
    mx = np.abs(a / b - 1).max()
    print(f'{what}: rel L2 {l2:.3g}, max rel {mx:.3g}')
    assert l2 <= tol and mx <= tol, what


def test_norm_span11_matches_oracle(torch, params):
    from nipet_amd.mmrnorm import NormSino
    comp, bk = _comp(1)
    n = NormSino(params, span=11)(comp, bk)
    ref = oracle.norm_sino(comp, bk, params, span=11)
    _cmp(n.cpu().numpy(), ref, 'norm span 11')


def test_norm_span1_matches_oracle(torch, params):
    from nipet_amd.mmrnorm import NormSino
    comp, bk = _comp(2)
    n = NormSino(params, span=1)(comp, bk)
    bins = np.arange(0, params['txLUT']['Naw'], 37)
    ref = oracle.norm_sino(comp, bk, params, span=1, bins=bins)
    _cmp(n[torch.from_numpy(bins).cuda()].cpu().numpy(), ref, 'norm span 1')


def test_get_norm_sino_layouts_and_osemone(torch, params):
    from nipet_amd import osemone
    from nipet_amd.mmraux import putgaps
    from nipet_amd.mmrnorm import NormSino, get_norm_sino
    from nipet_amd.prj import mmrsim
    from nipet_amd.prj.mmrprj import frwd_prj
    comp, bk = _comp(3)
    hst = {'buckets': bk}
    dev = get_norm_sino({}, params, hst, normcomp=comp, gpu_dim=True)
    sie = get_norm_sino({}, params, hst, normcomp=comp)
    n = NormSino(params)(comp, bk)
    assert np.array_equal(dev, n.cpu().numpy())
    assert sie.shape == (837, 252, 344)
    assert np.array_equal(sie, putgaps(n, params['txLUT'], params['Cnt']).cpu().numpy())
    # osemone uses the component normalisation when given components
    img = torch.from_numpy(mmrsim.ellipsoid_phantom(2)).cuda()
    psino = torch.poisson(frwd_prj(img, params, dev_out=True) * 0.05).cpu().numpy()
    mus = (np.zeros((127, 320, 320), np.float32),) * 2
    a = osemone({}, mus, {'psino': psino, 'buckets': bk}, params, recmod=0, itr=1,
                normcomp=comp)
    b = osemone({}, mus, {'psino': psino}, params, recmod=0, itr=1, normsino=dev)
    c = osemone({}, mus, {'psino': psino}, params, recmod=0, itr=1)
    rel = lambda x, y: np.linalg.norm((x - y).ravel()) / np.linalg.norm(y.ravel())  # noqa: E731
    assert rel(a.im, b.im) <= 1e-6
    assert rel(a.im, c.im) > 1e-3
    # with the frame times in the histogram the image carries the reference's
    # quantification factor qf / BF / dur * qf_loc (prj/mmrrec.py:324-331, 388),
    # and fwhm > 0 adds the smoothed image (prj/mmrrec.py:420-421)
    from nipet_amd.prj.mmrrec import quant_factor
    cq = dict(comp, qf=1.25e5, qf_loc=0.205)
    hq = {'psino': psino, 'buckets': bk, 't0': 30, 't1': 630, 'dur': 600}
    d = osemone({'lm_bf': '/data/LM/scan.bf', 'corepath': '/data'}, mus, hq, params,
                recmod=0, itr=1, normcomp=cq, fwhm=4.0)
    q = quant_factor(params['Cnt'], hq, cq)
    assert abs(q - 1.25e5 / 0.967 / 600 * 0.205) <= 1e-9 * q
    assert rel(d.im, a.im * np.float32(q)) <= 1e-6
    assert d.fpet == '/data/reconstructed/scan_t30-630sec_itr1.nii.gz'
    assert d.imsmo.shape == d.im.shape
    tv = lambda x: np.abs(np.diff(x, axis=2)).sum()  # noqa: E731
    assert tv(d.imsmo) < tv(d.im)
    assert abs(float(d.imsmo.sum(dtype=np.float64)) / float(d.im.sum(dtype=np.float64)) - 1) < 1e-4
    with pytest.raises(NotImplementedError):
        osemone({}, mus, hq, params, recmod=0, itr=1, store_img=True)


def test_norm_bad_arguments(torch, params):
    from nipet_amd import _lib
    from nipet_amd.mmrnorm import NormSino
    comp, bk = _comp(4)
    ns = NormSino(params)
    with pytest.raises(ValueError):
        ns(comp, bk[:100])
    bad = dict(comp)
    bad['ceff'] = comp['ceff'].T
    with pytest.raises(ValueError):
        ns(bad, bk)
    assert _lib.lib().nipet_norm_sino(None, None, None, None, None) != 0


def test_varon_device_and_host(torch):
    from nipet_amd.mmraux import varon
    rng = np.random.default_rng(5)
    xs = rng.normal(10.0, 3.0, (7, 320 * 320)).astype(np.float32)
    m1 = torch.zeros(xs.shape[1], device='cuda')
    m2 = torch.zeros_like(m1)
    h1 = np.zeros(xs.shape[1], np.float32)
    h2 = np.zeros_like(h1)
    o1, o2 = np.zeros(xs.shape[1]), np.zeros(xs.shape[1])
    for b, x in enumerate(xs):
        varon(m1, m2, torch.from_numpy(x).cuda(), b)
        varon(h1, h2, x, b)
        o1, o2 = oracle.varon(o1, o2, x, b)
    _cmp(m1.cpu().numpy(), o1, 'varon mean')
    assert np.abs(m2.cpu().numpy() - o2).max() <= 1e-4 * np.abs(o2).max()
    assert np.array_equal(h1, m1.cpu().numpy()) and np.array_equal(h2, m2.cpu().numpy())
    assert np.allclose(o2 / len(xs), xs.astype(np.float64).var(0), rtol=1e-9)
