"""Randoms from delayed fan-sums on the GPU (nipet_rnd_* through
nipet_amd.lm.mmrhist.Randoms) against the FP64 oracle (gpu_randoms,
lm/src/rnd.cu:180-420). The GPU computes in float32: singles maps and
randoms sinograms agree with the oracle within 1e-5 relative L2 and 1e-4
relative per element (float32 sums of up to 307 crystals x 121 rings)."""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

RTOL_L2, RTOL_MAX = 1e-5, 1e-4


@pytest.fixture(scope='module')
def torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


def _cmp(a, b, what):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    assert np.isfinite(a).all(), what
    l2 = np.linalg.norm(a - b) / np.linalg.norm(b)
    mx = np.abs(a - b).max() / np.abs(b).max()
    print(f'{what}: rel L2 {l2:.3g}, max rel {mx:.3g}')
    assert l2 <= RTOL_L2 and mx <= RTOL_MAX, what


def _fansums(params, seed=3, level=40.0):
    """Integer fan-sums of the expected delayeds of a smooth singles pattern
    (fan(c, r) = s(c, r) (F s)(c, r), F from the coincidence matrix and MRD)."""
    rng = np.random.default_rng(seed)
    Cnt, tx = params['Cnt'], params['txLUT']
    c = np.arange(448)[:, None] * 2 * np.pi / 448
    r = np.arange(64)[None, :] / 63.0
    s = level * (1 + 0.3 * np.sin(3 * c + rng.uniform(0, 6)) * (0.5 + r)
                 + 0.1 * rng.uniform(-1, 1, (448, 64)))
    C, R = oracle.rnd_fan_matrices(tx['cij'], Cnt['MRD'])
    fan = np.zeros((64, 504), np.uint32)
    crsri = np.asarray(tx['crsri'], np.int64)
    live = crsri >= 0
    fan[:, live] = np.rint(s * (C @ s @ R.T))[crsri[live]].T.astype(np.uint32)
    return fan


def test_randoms_span11_matches_oracle(torch, params):
    from nipet_amd.lm.mmrhist import Randoms
    fan = _fansums(params)
    rs, cmap = Randoms(params, span=11)(fan)
    torch.cuda.synchronize()
    ors, ocmap, _ = oracle.randoms(fan, params, span=11)
    _cmp(cmap.cpu().numpy(), ocmap, 'cmap')
    _cmp(rs.cpu().numpy(), ors, 'span-11 randoms')
    gaps = np.asarray(params['txLUT']['crsri']) < 0
    assert np.all(cmap.cpu().numpy()[gaps] == 0)


def test_randoms_span1_matches_oracle(torch, params):
    from nipet_amd.lm.mmrhist import Randoms
    fan = _fansums(params, seed=8, level=5.0)
    rs, cmap = Randoms(params, span=1)(fan)
    bins = np.arange(0, params['txLUT']['Naw'], 53)
    ors, ocmap, _ = oracle.randoms(fan, params, span=1, bins=bins)
    _cmp(cmap.cpu().numpy(), ocmap, 'cmap')
    _cmp(rs[torch.from_numpy(bins).cuda()].cpu().numpy(), ors, 'span-1 randoms')


def test_randoms_from_histogram_delayeds(torch, params):
    """Fan-sums of a histogrammed list-mode stream (delayed events), device
    resident; the host API returns the Siemens layout with zero gaps."""
    from nipet_amd.lm.lmsim import synth_lm
    from nipet_amd.lm.mmrhist import Randoms, hist_words, randoms
    from nipet_amd.mmraux import putgaps
    words = synth_lm(params, 10, 3000, seed=21)
    h = hist_words(words, params)
    fan_dev = h.out['fan']
    fan = fan_dev.cpu().numpy().astype(np.uint32)
    assert fan.sum() > 0
    rs, cmap = Randoms(params)(fan_dev)
    ors, ocmap, _ = oracle.randoms(fan, params)
    _cmp(cmap.cpu().numpy(), ocmap, 'cmap (histogram)')
    _cmp(rs.cpu().numpy(), ors, 'randoms (histogram)')
    rsn, sgl = randoms({'fansums': fan}, params)
    assert rsn.shape == (837, 252, 344) and sgl.shape == (504, 64)
    assert np.array_equal(rsn, putgaps(rs, params['txLUT'], params['Cnt']).cpu().numpy())
    rsd, _ = randoms({'fansums': fan}, params, gpu_dim=True)
    assert np.array_equal(rsd, rs.cpu().numpy())


def test_randoms_dead_crystals_and_empty(torch, params):
    """Crystals with zero fan-sums get zero singles; all-zero fan-sums give
    zeros (the guarded division), never NaN."""
    from nipet_amd.lm.mmrhist import Randoms
    R = Randoms(params)
    fan = _fansums(params, seed=4)
    fan[:, 100:104] = 0
    fan[10, :] = 0
    rs, cmap = R(fan)
    ors, ocmap, _ = oracle.randoms(fan, params)
    _cmp(cmap.cpu().numpy(), ocmap, 'cmap (dead crystals)')
    _cmp(rs.cpu().numpy(), ors, 'randoms (dead crystals)')
    rs0, c0 = R(np.zeros((64, 504), np.uint32))
    assert float(rs0.abs().max()) == 0 and float(c0.abs().max()) == 0


def test_randoms_bad_arguments(torch, params):
    from nipet_amd import _lib
    from nipet_amd.lm.mmrhist import Randoms
    with pytest.raises(ValueError):
        Randoms(params, span=0)
    with pytest.raises(ValueError):
        Randoms(params)(np.zeros((64, 500), np.uint32))
    assert _lib.lib().nipet_rnd_estimate(None, None, 10, None, None, None) != 0


def test_osemone_estimates_randoms_from_fansums(torch, params):
    """osemone without a randoms sinogram uses the fan-sum estimate
    (prj/mmrrec.py:244-250): the image equals the one given that estimate."""
    from nipet_amd import osemone
    from nipet_amd.lm.mmrhist import Randoms
    from nipet_amd.prj import mmrsim
    from nipet_amd.prj.mmrprj import frwd_prj
    img = torch.from_numpy(mmrsim.ellipsoid_phantom(1)).cuda()
    psino = torch.poisson(frwd_prj(img, params, dev_out=True) * 0.05).cpu().numpy()
    fan = _fansums(params, seed=6, level=0.5)
    mus = (np.zeros((127, 320, 320), np.float32), np.zeros((127, 320, 320), np.float32))
    a = osemone({}, mus, {'psino': psino, 'fansums': fan}, params, recmod=0, itr=1)
    rs, _ = Randoms(params)(fan)
    b = osemone({}, mus, {'psino': psino}, params, recmod=0, itr=1, randsino=rs.cpu().numpy())
    c = osemone({}, mus, {'psino': psino}, params, recmod=0, itr=1)
    rel = lambda x, y: np.linalg.norm((x - y).ravel()) / np.linalg.norm(y.ravel())  # noqa: E731
    print(f'osemone: rel L2 fan-sum estimate vs given {rel(a.im, b.im):.3g}, '
          f'vs no randoms {rel(a.im, c.im):.3g}')
    assert rel(a.im, b.im) <= 1e-6      # reruns differ only by float atomics
    assert rel(a.im, c.im) > 1e-3
