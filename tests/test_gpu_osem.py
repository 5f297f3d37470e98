"""GPU OSEM parity against the FP64 oracle OSEM (oracle/oracle.py::osem,
following prj/src/recon.cu:459-497) and MLEM invariants.

Bars: images within 1e-4 relative L2 of the oracle; MLEM count preservation
sum_S A_S x' = sum_S y to 1e-5 relative.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope='module')
def torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


def _mini(params, subsets, step=20):
    return [subsets[k][::step].astype(np.int32) for k in (0, 3, 7, 10)]


def _data(params, rows_all, seed=0, nsn=837):
    naw = params['txLUT']['Naw']
    rng = np.random.default_rng(seed)
    y = np.zeros((naw, nsn), np.uint16)
    a = np.ones((naw, nsn), np.float32)
    n = np.ones((naw, nsn), np.float32)
    r = np.zeros((naw, nsn), np.float32)
    y[rows_all] = rng.poisson(6.0, size=(len(rows_all), nsn))
    a[rows_all] = rng.uniform(0.3, 1.0, size=(len(rows_all), nsn))
    n[rows_all] = rng.uniform(0.8, 1.2, size=(len(rows_all), nsn))
    r[rows_all] = rng.uniform(0.0, 0.5, size=(len(rows_all), nsn))
    a[rows_all[:5], :3] = 0.0      # exercise the |a n| < eps guard
    return y, a, n, r


def test_osem_corr_guards(torch, params):
    from nipet_amd import _lib
    a = torch.tensor([0.5, 1e-7, 0.0, 2.0], device='cuda')
    n = torch.tensor([2.0, 1.0, 1.0, -0.25], device='cuda')
    r = torch.tensor([1.0, 3.0, 3.0, 1.0], device='cuda')
    s = torch.tensor([0.5, 0.0, 1.0, 0.0], device='cuda')
    rho, an = torch.empty_like(a), torch.empty_like(a)
    _lib.check(_lib.lib().nipet_osem_corr(rho.data_ptr(), an.data_ptr(), a.data_ptr(),
                                          n.data_ptr(), r.data_ptr(), s.data_ptr(), 4,
                                          _lib.stream_handle()))
    torch.cuda.synchronize()
    assert torch.allclose(an, a * n)
    assert rho.tolist() == [1.5, 0.0, 0.0, -2.0]


def _mask(params, kind):
    """'planes': every column (no segment window skips anything); 'disc': the
    reference's cylinder of radius 20 cm (the windows skip the segments outside)."""
    if kind == 'disc':
        import nipet_amd as nip
        return nip.get_cylinder(params['Cnt'], rad=20., unival=1, gpu_dim=True) > 0.9
    mask = np.ones((320, 320, 127), bool)
    mask[:, :, :3] = False
    return mask


@pytest.mark.parametrize('kind', ['planes', 'disc'])
def test_osem_mini_vs_oracle(torch, params, subsets, kind):
    from oracle import oracle as O

    from nipet_amd.prj.osem import osem_prepare, osem_run
    mini = _mini(params, subsets)
    rows_all = np.concatenate(mini)
    y, a, n, r = _data(params, rows_all)
    mask = _mask(params, kind)
    st = osem_prepare(params, y, a, n, r, None, subsets=mini, mask=mask)
    gpu = osem_run(st, 2).cpu().numpy().astype(np.float64)
    an = a.astype(np.float64) * n
    rho = np.where(np.abs(an) < 1e-6, 0.0, r / np.where(np.abs(an) < 1e-6, 1.0, an))
    sens = np.stack([O.bprj(params, an[s], s) for s in mini])
    ref = O.osem(params, y, rho, sens, mask, mini, np.ones((320, 320, 127)), 2)
    rel = np.linalg.norm(gpu - ref) / np.linalg.norm(ref)
    print(f'parity osem mini ({kind} mask): {rel:.3g} rel L2 (bar 1e-4)', flush=True)
    assert rel < 1e-4, rel
    assert np.all(gpu[~mask] == 0)


def test_osem_window_matches_host(torch, params):
    """nipet_osem_window: per bin, the first and the number of segments of its
    transaxial list whose columns hold a mask voxel (columns dilated by a box
    of half-width 8 for resolution modelling), equal to a NumPy scan of the
    same lists."""
    from scipy.ndimage import maximum_filter

    import nipet_amd as nip
    from nipet_amd.geom import get_geom
    from nipet_amd.prj.osem import mask_window
    g = get_geom(params)
    off, col = g.tx_lists()[:2]
    off = np.asarray(off, np.int64)
    for rad, psf in ((29., False), (29., True), (12.5, False)):
        m = nip.get_cylinder(params['Cnt'], rad=rad, unival=1, gpu_dim=True) > 0.9
        m2 = m.any(axis=2)
        if psf:
            m2 = maximum_filter(m2.astype(np.uint8), size=17, mode='constant') > 0
        inm = m2.ravel()[np.asarray(col)]
        want = np.zeros((g.naw, 2), np.int64)
        for aw in range(g.naw):
            k = np.flatnonzero(inm[off[aw]:off[aw + 1]])
            if k.size:
                want[aw] = (k[0], k[-1] - k[0] + 1)
        got = mask_window(g, torch.from_numpy(m.astype(np.uint8)).cuda(), psf).cpu().numpy()
        assert np.array_equal(got, want), (rad, psf)
        kept = want[:, 1].sum() / (off[1:] - off[:-1]).sum()
        print(f'mask radius {rad} cm, psf {psf}: windows keep {kept:.3f} of the segments',
              flush=True)
        # the lists end at the scanner's field of view (about the 29 cm disc):
        # only smaller masks shorten them
        assert kept <= 1.0 and (rad > 20. or kept < 0.5)


def test_osem_window_clamped_to_list(torch, params, subsets):
    """Segment windows reach the projectors clamped into each bin's list: a
    window that starts before the list and claims more segments than it holds
    gives the results of the whole list, as do the mask windows of an all-ones
    mask. One subset update from the same image: its FP ratios are
    bit-identical, the images agree within 1e-6 relative L2 (float atomics
    merge the BP's rays in any order)."""
    from nipet_amd.geom import get_geom
    from nipet_amd.prj.osem import osem_iteration, osem_prepare
    g = get_geom(params)
    off = np.asarray(g.tx_lists()[0], np.int64)
    full = np.stack([np.zeros(g.naw, np.int64), off[1:] - off[:-1]], 1).astype(np.int32)
    bad = np.tile(np.array([-7, 1 << 30], np.int32), (g.naw, 1))
    mini = _mini(params, subsets, step=4)[:1]
    rows_all = np.concatenate(mini)
    y, a, n, r = _data(params, rows_all, seed=5)
    mask = np.ones((320, 320, 127), bool)
    out = []
    for win in (None, full, bad):
        st = osem_prepare(params, y, a, n, r, None, subsets=mini, mask=mask)
        if win is None:   # the mask's own windows: the whole lists
            assert np.array_equal(st.win.cpu().numpy(), full)
        else:
            st.win = torch.from_numpy(win).cuda()
        img = torch.ones((320, 320, 127), dtype=torch.float32, device='cuda')
        osem_iteration(st, img, first_full=False)
        torch.cuda.synchronize()
        out.append((img.double(), st.work_sino[:len(mini[0])].clone()))
    ref_img, ref_ratio = out[0]
    for img, ratio in out[1:]:
        assert torch.equal(ratio, ref_ratio)
        rel = float((img - ref_img).norm() / ref_img.norm())
        assert rel < 1e-6, rel


def test_bins_longest_first_device(torch, params, subsets):
    """nipet_bins_longest_first (device subsets are ordered without a host
    copy): the same bins as the host ordering, segment counts non-increasing,
    bins outside the sinogram last."""
    from nipet_amd.geom import get_geom
    from nipet_amd.prj.osem import _longest_first
    g = get_geom(params)
    segc = g.seg_counts()
    rows = np.concatenate([subsets[3], [-1, g.naw + 5], subsets[3][:7]]).astype(np.int32)
    got = _longest_first(g, torch.from_numpy(rows).cuda()).cpu().numpy()
    host = _longest_first(g, rows).cpu().numpy()
    assert np.array_equal(np.sort(got), np.sort(rows))
    valid = (got >= 0) & (got < g.naw)
    assert not valid[-2:].any() and valid[:-2].all()
    key = segc[got[:-2]]
    assert (np.diff(key) <= 0).all()
    assert np.array_equal(key, segc[host[:-2]])   # same counts in the same order


def test_osem_span1_mini_vs_oracle(torch, params, subsets):
    """OSEM on span-1 data (4084 sinograms, one LOR per bin; the sensitivity
    images are made before the loop in this span) against the FP64 oracle
    OSEM: 2 iterations of two sub-sampled subsets within 1e-4 relative L2."""
    from oracle import oracle as O

    from nipet_amd.prj.osem import osem_prepare, osem_run
    Cnt = dict(params['Cnt'])
    Cnt['SPN'] = 1
    p1 = {'Cnt': Cnt, 'txLUT': params['txLUT'], 'axLUT': params['axLUT']}
    mini = [subsets[k][::40].astype(np.int32) for k in (2, 9)]
    rows_all = np.concatenate(mini)
    y, a, n, r = _data(p1, rows_all, seed=5, nsn=4084)
    mask = np.ones((320, 320, 127), bool)
    mask[:, :, 124:] = False
    st = osem_prepare(p1, y, a, n, r, None, subsets=mini, mask=mask)
    assert not st.pending
    gpu = osem_run(st, 2).cpu().numpy().astype(np.float64)
    del st
    rho = np.zeros(y.shape, np.float64)     # only the subsets' rows are read
    an = a[rows_all].astype(np.float64) * n[rows_all]
    rho[rows_all] = np.where(np.abs(an) < 1e-6, 0.0,
                             r[rows_all] / np.where(np.abs(an) < 1e-6, 1.0, an))
    sens = np.stack([O.bprj(p1, a[s].astype(np.float64) * n[s], s) for s in mini])
    ref = O.osem(p1, y, rho, sens, mask, mini, np.ones((320, 320, 127)), 2)
    rel = np.linalg.norm(gpu - ref) / np.linalg.norm(ref)
    assert rel < 1e-4, rel
    assert np.all(gpu[~mask] == 0)


def test_first_update_pair_matches_separate_sensitivity(torch, params, subsets):
    """Sensitivity images back-projected inside the first subset updates
    (nipet_osem_subset_sens: one two-image BP with the ratio) equal those of
    the separate pass (one-image BP before the loop, prj/mmrrec.py:291-298),
    and so do the images after two iterations. Float atomics merge rays in
    either case: sensitivity within 1e-6, images within 1e-5 relative L2."""
    from nipet_amd.prj.osem import osem_prepare, osem_run
    mini = _mini(params, subsets, step=8)
    rows_all = np.concatenate(mini)
    y, a, n, r = _data(params, rows_all, seed=3)
    mask = np.ones((320, 320, 127), bool)
    mask[:, :, 120:] = False
    lazy = osem_prepare(params, y, a, n, r, None, subsets=mini, mask=mask)
    assert lazy.pending == set(range(len(mini)))
    sep = osem_prepare(params, y, a, n, r, None, subsets=mini, mask=mask)
    sens_sep = sep.sens.clone()        # completes every pending image with the one-image BP
    assert not sep.pending
    x_lazy = osem_run(lazy, 2)
    assert not lazy.pending and lazy._pair is None
    x_sep = osem_run(sep, 2)
    for k in range(len(mini)):
        ref = sens_sep[k].double()
        assert ref.norm() > 0
        rel = float((lazy.sens[k].double() - ref).norm() / ref.norm())
        assert rel < 1e-6, (k, rel)
    rel = float((x_lazy.double() - x_sep.double()).norm() / x_sep.double().norm())
    assert rel < 1e-5, rel
    assert bool((x_lazy[~torch.from_numpy(mask).cuda()] == 0).all())


def test_cylinder_mask_equals_reference_mask(torch, params):
    """The device mask built from one transaxial disc equals the reference's
    get_cylinder(..., gpu_dim=True) > 0.9 (prj/mmrrec.py:303), voxel for voxel."""
    import nipet_amd as nip
    from nipet_amd.prj.osem import cylinder_mask
    Cnt = params['Cnt']
    for rad in (29., 12.5):
        ref = nip.get_cylinder(Cnt, rad=rad, xo=0, yo=0, unival=1, gpu_dim=True) > 0.9
        m = cylinder_mask(Cnt, rad)
        assert m.dtype == torch.uint8 and tuple(m.shape) == ref.shape
        assert np.array_equal(m.cpu().numpy().astype(bool), ref)
        assert cylinder_mask(Cnt, rad) is m


def test_mlem_count_preservation(torch, params, subsets):
    import nipet_amd as nip
    from nipet_amd.prj.osem import osem_prepare, osem_run
    sub = subsets[4][::4].astype(np.int32)
    naw = params['txLUT']['Naw']
    y = np.zeros((naw, 837), np.uint16)
    y[sub] = np.random.default_rng(1).poisson(4.0, size=(len(sub), 837))
    ones = np.ones((naw, 837), np.float32)
    st = osem_prepare(params, y, ones, ones, None, None, subsets=[sub],
                      mask=np.ones((320, 320, 127), bool))
    x1 = osem_run(st, 1)
    fp = nip.frwd_prj(x1, params, isub=sub, dev_out=True, fullsino_out=False)
    tot = float(fp.sum(dtype=torch.float64))
    ysum = float(y[sub].sum(dtype=np.float64))
    assert abs(tot - ysum) / ysum < 1e-5


def test_osem_4x14_full_vs_golden(torch, params, subsets):
    """Full config: 4 iterations x 14 subsets, span-11, 320x320x127, on the
    synthetic inputs of tests/golden/osem_data.py; the oracle image (FP64)
    is stored as 200,000 sampled voxels in osem4x14_sample.npz."""
    import sys
    sys.path.insert(0, os.path.join(HERE, 'golden'))
    from osem_data import checksum, make_osem_data

    import nipet_amd as nip
    from nipet_amd.prj.osem import osem_prepare, osem_run
    fx = os.path.join(HERE, 'golden', 'osem4x14_sample.npz')
    if not os.path.exists(fx):
        pytest.skip('golden fixture not generated (tests/golden/make_osem_golden.py)')
    g = np.load(fx)
    d = make_osem_data(params)
    assert tuple(checksum(d['prompts'])) == tuple(int(v) for v in g['checksum']), \
        'synthetic inputs differ from the ones the fixture was made from'
    mask = nip.get_cylinder(params['Cnt'], rad=29., unival=1, gpu_dim=True) > 0.9
    st = osem_prepare(params, d['prompts'], d['attn'], d['norm'], d['randoms'], d['scatter'],
                      subsets=subsets, mask=mask)
    img = osem_run(st, int(g['niter'])).cpu().numpy().astype(np.float64).ravel()
    ref = g['val']
    got = img[g['idx']]
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert rel < 1e-4, rel
    assert abs(img.sum() - float(g['total'])) / float(g['total']) < 1e-4
    assert abs(np.sqrt((img * img).sum()) - float(g['l2'])) / float(g['l2']) < 1e-4


def test_psf_conv_vs_oracle(torch, params):
    """nipet_psf_conv of one image and of a 3-image batch against the FP64
    oracle convolution (1e-6 relative)."""
    from oracle import oracle as O

    from nipet_amd.prj.mmrrec import psf_config
    from nipet_amd.prj.osem import psf_conv
    k = psf_config([4.5, 3.0, 6.0], params['Cnt'])
    rng = np.random.default_rng(2)
    x = rng.random((320, 320, 127)).astype(np.float32)
    x[:, :, :5] = 0
    g = psf_conv(torch.from_numpy(x).cuda(), k).cpu().numpy()
    ref = O.psf_conv(x, k)
    assert np.abs(g - ref).max() <= 1e-6 * np.abs(ref).max()
    xb = np.stack([x, 2 * x[::-1], x[:, ::-1]], axis=3).astype(np.float32)
    gb = psf_conv(torch.from_numpy(np.ascontiguousarray(xb)).cuda(), k).cpu().numpy()
    for b in range(3):
        rb = O.psf_conv(xb[..., b], k)
        assert np.abs(gb[..., b] - rb).max() <= 1e-6 * np.abs(rb).max()


@pytest.mark.parametrize('kind', ['planes', 'disc'])
def test_osem_psf_mini_vs_oracle(torch, params, subsets, kind):
    """OSEM with resolution modelling (convolved sensitivity, H x projected,
    back projection convolved) against the FP64 oracle: 1e-4 relative L2 (with
    a disc mask the segment windows are dilated by the kernel's reach)."""
    from oracle import oracle as O

    from nipet_amd.prj.mmrrec import psf_config
    from nipet_amd.prj.osem import osem_prepare, osem_run
    mini = _mini(params, subsets)
    rows_all = np.concatenate(mini)
    y, a, n, r = _data(params, rows_all, seed=3)
    mask = _mask(params, kind)
    k = psf_config(4.0, params['Cnt'])
    st = osem_prepare(params, y, a, n, r, None, subsets=mini, mask=mask, psf=k)
    gpu = osem_run(st, 2).cpu().numpy().astype(np.float64)
    an = a.astype(np.float64) * n
    rho = np.where(np.abs(an) < 1e-6, 0.0, r / np.where(np.abs(an) < 1e-6, 1.0, an))
    sens = np.stack([O.psf_conv(O.bprj(params, an[s], s), k) for s in mini])
    ref = O.osem(params, y, rho, sens, mask, mini, np.ones((320, 320, 127)), 2, krnl=k)
    rel = np.linalg.norm(gpu - ref) / np.linalg.norm(ref)
    print(f'parity osem psf mini ({kind} mask): {rel:.3g} rel L2 (bar 1e-4)', flush=True)
    assert rel < 1e-4, rel


@pytest.mark.parametrize('psf', [False, True])
def test_petprj_osem_one_call_vs_oracle(torch, params, subsets, psf):
    """petprj.osem through the one-call C entry nipet_osem_iter (reference
    argument list: prompts, randoms, scatter, norm, attenuation, the (Nsub,
    Nprj) subset table with ragged counts, unconvolved sensitivity images,
    mask, PSF kernel) against the FP64 oracle OSEM: 2 calls (iterations) of 3
    sub-sampled subsets within 1e-4 relative L2, NumPy arrays updated in place."""
    from oracle import oracle as O

    from nipet_amd.prj import petprj
    from nipet_amd.prj.mmrrec import psf_config
    mini = [subsets[k][::s].astype(np.int32) for k, s in ((1, 25), (6, 31), (12, 40))]
    rows_all = np.concatenate(mini)
    y, a, n, r = _data(params, rows_all, seed=3)
    sc = np.zeros_like(r)
    sc[rows_all] = np.random.default_rng(4).uniform(0.0, 0.3, (len(rows_all), 837))
    mask = _mask(params, 'disc')
    mask[:, :, 120:] = False
    an = a.astype(np.float64) * n
    rs = r.astype(np.float64) + sc
    rho = np.where(np.abs(an) < 1e-6, 0.0, rs / np.where(np.abs(an) < 1e-6, 1.0, an))
    sens = np.stack([O.bprj(params, an[s], s) for s in mini])
    krnl = psf_config(4.0 if psf else None, params['Cnt'])
    table = petprj.subset_table(mini)
    assert table.shape == (3, max(len(s) for s in mini) + 1)
    img = np.ones((320, 320, 127), np.float32)
    for it in range(2):
        print(f'petprj.osem psf={psf} call {it}', flush=True)
        petprj.osem(img, y, r, sc, n, a, table, sens.astype(np.float32), mask, krnl,
                    params['txLUT'], params['axLUT'], params['Cnt'])
    if psf:
        sens_rm = np.stack([O.psf_conv(x, krnl) for x in sens])
        ref = O.osem(params, y, rho, sens_rm, mask, mini, np.ones((320, 320, 127)), 2, krnl=krnl)
    else:
        ref = O.osem(params, y, rho, sens, mask, mini, np.ones((320, 320, 127)), 2)
    rel = np.linalg.norm(img - ref) / np.linalg.norm(ref)
    print(f'parity petprj.osem psf={psf}: {rel:.3g} rel L2 (bar 1e-4)', flush=True)
    assert rel < 1e-4, rel
    assert np.all(img[~mask] == 0)
