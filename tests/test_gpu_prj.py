"""GPU parity of the projectors against the FP64 oracle (oracle/nipet_oracle.c).

Bars: transaxial intersection lists bit-exact (voxel indices equal, t and
length equal to the float32 rounding of the oracle's doubles); forward
projections within 1e-5 relative error per bin and back projections within
1e-5 relative error per voxel, each relative to max(|ref|, 1e-6 max|ref|)
(so bins and voxels down to a millionth of the largest are held to 1e-5 of
their own value); adjoint identity <Ax,y> = <x,A^T y> to 1e-6 relative.
Every check prints its measured margin.
"""
import numpy as np
import pytest

from conftest import margin, phantom, rel_err, worst_elem

pytestmark = pytest.mark.gpu

FP_RTOL = 1e-5
BP_RTOL = 1e-5


def _rows(subset, n=192, seed=0):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(subset, size=n, replace=False)).astype(np.int32)


@pytest.fixture(scope='module')
def torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


@pytest.fixture(scope='module')
def img():
    return phantom(1)


def _ok(name, gpu, ref, bar=FP_RTOL):
    gpu = gpu.cpu().numpy() if hasattr(gpu, 'cpu') else np.asarray(gpu)
    idx, lvl, _ = worst_elem(gpu, ref)
    print(f'  worst element of {name}: index {tuple(int(i) for i in idx)}, |ref| = {lvl:.3g} of max',
          flush=True)
    return margin(name, rel_err(gpu, ref), bar)


def test_tx_lists_bit_exact(torch, params):
    from oracle import oracle as O

    from nipet_amd.geom import get_geom
    g = get_geom(params)
    off, col, t, ln = g.tx_lists()
    rays = np.arange(params['txLUT']['Naw'], dtype=np.int32)
    ooff, ocol, ot, oln = O.tx_lists(params, rays)[:4]
    assert np.array_equal(off, ooff)
    assert np.array_equal(col, ocol)
    assert np.array_equal(t, ot.astype(np.float32))
    assert np.array_equal(ln, oln.astype(np.float32))


def test_fprj_emission_vs_oracle(torch, params, subsets, img):
    from oracle import oracle as O

    import nipet_amd as nip
    sub = subsets[0]
    gpu = nip.frwd_prj(img, params, isub=sub, dev_out=True, fullsino_out=False)
    assert gpu.shape == (len(sub), 837)
    pick = np.random.default_rng(3).choice(len(sub), 160, replace=False)
    ref = O.fprj(params, img, sub[pick])
    assert _ok('fp span-11 subset 0', gpu[pick], ref)


def test_fprj_all_bins_default_isub(torch, params, img):
    from oracle import oracle as O

    import nipet_amd as nip
    gpu = nip.frwd_prj(img, params, dev_out=True)
    assert gpu.shape == (68516, 837)
    rows = np.random.default_rng(4).choice(68516, 96, replace=False).astype(np.int32)
    assert _ok('fp span-11 all bins', gpu[rows], O.fprj(params, img, rows))


def test_fprj_attenuation_vs_oracle(torch, params, subsets):
    from oracle import oracle as O

    import nipet_amd as nip
    mu = 0.096 * (phantom(2, nsph=0) > 0).astype(np.float32)
    sub = subsets[5]
    gpu = nip.frwd_prj(mu, params, isub=sub, dev_out=True, fullsino_out=False, attenuation=True)
    pick = np.random.default_rng(5).choice(len(sub), 128, replace=False)
    ref = O.fprj(params, mu, sub[pick], att=True)
    assert np.all(gpu <= 1.0 + 1e-6) and np.all(gpu > 0)
    assert _ok('fp attenuation span-11', gpu[pick], ref)
    # mu = 0 gives attenuation factors of exactly 1
    one = nip.frwd_prj(np.zeros_like(mu), params, isub=sub[:64], dev_out=True,
                       fullsino_out=False, attenuation=True)
    assert np.abs(one - 1).max() < 1e-6


def test_bprj_vs_oracle(torch, params, subsets):
    from oracle import oracle as O

    import nipet_amd as nip
    rows = _rows(subsets[2], 160, seed=7)
    y = np.random.default_rng(8).random((len(rows), 837)).astype(np.float32)
    gpu = nip.back_prj(y, params, isub=rows, dev_out=True)
    ref = O.bprj(params, y.astype(np.float64), rows)
    assert _ok('bp span-11 per voxel', gpu, ref, BP_RTOL)
    assert abs(gpu.sum(dtype=np.float64) - ref.sum()) / ref.sum() < 1e-6


def test_adjoint_identity_gpu(torch, params, subsets, img):
    import nipet_amd as nip
    sub = subsets[0]
    x = torch.from_numpy(img).cuda()
    y = torch.rand((len(sub), 837), device='cuda', generator=torch.Generator('cuda').manual_seed(0))
    Ax = nip.frwd_prj(x, params, isub=sub, dev_out=True, fullsino_out=False)
    Aty = nip.back_prj(y, params, isub=sub, dev_out=True)
    a = float((Ax.double() * y.double()).sum())
    b = float((x.double() * Aty.double()).sum())
    assert margin('adjoint span-11 subset 0', abs(a - b) / abs(a), 1e-6)


def test_gaps_roundtrip_gpu(torch, params):
    import nipet_amd as nip
    Cnt, tx = params['Cnt'], params['txLUT']
    s = torch.rand((tx['Naw'], 837), device='cuda')
    full = nip.putgaps(s, tx, Cnt)
    assert tuple(full.shape) == (837, 252, 344)
    back = nip.remgaps(full, tx, Cnt)
    assert torch.equal(back, s)
    # against the host restatement
    host = nip.putgaps(s[:, :].cpu().numpy(), tx, Cnt)
    assert np.array_equal(host, full.cpu().numpy())
    u16 = torch.randint(0, 1000, (837, 252, 344), dtype=torch.int32, device='cuda')
    u16 = (u16 * torch.from_numpy(tx['msino'].T.copy()).cuda()[None]).to(torch.uint16)
    r = nip.remgaps(u16, tx, Cnt)
    ref = nip.remgaps(u16.cpu().numpy(), tx, Cnt)
    assert np.array_equal(r.cpu().numpy(), ref.astype(np.float32))


def test_operator_shapes(torch, params, img):
    import nipet_amd as nip
    Cnt = params['Cnt']
    e7 = nip.convert2e7(img, Cnt)
    sino = nip.frwd_prj(e7, params, isub=np.arange(100, 200, dtype=np.int32))
    assert sino.shape == (837, 252, 344)
    bim = nip.back_prj(sino, params)
    assert bim.shape == (127, 344, 344)
    with pytest.raises(ValueError):
        nip.frwd_prj(np.zeros((10, 10, 10), np.float32), params)


def test_ragged_and_invalid_rows_through_c_abi(torch, params, img):
    """Edge cases through the C-ABI: the shortest and the longest transaxial
    bin projected alone (nprj = 1) against the oracle; bin indices outside
    0..naw-1 in a subset list give zero FP rows and add nothing to the BP;
    argument errors come back as status codes with a message."""
    from oracle import oracle as O

    from nipet_amd import _lib
    from nipet_amd.geom import get_geom
    g = get_geom(params)
    naw = params['txLUT']['Naw']
    segc = g.seg_counts()
    L, st = _lib.lib(), _lib.stream_handle()
    x = torch.from_numpy(img).cuda()
    for aw in (int(np.argmin(segc)), int(np.argmax(segc))):
        print(f'single bin {aw} ({segc[aw]} segments)', flush=True)
        sub = torch.tensor([aw], dtype=torch.int32, device='cuda')
        out = torch.full((1, 837), 7.0, device='cuda')
        _lib.check(L.nipet_fprj(g.handle, out.data_ptr(), x.data_ptr(), sub.data_ptr(), 1, 0, st))
        ref = O.fprj(params, img, np.array([aw], np.int32))
        if np.abs(ref).max() == 0:
            assert float(out.abs().max()) == 0.0
        else:
            assert _ok(f'fp single bin {aw}', out.cpu().numpy(), ref)
    rows = np.array([5, -1, 40000, naw, 123], np.int32)
    sub = torch.from_numpy(rows).cuda()
    out = torch.full((5, 837), 7.0, device='cuda')
    _lib.check(L.nipet_fprj(g.handle, out.data_ptr(), x.data_ptr(), sub.data_ptr(), 5, 0, st))
    o = out.cpu().numpy()
    assert np.all(o[[1, 3]] == 0)
    assert _ok('fp valid rows among invalid', o[[0, 2, 4]],
               O.fprj(params, img, rows[[0, 2, 4]]))
    y = torch.rand((5, 837), device='cuda', generator=torch.Generator('cuda').manual_seed(21))
    b_all = torch.empty_like(x)
    _lib.check(L.nipet_bprj(g.handle, b_all.data_ptr(), y.data_ptr(), sub.data_ptr(), 5, st))
    keep = torch.tensor([0, 2, 4], device='cuda')
    yv, sv = y[keep].contiguous(), sub[keep].contiguous()
    b_valid = torch.empty_like(x)
    _lib.check(L.nipet_bprj(g.handle, b_valid.data_ptr(), yv.data_ptr(), sv.data_ptr(), 3, st))
    assert float((b_all - b_valid).abs().max() / b_valid.abs().max()) < 1e-6
    assert L.nipet_fprj(g.handle, out.data_ptr(), x.data_ptr(), None, 5, 0, st) == -1
    assert b'nprj=5' in L.nipet_last_error()
    assert L.nipet_bprj(g.handle, b_all.data_ptr(), y.data_ptr(), sub.data_ptr(), 0, st) == -1
    assert L.nipet_fprj_b(g.handle, out.data_ptr(), x.data_ptr(), sub.data_ptr(), 5, 65, 0,
                          st) == -1
    assert b'batch 65' in L.nipet_last_error()
    assert L.nipet_fprj(g.handle, out.data_ptr(), x.data_ptr(), sub.data_ptr(), 5, 2, st) == -1


def test_steepest_bins_vs_oracle(torch, params):
    """The bins whose longest segment is largest against the chord (where the
    steepest LORs cross two z boundaries in one segment, so the BP's base plane
    steps by 2 and its three-piece path runs): FP and BP against the oracle,
    and a check that such two-boundary segments exist in the selection."""
    from oracle import oracle as O

    import nipet_amd as nip
    from nipet_amd.geom import get_geom
    g = get_geom(params)
    off, _, _, ln = g.tx_lists()
    naw = params['txLUT']['Naw']
    crs = params['txLUT']['crs'].astype(np.float64)
    c = params['txLUT']['s2c'].astype(np.int64)
    pa = 0.5 * (crs[c[:, 0], :2] + crs[c[:, 0], 2:])
    pb = 0.5 * (crs[c[:, 1], :2] + crs[c[:, 1], 2:])
    atn = np.hypot(*(pb - pa).T)
    ok = np.diff(off) > 0
    lmax = np.zeros(naw)
    lmax[ok] = np.maximum.reduceat(ln.astype(np.float64), off[:-1][ok])
    dmax = int(np.abs(np.diff(params['axLUT']['li2rno'].astype(np.int64), axis=1)).max())
    ratio = 2.0 * dmax * lmax / atn          # z-boundary spacings per longest segment
    rows = np.sort(np.argsort(-ratio)[:48]).astype(np.int32)
    assert ratio[rows].max() > 1.0, ratio[rows].max()   # two boundaries fit in one segment
    # these chords graze the edge of the field of view: a random image covers it
    xr = np.random.default_rng(30).random((320, 320, 127), dtype=np.float32)
    fp = nip.frwd_prj(xr, params, isub=rows, dev_out=True, fullsino_out=False)
    ref = O.fprj(params, xr, rows)
    assert np.abs(ref).max() > 0
    assert _ok('fp steepest bins', fp, ref)
    y = np.random.default_rng(31).random((len(rows), 837)).astype(np.float32)
    bimg = nip.back_prj(y, params, isub=rows, dev_out=True)
    ref = O.bprj(params, y.astype(np.float64), rows)
    assert _ok('bp steepest bins per voxel', bimg, ref, BP_RTOL)


@pytest.fixture(scope='module')
def params1(params):
    """The same scanner in span 1 (4084 sinograms, one LOR per bin)."""
    Cnt = dict(params['Cnt'])
    Cnt['SPN'] = 1
    return {'Cnt': Cnt, 'txLUT': params['txLUT'], 'axLUT': params['axLUT']}


def test_span1_projectors_vs_oracle(torch, params1, subsets, img):
    """Span-1 geometry (li2sn1): emission and attenuation FP within 1e-5, BP
    within 1e-5 of the image maximum, adjoint identity to 1e-6."""
    from oracle import oracle as O

    import nipet_amd as nip
    sub = subsets[3]
    gpu = nip.frwd_prj(img, params1, isub=sub, dev_out=True, fullsino_out=False)
    assert gpu.shape == (len(sub), 4084)
    pick = np.random.default_rng(11).choice(len(sub), 96, replace=False)
    assert _ok('fp span-1', gpu[pick], O.fprj(params1, img, sub[pick]))
    mu = 0.096 * (phantom(2, nsph=0) > 0).astype(np.float32)
    att = nip.frwd_prj(mu, params1, isub=sub, dev_out=True, fullsino_out=False, attenuation=True)
    assert _ok('fp attenuation span-1', att[pick], O.fprj(params1, mu, sub[pick], att=True))
    rows = _rows(sub, 96, seed=12)
    y = np.random.default_rng(13).random((len(rows), 4084)).astype(np.float32)
    bimg = nip.back_prj(y, params1, isub=rows, dev_out=True)
    ref = O.bprj(params1, y.astype(np.float64), rows)
    assert _ok('bp span-1 per voxel', bimg, ref, BP_RTOL)
    x = torch.from_numpy(img).cuda()
    yt = torch.rand((len(sub), 4084), device='cuda',
                    generator=torch.Generator('cuda').manual_seed(14))
    ax = nip.frwd_prj(x, params1, isub=sub, dev_out=True, fullsino_out=False)
    aty = nip.back_prj(yt, params1, isub=sub, dev_out=True)
    a = float((ax.double() * yt.double()).sum())
    b = float((x.double() * aty.double()).sum())
    assert margin('adjoint span-1 subset 3', abs(a - b) / abs(a), 1e-6)
