"""Geometry / LUT invariants (CPU). Cross-checks tables that the reference
defines twice, and the sizes fixed in niftypet/nipet/include/def.h."""
import numpy as np

from nipet_amd import mmraux
from nipet_amd.img import mmrimg


def test_sizes(params):
    Cnt, tx, ax = params['Cnt'], params['txLUT'], params['axLUT']
    assert tx['Naw'] == 68516 == Cnt['Naw']            # AW, include/def.h:66
    assert len(ax['li2rno']) == 2074                    # NLI2R, include/def.h:67
    assert len(ax['sn1_sn11']) == 4084 and ax['Msn11'].max() == 836
    assert tx['msino'].sum() == 68516


def test_span11_two_definitions_agree(params):
    # axial_lut (mmraux.py:339-544) vs span1_span11 (src/scanner_0.cu:44-87):
    # pins the MNRD/MXRD segment order of the un-vendored constants.
    li2s11, nos = mmraux.span1_span11(params['Cnt'])
    ax = params['axLUT']
    assert np.array_equal(li2s11, ax['sn1_sn11'])
    assert np.array_equal(np.bincount(ax['sn1_sn11'], minlength=837), nos)
    # span-11 sinograms per segment follow SEG
    seg_of_s11 = np.repeat(np.arange(11), params['Cnt']['SEG'])
    s11 = ax['sn1_sn11']
    rd = ax['sn1_rno'][:, 1].astype(int) - ax['sn1_rno'][:, 0].astype(int)
    for iseg, (lo, hi) in enumerate(zip(params['Cnt']['MNRD'], params['Cnt']['MXRD'])):
        # sn1_sn11 uses Msn[r0, r1], i.e. the ring difference r0 - r1
        sel = (-rd >= lo) & (-rd <= hi)
        assert np.all(seg_of_s11[s11[sel]] == iseg)


def test_li_tables(params):
    ax = params['axLUT']
    r = ax['li2rno'].astype(int)
    assert np.all(r[:, 1] >= r[:, 0]) and np.all(r[:64, 0] == r[:64, 1])
    assert np.all(np.diff(r[:, 1] - r[:, 0]) >= 0)          # ring-difference major
    assert np.abs(r[:, 1] - r[:, 0]).max() == 60
    z = ax['li2rng']
    assert np.allclose(z, (r - 32) * 0.40625)
    # every span-1 LOR (both directions) is covered exactly once
    cnt = np.zeros((64, 64), int)
    for (a, b) in r:
        cnt[a, b] += 1
        if a != b:
            cnt[b, a] += 1
    assert np.array_equal(cnt, (np.abs(np.subtract.outer(np.arange(64), np.arange(64))) <= 60))
    # li2sn / li2nos consistent with Msn11 / Mnos
    M, Mn = ax['Msn11'], ax['Mnos']
    assert np.array_equal(ax['li2sn'][:, 0], M[r[:, 0], r[:, 1]])
    assert np.array_equal(ax['li2sn'][:, 1], M[r[:, 1], r[:, 0]])
    assert np.array_equal(ax['li2nos'], Mn[r[:, 1], r[:, 0]])
    # sinos per span-11 bin = number of (ordered) LORs mapped to it
    hits = np.bincount(np.concatenate([ax['li2sn'][:, 0], ax['li2sn'][64:, 1]]), minlength=837)
    assert np.array_equal(hits[ax['li2sn'][:, 0]], ax['li2nos'])


def test_crystals(params):
    crs = params['txLUT']['crs'].astype(np.float64)
    gap = (np.arange(504) % 9) == 8
    assert np.all(crs[gap] == 0)
    w = np.hypot(crs[~gap, 2] - crs[~gap, 0], crs[~gap, 3] - crs[~gap, 1])
    assert np.allclose(w, 3.209 / 8, atol=1e-5)
    r = np.hypot(0.5 * (crs[~gap, 0] + crs[~gap, 2]), 0.5 * (crs[~gap, 1] + crs[~gap, 3]))
    assert np.all((r > 33.47 - 1e-4) & (r < 33.47 + 0.2))


def test_sino_crystal_maps(params):
    tx = params['txLUT']
    s2c = tx['s2c'].astype(int)
    assert np.all((s2c % 9) != 8)                            # no gap crystal in active bins
    aw2sn = tx['aw2sn'].astype(int)
    assert np.array_equal(tx['aw2ali'], aw2sn[:, 1] + 344 * aw2sn[:, 0])
    iw = np.nonzero(tx['msino'])[0]
    assert np.array_equal(aw2sn[:, 1], iw)                   # radial-bin major order
    assert np.array_equal(tx['c2sF'][s2c[:, 0], s2c[:, 1]],
                          aw2sn[:, 0] + 252 * aw2sn[:, 1])


def test_subsets_partition(params):
    from nipet_amd import get_subsets14
    subs = [get_subsets14(n, params)[0] for n in range(14)]
    assert all(len(s) == 4894 for s in subs)
    allb = np.concatenate(subs)
    assert np.array_equal(np.sort(allb), np.arange(68516))
    S = get_subsets14(0, params)[1]
    assert np.array_equal(np.sort(S.ravel()), np.arange(252))


def test_gaps_host_roundtrip(params):
    Cnt, tx = params['Cnt'], params['txLUT']
    s = np.random.default_rng(0).random((tx['Naw'], 837)).astype(np.float32)
    full = mmraux.putgaps(s, tx, Cnt)
    assert full.shape == (837, 252, 344)
    assert np.all(full[:, tx['msino'].T == 0] == 0)
    assert np.array_equal(mmraux.remgaps(full, tx, Cnt), s)


def test_image_layouts(params):
    Cnt = params['Cnt']
    im = np.random.default_rng(1).random((320, 320, 127)).astype(np.float32)
    e7 = mmrimg.convert2e7(im, Cnt)
    assert e7.shape == (127, 344, 344)
    assert np.array_equal(mmrimg.convert2dev(e7, Cnt), im)
    assert e7[:, :12].sum() == 0 and e7[:, :, -12:].sum() == 0


def test_cylinder_mask(params):
    Cnt = params['Cnt']
    m = mmrimg.get_cylinder(Cnt, rad=29., unival=1, gpu_dim=True)
    assert m.shape == (320, 320, 127)
    area = m[:, :, 0].sum() * Cnt['SZ_VOXY']**2
    assert abs(area - np.pi * 29.**2) / (np.pi * 29.**2) < 0.01
    assert np.array_equal(m[:, :, 0], m[:, :, 100])
