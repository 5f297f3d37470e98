"""Pins of the FP64 normalisation oracle (oracle.norm_sino, following dev_norm
src/norm.cu:13-51), of the online variance (oracle.varon, src/auxmath.cu:15-22)
and of the component-file reader (nipet_amd.mmrnorm.get_components,
mmrnorm.py:19-112), by facts that do not restate them:

* unit components and zero singles give unit factors in span 1 and span 11
  (span 11 is a mean over its span-1 LORs);
* the crystals of each bin are the histogrammer's crystal pair (txLUT s2c),
  first crystal with the first ring;
* uniform singles with constant dead-time parameters give the closed form
  (exp(q p / (1 + q n)) / (1 + q n))^2;
* one live bucket changes exactly the LORs with an end in its 18 crystals x
  8 rings (2 x 1 detector blocks of 9 slots x 8 rings);
* separable crystal efficiencies u[c] v[r] factor out; span 11 = sum over
  members of span-1 / (axf1 nos);
* Welford's recursion reproduces the mean and variance of the realisations.
"""
import os

import numpy as np
import pytest

from oracle import oracle


def _ones(params):
    return {'geo': np.ones((127, 344)), 'cinf': np.ones((344, 9)), 'ceff': np.ones((64, 504)),
            'axe1': np.ones(837), 'axf1': np.ones(4084), 'dtp': np.zeros(64),
            'dtnp': np.zeros(64)}


def _random_comp(seed=1):
    rng = np.random.default_rng(seed)
    return {'geo': rng.uniform(0.8, 1.2, (127, 344)), 'cinf': rng.uniform(0.9, 1.1, (344, 9)),
            'ceff': rng.uniform(0.7, 1.3, (64, 504)), 'axe1': rng.uniform(0.9, 1.1, 837),
            'axf1': rng.uniform(0.75, 1.17, 4084), 'dtp': rng.uniform(1e-7, 1e-6, 64),
            'dtnp': rng.uniform(1e-7, 1e-6, 64)}


BINS = np.arange(0, 68516, 211)


def test_unit_components_give_ones(params):
    z = np.zeros(224)
    for span in (1, 11):
        n = oracle.norm_sino(_ones(params), z, params, span=span, bins=BINS)
        assert np.allclose(n, 1.0, rtol=1e-14, atol=0)


def test_crystals_are_histogram_pairs(params):
    """A bin's first crystal takes the first ring of the sinogram: with
    efficiencies 2 on ring 0 of crystal c and 1 elsewhere, exactly the span-1
    LORs (aw, si) with (s2c[aw][0], r0) or (s2c[aw][1], r1) = (c, 0) change."""
    s2c = np.asarray(params['txLUT']['s2c'], np.int64)
    rno = np.asarray(params['axLUT']['sn1_rno'], np.int64)
    c = int(s2c[BINS[5], 0])
    comp = _ones(params)
    comp['ceff'][0, c] = 2.0
    n = oracle.norm_sino(comp, np.zeros(224), params, span=1, bins=BINS)
    hit = ((s2c[BINS, 0][:, None] == c) & (rno[None, :, 0] == 0)) | \
          ((s2c[BINS, 1][:, None] == c) & (rno[None, :, 1] == 0))
    assert hit.any()
    assert np.array_equal(n != 1.0, hit)


def test_dead_time_closed_form(params):
    beta, p, q = 4.0e4, 3e-7, 5e-7
    comp = _ones(params)
    comp['dtp'][:] = p
    comp['dtnp'][:] = q
    h = 0.5 * beta
    d = np.exp(h * p / (1 + h * q)) / (1 + h * q)
    n = oracle.norm_sino(comp, np.full(224, beta), params, span=11, bins=BINS)
    assert np.allclose(n, d * d, rtol=1e-13)


def test_single_bucket_footprint(params):
    b = 28 * 3 + 11
    bk = np.zeros(224)
    bk[b] = 5e4
    comp = _ones(params)
    comp['dtp'][:] = 1e-6
    n = oracle.norm_sino(comp, bk, params, span=1, bins=BINS)
    s2c = np.asarray(params['txLUT']['s2c'], np.int64)[BINS]
    rno = np.asarray(params['axLUT']['sn1_rno'], np.int64)

    def inb(c, r):
        return (c // 9 // 2 == b % 28) & (r // 8 == b // 28)
    hit = inb(s2c[:, 0][:, None], rno[None, :, 0]) | inb(s2c[:, 1][:, None], rno[None, :, 1])
    assert hit.any() and (~hit).any()
    assert np.array_equal(n != 1.0, hit)


def test_separable_efficiencies_and_span_relation(params):
    rng = np.random.default_rng(4)
    comp = _random_comp(3)
    u, v = rng.uniform(0.5, 1.5, 504), rng.uniform(0.5, 1.5, 64)
    comp['ceff'] = v[:, None] * u[None, :]
    comp['dtp'][:] = 0
    bk = rng.uniform(1e4, 1e5, 224)
    n1 = oracle.norm_sino(comp, bk, params, span=1, bins=BINS)
    ax, tx = params['axLUT'], params['txLUT']
    s2c = np.asarray(tx['s2c'], np.int64)[BINS]
    rno = np.asarray(ax['sn1_rno'], np.int64)
    s11 = np.asarray(ax['sn1_sn11'], np.int64)
    ali = np.asarray(tx['aw2ali'], np.int64)[BINS]
    w, a = ali % 344, ali // 344
    # the efficiencies factor out of the same components with unit efficiencies
    ref = comp.copy()
    ref['ceff'] = np.ones((64, 504))
    n1u = oracle.norm_sino(ref, bk, params, span=1, bins=BINS)
    eff = u[s2c[:, 0]][:, None] * v[rno[:, 0]][None, :] * u[s2c[:, 1]][:, None] * \
        v[rno[:, 1]][None, :]
    assert np.allclose(n1, n1u * eff, rtol=1e-12)
    # geometric and interference factors: divide out with everything else unit
    g = _ones(params)
    g['geo'], g['cinf'] = comp['geo'], comp['cinf']
    ng = oracle.norm_sino(g, np.zeros(224), params, span=1, bins=BINS)
    assert np.allclose(ng, (comp['geo'][0][w] * comp['cinf'][w, a % 9])[:, None], rtol=1e-13)
    # span 11 from span 1
    n11 = oracle.norm_sino(comp, bk, params, span=11, bins=BINS)
    nos = np.asarray(ax['sn1_sn11no'], np.float64)
    want = np.zeros_like(n11)
    for si in range(4084):
        want[:, s11[si]] += n1[:, si] / (comp['axf1'][si] * nos[si])
    assert np.allclose(n11, want, rtol=1e-12)


def test_only_first_geo_row_used(params):
    comp = _random_comp(5)
    a = oracle.norm_sino(comp, np.full(224, 3e4), params, span=11, bins=BINS[:50])
    comp['geo'][1:] = 7.0
    b = oracle.norm_sino(comp, np.full(224, 3e4), params, span=11, bins=BINS[:50])
    assert np.array_equal(a, b)


def test_varon_oracle_mean_variance():
    rng = np.random.default_rng(0)
    xs = rng.normal(3.0, 2.0, (12, 1000))
    m1, m2 = np.zeros(1000), np.zeros(1000)
    for b, x in enumerate(xs):
        m1, m2 = oracle.varon(m1, m2, x, b)
    assert np.allclose(m1, xs.mean(0), rtol=1e-12)
    assert np.allclose(m2 / len(xs), xs.var(0), rtol=1e-10)


def test_get_components_reads_files(tmp_path, params):
    from nipet_amd.mmrnorm import get_components
    Cnt = params['Cnt']
    rng = np.random.default_rng(9)
    parts = {'geo': (127, 344), 'cinf': (344, 9), 'ceff': (64, 504), 'axe1': (837,),
             'dtp': (64,), 'dtnp': (64,), 'dtc': (9,), 'axe2': (837,)}
    arrs = {k: rng.uniform(0, 1, s).astype(np.float32) for k, s in parts.items()}
    fbf = tmp_path / 'norm.n'
    with open(fbf, 'wb') as f:
        for k in parts:
            arrs[k].tofile(f)
    hdr = ('!INTERFILE:=\n%comment:=synthetic\n'
           'scanner quantification factor (Bq*s/ECAT counts):=3.21987e+07\n!END OF INTERFILE:=\n')
    fhdr = tmp_path / 'norm.dcm'
    fhdr.write_bytes(b'\x00' * 128 + b'DICM' + b'\x12\x34' * 20 + hdr.encode() + b'\x00\x00')
    aux = tmp_path / 'aux'
    aux.mkdir()
    axf1 = rng.uniform(0.7, 1.2, 4084).astype(np.float32)
    np.save(aux / 'AxialFactorForSpan1.npy', axf1)
    rng.uniform(0, 1, 837).astype(np.float32).tofile(
        aux / 'RelativeScaleFactors_scatter_axial_ssrTOspan11.f32')
    rng.uniform(0, 1, 4084).astype(np.float32).tofile(
        aux / 'RelativeScaleFactors_scatter_axial_ssrTOspan1.f32')
    comp, txt = get_components({'nrm_bf': str(fbf), 'nrm_dcm': str(fhdr)}, Cnt, auxdata=str(aux))
    for k in parts:
        assert np.array_equal(comp[k], arrs[k]), k
    assert np.array_equal(comp['axf1'], axf1)
    assert comp['qf'] == pytest.approx(3.21987e7) and comp['qf_loc'] == 0.205
    assert txt.startswith('!INTERFILE')
    with pytest.raises(NameError):
        get_components({}, Cnt)
    os.remove(aux / 'AxialFactorForSpan1.npy')
    with pytest.raises(FileNotFoundError):
        get_components({'nrm_bf': str(fbf), 'nrm_ima': str(fhdr)}, Cnt, auxdata=str(aux))
