"""Pinning the FP64 oracle (oracle/nipet_oracle.c) with checks the mathematics
fixes. The reference ships no golden vectors for this path (its only fixtures
are dataset-bound, tests/test_amyloid_pvc.py:80), so every oracle function is
checked against an independent closed form or brute force here:

* transaxial lists vs. brute-force point sampling of the chord (pins u/v
  conventions, stepping signs, skipped end pieces);
* single-plane images vs. the closed-form z-overlap of every LOR and mirror
  (pins the z-splitting, sqrt(1+(az/atn)^2), li2sn mapping and mirror index);
* uniform image, constant mu-map, adjoint identity, bin sums.
"""
import numpy as np
import pytest

from oracle import oracle as O

VOX = float(np.float32(0.208626))
VOXZ = 0.203125


@pytest.fixture(scope='module')
def rays():
    return np.array([0, 1, 5000, 12345, 30000, 34257, 34258, 50000, 68515, 40001], np.int32)


def _chord(params, aw):
    crs = params['txLUT']['crs'].astype(np.float64)
    c1, c2 = params['txLUT']['s2c'][aw].astype(int)
    p = 0.5 * (crs[c1, :2] + crs[c1, 2:])
    q = 0.5 * (crs[c2, :2] + crs[c2, 2:])
    atn = np.hypot(*(q - p))
    a = (q - p) / atn
    b = p @ a
    disc = b * b - (p @ p - 890.0)
    return p, a, atn, -b - np.sqrt(disc), -b + np.sqrt(disc)


def test_tx_lists_brute_force(params, rays):
    off, col, ts, ln, atn, t12 = O.tx_lists(params, rays)
    for i, aw in enumerate(rays):
        p, a, at, t1, t2 = _chord(params, aw)
        assert abs(at - atn[i]) < 1e-12 and abs(t1 - t12[i, 0]) < 1e-9
        s, e = off[i], off[i + 1]
        tt, ll, cc = ts[s:e], ln[s:e], col[s:e]
        # first segment starts on a grid line, after t1; contiguous segments
        assert tt[0] > t1 and tt[-1] + ll[-1] < t2
        assert np.allclose(tt[1:], tt[:-1] + ll[:-1], rtol=0, atol=1e-12)
        x0, y0 = p + a * tt[0]
        assert min(abs(x0 / VOX - round(x0 / VOX)), abs(y0 / VOX - round(y0 / VOX))) < 1e-9
        # the partial pieces [t1, first) and (last, t2] are shorter than a voxel diagonal
        assert tt[0] - t1 < VOX * np.sqrt(2) + 1e-9 and t2 - (tt[-1] + ll[-1]) < VOX * 1.5
        # brute force: voxel of sample points inside each segment
        rng = np.random.default_rng(aw)
        for _ in range(4):
            u01 = rng.random(len(tt)) * 0.98 + 0.01
            t = tt + u01 * ll
            x, y = p[0] + a[0] * t, p[1] + a[1] * t
            u = 160 + np.floor(x / VOX).astype(int)
            v = 160 - np.ceil(y / VOX).astype(int)
            assert np.array_equal(u + 320 * v, cc)
        # every grid crossing in (t1, t2) is a segment boundary
        nx = len(np.arange(np.ceil(min(p[0] + a[0] * t1, p[0] + a[0] * t2) / VOX),
                           np.floor(max(p[0] + a[0] * t1, p[0] + a[0] * t2) / VOX) + 1))
        ny = len(np.arange(np.ceil(min(p[1] + a[1] * t1, p[1] + a[1] * t2) / VOX),
                           np.floor(max(p[1] + a[1] * t1, p[1] + a[1] * t2) / VOX) + 1))
        assert len(tt) == nx + ny - 1


def _lor_tables(params, atn):
    ax = params['axLUT']
    z = ax['li2rng'].astype(np.float64) + 0.5 * 0.40625
    zeta = z / VOXZ + 63.5
    m = (zeta[:, 1] - zeta[:, 0]) / atn
    sf = np.sqrt(1 + ((z[:, 1] - z[:, 0]) / atn)**2)
    return zeta, m, sf


def _overlap(lo, hi, a, b):
    return np.clip(np.minimum(hi, b) - np.maximum(lo, a), 0, None)


@pytest.mark.parametrize('plane', [0, 1, 40, 63, 64, 101, 126])
def test_fprj_single_plane_closed_form(params, rays, plane):
    img = np.zeros((320, 320, 127), np.float32)
    img[:, :, plane] = 1.0
    fp = O.fprj(params, img, rays)
    off, col, ts, ln, atn, _ = O.tx_lists(params, rays)
    ax = params['axLUT']
    for i in range(len(rays)):
        ta, tb = ts[off[i]], ts[off[i + 1] - 1] + ln[off[i + 1] - 1]
        zeta, m, sf = _lor_tables(params, atn[i])
        exp = np.zeros(837)
        for li in range(len(zeta)):
            if m[li] == 0:
                L = (tb - ta) if int(np.floor(zeta[li, 0])) == plane else 0.0
                exp[ax['li2sn'][li, 0]] += L
                continue
            # LOR a: zeta = zeta0 + m t in [plane, plane+1)
            la = _overlap((plane - zeta[li, 0]) / m[li], (plane + 1 - zeta[li, 0]) / m[li], ta, tb)
            # mirror b: zeta = zeta1 - m t
            lb = _overlap((zeta[li, 1] - plane - 1) / m[li], (zeta[li, 1] - plane) / m[li], ta, tb)
            exp[ax['li2sn'][li, 0]] += la * sf[li]
            exp[ax['li2sn'][li, 1]] += lb * sf[li]
        assert np.abs(fp[i] - exp).max() < 1e-9 * max(1.0, exp.max())


def test_fprj_uniform_and_attenuation_closed_form(params, rays):
    off, col, ts, ln, atn, _ = O.tx_lists(params, rays)
    ax = params['axLUT']
    ones = np.ones((320, 320, 127), np.float32)
    fp = O.fprj(params, ones, rays)
    mu = 0.0961
    att = O.fprj(params, np.full((320, 320, 127), mu, np.float32), rays, att=True)
    for i in range(len(rays)):
        Ltx = ts[off[i + 1] - 1] + ln[off[i + 1] - 1] - ts[off[i]]
        _, _, sf = _lor_tables(params, atn[i])
        e = np.zeros(837)
        np.add.at(e, ax['li2sn'][:, 0], Ltx * sf)
        np.add.at(e, ax['li2sn'][64:, 1], Ltx * sf[64:])
        assert np.abs(fp[i] - e).max() < 1e-10 * e.max()
        ea = np.zeros(837)
        w = np.exp(-np.float64(np.float32(mu)) * Ltx * sf) / ax['li2nos']
        np.add.at(ea, ax['li2sn'][:, 0], w)
        np.add.at(ea, ax['li2sn'][64:, 1], w[64:])
        assert np.abs(att[i] - ea).max() < 1e-12


def test_adjoint_identity(params):
    rows = np.arange(3000, 68516, 1931, dtype=np.int32)
    rng = np.random.default_rng(0)
    x = rng.random((320, 320, 127)).astype(np.float32)
    y = rng.random((len(rows), 837))
    a = (O.fprj(params, x, rows) * y).sum()
    b = (x.astype(np.float64) * O.bprj(params, y, rows, nthreads=4)).sum()
    assert abs(a - b) / abs(a) < 1e-12


def test_bprj_single_bin(params):
    # back projection of a single span-11 bin = sum of its LORs' voxel weights,
    # whose total is the bin's closed-form uniform projection
    rows = np.array([34000], np.int32)
    y = np.zeros((1, 837))
    y[0, 400] = 1.0
    bp = O.bprj(params, y, rows, nthreads=2)
    fp1 = O.fprj(params, np.ones((320, 320, 127), np.float32), rows)
    assert abs(bp.sum() - fp1[0, 400]) < 1e-9 * fp1[0, 400]
    assert np.count_nonzero(bp) > 0


def test_term_count_formula(params):
    # the closed-form term count used by bench.py's byte model equals the
    # oracle's piece count
    from nipet_amd.geom import count_terms
    rays = np.array([0, 777, 12000, 34257, 50000, 68000], np.int32)
    tx = O.tx_lists(params, np.arange(68516, dtype=np.int32))[:4]
    ref = [O.fprj(params, np.zeros((320, 320, 127), np.float32), rays[i:i + 1],
                  return_terms=True)[1] for i in range(len(rays))]
    assert np.array_equal(count_terms(params, rays, tx), ref)


def test_psf_conv_pins():
    """The FP64 PSF convolution: a delta gives the outer product of the three
    kernels (closed form); it equals scipy.ndimage.correlate1d chained along
    z, x, y with zero boundaries (library routine); normalised kernels keep
    the sum of an image away from the edges; planting a z/y kernel swap is
    caught by an anisotropic kernel."""
    import numpy as np
    from scipy.ndimage import correlate1d

    from nipet_amd import get_mmrparams
    from nipet_amd.prj.mmrrec import psf_config
    from oracle import oracle as O
    Cnt = get_mmrparams()['Cnt']
    k = psf_config([4.5, 3.0, 6.0], Cnt).astype(np.float64)   # x, y, z FWHM [mm]
    img = np.zeros((320, 320, 127))
    img[100, 150, 60] = 1.0
    out = O.psf_conv(img, k)
    ref = np.zeros_like(img)
    ref[92:109, 142:159, 52:69] = (k[2][:, None, None] * k[1][None, :, None] *
                                   k[0][None, None, :])
    assert np.allclose(out, ref, rtol=0, atol=1e-15)
    rng = np.random.default_rng(0)
    x = rng.random((320, 320, 127))
    lib = correlate1d(correlate1d(correlate1d(x, k[0], axis=2, mode='constant'),
                                  k[1], axis=1, mode='constant'), k[2], axis=0, mode='constant')
    assert np.allclose(O.psf_conv(x, k), lib, rtol=1e-12, atol=1e-12)
    inner = np.zeros_like(x)
    inner[40:280, 40:280, 20:107] = x[40:280, 40:280, 20:107]
    kn = k / k.sum(axis=1, keepdims=True)        # float64 normalisation
    assert abs(O.psf_conv(inner, kn).sum() - inner.sum()) < 1e-9 * inner.sum()
    swapped = k[[2, 1, 0]]
    assert not np.allclose(O.psf_conv(img, swapped), ref, atol=1e-6)
