"""Pins of the FP64 randoms oracle (oracle.randoms, following gpu_randoms,
lm/src/rnd.cu:180-420) by facts that do not restate it:

* the reference's crystal walk (start of the run in cij + up to 320 crystals)
  visits exactly the crystals of the coincidence matrix cij;
* the fan-sum operator equals the fan-sums of the expected delayeds
  d(LOR) = s(c0, r0) s(c1, r1) summed over every LOR of the sinogram LUTs
  (s2cr bins x sn1_rno ring pairs), a count built from different tables;
* uniform singles a: the estimate is a after the sqrt start, randoms are
  a^2 times the number of span-1 sinograms of each output sinogram;
* the damped iteration converges to the singles that generated the fan-sums
  (the undamped one does not), and scaling the fan-sums by k^2 scales every
  iterate by k (the sqrt start).
"""
import numpy as np
import pytest

from oracle import oracle


@pytest.fixture(scope='module')
def fanop(params):
    Cnt = params['Cnt']
    return oracle.rnd_fan_matrices(params['txLUT']['cij'], Cnt['MRD'], Cnt['NRNG'])


def _fan_of_singles(params, s):
    """Expected delayed fan-sums [64][504] for crystal singles s [448][64],
    accumulated LOR by LOR over the transaxial bins and span-1 ring pairs."""
    tx, ax, Cnt = params['txLUT'], params['axLUT'], params['Cnt']
    s2cr = np.asarray(tx['s2cr'], np.int64)
    rno = np.asarray(ax['sn1_rno'], np.int64)
    M = np.zeros((64, 64))
    np.add.at(M, (rno[:, 0], rno[:, 1]), 1.0)          # ring pairs (r0, r1) of the sinograms
    S0, S1 = s[s2cr[:, 0]], s[s2cr[:, 1]]              # [Naw][64]
    fan = np.zeros((Cnt['NCRSR'], 64))
    np.add.at(fan, s2cr[:, 0], S0 * (S1 @ M.T))        # end 0 of every LOR
    np.add.at(fan, s2cr[:, 1], S1 * (S0 @ M))          # end 1
    out = np.zeros((64, Cnt['NCRS']))
    crsri = np.asarray(tx['crsri'], np.int64)
    live = crsri >= 0
    out[:, live] = fan[crsri[live]].T
    return out


def _smooth_singles(seed=2):
    rng = np.random.default_rng(seed)
    c = np.arange(448)[:, None] * 2 * np.pi / 448
    r = np.arange(64)[None, :] / 63.0
    return (1.0 + 0.25 * np.sin(c + rng.uniform(0, 6)) * np.cos(2 * r)
            + 0.05 * rng.uniform(-1, 1, (448, 64)))


def test_crystal_walk_is_cij(params, fanop):
    cij = params['txLUT']['cij']
    crange, _ = oracle.rnd_ranges(cij, params['Cnt']['MRD'])
    assert crange[3].max() <= oracle.RND_CFOR * oracle.RND_ROWS
    C, R = fanop
    assert np.array_equal(C, cij.astype(np.float64))
    assert np.array_equal(C, C.T)
    band = np.abs(np.arange(64)[:, None] - np.arange(64)[None, :]) <= params['Cnt']['MRD']
    assert np.array_equal(R, band.astype(np.float64))


def test_fan_operator_matches_lor_sums(params, fanop):
    s = _smooth_singles()
    C, R = fanop
    fan = _fan_of_singles(params, s)
    crsri = np.asarray(params['txLUT']['crsri'], np.int64)
    live = crsri >= 0
    pred = (s * (C @ s @ R.T))[crsri[live]].T
    assert np.allclose(fan[:, live], pred, rtol=1e-12, atol=0)
    assert np.all(fan[:, ~live] == 0)


def test_uniform_singles_closed_form(params, fanop):
    a = 3.0
    C, R = fanop
    ones = np.ones((448, 64))
    fan = _fan_of_singles(params, a * ones)
    bins = np.arange(0, params['txLUT']['Naw'], 97)
    for span in (11, 1):
        rs, cmap, s = oracle.randoms(fan, params, span=span, bins=bins)
        assert np.allclose(s, a, rtol=1e-13)
        nsn = params['Cnt']['NSN11'] if span == 11 else params['Cnt']['NSN1']
        mult = np.bincount(np.asarray(params['axLUT']['sn1_sn11'], np.int64), minlength=nsn) \
            if span == 11 else np.ones(nsn)
        assert rs.shape == (len(bins), nsn)
        assert np.allclose(rs, a * a * mult[None, :], rtol=1e-13)
    live = np.asarray(params['txLUT']['crsri']) >= 0
    assert np.allclose(cmap[live], a) and np.all(cmap[~live] == 0)


def test_damped_iteration_converges_to_generating_singles(params):
    s_true = _smooth_singles(5)
    fan = _fan_of_singles(params, s_true)
    errs = []
    for k in (0, 2, 5, 10, 60):
        _, _, s = oracle.randoms(fan, params, niter=k, sino=False)
        errs.append(np.abs(s / s_true - 1).max())
    assert all(b < a for a, b in zip(errs, errs[1:]))
    assert errs[-1] < 1e-9
    # the same map without the damping oscillates about the solution
    C, R = oracle.rnd_fan_matrices(params['txLUT']['cij'], params['Cnt']['MRD'])
    crsri = np.asarray(params['txLUT']['crsri'], np.int64)
    f = fan[:, crsri >= 0].T
    s = np.sqrt(f / (C @ np.ones_like(f) @ R.T))
    for _ in range(40):
        s = f / (C @ s @ R.T)
    assert np.abs(s / s_true - 1).max() > 1e-3


def test_scaling_fansums_scales_singles(params):
    rng = np.random.default_rng(7)
    fan = rng.integers(50, 500, size=(64, 504)).astype(np.float64)
    _, c1, _ = oracle.randoms(fan, params, sino=False)
    _, c2, _ = oracle.randoms(4.0 * fan, params, sino=False)
    assert np.allclose(c2, 2.0 * c1, rtol=1e-12, atol=0)


def test_zero_fansums_guard(params):
    fan = np.zeros((64, 504))
    rs, cmap, _ = oracle.randoms(fan, params, bins=np.arange(100))
    assert np.all(rs == 0) and np.all(cmap == 0)
    _, c, _ = oracle.randoms(fan, params, guard=False, sino=False)
    assert np.isnan(c[np.asarray(params['txLUT']['crsri']) >= 0]).all()
