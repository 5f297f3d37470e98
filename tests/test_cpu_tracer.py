"""The float32 OpenMP CPU baseline (oracle/cpu_tracer.cpp): the reference's
tracer, OSEM subset update and histogram loop, timed by bench.py. These
tests make sure the timed code computes what the reference computes: walks
stay inside the image, FP agrees with the FP64 oracle within the float32
tracer's error, BP is FP's adjoint, the OSEM glue routes the subset's rows,
and the histogram equals the list-mode oracle where their time-tag rules
agree."""
import numpy as np
import pytest

from conftest import phantom


@pytest.fixture(scope='module')
def tracer(params):
    from oracle import cpu_tracer as C
    return C.Tracer(params)


def test_tables_and_walks_in_bounds(tracer, subsets):
    assert tracer.rec.shape == (68516, 10)
    assert int(tracer.rec[:, 9].max()) < 1807
    # subset 0 is the bench's sample; 7 and 13 are the other ends of the angles
    for k in (0, 7, 13):
        assert tracer.check(subsets[k]) == 0


def test_fprj_against_fp64_oracle(params, tracer, subsets):
    """The reference's float32 tracer (single z split per transaxial step,
    z crossings spaced (tz2 - tz1) / nz) against the exact FP64 crossing walk:
    measured 3.0e-5 of the largest bin; bar 2e-4."""
    from oracle import oracle as O
    rows = np.concatenate([subsets[0][::60], subsets[9][::97]]).astype(np.int32)
    img = phantom(1)
    got = tracer.fprj(img, rows)
    ref = O.fprj(params, img, rows)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f'cpu baseline FP vs FP64 oracle: {err:.3g} of max (bar 2e-4)')
    assert err < 2e-4
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 2e-5
    att = tracer.fprj(phantom(2) * 0.01, rows, att=True)
    aref = O.fprj(params, phantom(2) * 0.01, rows, att=True)
    assert np.abs(att - aref).max() < 1e-4


def test_bprj_is_adjoint(tracer, subsets):
    rng = np.random.default_rng(3)
    rows = subsets[4][::11]
    x = phantom(4)
    y = rng.random((len(rows), 837)).astype(np.float32)
    lhs = float((tracer.fprj(x, rows).astype(np.float64) * y).sum())
    rhs = float((x.astype(np.float64) * tracer.bprj(y, rows)).sum())
    assert abs(lhs - rhs) / abs(lhs) < 1e-6


def test_osem_subset_glue(tracer, subsets):
    """cb_osem_subset = the reference's subset body over the baseline's own
    projectors: rows of prompts / rho taken at the subset's bins, 0 where
    |e| < 1e-6 or |sens| < 1e-6, masked voxels set to 0."""
    rng = np.random.default_rng(5)
    rows = subsets[2][::23]
    naw = 68516
    x = phantom(1) + 0.1
    prompts = rng.integers(0, 20, (naw, 837)).astype(np.uint16)
    rho = rng.uniform(0, 0.5, (naw, 837)).astype(np.float32)
    rho[rows[0], :5] = 0            # e = FP + rho may vanish
    sens = tracer.bprj(np.ones((len(rows), 837), np.float32), rows)
    sens[:, :, :3] = 0              # |sens| < 1e-6
    mask = np.ones((320, 320, 127), np.uint8)
    mask[:40] = 0
    got = x.copy()
    tracer.osem_subset(got, prompts, rho, sens, mask, rows)
    e = tracer.fprj(x, rows) + rho[rows]
    q = np.where(np.abs(e) < 1e-6, 0, prompts[rows] / np.where(e == 0, 1, e)).astype(np.float32)
    b = tracer.bprj(q, rows)
    b = np.where(np.abs(sens) < 1e-6, 0, b / np.where(sens == 0, 1, sens))
    ref = np.where(mask > 0, x * b, 0)
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()
    assert np.all(got[:40] == 0)


def _stream(params, seconds, per_ms, seed):
    """Host word stream: a time tag every millisecond followed by per_ms - 1
    events (85 % prompts) on active bins; 512 words per tag, so every
    4096-word patch starts with a tag."""
    rng = np.random.default_rng(seed)
    n = seconds * 1000
    aw2ali = np.asarray(params['txLUT']['aw2ali'], np.int64)
    si = rng.integers(0, 4084, (n, per_ms - 1))
    aw = aw2ali[rng.integers(0, len(aw2ali), (n, per_ms - 1))]
    ev = si * 86688 + aw
    ev |= (rng.random((n, per_ms - 1)) < 0.85).astype(np.int64) << 30
    tags = (np.int64(-(1 << 31)) | (1000 + np.arange(n, dtype=np.int64)))[:, None]
    return np.concatenate([tags, ev], axis=1).reshape(-1).astype(np.int32)


def test_hst_against_lm_oracle(params):
    """Every 4096-word patch starts with a time tag, so the reference's
    patch-first-tag rule and the oracle's preceding-tag rule agree: all
    outputs bit-exact (delayeds are the upper 16 bits of the sinogram)."""
    from oracle import cpu_tracer as C
    from oracle import oracle as O
    w = _stream(params, 6, 512, 9)
    got = C.hst(w, params, toff=1000, nitag=6)
    ref = O.lm_hist(w, params, 1000, 6)
    assert got['counted'] == ref['counted'] == 6 * 1000 * 511
    assert np.array_equal(got['psino'] & 0xffff, ref['psino'][0])
    assert np.array_equal(got['psino'] >> 16, ref['dsino'][0])
    for a, b in (('ssrb', 'ssr'), ('zR', 'zR'), ('zM', 'zM'), ('snview', 'pvs'),
                 ('fansums', 'fan')):
        assert np.array_equal(got[a], ref[b]), a
    assert np.array_equal(got['rprmt'], ref['phc'][0])
    assert np.array_equal(got['rdlyd'], ref['dhc'][0])
    # frame window: seconds 2-4 only
    got = C.hst(w, params, toff=1000, nitag=6, tstart=2, tstop=4)
    ref = O.lm_hist(w, params, 1000, 6, tstart=2, tstop=4)
    assert np.array_equal(got['psino'] & 0xffff, ref['psino'][0])
    assert np.array_equal(got['rprmt'], ref['phc'][0])


def test_timers(params, tracer, subsets):
    from oracle import cpu_tracer as C
    fp, bp = tracer.time_pair(phantom(1), subsets[0][::40], runs=2)
    assert fp.shape == bp.shape == (2,) and np.all(fp > 0) and np.all(bp > 0)
    secs, n = C.time_hst(_stream(params, 2, 256, 1), params, 1000, 2, runs=2)
    assert n == 2 * 1000 * 255 and np.all(secs > 0)
    info = C.host_info()
    assert info['omp_threads'] >= 1 and info['cpu_model']
