"""Pinning the oracle OSEM (oracle/oracle.py::osem, restating the subset loop
of prj/src/recon.cu:459-497 with the corrections of :381-398) on the CPU.

The reference computes, per subset S, x' = mask * x / sens_S * A_S^T(y / (A_S x
+ rho)) with rho = (r + s) / (a n) and sens_S = A_S^T(a n). For the data model
ybar = a n (A x + rho) this is exactly MLEM, which fixes:

* a fixed point: noise-free y = a n (A x_true + rho) leaves x_true unchanged
  on the subset's support (pins the place and sign of rho, and that sens is
  the back projection of a n and not of ones);
* a generalised count identity: sum_i a_i n_i (A x')_i = sum_i y_i (A x)_i /
  ((A x)_i + rho_i) over the subset's bins; with rho = 0, a = n = 1 this is
  the Shepp-Vardi count preservation sum_i (A x')_i = sum_i y_i;
* the 1e-6 guards: a zero image stays zero (no 0/0).
Small ray sets keep the FP64 projections fast; each planted mistake below
(rho dropped, sens of ones, rho on the wrong side) must break the check.
"""
import numpy as np
import pytest

from conftest import phantom
from oracle import oracle as O


@pytest.fixture(scope='module')
def case(params, subsets):
    rng = np.random.default_rng(17)
    rows = [np.sort(rng.choice(subsets[k], 48, replace=False)).astype(np.int32) for k in (0, 7)]
    naw, nsn = params['txLUT']['Naw'], 837
    an = np.zeros((naw, nsn))
    rho = np.zeros((naw, nsn))
    for r in rows:
        an[r] = rng.uniform(0.3, 0.9, (len(r), nsn))
        rho[r] = rng.uniform(0.5, 4.0, (len(r), nsn))
    xt = phantom(1, nsph=6).astype(np.float64)
    return dict(rows=rows, an=an, rho=rho, xt=xt)


def _sens(params, an, rows):
    return np.stack([O.bprj(params, an[r], r) for r in rows])


def _mask():
    return np.ones((320, 320, 127), bool)


def test_noise_free_fixed_point(params, case):
    rows, an, rho, xt = case['rows'][:1], case['an'], case['rho'], case['xt']
    y = np.zeros_like(an)
    r = rows[0]
    y[r] = an[r] * (O.fprj(params, xt, r) + rho[r])
    sens = _sens(params, an, rows)
    x1 = O.osem(params, y, rho, sens, _mask(), rows, xt, 2)
    sup = sens[0] > 1e-6
    assert sup.sum() > 10000
    err = np.abs(x1[sup] - xt[sup]).max() / xt.max()
    assert err < 1e-10, err
    # planted mistakes break the fixed point
    x_norho = O.osem(params, y, np.zeros_like(rho), sens, _mask(), rows, xt, 1)
    assert np.abs(x_norho[sup] - xt[sup]).max() / xt.max() > 1e-2
    ones = _sens(params, np.where(an > 0, 1.0, 0.0), rows)
    x_ones = O.osem(params, y, rho, ones, _mask(), rows, xt, 1)
    assert np.abs(x_ones[sup] - xt[sup]).max() / xt.max() > 1e-2


def test_count_identity_two_subsets(params, case):
    """sum_i an_i (A x')_i = sum_i y_i (A x)_i / ((A x)_i + rho_i) for each
    subset update, over two subsets in sequence (x from a noisy start)."""
    rows, an, rho = case['rows'], case['an'], case['rho']
    rng = np.random.default_rng(5)
    xt = case['xt']
    y = np.zeros_like(an)
    for r in rows:
        y[r] = rng.poisson(an[r] * (O.fprj(params, xt, r) + rho[r]))
    sens = _sens(params, an, rows)
    x = phantom(3, nsph=4).astype(np.float64) + 0.5
    for k, r in enumerate(rows):
        ax = O.fprj(params, x, r)
        x_new = O.osem(params, y, rho, sens[k:k + 1], _mask(), [r], x, 1)
        lhs = float((an[r] * O.fprj(params, x_new, r)).sum())
        rhs = float((y[r] * ax / (ax + rho[r])).sum())
        assert abs(lhs - rhs) / rhs < 1e-10, (k, lhs, rhs)
        # rho added to the prompts instead of the projection breaks it
        x_bad = O.osem(params, y + rho, np.zeros_like(rho), sens[k:k + 1], _mask(), [r], x, 1)
        bad = float((an[r] * O.fprj(params, x_bad, r)).sum())
        assert abs(bad - rhs) / rhs > 1e-3
        x = x_new


def test_mlem_count_preservation_and_guards(params, case):
    r = case['rows'][0]
    naw = params['txLUT']['Naw']
    one = np.zeros((naw, 837))
    one[r] = 1.0
    y = np.zeros((naw, 837))
    y[r] = np.random.default_rng(9).poisson(O.fprj(params, case['xt'], r))
    sens = _sens(params, one, [r])
    x = O.osem(params, y, np.zeros_like(y), sens, _mask(), [r], np.ones((320, 320, 127)), 1)
    assert abs(O.fprj(params, x, r).sum() - y[r].sum()) / y[r].sum() < 1e-10
    # zero image: e = 0 -> ratio 0 (guard), sens = 0 off the rays -> 0; no NaN
    z = O.osem(params, y, np.zeros_like(y), sens, _mask(), [r], np.zeros((320, 320, 127)), 1)
    assert np.isfinite(z).all() and not z.any()
    # the mask zeroes what it excludes
    m = _mask()
    m[:, :160] = False
    xm = O.osem(params, y, np.zeros_like(y), sens, m, [r], np.ones((320, 320, 127)), 1)
    assert not xm[:, :160].any() and xm[:, 160:].any()
