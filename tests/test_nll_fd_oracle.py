"""Pins the FP64 finite-difference checker of the Poisson log-likelihood
gradient (oracle orc_nll_fd) against the analytic gradients built from the
FP64 oracle projectors:
  dL/dx  = A^T (a n (1 - y/ybar)),
  dL/dmu = -A1^T (exp(-A1 mu) (1 - y_b/ybar_b) n_b p_b / n_LOR(b)),
A1 the span-1 projector (each span-11 attenuation bin averages exp(-l) over
its n_LOR(b) span-1 LORs, prj/src/prjf.cu:88-92, 193-200). The two sides
share no code: one differentiates by evaluation, the other by adjoint."""
import numpy as np
import pytest

from conftest import nll_case


@pytest.fixture(scope='module')
def case(params):
    return nll_case(params, [5000, 33000, 61000])


def _span1(params):
    p1 = dict(params)
    p1['Cnt'] = dict(params['Cnt'], SPN=1)
    ax = params['axLUT']
    li2sn1, li2sn = np.asarray(ax['li2sn1'], np.int64), np.asarray(ax['li2sn'], np.int64)
    nos = np.asarray(ax['li2nos'], np.float64)
    s11, nl = np.zeros(4084, np.int64), np.zeros(4084)
    for k in range(2):
        s11[li2sn1[:, k]] = li2sn[:, k]
        nl[li2sn1[:, k]] = nos
    return p1, s11, nl


def test_fd_image_matches_adjoint(params, case):
    from oracle import oracle as O
    c = case
    a = O.fprj(params, c['mu'], c['rays'], att=True)
    p = O.fprj(params, c['x'], c['rays'])
    ybar = a * c['n'] * p + c['rs']
    g = O.bprj(params, a * c['n'] * (1 - c['y'] / ybar), c['rays']).ravel()[c['vox']]
    fd = O.nll_fd(params, c['x'], c['mu'], c['rays'], c['y'], c['n'], c['rs'], c['vox'], 0, 1e-3)
    assert np.abs(g).max() > 1.0
    assert np.abs(fd - g).max() <= 1e-6 * np.abs(g).max()


def test_fd_mumap_matches_adjoint(params, case):
    from oracle import oracle as O
    c = case
    a = O.fprj(params, c['mu'], c['rays'], att=True)
    p = O.fprj(params, c['x'], c['rays'])
    ybar = a * c['n'] * p + c['rs']
    G = (1 - c['y'] / ybar) * c['n'] * p                     # dL/da per span-11 bin
    p1, s11, nl = _span1(params)
    t = np.exp(-O.fprj(p1, c['mu'], c['rays']))             # span-1 transmissions
    # span-1 mean of exp reproduces the span-11 attenuation factors
    a1 = np.zeros_like(a)
    for i in range(len(c['rays'])):
        np.add.at(a1[i], s11, t[i] / nl)
    assert np.allclose(a1, a, rtol=1e-12, atol=0)
    g = -O.bprj(p1, t * G[:, s11] / nl, c['rays']).ravel()[c['vox']]
    fd = O.nll_fd(params, c['x'], c['mu'], c['rays'], c['y'], c['n'], c['rs'], c['vox'], 1, 1e-4)
    assert np.abs(g).max() > 1.0
    assert np.abs(fd - g).max() <= 1e-6 * np.abs(g).max()


def test_fd_detects_dropped_norm(params, case):
    """Planted mistake on the oracle side: leaving the norm factor out of the
    analytic gradient is caught by the finite differences."""
    from oracle import oracle as O
    c = case
    a = O.fprj(params, c['mu'], c['rays'], att=True)
    p = O.fprj(params, c['x'], c['rays'])
    ybar = a * c['n'] * p + c['rs']
    wrong = O.bprj(params, a * (1 - c['y'] / ybar), c['rays']).ravel()[c['vox']]
    fd = O.nll_fd(params, c['x'], c['mu'], c['rays'], c['y'], c['n'], c['rs'], c['vox'], 0, 1e-3)
    assert np.abs(fd - wrong).max() > 1e-3 * np.abs(fd).max()
