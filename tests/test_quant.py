"""Decay correction and quantification factor of osemone
(prj/mmrrec.py:308-331, 384-388), checked against what the mathematics fixes:
with DCYCRR the decay factor is exp(l tref) over the mean of exp(-l t) across
the frame (an integral, evaluated by quadrature here); without DCYCRR or
without frame times it reduces to qf * qf_loc or to 1."""
import numpy as np
import pytest
from scipy.integrate import quad

from nipet_amd.prj.mmrrec import RI_LUT, quant_factor


def _cnt(iso='F18', dcy=False):
    return {'ISOTOPE': iso, 'DCYCRR': dcy}


NCMP = {'qf': 3.0e4, 'qf_loc': 0.205}


def test_quant_without_decay_correction():
    h = {'t0': 100, 't1': 700, 'dur': 600}
    for iso in RI_LUT:
        q = quant_factor(_cnt(iso), h, NCMP)
        assert q == pytest.approx(NCMP['qf'] / RI_LUT[iso]['BF'] / 600 * NCMP['qf_loc'], rel=1e-15)


@pytest.mark.parametrize('iso,t0,dur,tref', [('F18', 0, 600, None), ('C11', 300, 1200, None),
                                             ('O15', 60, 120, 10), ('Ga68', 3600, 1800, 0)])
def test_decay_correction_is_frame_average(iso, t0, dur, tref):
    h = {'t0': t0, 't1': t0 + dur, 'dur': dur}
    lam = np.log(2) / RI_LUT[iso]['thalf']
    ref_t = t0 if tref is None else tref
    mean_decay = quad(lambda t: np.exp(-lam * t), 0, dur, epsabs=0, epsrel=1e-13)[0] / dur
    expect = np.exp(lam * ref_t) / mean_decay
    plain = quant_factor(_cnt(iso), h, NCMP)
    q = quant_factor(_cnt(iso, True), h, NCMP, decay_ref_time=tref)
    assert q / plain == pytest.approx(expect, rel=1e-10)
    assert q / plain >= 1.0


def test_quant_needs_components_and_frame_times():
    assert quant_factor(_cnt(), {'t0': 0, 'dur': 10}, None) == 1.0
    assert quant_factor(_cnt(), {'dur': 10}, NCMP) == 1.0
    assert quant_factor(_cnt(), {'t0': 0}, NCMP) == 1.0


def test_half_lives_in_seconds():
    # F18 ~110 min, C11 ~20 min, O15 ~2 min, Ga68 ~68 min, Ge68 ~271 days
    hl = {k: v['thalf'] for k, v in RI_LUT.items()}
    assert 6500 < hl['F18'] < 6700 and 1200 < hl['C11'] < 1250 and 120 < hl['O15'] < 125
    assert 4000 < hl['Ga68'] < 4100 and 2.3e7 < hl['Ge68'] < 2.4e7
    assert all(0.85 < v['BF'] <= 1.0 for v in RI_LUT.values())
