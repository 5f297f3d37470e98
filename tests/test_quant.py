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


def test_prompts_u16_saturates_for_every_input_type():
    """osemone's prompts reach the device as uint16 by one rule whatever their
    container: counts saturate at 0 and 65535, fractions truncate (NumPy
    arrays and torch tensors, integer and float)."""
    import torch

    from nipet_amd.prj.osem import prompts_u16
    vals = [0, 1, 65535, 65536, 70000, 2**31 - 1, -3]
    exp = [0, 1, 65535, 65535, 65535, 65535, 0]
    for dt in (np.int64, np.int32, np.uint32):
        v = np.array([x for x in vals if dt != np.uint32 or x >= 0], dt)
        e = [y for x, y in zip(vals, exp) if dt != np.uint32 or x >= 0]
        assert list(prompts_u16(v)) == e
        assert prompts_u16(v).dtype == np.uint16
        t = prompts_u16(torch.from_numpy(v.astype(np.int64)))
        assert t.dtype == torch.uint16 and t.to(torch.int64).tolist() == e
    f = np.array([12.7, 65535.9, 1e9, -0.5], np.float32)
    assert list(prompts_u16(f)) == [12, 65535, 65535, 0]
    assert prompts_u16(torch.from_numpy(f)).to(torch.int64).tolist() == [12, 65535, 65535, 0]
    assert list(prompts_u16(np.array([5, 6], np.int16))) == [5, 6]
    assert prompts_u16(torch.tensor([5, -6], dtype=torch.int16)).to(torch.int64).tolist() == [5, 0]


def test_frame_definitions_and_skip_rule(params):
    """mmrchain's frame forms (img/pipe.py:95-125, dynamic_timings
    lm/mmrhist.py:492-556) and the skip of frames whose delayeds exceed 0.99
    of the prompts (img/pipe.py:317-324), which needs no reconstruction."""
    import nipet_amd as nip
    from nipet_amd.img.pipe import frame_timings
    assert frame_timings(['fluid', [0, 10], [10, 40]]) == [[0, 10], [10, 40]]
    assert frame_timings(['timings', [5, 6]]) == [[5, 6]]
    assert frame_timings(['def', [2, 15], [1, 30]]) == [[0, 15], [15, 30], [30, 60]]
    assert frame_timings([10, 20, 30]) == [[0, 10], [10, 30], [30, 60]]
    dt = nip.dynamic_timings([15, 30], offset=10)
    assert dt['timings'] == ['timings', [0, 10], [10, 25], [25, 55]] and dt['total'] == 55
    assert list(dt['frames']) == [15, 30]
    assert nip.dynamic_timings(['def', [2, 15]], offset=10)['timings'][1:] == [[0, 15], [15, 30]]
    with pytest.raises(TypeError):
        frame_timings(['def', [1, 2, 3]])
    hst = {'psino': np.zeros((837, 252, 344), np.uint16), 't0': 0, 't1': 10,
           'phc': np.full(10, 100), 'dhc': np.full(10, 100)}
    out = nip.mmrchain({}, params, histo=hst, itr=1)
    assert out['frames'] == [[0, 10]] and out['#frames'] == 1 and out['recmod'] == 0
    assert out['im'].shape == (127, 344, 344) and not out['im'].any()
    assert 'tuple' not in out
