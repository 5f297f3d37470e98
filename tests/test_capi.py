"""The C-ABI library loads and exports every symbol declared in include/*.h;
argument errors come back as status codes (no GPU needed: nothing here
launches a kernel)."""
import ctypes as ct
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, 'include', '*.h')):
        src = open(h).read()
        src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
        names |= set(re.findall(r'\b(nipet_\w+)\s*\(', src))
    return sorted(names)


@pytest.fixture(scope='module')
def L():
    from nipet_amd import _lib
    if not os.path.exists(_lib.LIBPATH):
        pytest.fail('libnipet_hip.so is not built (run __graft_entry__.build())')
    return _lib.lib()


def test_exports_every_declared_symbol(L):
    names = _declared()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_status_codes_without_gpu(L):
    from nipet_amd import _lib
    rc = L.nipet_fprj(None, None, None, None, 10, 0, None)
    assert rc == -1
    assert b'NULL geometry' in L.nipet_last_error()
    h = ct.c_void_p()
    assert L.nipet_geom_create(None, ct.byref(h)) == -1
    assert L.nipet_osem_corr(None, None, None, None, None, None, 10, None) == -1
    assert b'nipet_osem_corr' in L.nipet_last_error()
    assert L.nipet_osem_subset_sens(None, None, None, None, None, None, None, None, None, 10,
                                    None, None, None, None) == -1
    assert b'nipet_osem_subset_sens: NULL geometry' in L.nipet_last_error()
    assert L.nipet_osem_window(None, None, 0, None, None) == -1
    assert b'nipet_osem_window: NULL geometry' in L.nipet_last_error()
    assert L.nipet_bins_longest_first(None, None, 5, None, None) == -1
    assert b'nipet_bins_longest_first: bad argument' in L.nipet_last_error()
    assert L.nipet_osem_iter(None, None, None, None, None, None, None, None, 1, 2, None, None,
                             None, None) == -1
    assert b'nipet_osem_iter: NULL geometry' in L.nipet_last_error()
    # the allocation-free gap removal refuses layouts beyond its static array
    # before touching the device
    buf = ct.c_float()
    assert L.nipet_remgaps_noalloc(ct.byref(buf), ct.byref(buf), 0, ct.byref(buf), 4, 4,
                                   252 * 344 + 1, None) == -1
    assert b'exceeds the static scratch' in L.nipet_last_error()
    assert L.nipet_version().startswith(b'nipet_amd')
    with pytest.raises(_lib.NipetError):
        _lib.check(-1, 'x')


def test_dev_info_quiet(capsys):
    """dev_info lists devices without printing unless asked (the reference's
    tests/test_basic.py:6-12)."""
    import nipet_amd as nip
    devs = nip.dev_info()
    out, err = capsys.readouterr()
    assert not out and not err
    nip.dev_info(showprop=True)
    out, err = capsys.readouterr()
    assert not err
    assert not devs or out


def test_new_entry_points_reject_bad_arguments(L, params):
    """Randoms, normalisation, online variance and PSF entry points validate
    their descriptions before touching the device (status codes, messages)."""
    import numpy as np

    from nipet_amd import _lib
    tx, ax = params['txLUT'], params['axLUT']
    h = ct.c_void_p()
    assert L.nipet_rnd_create(None, ct.byref(h)) == -1
    keep = [np.ascontiguousarray(tx['cij'], np.int8), np.ascontiguousarray(tx['crsri'], np.int16),
            np.ascontiguousarray(tx['s2cr'], np.int16), np.ascontiguousarray(ax['sn1_rno'], np.int16),
            np.ascontiguousarray(ax['sn1_sn11'], np.int16)]
    d = _lib.NipetRndDesc(tx['Naw'], 500, 4084, 60, *[k.ctypes.data for k in keep])  # bad nsn
    assert L.nipet_rnd_create(ct.byref(d), ct.byref(h)) == -1
    assert b'span 1 or 11' in L.nipet_last_error()
    bad_cij = np.zeros((448, 448), np.int8)                                          # no runs
    d = _lib.NipetRndDesc(tx['Naw'], 837, 4084, 60, bad_cij.ctypes.data,
                          *[k.ctypes.data for k in keep[1:]])
    assert L.nipet_rnd_create(ct.byref(d), ct.byref(h)) == -1
    assert b'circular run' in L.nipet_last_error()
    assert L.nipet_norm_create(None, ct.byref(h)) == -1
    nk = [np.ascontiguousarray(tx['aw2ali'], np.int32), np.ascontiguousarray(ax['sn1_rno'], np.int16),
          np.ascontiguousarray(ax['sn1_sn11'], np.int16),
          np.ascontiguousarray(ax['sn1_sn11no'], np.int8)]
    nd = _lib.NipetNormDesc(tx['Naw'], 3, *[k.ctypes.data for k in nk])               # bad span
    assert L.nipet_norm_create(ct.byref(nd), ct.byref(h)) == -1
    bad_ali = nk[0].copy()
    bad_ali[5] = 252 * 344
    nd = _lib.NipetNormDesc(tx['Naw'], 11, bad_ali.ctypes.data, *[k.ctypes.data for k in nk[1:]])
    assert L.nipet_norm_create(ct.byref(nd), ct.byref(h)) == -1
    assert b'aw2ali entry 5' in L.nipet_last_error()
    assert L.nipet_norm_sino(None, None, None, None, None) == -1
    assert L.nipet_varon(None, None, None, 0, 10, None) == -1
    assert L.nipet_rnd_estimate(None, None, 10, None, None, None) == -1
    assert L.nipet_psf_conv(None, None, None, None, 1, None) == -1


def test_lm_create_rejects_frame_beyond_acquisition(L, params):
    """nipet_lm_create refuses tstop > nitag (the head curves and projection
    views hold nitag seconds) and unknown bootstrap types, before touching the
    device; LmHistogram checks the frame the same way."""
    import numpy as np

    from nipet_amd import _lib
    from nipet_amd.lm.mmrhist import LmHistogram
    tx, ax = params['txLUT'], params['axLUT']
    keep = [np.ascontiguousarray(ax['sn1_sn11'], np.int16),
            np.ascontiguousarray(ax['sn1_ssrb'], np.int16),
            np.ascontiguousarray(ax['sn1_rno'], np.int16),
            np.ascontiguousarray(tx['s2cF'], np.int16)]
    h = ct.c_void_p()
    for tstart, tstop, btp in ((0, 11, 0), (5, 4, 0), (0, 10, 3)):
        cfg = _lib.NipetLmConfig(0, 10, tstart, tstop, 11, btp, 1.0, 0, 1, 0,
                                 *[k.ctypes.data for k in keep])
        assert L.nipet_lm_create(ct.byref(cfg), ct.byref(h)) == -1
        assert b'bad configuration' in L.nipet_last_error()
    with pytest.raises(ValueError, match='outside the acquisition'):
        LmHistogram(params, 0, 10, 0, 11)
