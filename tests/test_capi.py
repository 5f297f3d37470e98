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
