"""nipet_remgaps_noalloc (gap removal with its inverse bin map in a static
device array, no allocation) through the C ABI against nipet_remgaps, on two
calls in a row: the full mMR span-11 conversion, then a smaller synthetic one
(fewer Siemens bins and sinograms, uint16 input) that must not see what the
first call left in the shared array. Bit-exact (a gather of float values)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


def _both(torch, L, _lib, x, dtype, ali, naw, nsn, nbinang):
    st = _lib.stream_handle()
    ref = torch.full((naw, nsn), -7.0, dtype=torch.float32, device='cuda')
    got = torch.full((naw, nsn), -9.0, dtype=torch.float32, device='cuda')
    _lib.check(L.nipet_remgaps(_lib.ptr(ref), _lib.ptr(x), dtype, _lib.ptr(ali), naw, nsn,
                               nbinang, st), 'nipet_remgaps')
    _lib.check(L.nipet_remgaps_noalloc(_lib.ptr(got), _lib.ptr(x), dtype, _lib.ptr(ali), naw,
                                       nsn, nbinang, st), 'nipet_remgaps_noalloc')
    torch.cuda.synchronize()
    return got, ref


def test_noalloc_remgaps_two_calls(torch, params):
    from nipet_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(12)
    # call 1: full mMR span-11 Siemens sinogram (837 x 252*344) float32
    ali = np.asarray(params['txLUT']['aw2ali'], np.int32)
    naw, nsn, nbinang = len(ali), 837, 252 * 344
    print(f'noalloc call 1: naw={naw} nsn={nsn} nbinang={nbinang} float32', flush=True)
    x = torch.from_numpy(rng.random((nsn, nbinang), dtype=np.float32)).cuda()
    ali_d = torch.from_numpy(ali).cuda()
    got, ref = _both(torch, L, _lib, x, 0, ali_d, naw, nsn, nbinang)
    assert torch.equal(got, ref)
    assert torch.equal(ref[:, 5], x[5, ali_d.long()])
    # call 2: smaller synthetic layout, uint16 input; the array still holds
    # call 1's map, of which only the first nbinang2 entries are rewritten
    nbinang2, naw2, nsn2 = 20011, 15003, 11
    ali2 = np.sort(rng.choice(nbinang2, naw2, replace=False)).astype(np.int32)[::-1].copy()
    print(f'noalloc call 2: naw={naw2} nsn={nsn2} nbinang={nbinang2} uint16', flush=True)
    x2 = torch.from_numpy(rng.integers(0, 65535, (nsn2, nbinang2)).astype(np.uint16)).cuda()
    ali2_d = torch.from_numpy(ali2).cuda()
    got2, ref2 = _both(torch, L, _lib, x2, 1, ali2_d, naw2, nsn2, nbinang2)
    assert torch.equal(got2, ref2)
    want = x2.to(torch.int32)[:, ali2_d.long()].t().to(torch.float32)
    assert torch.equal(ref2, want)

