"""ctypes binding of ``libnipet_hip.so`` (C ABI declared in ``include/nipet_hip.h``).

The library is loaded from this package directory only. There is no CPU
fallback: if the library is missing or the GPU is unavailable every entry
point raises.
"""
import ctypes as ct
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, 'libnipet_hip.so')
_lock = threading.Lock()
_lib = None

PROF_FPRJ, PROF_BPRJ, PROF_OSEM_UPDATE, PROF_FILL, PROF_HIST, PROF_OTHER, PROF_BPRJ_PAIR = range(7)


class NipetError(RuntimeError):
    pass


class NipetLmConfig(ct.Structure):
    _fields_ = [
        ('toff', ct.c_int), ('nitag', ct.c_int), ('tstart', ct.c_int), ('tstop', ct.c_int),
        ('span', ct.c_int), ('btp', ct.c_int), ('btprt', ct.c_double), ('seed', ct.c_uint64),
        ('nrep', ct.c_int), ('rep0', ct.c_int), ('sn1_sn11', ct.c_void_p), ('sn1_ssrb', ct.c_void_p),
        ('sn1_rno', ct.c_void_p), ('s2cF', ct.c_void_p)]


class NipetLmOut(ct.Structure):
    _fields_ = [(k, ct.c_void_p) for k in
                ('psino', 'dsino', 'ssr', 'phc', 'dhc', 'zR', 'zM', 'pvs', 'fan', 'bck')]


class NipetRndDesc(ct.Structure):
    _fields_ = [('naw', ct.c_int), ('nsn', ct.c_int), ('nsn1', ct.c_int), ('mrd', ct.c_int),
                ('cij', ct.c_void_p), ('crsri', ct.c_void_p), ('s2cr', ct.c_void_p),
                ('sn1_rno', ct.c_void_p), ('sn1_sn11', ct.c_void_p)]


class NipetNormDesc(ct.Structure):
    _fields_ = [('naw', ct.c_int), ('span', ct.c_int), ('aw2ali', ct.c_void_p),
                ('sn1_rno', ct.c_void_p), ('sn1_sn11', ct.c_void_p), ('sn1_sn11no', ct.c_void_p)]


class NipetNormComp(ct.Structure):
    _fields_ = [(k, ct.c_void_p) for k in ('geo', 'cinf', 'ceff', 'axe1', 'axf1', 'dtp', 'dtnp')]


class NipetGeomDesc(ct.Structure):
    _fields_ = [
        ('crs', ct.c_void_p), ('ncrs', ct.c_int), ('s2c', ct.c_void_p), ('naw', ct.c_int),
        ('li2rng', ct.c_void_p), ('li2sn', ct.c_void_p), ('li2nos', ct.c_void_p),
        ('nli', ct.c_int), ('nsn', ct.c_int), ('ring_pitch', ct.c_float),
        ('vox_xy', ct.c_float), ('vox_z', ct.c_float), ('tfov2', ct.c_float)]


def lib():
    """Load the shared library (raises NipetError when it is not built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIBPATH):
            raise NipetError(f'{LIBPATH} is not built: run __graft_entry__.build() or '
                             '`python -m nipet_amd._build`')
        L = ct.CDLL(LIBPATH)
        P, I, I64 = ct.c_void_p, ct.c_int, ct.c_int64
        sig = {
            'nipet_geom_create': ([ct.POINTER(NipetGeomDesc), ct.POINTER(P)], I),
            'nipet_geom_destroy': ([P], None),
            'nipet_geom_info': ([P, P, P, P, P], I),
            'nipet_tx_lists': ([P, P, P, P, P], I),
            'nipet_fprj': ([P, P, P, P, I, I, P], I),
            'nipet_bprj': ([P, P, P, P, I, P], I),
            'nipet_osem_corr': ([P, P, P, P, P, P, I64, P], I),
            'nipet_osem_sens': ([P, P, P, P, I, P], I),
            'nipet_remgaps': ([P, P, I, P, I, I, I, P], I),
            'nipet_remgaps_noalloc': ([P, P, I, P, I, I, I, P], I),
            'nipet_putgaps': ([P, P, P, I, I, I, P], I),
            'nipet_remgaps_u16b': ([P, P, P, I, I, I, I, I, I, I, I, P], I),
            'nipet_tune_set': ([I, I], I),
            'nipet_psf_conv': ([P, P, P, P, I, P], I),
            'nipet_osem_subset_rm_b': ([P, P, P, P, I, P, I, P, P, P, I, I, P, P, P, P, P], I),
            'nipet_tune_get': ([I, P], I),
            'nipet_tune_set_run': ([I, I], I),
            'nipet_tune_get_run': ([I, P, P], I),
            'nipet_osem_subset': ([P, P, P, P, P, P, P, P, I, P, P, P], I),
            'nipet_osem_window': ([P, P, I, P, P], I),
            'nipet_bins_longest_first': ([P, P, I, P, P], I),
            'nipet_osem_subset_sens': ([P, P, P, P, P, P, P, P, P, I, P, P, P, P], I),
            'nipet_osem_iter': ([P, P, P, P, P, P, P, P, I, I, P, P, P, P], I),
            'nipet_fprj_b': ([P, P, P, P, I, I, I, P], I),
            'nipet_bprj_b': ([P, P, P, P, I, I, P], I),
            'nipet_osem_subset_b': ([P, P, P, P, I, P, I, P, P, P, I, I, P, P, P], I),
            'nipet_lm_create': ([ct.POINTER(NipetLmConfig), ct.POINTER(P)], I),
            'nipet_lm_chunk': ([P, P, I64, ct.POINTER(NipetLmOut), P], I),
            'nipet_lm_status': ([P, P, P, P], I),
            'nipet_lm_finish': ([P, ct.POINTER(NipetLmOut), P], I),
            'nipet_lm_destroy': ([P], None),
            'nipet_lm_tag_span': ([P, I64, P, P], I),
            'nipet_lm_second_index': ([P, I64, I, I, P, P], I),
            'nipet_rnd_create': ([ct.POINTER(NipetRndDesc), ct.POINTER(P)], I),
            'nipet_rnd_estimate': ([P, P, I, P, P, P], I),
            'nipet_rnd_destroy': ([P], None),
            'nipet_norm_create': ([ct.POINTER(NipetNormDesc), ct.POINTER(P)], I),
            'nipet_norm_sino': ([P, ct.POINTER(NipetNormComp), P, P, P], I),
            'nipet_norm_destroy': ([P], None),
            'nipet_varon': ([P, P, P, I, I64, P], I),
            'nipet_philox4x32': ([P, P, P], I),
            'nipet_prof_enable': ([I], I),
            'nipet_prof_reset': ([], I),
            'nipet_prof_read': ([I, P, P], I),
            'nipet_last_error': ([], ct.c_char_p),
            'nipet_version': ([], ct.c_char_p),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
        return L


def check(rc, what=''):
    if rc != 0:
        msg = lib().nipet_last_error().decode(errors='replace')
        raise NipetError(f'{what}: {msg} (code {rc})')


def ptr(t):
    """Device (or host) address of a torch tensor / numpy array, or None."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    return t.data_ptr()


def stream_handle(device=None):
    import torch
    return torch.cuda.current_stream(device).cuda_stream


class KernelTimer:
    """Context manager: per-kernel-class HIP-event timing inside the library.
    ``times[k] = (total_ms, launches)`` after exit."""

    def __enter__(self):
        L = lib()
        check(L.nipet_prof_reset(), 'nipet_prof_reset')
        check(L.nipet_prof_enable(1), 'nipet_prof_enable')
        return self

    def __exit__(self, *exc):
        L = lib()
        check(L.nipet_prof_enable(0), 'nipet_prof_enable')
        self.times = {}
        for k in range(PROF_BPRJ_PAIR + 1):
            ms, n = ct.c_double(), ct.c_int64()
            check(L.nipet_prof_read(k, ct.byref(ms), ct.byref(n)), 'nipet_prof_read')
            self.times[k] = (ms.value, n.value)
        return False

    def mean_ms(self, k):
        ms, n = self.times[k]
        return ms / n if n else float('nan')
