"""nipet_amd: MI355X (gfx950) implementation of NIPET's projection,
reconstruction and list-mode histogramming path for the Siemens mMR.

The public names mirror ``niftypet.nipet`` (``niftypet/nipet/__init__.py:43-54``):
``get_mmrparams``, ``frwd_prj``, ``back_prj``, ``osemone``, ``mmrchain``,
``mmrhist``, ``dynamic_timings``, ``randoms``, ``get_norm_sino``, ``sino2ssr``,
``im_e72dev`` / ``im_dev2e7``, ``get_subsets14``, ``psf_config``,
``convert2e7``, ``convert2dev``, ``get_cylinder``, ``putgaps``, ``remgaps``. The compute runs in hand-written
HIP kernels (``nipet_amd/libnipet_hip.so``); there is no CPU fallback.
"""
from .img.mmrimg import convert2dev, convert2e7, get_cylinder
from .img.mmrimg import convert2dev as im_e72dev
from .img.mmrimg import convert2e7 as im_dev2e7
from .mmraux import (axial_lut, get_mmr_constants, get_mmrparams, mmrinit, putgaps, remgaps,
                     sino2ssr, span1_span11, transaxial_lut)
from .prj.mmrprj import back_prj, frwd_prj, trnx_prj
from .prj.mmrrec import get_subsets14, osemone, psf_config

__version__ = '0.1.0'

__all__ = [
    'get_mmrparams', 'get_mmr_constants', 'mmrinit', 'transaxial_lut', 'axial_lut',
    'span1_span11', 'putgaps', 'remgaps', 'convert2e7', 'convert2dev', 'get_cylinder',
    'frwd_prj', 'back_prj', 'trnx_prj', 'get_subsets14', 'psf_config', 'osemone', 'mmrhist',
    'lminfo', 'dev_info', 'im_e72dev', 'im_dev2e7', 'sino2ssr', 'mmrchain', 'dynamic_timings',
    'randoms', 'get_norm_sino']


def dev_info(showprop=False):
    """HIP devices visible to this process as a list of (name, total memory
    [MiB], gfx architecture, compute units), like ``niftypet.nipet.dev_info``
    (``ninst.dinf.dev_info``); prints the properties when ``showprop``."""
    try:
        import torch
        n = torch.cuda.device_count() if torch.cuda.is_available() else 0
    except Exception:
        n = 0
    out = []
    for i in range(n):
        pr = torch.cuda.get_device_properties(i)
        out.append((pr.name, pr.total_memory >> 20, getattr(pr, 'gcnArchName', ''),
                    pr.multi_processor_count))
        if showprop:
            print(f'device {i}: {pr.name}, {pr.total_memory >> 20} MiB, '
                  f'{getattr(pr, "gcnArchName", "")}, {pr.multi_processor_count} CUs')
    return out


def __getattr__(name):
    # list-mode processing is imported lazily (it pulls in the histogram binding)
    if name in ('mmrhist', 'lminfo', 'dynamic_timings', 'randoms'):
        from .lm import mmrhist as _m
        return getattr(_m, name)
    if name == 'mmrchain':
        from .img.pipe import mmrchain
        return mmrchain
    if name == 'get_norm_sino':
        from .mmrnorm import get_norm_sino
        return get_norm_sino
    raise AttributeError(name)
