"""Siemens Biograph mMR geometry: scanner constants and the transaxial/axial
look-up tables used by the projectors, the OSEM loop and the list-mode
histogrammer.

The tables keep the reference's names, dtypes and layouts so that code written
against NIPET's ``scanner_params`` dictionary works unchanged:

* ``get_mmr_constants`` -- the ``Cnt`` dictionary (reference: ``niftypet.ninst``
  ``resources.get_mmr_constants``, not vendored; values follow
  ``niftypet/nipet/include/def.h:52-126`` and ``sct/src/sct.h:18-41``).
* ``transaxial_lut`` -- ``niftypet/nipet/mmraux.py:620-814``.
* ``axial_lut``      -- ``niftypet/nipet/mmraux.py:339-544``.
* ``span1_span11``   -- ``niftypet/nipet/src/scanner_0.cu:44-87``.
* ``putgaps`` / ``remgaps`` -- ``niftypet/nipet/mmraux.py:1120-1160``
  (kernels ``src/scanner_0.cu:94-166``).
* ``get_mmrparams``  -- ``niftypet/nipet/mmraux.py:1163-1208`` (without the
  hardware mu-map discovery, which needs scanner data).
* ``varon``          -- ``mmr_auxe.varon`` (``niftypet/nipet/src/aux_module.cu:491-556``):
  online mean / variance update over realisations (kernel ``nipet_varon``).
"""
import logging
from math import pi

import numpy as np

log = logging.getLogger(__name__)

__all__ = [
    'get_mmr_constants', 'transaxial_lut', 'axial_lut', 'span1_span11', 'mmrinit',
    'get_mmrparams', 'putgaps', 'remgaps', 'sino2ssr', 'varon']


def get_mmr_constants():
    """Scanner constants of the Siemens Biograph mMR (the ``Cnt`` dictionary).

    Sizes: 64 rings x 504 crystals (448 active: every 9th transaxial slot is a
    gap), 344 radial bins x 252 angles, 4084 span-1 / 837 span-11 sinograms,
    device image (y=320, x=320, z=127) with 0.208626 cm x 0.208626 cm x
    0.203125 cm voxels, transaxial FOV radius^2 = 890 cm^2
    (``include/def.h:52-126``, ``sct/src/sct.h:37``).
    """
    Cnt = {
        'ISOTOPE': 'F18', 'DCYCRR': False,
        # scanner
        'NRNG': 64, 'NCRS': 504, 'NCRSR': 448, 'NBCKT': 224, 'NSBINS': 344, 'NSANGLES': 252,
        'NSBINANG': 344 * 252, 'NSN1': 4084, 'NSN11': 837, 'NSN64': 4096, 'NSEG0': 127,
        'MRD': 60, 'SPN': 11, 'TGAP': 9, 'OFFGAP': 1, 'AXR': 0.40625, 'R_RING': 33.47,
        'R': 32.8, 'SEG': [127, 115, 115, 93, 93, 71, 71, 49, 49, 27, 27],
        # minimum and maximum ring difference per span-11 segment (negative segment first)
        'MNRD': [-5, -16, 6, -27, 17, -38, 28, -49, 39, -60, 50],
        'MXRD': [5, -6, 16, -17, 27, -28, 38, -39, 49, -50, 60],
        # Siemens (E7) image
        'SO_IMZ': 127, 'SO_IMY': 344, 'SO_IMX': 344,
        'SO_VXZ': 0.203125, 'SO_VXY': 0.208626, 'SO_VXX': 0.208626,
        # device image
        'SZ_IMZ': 127, 'SZ_IMY': 320, 'SZ_IMX': 320,
        'SZ_VOXZ': 0.203125, 'SZ_VOXY': 0.208626, 'SZ_VOXZi': 4.923077,
        'TFOV2': 890.0, 'RSZ_PSF_KRNL': 8,
        # projector tables of the reference tracer (kept for API completeness)
        'NTV': 1807, 'NTT': 10,
        # list-mode
        'BPE': 4, 'LMOFF': 0, 'BTP': 0, 'BTPRT': 1.0, 'ITIME': 1000, 'BTPTIME': 100,
        'VTIME': 2, 'MXNITAG': 5400,
        # ring range (full scanner)
        'RNG_STRT': 0, 'RNG_END': 64,
        'DEVID': 0, 'LOG': logging.WARNING,
    }
    return Cnt


def transaxial_lut(Cnt, visualisation=False):
    """Transaxial crystal geometry and sinogram <-> crystal LUTs.

    Follows ``niftypet/nipet/mmraux.py:620-814``. Crystal faces: 56 blocks of
    8 crystals (plus a gap slot, left as a zero row) laid along the block face
    tangent to the ring of radius ``R_RING``. The arithmetic is done in the same
    order as the reference so that the float32 ``crs`` table is the same.

    Returns a dict with ``crs`` (504,4) f32 [x0,y0,x1,y1] in cm, ``crsri``,
    ``msino`` (344,252) i8, ``s2cF`` (86688,2) i16, ``c2sF``, ``c2sFw``,
    ``Naw``, ``s2c`` (Naw,2) i16, ``s2cr``, ``cr2s``, ``aw2sn`` (Naw,2) i16
    [angle, bin], ``aw2ali`` (Naw,) i32 = bin + 344*angle, ``cij``.
    """
    NCRS, NSB, NSA = Cnt['NCRS'], Cnt['NSBINS'], Cnt['NSANGLES']
    TGAP, OFFGAP = Cnt['TGAP'], Cnt['OFFGAP']

    bw = 3.209      # block width [cm]
    ntblk = 56      # transaxial blocks
    alpha = 0.1122  # angular block pitch [rad]
    crs = np.zeros((NCRS, 4), dtype=np.float32)
    phi = 0.5*pi - alpha/2 - 0.001
    for bi in range(ntblk):
        y = Cnt['R_RING'] * np.sin(phi)
        x = Cnt['R_RING'] * np.cos(phi)
        face = np.array([-y, x])
        face /= np.sum(face**2)**.5
        phi -= alpha
        xe = x + (bw/2) * face[0]
        ye = y + (bw/2) * face[1]
        for n in range(1, 9):
            c = bi*9 + n - 1
            xn = x + (bw/2 - n*bw/8) * face[0]
            yn = y + (bw/2 - n*bw/8) * face[1]
            crs[c] = (xe, ye, xn, yn)
            xe, ye = xn, yn
    out = {'crs': crs}

    # crystal index with the gap slots removed (-1 for a gap)
    live = ((np.arange(NCRS) + OFFGAP) % TGAP) > 0
    crsri = -np.ones(NCRS, dtype=np.int16)
    crsri[live] = np.arange(int(live.sum()), dtype=np.int16)
    out['crsri'] = crsri

    # sinogram (bin iw, angle ia) -> crystal pair, iw-major
    iw = np.repeat(np.arange(NSB), NSA)
    ia = np.tile(np.arange(NSA), NSB)
    c0 = np.floor((ia + 0.5 * (NCRS - 2 + NSB/2 - iw)) % NCRS).astype(np.int64)
    c1 = np.floor((ia + 0.5 * (2*NCRS - 2 - NSB/2 + iw)) % NCRS).astype(np.int64)
    s2cF = np.stack([c0, c1], axis=1).astype(np.int16)
    lin = (ia + iw*NSA).astype(np.int32)
    linw = (iw + ia*NSB).astype(np.int32)
    c2sF = -np.ones((NCRS, NCRS), dtype=np.int32)
    c2sF[c1, c0] = lin
    c2sF[c0, c1] = lin
    c2sFw = -np.ones((NCRS, NCRS), dtype=np.int32)
    c2sFw[c1, c0] = linw
    c2sFw[c0, c1] = linw
    ok = (((c0 + OFFGAP) % TGAP) * ((c1 + OFFGAP) % TGAP)) > 0
    msino = ok.reshape(NSB, NSA).astype(np.int8)
    out.update(s2cF=s2cF, c2sF=c2sF, c2sFw=c2sFw, msino=msino)
    Naw = int(ok.sum())
    out['Naw'] = Naw

    # active bins only, same iw-major / ia-minor order
    s2c = s2cF[ok].copy()
    s2cr = crsri[s2c.astype(np.int64)].astype(np.int16)
    aw = np.arange(Naw, dtype=np.int32)
    cr2s = np.zeros((Cnt['NCRSR'], Cnt['NCRSR']), dtype=np.int32)
    cr2s[s2cr[:, 1], s2cr[:, 0]] = aw
    cr2s[s2cr[:, 0], s2cr[:, 1]] = aw
    aw2sn = np.stack([ia[ok], iw[ok]], axis=1).astype(np.int16)
    aw2ali = (iw[ok] + NSB * ia[ok]).astype(np.int32)
    cij = np.zeros((Cnt['NCRSR'], Cnt['NCRSR']), dtype=np.int8)
    cij[s2cr[:, 0], s2cr[:, 1]] = 1
    cij[s2cr[:, 1], s2cr[:, 0]] = 1
    out.update(s2c=s2c, s2cr=s2cr, cr2s=cr2s, aw2sn=aw2sn, aw2ali=aw2ali, cij=cij)
    return out


def _ring_table(Cnt):
    """Axial ring extents [z_start, z_end) in cm, centred on the scanner."""
    NRNG = Cnt['NRNG']
    rng = np.zeros((NRNG, 2), dtype=np.float32)
    z = -.5 * NRNG * Cnt['AXR']
    for i in range(NRNG):
        rng[i, 0] = z
        z += Cnt['AXR']
        rng[i, 1] = z
    return rng


def axial_lut(Cnt):
    """Michelogram LUTs for span-1/span-11 (``niftypet/nipet/mmraux.py:339-544``).

    Ring pairs (r0, r1) with |r1-r0| <= MRD. Span-11 sinogram numbers are given
    segment by segment (segments ordered as ``MNRD``/``MXRD``), and within a
    segment by the axial position r0+r1. The linear Michelogram index ``li``
    (0..2073) walks the diagonals ro = r1-r0 = 0..60 with r0 ascending;
    li < 64 are the direct planes.
    """
    NRNG = Cnt['NRNG']
    if Cnt['SPN'] == 1:
        NRNG_c = Cnt['RNG_END'] - Cnt['RNG_STRT']
        NSN1_c = NRNG_c**2
        if NRNG_c == 64:
            NSN1_c -= 12
    else:
        NRNG_c = NRNG
        NSN1_c = Cnt['NSN1']
        if Cnt['RNG_END'] != NRNG or Cnt['RNG_STRT'] != 0:
            raise ValueError('the reduced axial FOV only works in span-1')
    rs, re = Cnt['RNG_STRT'], Cnt['RNG_END']
    rng = _ring_table(Cnt)

    # segment of each ring difference; ssrb plane of each ring pair
    rdiff = np.arange(-Cnt['MRD'], Cnt['MRD'] + 1)
    seg_of_rd = {}
    for rd in rdiff:
        for iseg, (lo, hi) in enumerate(zip(Cnt['MNRD'], Cnt['MXRD'])):
            if lo <= rd <= hi:
                seg_of_rd[int(rd)] = iseg
    # matrices indexed [r1, r0]
    Mssrb = -np.ones((NRNG, NRNG), dtype=np.int32)
    Mseg = -np.ones((NRNG, NRNG), dtype=np.int32)
    for r1 in range(rs, re):
        for r0 in range(rs, re):
            if abs(r1 - r0) > Cnt['MRD']:
                continue
            Mssrb[r1, r0] = r0 + r1
            Mseg[r1, r0] = seg_of_rd[r1 - r0]

    Msn = -np.ones((NRNG, NRNG), dtype=np.int32)
    Mnos = -np.ones((NRNG, NRNG), dtype=np.int32)
    isn = 0
    for iseg in range(len(Cnt['SEG'])):
        inseg = Mseg == iseg
        for pos in np.unique(Mssrb[inseg]):
            sel = inseg & (Mssrb == pos)
            Msn[sel] = isn
            Mnos[sel] = int(sel.sum())
            isn += 1

    # span-1 sinogram order: ring difference 0, then +1/-1 diagonals, ...
    sn1_rno = np.zeros((NSN1_c, 2), dtype=np.int16)
    sn1_ssrb = np.zeros(NSN1_c, dtype=np.int16)
    sn1_sn11 = np.zeros(NSN1_c, dtype=np.int16)
    sn1_sn11no = np.zeros(NSN1_c, dtype=np.int8)
    Msn1 = -np.ones((NRNG, NRNG), dtype=np.int16)
    sni = 0
    for ro in range(NRNG):
        for m in range(1 if ro == 0 else 2):
            for k in range(rs, re - ro):
                # m == 0: (r0, r1) = (k, k+ro); m == 1: (r0, r1) = (k+ro, k)
                if m == 0:
                    r0, r1 = k, k + ro
                else:
                    r0, r1 = k + ro, k
                if Msn[r1, r0] < 0:
                    continue
                sn1_rno[sni] = (r0, r1)
                sn1_ssrb[sni] = Mssrb[r1, r0]
                sn1_sn11[sni] = Msn[r0, r1]
                sn1_sn11no[sni] = Mnos[r0, r1]
                Msn1[r0, r1] = sni
                sni += 1

    NSN11 = Cnt['NSN11']
    sn11_ssrb = -np.ones(NSN11, dtype=np.int32)
    sn1_ssrno = np.zeros(Cnt['NSEG0'], dtype=np.int8)
    for i in range(NSN1_c):
        sn11_ssrb[sn1_sn11[i]] = sn1_ssrb[i]
        sn1_ssrno[sn1_ssrb[i]] += 1
    sn11_ssrno = np.zeros(Cnt['NSEG0'], dtype=np.int8)
    for i in range(NSN11):
        if sn11_ssrb[i] > 0:
            sn11_ssrno[sn11_ssrb[i]] += 1
    used = np.unique(sn1_ssrb)
    sn1_ssrno = sn1_ssrno[used]
    sn11_ssrno = sn11_ssrno[used]
    sn11_ssrb = sn11_ssrb[sn11_ssrb >= 0]

    # Michelogram diagonals (r1 >= r0) -> rings, z, sinogram numbers
    NLI2R_c = int(NRNG_c**2 / 2. + NRNG_c/2.)
    if NRNG_c == NRNG:
        NLI2R_c -= 6
    li2r = np.zeros((NLI2R_c, 2), dtype=np.int8)
    li2sn = np.zeros((NLI2R_c, 2), dtype=np.int16)
    li2sn1 = np.zeros((NLI2R_c, 2), dtype=np.int16)
    li2rng = np.zeros((NLI2R_c, 2), dtype=np.float32)
    li2nos = np.zeros(NLI2R_c, dtype=np.int8)
    dli = 0
    for ro in range(NRNG_c):
        for k in range(rs, re - ro):
            r0, r1 = k, k + ro
            if Msn[r1, r0] < 0:
                continue
            li2r[dli] = (r0, r1)
            li2rng[dli] = (rng[r0, 0], rng[r1, 0])
            li2sn[dli] = (Msn[r0, r1], Msn[r1, r0])
            li2sn1[dli] = (Msn1[r0, r1], Msn1[r1, r0])
            li2nos[dli] = Mnos[r1, r0]
            dli += 1

    return {
        'li2rno': li2r, 'li2sn': li2sn, 'li2sn1': li2sn1, 'li2nos': li2nos, 'li2rng': li2rng,
        'sn1_rno': sn1_rno, 'sn1_ssrb': sn1_ssrb, 'sn1_sn11': sn1_sn11, 'sn1_sn11no': sn1_sn11no,
        'sn11_ssrb': sn11_ssrb, 'sn1_ssrno': sn1_ssrno, 'sn11_ssrno': sn11_ssrno, 'Msn11': Msn,
        'Msn1': Msn1, 'Mnos': Mnos, 'rng': rng}


def span1_span11(Cnt):
    """Span-1 -> span-11 sinogram index and sinograms per span-11 plane.

    An independent (closed-form) definition of the same map as
    ``axial_lut()['sn1_sn11']``, following ``src/scanner_0.cu:44-87``: walk
    the span-1 sinograms segment by segment (ring difference rd), split each
    diagonal pair into its -rd and +rd halves, and place the axial position
    inside span-11 segment ``(rd+5)//11``.
    Returns (li2s11 (NSN1,) i16, nsinos (NSN11,) i8).
    """
    seg = Cnt['SEG']
    cum = np.concatenate([[0], np.cumsum(seg[:-1])])
    NRNG = Cnt['NRNG']
    li2s11 = np.zeros(Cnt['NSN1'], dtype=np.int16)
    nos = np.zeros(Cnt['NSN11'], dtype=np.int8)
    end = NRNG      # one past the last span-1 index of ring difference rd
    rd = 0
    for si in range(Cnt['NSN1']):
        while end <= si:
            rd += 1
            end += 2 * (NRNG - rd)
        split = end - (NRNG - rd)   # first index of the +rd half
        if si >= split:
            pos = 2 * (si - split) + rd
            minus = 0
        else:
            pos = 2 * (si - end + 2 * (NRNG - rd)) + rd
            minus = 1
        iseg = (rd + 5) // 11
        off = (127 - seg[2 * iseg]) // 2
        ci = 2 * iseg - minus * (iseg > 0)
        s11 = cum[ci] + pos - off
        li2s11[si] = s11
        nos[s11] += 1
    return li2s11, nos


def mmrinit():
    Cnt = get_mmr_constants()
    txLUT = transaxial_lut(Cnt)
    Cnt['Naw'] = txLUT['Naw']
    axLUT = axial_lut(Cnt)
    return Cnt, txLUT, axLUT


def get_mmrparams(hmu_dir=None):
    """All scanner parameters: ``{'Cnt', 'txLUT', 'axLUT'}``
    (``niftypet/nipet/mmraux.py:1177-1208``; hardware mu-maps are not part of
    this path)."""
    Cnt, txLUT, axLUT = mmrinit()
    if hmu_dir is not None:
        Cnt['HMUDIR'] = hmu_dir
    return {'Cnt': Cnt, 'txLUT': txLUT, 'axLUT': axLUT}


def _nsinos(Cnt):
    if Cnt['SPN'] == 1:
        nrng = Cnt['RNG_END'] - Cnt['RNG_STRT']
        return nrng**2 - (12 if nrng == 64 else 0)
    if Cnt['SPN'] == 11:
        return Cnt['NSN11']
    if Cnt['SPN'] == 0:
        return Cnt['NSEG0']
    raise ValueError(f"unrecognised span: {Cnt['SPN']}")


def putgaps(s, txLUT, Cnt, sino_no=0):
    """Device sinogram (Naw, nsinos) -> Siemens sinogram (nsinos, 252, 344)
    with zero gaps (``niftypet/nipet/mmraux.py:1120-1143``).

    Runs on the GPU when ``s`` is a CUDA tensor (kernel ``nipet_putgaps``), on
    the host otherwise (pure data movement: a scatter through ``aw2ali``)."""
    nsinos = sino_no if sino_no > 0 else _nsinos(Cnt)
    if _is_cuda(s):
        from .prj import mmrprj
        return mmrprj._gaps_dev(s, txLUT, Cnt, nsinos, put=True)
    s = np.asarray(s)
    if s.shape != (txLUT['Naw'], nsinos):
        raise ValueError(f'expected device sinogram of shape {(txLUT["Naw"], nsinos)}, '
                         f'got {s.shape}')
    sino = np.zeros((Cnt['NSANGLES'] * Cnt['NSBINS'], nsinos), dtype=np.float32)
    sino[txLUT['aw2ali']] = s
    sino = sino.reshape(Cnt['NSANGLES'], Cnt['NSBINS'], nsinos).transpose(2, 0, 1)
    return np.ascontiguousarray(sino).astype(s.dtype)


def remgaps(sino, txLUT, Cnt, output=None):
    """Siemens sinogram (nsinos, 252, 344) -> device sinogram (Naw, nsinos)
    (``niftypet/nipet/mmraux.py:1146-1160``). The result keeps the input dtype."""
    if _is_cuda(sino):
        from .prj import mmrprj
        return mmrprj._gaps_dev(sino, txLUT, Cnt, sino.shape[0], put=False, output=output)
    sino = np.asarray(sino)
    if sino.ndim != 3 or sino.shape[1:] != (Cnt['NSANGLES'], Cnt['NSBINS']):
        raise ValueError(f'expected a Siemens sinogram (n, 252, 344), got {sino.shape}')
    flat = sino.reshape(sino.shape[0], -1)
    out = np.ascontiguousarray(flat[:, txLUT['aw2ali']].T)
    if output is not None:
        output[...] = out
        return output
    return out


def sino2ssr(sino, axLUT, Cnt):
    """Single-slice rebinning of a Siemens-shaped sinogram
    (``niftypet/nipet/mmraux.py:547-563``)."""
    if Cnt['SPN'] == 1:
        slut, snno = axLUT['sn1_ssrb'], Cnt['NSN1']
    elif Cnt['SPN'] == 11:
        slut, snno = axLUT['sn11_ssrb'], Cnt['NSN11']
    else:
        raise ValueError(f"unrecognised span: {Cnt['SPN']}")
    ssr = np.zeros((Cnt['NSEG0'], Cnt['NSANGLES'], Cnt['NSBINS']), dtype=np.float32)
    for i in range(snno):
        ssr[slut[i]] += sino[i]
    return ssr


def _is_cuda(x):
    return getattr(x, 'is_cuda', False)


def varon(M1, M2, X, b, Cnt=None):
    """Update the running mean ``M1`` and sum of squared deviations ``M2`` in
    place with realisation ``X`` (0-based index ``b``), Welford's method as
    ``src/auxmath.cu:15-22``; the variance of b+1 realisations is
    M2 / (b + 1). float32 NumPy arrays or CUDA tensors (then asynchronous)."""
    from . import _lib
    import torch
    if b < 0:
        raise ValueError('varon: realisation index b must be >= 0')
    if all(isinstance(a, torch.Tensor) and a.is_cuda for a in (M1, M2, X)):
        for a in (M1, M2):
            if a.dtype != torch.float32 or not a.is_contiguous():
                raise ValueError('varon: M1 and M2 must be contiguous float32 tensors')
        x = X.to(torch.float32).contiguous()
        if not (M1.numel() == M2.numel() == x.numel()):
            raise ValueError('varon: sizes differ')
        _lib.check(_lib.lib().nipet_varon(M1.data_ptr(), M2.data_ptr(), x.data_ptr(), int(b),
                                          x.numel(), _lib.stream_handle()), 'nipet_varon')
        return M1, M2
    for a in (M1, M2):
        if not (isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags.c_contiguous):
            raise ValueError('varon: M1 and M2 must be contiguous float32 arrays')
    x = np.ascontiguousarray(np.asarray(X, np.float32))
    if not (M1.size == M2.size == x.size):
        raise ValueError('varon: sizes differ')
    d = [torch.from_numpy(a.reshape(-1)).cuda() for a in (M1, M2, x)]
    _lib.check(_lib.lib().nipet_varon(d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), int(b),
                                      x.size, _lib.stream_handle()), 'nipet_varon')
    M1.reshape(-1)[:] = d[0].cpu().numpy()
    M2.reshape(-1)[:] = d[1].cpu().numpy()
    return M1, M2
