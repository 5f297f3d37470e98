"""ctypes wrapper of the FP64 CPU oracle (``oracle/nipet_oracle.c``).

Test infrastructure only: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg, never by the ``nipet_amd`` package.
"""
import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

NZ = 127
NVOX = 320 * 320 * 127


def build():
    import subprocess
    subprocess.run(['make', '-s', '-C', _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, 'liboracle.so')
        if not os.path.exists(path):
            build()
        L = ct.CDLL(path)
        P = ct.c_void_p
        L.orc_tx_lists.argtypes = [P, P, P, ct.c_int, P, P, P, P, P, P]
        L.orc_fprj.argtypes = [P, P, P, ct.c_int, P, P, P, P, P, ct.c_int, ct.c_int, ct.c_int, P]
        L.orc_fprj_d.argtypes = L.orc_fprj.argtypes
        L.orc_fprj_d.restype = ct.c_int
        L.orc_bprj.argtypes = [P, P, P, ct.c_int, P, P, P, P, ct.c_int, ct.c_int, ct.c_int]
        for f in (L.orc_tx_lists, L.orc_fprj, L.orc_bprj):
            f.restype = ct.c_int
        L.orc_max_threads.restype = ct.c_int
        L.orc_nll_fd.argtypes = [P, P, P, ct.c_int, P, P, P, P, P, P, P, P, ct.c_int, ct.c_int,
                                 P, ct.c_int, ct.c_int, ct.c_double, P]
        L.orc_nll_fd.restype = ct.c_int
        L.orc_philox4x32.argtypes = [P, P, P]
        L.orc_poisson_cdf.argtypes = [ct.c_double, P, ct.c_int]
        L.orc_lm_hist.argtypes = [P, ct.c_int64, ct.c_int64, ct.c_int, ct.c_int, ct.c_int,
                                  ct.c_int, ct.c_int, P, P, P, P, ct.c_int, P, ct.c_int,
                                  ct.c_uint64, ct.c_int, ct.c_int64] + [P] * 10
        L.orc_lm_hist.restype = ct.c_int64
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


def _geom(params):
    tx, ax = params['txLUT'], params['axLUT']
    crs = np.ascontiguousarray(tx['crs'], dtype=np.float32)
    s2c = np.ascontiguousarray(tx['s2c'], dtype=np.int16)
    li2rng = np.ascontiguousarray(ax['li2rng'], dtype=np.float32)
    span = params['Cnt']['SPN']
    li2sn = np.ascontiguousarray(ax['li2sn'] if span == 11 else ax['li2sn1'], dtype=np.int16)
    li2nos = np.ascontiguousarray(ax['li2nos'] if span == 11 else np.ones(len(li2rng)),
                                  dtype=np.int8)
    nsn = params['Cnt']['NSN11'] if span == 11 else params['Cnt']['NSN1']
    return crs, s2c, li2rng, li2sn, li2nos, nsn


def tx_lists(params, rays):
    """Segment lists of the given transaxial bins.
    Returns (offsets int64, col int32, t double, len double, atn double, t12 double)."""
    crs, s2c = _geom(params)[:2]
    rays = np.ascontiguousarray(rays, dtype=np.int32)
    n = len(rays)
    off = np.zeros(n + 1, dtype=np.int64)
    atn = np.zeros(n)
    t12 = np.zeros((n, 2))
    rc = lib().orc_tx_lists(_p(crs), _p(s2c), _p(rays), n, _p(off), None, None, None, _p(atn),
                            _p(t12))
    assert rc == 0
    m = int(off[-1])
    col = np.zeros(m, dtype=np.int32)
    ts = np.zeros(m)
    ln = np.zeros(m)
    rc = lib().orc_tx_lists(_p(crs), _p(s2c), _p(rays), n, _p(off), _p(col), _p(ts), _p(ln),
                            None, None)
    assert rc == 0
    return off, col, ts, ln, atn, t12


def fprj(params, img, rays, att=False, return_terms=False):
    """FP64 forward projection of device image ``img`` (320,320,127; float32 or
    float64) over the
    transaxial bins ``rays``; returns (len(rays), nsinos) float64."""
    crs, s2c, li2rng, li2sn, li2nos, nsn = _geom(params)
    dbl = np.asarray(img).dtype == np.float64
    img = np.ascontiguousarray(img, dtype=np.float64 if dbl else np.float32)
    assert img.size == NVOX
    rays = np.ascontiguousarray(rays, dtype=np.int32)
    out = np.zeros((len(rays), nsn))
    nt = np.zeros(1, dtype=np.int64)
    f = lib().orc_fprj_d if dbl else lib().orc_fprj
    rc = f(_p(img), _p(out), _p(rays), len(rays), _p(crs), _p(s2c), _p(li2rng),
           _p(li2sn), _p(li2nos), len(li2rng), nsn, int(bool(att)), _p(nt))
    assert rc == 0
    return (out, int(nt[0])) if return_terms else out


def bprj(params, sino, rays, nthreads=0):
    """FP64 back projection of ``sino`` (len(rays), nsinos) over ``rays``;
    returns the device image (320,320,127) float64."""
    crs, s2c, li2rng, li2sn, li2nos, nsn = _geom(params)
    rays = np.ascontiguousarray(rays, dtype=np.int32)
    sino = np.ascontiguousarray(sino, dtype=np.float64)
    assert sino.shape == (len(rays), nsn)
    img = np.zeros((320, 320, NZ))
    rc = lib().orc_bprj(_p(img), _p(sino), _p(rays), len(rays), _p(crs), _p(s2c), _p(li2rng),
                        _p(li2sn), len(li2rng), nsn, int(nthreads))
    assert rc == 0
    return img


def psf_conv(img, krnl):
    """FP64 separable convolution of a device-layout image (320 y, 320 x,
    127 z) with krnl (3, 17) = taps for z, x, y, following d_conv
    (prj/src/recon.cu:49-215): out[i] = sum_{j=-8..8} k[8-j] in[i+j], zero
    outside the image, applied along z, then x, then y."""
    x = np.asarray(img, dtype=np.float64)
    k = np.asarray(krnl, dtype=np.float64)
    R = (k.shape[1] - 1) // 2
    for axis, taps in ((2, k[0]), (1, k[1]), (0, k[2])):
        n = x.shape[axis]
        pad = [(0, 0)] * 3
        pad[axis] = (R, R)
        xp = np.pad(x, pad)
        out = np.zeros_like(x)
        for j in range(-R, R + 1):
            sl = [slice(None)] * 3
            sl[axis] = slice(R + j, R + j + n)
            out += taps[R - j] * xp[tuple(sl)]
        x = out
    return x


def osem(params, prompts, rho, sens, mask, subsets, img0, niter, eps=1e-6, krnl=None):
    """FP64 OSEM following prj/src/recon.cu:459-497: per subset
    e = FP(x) + rho; e = y / e (0 where |e| < eps); b = BP(e);
    b = b / sens (0 where |sens| < eps); x = mask ? x*b : 0.
    With ``krnl`` (resolution modelling, prj/src/recon.cu:468-495) the
    projected image is H x and the back projection is convolved by H (the
    caller passes sensitivity images already convolved).
    prompts/rho: (Naw, nsn) indexed by transaxial bin; sens: (nsub, 320,320,127);
    subsets: list of int arrays of transaxial bins."""
    x = np.array(img0, dtype=np.float64)
    for _ in range(niter):
        for s, rows in enumerate(subsets):
            xf = x if krnl is None else psf_conv(x, krnl)
            e = fprj(params, xf, rows) + rho[rows]
            y = prompts[rows].astype(np.float64)
            ratio = np.where(np.abs(e) < eps, 0.0, y / np.where(np.abs(e) < eps, 1.0, e))
            b = bprj(params, ratio, rows)
            if krnl is not None:
                b = psf_conv(b, krnl)
            sv = sens[s]
            b = np.where(np.abs(sv) < eps, 0.0, b / np.where(np.abs(sv) < eps, 1.0, sv))
            x = np.where(mask, x * b, 0.0)
    return x


def nll_fd(params, x, mu, rays, y, nrm, rs, vox, wrt, h):
    """FP64 central finite differences of the Poisson negative log-likelihood
    sum_b ybar_b - y_b log ybar_b, ybar = a(mu) n p(x) + rs over the transaxial
    bins ``rays`` (y, nrm, rs: (len(rays), nsinos)), with respect to voxels
    ``vox`` (flat indices of the (320,320,127) image) of x (wrt = 0) or of mu
    (wrt = 1), step h."""
    crs, s2c, li2rng, li2sn, li2nos, nsn = _geom(params)
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    mu = np.ascontiguousarray(mu, dtype=np.float64).ravel()
    rays = np.ascontiguousarray(rays, dtype=np.int32)
    y, nrm, rs = (np.ascontiguousarray(a, dtype=np.float64) for a in (y, nrm, rs))
    for a in (y, nrm, rs):
        assert a.shape == (len(rays), nsn)
    vox = np.ascontiguousarray(vox, dtype=np.int64)
    assert vox.min() >= 0 and vox.max() < NVOX
    g = np.zeros(len(vox))
    rc = lib().orc_nll_fd(_p(x), _p(mu), _p(rays), len(rays), _p(y), _p(nrm), _p(rs), _p(crs),
                          _p(s2c), _p(li2rng), _p(li2sn), _p(li2nos), len(li2rng), nsn, _p(vox),
                          len(vox), int(wrt), float(h), _p(g))
    assert rc == 0
    return g


def philox4x32(ctr, key):
    """Philox4x32-10 of one 4-word counter and 2-word key."""
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32(_p(c), _p(k), _p(out))
    return out


def poisson_cdf(lam, K=40):
    cdf = np.zeros(K, np.uint32)
    lib().orc_poisson_cdf(float(lam), _p(cdf), K)
    return cdf


def lm_hist(words, params, toff, nitag, tstart=0, tstop=None, span=11, btp=0, btprt=1.0,
            seed=0, nrep=1, base=0):
    """Sequential histogram of an mMR list-mode word stream (see lm_oracle.c).
    Returns a dict of uint32 / uint64 arrays (prompt and delayed counts separate)."""
    Cnt, tx, ax = params['Cnt'], params['txLUT'], params['axLUT']
    words = np.ascontiguousarray(words, dtype=np.int32)
    tstop = nitag if tstop is None else tstop
    nbins = {11: Cnt['NSN11'], 1: Cnt['NSN1'], 0: Cnt['NSEG0']}[span] * 86688
    tn = (min(nitag, 5400) + 3) // 4
    out = dict(psino=np.zeros((nrep, nbins), np.uint32), dsino=np.zeros((nrep, nbins), np.uint32),
               ssr=np.zeros(127 * 86688, np.uint32), phc=np.zeros((nrep, nitag), np.uint32),
               dhc=np.zeros((nrep, nitag), np.uint32), zR=np.zeros(nitag, np.uint64),
               zM=np.zeros(nitag, np.uint64), pvs=np.zeros(tn * 127 * 344, np.uint32),
               fan=np.zeros(64 * 504, np.uint32), bck=np.zeros(2 * nitag * 224, np.uint32))
    cdf = poisson_cdf(btprt)
    sn1_sn11 = np.ascontiguousarray(ax['sn1_sn11'], np.int16)
    sn1_ssrb = np.ascontiguousarray(ax['sn1_ssrb'], np.int16)
    sn1_rno = np.ascontiguousarray(ax['sn1_rno'], np.int16)
    s2cF = np.ascontiguousarray(tx['s2cF'], np.int16)
    n = lib().orc_lm_hist(_p(words), len(words), base, toff, nitag, tstart, tstop, span,
                          _p(sn1_sn11), _p(sn1_ssrb), _p(sn1_rno), _p(s2cF), btp, _p(cdf),
                          len(cdf), seed, nrep, nbins,
                          *[_p(out[k]) for k in ('psino', 'dsino', 'ssr', 'phc', 'dhc', 'zR', 'zM',
                                                 'pvs', 'fan', 'bck')])
    assert n >= 0, 'bad list-mode address'
    out['counted'] = n
    return out
