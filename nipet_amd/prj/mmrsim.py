"""Synthetic data on the mMR geometry (``niftypet/nipet/prj/mmrsim.py``).

* ``simulate_sino`` -- ``prj/mmrsim.py:18-116``: attenuation factors of a
  mu-map times the emission projection of a PET image (3D, mu-map input).
* ``ellipsoid_phantom`` / ``mumap_phantom`` -- seeded device-layout images:
  water ellipsoid (semi-axes 9, 7, 8 cm) of activity 1 with spheres of radius
  U(0.5, 2.5) cm and activity U(0.5, 4); mu-map of the same ellipsoid in water
  (0.096 /cm) with a 0.4 cm bone shell (0.15 /cm).
* ``osem_inputs`` -- a full span-11 OSEM data set: prompts =
  Poisson(scale (a n FP(phantom)) + randoms + scatter) with flat randoms
  (12 % of the trues), smooth scatter (25 % of the trues), norm U(0.8, 1.2),
  scaled to a given number of expected counts. Generated on the GPU.
* ``lm_workload`` -- a list-mode acquisition of the same data model: prompt
  bins drawn from the expected prompts of ``osem_inputs`` (trues + randoms +
  scatter), with the randoms + scatter sinogram at the stream's count level.
"""
import numpy as np

VOX, VOXZ = 0.208626, 0.203125
AXES = (9.0, 7.0, 8.0)

__all__ = ['simulate_sino', 'ellipsoid_phantom', 'mumap_phantom', 'osem_inputs', 'lm_workload']


def _grid_axes():
    y = (160 - np.arange(320) - 0.5) * VOX     # row v -> y
    x = (np.arange(320) - 160 + 0.5) * VOX     # column u -> x
    z = (np.arange(127) - 63) * VOXZ           # plane w -> z
    return y, x, z


def _ellipsoid(scale=1.0):
    y, x, z = _grid_axes()
    ax = np.array(AXES) * scale
    r = (x[None, :, None] / ax[0])**2 + (y[:, None, None] / ax[1])**2 + \
        (z[None, None, :] / ax[2])**2
    return r <= 1.0


def ellipsoid_phantom(seed=1, nsph=20):
    """Activity phantom in the device layout (320, 320, 127) float32."""
    rng = np.random.default_rng(seed)
    y, x, z = _grid_axes()
    img = _ellipsoid().astype(np.float32)
    ax = np.array(AXES)
    for _ in range(nsph):
        while True:
            c = rng.uniform(-1, 1, 3) * ax
            if ((c / ax)**2).sum() <= 1:
                break
        rad = rng.uniform(0.5, 2.5)
        act = rng.uniform(0.5, 4.0)
        # bounding box of the sphere
        iv = np.nonzero(np.abs(y - c[1]) <= rad)[0]
        iu = np.nonzero(np.abs(x - c[0]) <= rad)[0]
        iw = np.nonzero(np.abs(z - c[2]) <= rad)[0]
        if len(iv) == 0 or len(iu) == 0 or len(iw) == 0:
            continue
        sub = ((x[iu][None, :, None] - c[0])**2 + (y[iv][:, None, None] - c[1])**2 +
               (z[iw][None, None, :] - c[2])**2) <= rad * rad
        blk = img[iv[0]:iv[-1] + 1, iu[0]:iu[-1] + 1, iw[0]:iw[-1] + 1]
        blk[sub] = act
    return img


def mumap_phantom(seed=2):
    """mu-map [1/cm] (320, 320, 127) float32: water ellipsoid with a bone shell."""
    inner = _ellipsoid(1.0 - 0.4 / min(AXES))
    outer = _ellipsoid()
    mu = np.zeros((320, 320, 127), np.float32)
    mu[outer] = 0.15
    mu[inner] = 0.096
    return mu


def simulate_sino(petim, ctim, scanner_params, simulate_3d=True, slice_idx=-1, mu_input=True):
    """Attenuated emission sinogram attsino * emisino (Siemens shape), as
    ``prj/mmrsim.py:18-116`` for the 3D case with a mu-map input."""
    from .mmrprj import frwd_prj
    if not simulate_3d:
        raise NotImplementedError('the 2D reduced-ring simulation is not part of this path')
    if not mu_input:
        raise NotImplementedError('CT (HU) input needs a HU -> mu conversion; pass a mu-map')
    if petim.shape != ctim.shape:
        raise ValueError('The shapes of the PET and CT images are inconsistent.')
    mui = np.array(ctim, dtype=np.float32)
    mui[mui < 0] = 0
    attsino = frwd_prj(mui, scanner_params, attenuation=True)
    emisino = frwd_prj(petim, scanner_params, attenuation=False)
    return attsino * emisino


def _gauss_matrix(n, sigma, periodic):
    """(n, n) row-normalised Gaussian smoothing matrix (edge rows renormalised,
    or wrapped when periodic)."""
    i = np.arange(n)
    dist = np.abs(i[:, None] - i[None, :]).astype(np.float64)
    if periodic:
        dist = np.minimum(dist, n - dist)
    k = np.exp(-0.5 * (dist / sigma)**2)
    return (k / k.sum(1, keepdims=True)).astype(np.float32)


def _smooth_scatter(trues_dev, txLUT, Cnt, sigma_bins=10.0, sigma_ang=3.0):
    """Gaussian-smoothed copy of a device sinogram over the radial (344) and
    angular (252, periodic) axes of the Siemens layout, as two small dense
    matrix products on the GPU."""
    import torch

    from .mmrprj import _gaps_dev
    nsn = trues_dev.shape[1]
    sino = _gaps_dev(trues_dev, txLUT, Cnt, nsn, put=True)        # (nsn, 252, 344)
    Gr = torch.from_numpy(_gauss_matrix(Cnt['NSBINS'], sigma_bins, False)).cuda()
    Ga = torch.from_numpy(_gauss_matrix(Cnt['NSANGLES'], sigma_ang, True)).cuda()
    x = torch.matmul(sino, Gr.T)                                   # radial
    x = torch.matmul(Ga, x)                                        # angular (batched over nsn)
    return _gaps_dev(x.contiguous(), txLUT, Cnt, nsn, put=False)


def osem_inputs(scanner_params, counts=3e8, seed=4, img=None, mu=None):
    """Device tensors of a span-11 OSEM data set (Naw, 837): prompts (uint16),
    attenuation, norm, randoms, scatter, plus the phantom and mu-map used."""
    import torch

    from .mmrprj import frwd_prj
    Cnt, txLUT = scanner_params['Cnt'], scanner_params['txLUT']
    img = ellipsoid_phantom(1) if img is None else img
    mu = mumap_phantom(2) if mu is None else mu
    emis = frwd_prj(torch.from_numpy(img).cuda(), scanner_params, dev_out=True)
    attn = frwd_prj(torch.from_numpy(mu).cuda(), scanner_params, dev_out=True, attenuation=True)
    g = torch.Generator(device='cuda').manual_seed(seed)
    norm = 0.8 + 0.4 * torch.rand(emis.shape, generator=g, device='cuda')
    trues = attn * norm * emis
    tsum = float(trues.sum(dtype=torch.float64))
    sct = _smooth_scatter(emis, txLUT, Cnt)
    sct *= 0.25 * tsum / float(sct.sum(dtype=torch.float64))
    rnd = torch.full_like(trues, 0.12 * tsum / trues.numel())
    scale = counts / (tsum * 1.37)
    lam = scale * (trues + sct + rnd)
    prompts = torch.poisson(lam, generator=g).clamp_(0, 65535).to(torch.int32).to(torch.uint16)
    return {'prompts': prompts, 'attn': attn, 'norm': norm, 'randoms': scale * rnd,
            'scatter': scale * sct, 'img': img, 'mu': mu, 'scale': scale}


def lm_workload(scanner_params, d, seconds=600, per_ms=1000, seed=5, prompt_frac=0.85):
    """List-mode stream of ``seconds`` s (1 ms tags, ``per_ms`` events per ms,
    ``prompt_frac`` prompts) whose prompt bins follow the expected prompts of
    the ``osem_inputs`` data set ``d`` (a n FP(phantom) scaled, + randoms +
    scatter). Returns (device int32 words, randoms + scatter (Naw, 837) at the
    stream's prompt count: the additive term of a reconstruction of the whole
    stream, or of a bootstrap replicate of it; a frame of duration T takes
    T / ``seconds`` of it)."""
    import torch

    from ..lm.lmsim import synth_lm
    from .mmrprj import frwd_prj
    emis = frwd_prj(torch.from_numpy(d['img']).cuda(), scanner_params, dev_out=True)
    rs = d['randoms'] + d['scatter']
    lam = d['scale'] * d['attn'] * d['norm'] * emis + rs
    del emis
    words = synth_lm(scanner_params, seconds, per_ms, weights=lam, seed=seed,
                     prompt_frac=prompt_frac)
    nprompt = int(((words > 0) & (((words >> 30) & 1) == 1)).sum(dtype=torch.int64))
    rs = rs * (nprompt / float(lam.sum(dtype=torch.float64)))
    return words, rs.contiguous()
