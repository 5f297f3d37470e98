"""Differentiable mMR projectors for PyTorch autograd.

The reference exposes the Siddon projector only as forward / back operators
(``prj/mmrprj.py:46-221``) and uses their pairing inside OSEM
(``prj/src/recon.cu:459-497``). Here the same operators become autograd
functions on device tensors:

* ``emission_fp(x)``: span-11 emission forward projection A x, whose
  vector-Jacobian product is the back projection A^T g (the back projector is
  the exact adjoint of the emission forward projector, prj/src/prjb.cu:95-96);
* ``attenuation_fp(mu)``: span-11 attenuation factors
  a_b = (1/n_b) sum_{LOR in b} exp(-l_LOR(mu)), the mean over the bin's n_b
  span-1 LORs (prj/src/prjf.cu:88-92, 193-200). Its vector-Jacobian product is
  J^T g = -A1^T (t * g[b(LOR)] / n_b(LOR)), with l and t = exp(-l) the span-1
  line integrals and transmissions recomputed by a span-1 forward projection
  and A1^T the span-1 back projection.

Both run on the HIP kernels of ``libnipet_hip.so``; the functions here only
allocate outputs and sequence the calls. On top of them: the Poisson negative
log-likelihood of the mMR data model y ~ Poisson(a n A x + r + s), and a
gradient MAP reconstruction with a quadratic neighbourhood prior whose
per-subset likelihood terms are recomputed in the backward pass
(``torch.utils.checkpoint``) instead of holding every subset's sinograms.
"""
import numpy as np

from .. import _lib
from ..geom import get_geom

__all__ = ['Projector', 'emission_fp', 'attenuation_fp', 'poisson_nll', 'quadratic_prior',
           'map_recon']


def _torch():
    import torch
    return torch


class Projector:
    """Span-11 projector over the transaxial bins ``isub`` (all bins when
    ``None``) of the scanner parameters, with the span-1 geometry needed by the
    attenuation adjoint built on first use."""

    def __init__(self, scanner_params, isub=None, device=None):
        torch = _torch()
        Cnt = scanner_params['Cnt']
        if Cnt['SPN'] != 11:
            raise ValueError('the differentiable projector works on span-11 sinograms')
        self.params = scanner_params
        self.geom = get_geom(scanner_params, device)
        self.nsn = self.geom.nsn
        if isub is None:
            self.subs, self.nprj = None, self.geom.naw
        else:
            s = isub if isinstance(isub, torch.Tensor) else torch.from_numpy(
                np.ascontiguousarray(isub, np.int32))
            s = s.to(device='cuda', dtype=torch.int32).contiguous()
            if s.numel() == 0 or int(s.min()) < 0 or int(s.max()) >= self.geom.naw:
                raise ValueError(f'isub entries must be in [0, {self.geom.naw})')
            self.subs, self.nprj = s, s.numel()
        self._g1 = None
        self._s1map = None

    # -- raw kernels (no autograd) -------------------------------------------------
    def fp(self, img, att=False, geom=None):
        torch = _torch()
        g = self.geom if geom is None else geom
        out = torch.empty((self.nprj, g.nsn), dtype=torch.float32, device=img.device)
        _lib.check(_lib.lib().nipet_fprj(g.handle, _lib.ptr(out), _lib.ptr(img),
                                         _lib.ptr(self.subs), self.nprj, int(bool(att)),
                                         _lib.stream_handle()), 'nipet_fprj')
        return out

    def bp(self, sino, geom=None):
        torch = _torch()
        g = self.geom if geom is None else geom
        out = torch.empty((320, 320, 127), dtype=torch.float32, device=sino.device)
        _lib.check(_lib.lib().nipet_bprj(g.handle, _lib.ptr(out), _lib.ptr(sino),
                                         _lib.ptr(self.subs), self.nprj, _lib.stream_handle()),
                   'nipet_bprj')
        return out

    def span1(self):
        """(span-1 geometry, span-11 bin of each span-1 sinogram (long),
        1/n_b of that bin (float32)), built once."""
        if self._g1 is None:
            torch = _torch()
            p1 = dict(self.params)
            p1['Cnt'] = dict(self.params['Cnt'], SPN=1)
            self._g1 = get_geom(p1, self.geom.device)
            ax = self.params['axLUT']
            li2sn1 = np.asarray(ax['li2sn1'], np.int64)
            li2sn = np.asarray(ax['li2sn'], np.int64)
            nos = np.asarray(ax['li2nos'], np.float64)
            s11 = np.full(self._g1.nsn, -1, np.int64)
            inv = np.zeros(self._g1.nsn, np.float32)
            for k in range(2):
                s11[li2sn1[:, k]] = li2sn[:, k]
                inv[li2sn1[:, k]] = (1.0 / nos).astype(np.float32)
            if (s11 < 0).any():
                raise _lib.NipetError('span-1 to span-11 map is incomplete')
            self._s1map = (torch.from_numpy(s11).cuda(), torch.from_numpy(inv).cuda())
        return self._g1, self._s1map[0], self._s1map[1]

    def attenuation_vjp(self, mu, g):
        """J^T g for a = attenuation_fp(mu): -A1^T(exp(-A1 mu) * g[b] / n_b)."""
        g1, s11, inv = self.span1()
        t = self.fp(mu, att=False, geom=g1)           # span-1 line integrals (nprj, 4084)
        t.neg_().exp_()
        t.mul_(g.index_select(1, s11)).mul_(inv)
        return self.bp(t, geom=g1).neg_()


def _autograd():
    torch = _torch()

    class EmissionFP(torch.autograd.Function):
        @staticmethod
        def forward(ctx, img, prj):
            ctx.prj = prj
            return prj.fp(img.contiguous(), att=False)

        @staticmethod
        def backward(ctx, g):
            return ctx.prj.bp(g.contiguous()), None

    class AttenuationFP(torch.autograd.Function):
        @staticmethod
        def forward(ctx, mu, prj):
            mu = mu.contiguous()
            ctx.prj = prj
            ctx.save_for_backward(mu)
            return prj.fp(mu, att=True)

        @staticmethod
        def backward(ctx, g):
            (mu,) = ctx.saved_tensors
            return ctx.prj.attenuation_vjp(mu, g.contiguous()), None

    return EmissionFP, AttenuationFP


_FUNCS = None


def _funcs():
    global _FUNCS
    if _FUNCS is None:
        _FUNCS = _autograd()
    return _FUNCS


def _check_img(x, name):
    torch = _torch()
    if not isinstance(x, torch.Tensor) or not x.is_cuda or x.dtype != torch.float32 or \
            tuple(x.shape) != (320, 320, 127):
        raise ValueError(f'{name} must be a float32 CUDA tensor of shape (320, 320, 127)')


def emission_fp(img, prj):
    """Differentiable A x: (nprj, 837) span-11 emission sinogram of device
    image ``img`` (320,320,127) float32 over ``prj``'s transaxial bins."""
    _check_img(img, 'img')
    return _funcs()[0].apply(img, prj)


def attenuation_fp(mu, prj):
    """Differentiable span-11 attenuation factors (nprj, 837) of mu-map ``mu``
    (320,320,127) float32 [1/cm]."""
    _check_img(mu, 'mu')
    return _funcs()[1].apply(mu, prj)


def poisson_nll(img, mu, prompts, norm, rs, prj):
    """Poisson negative log-likelihood sum_b (ybar_b - y_b log ybar_b) of the
    prompts under ybar = attenuation_fp(mu) * norm * emission_fp(img) + rs
    (all sinograms (nprj, 837) over ``prj``'s bins; bins with ybar <= 0 and
    y = 0 contribute ybar). Differentiable in ``img`` and ``mu``; returned in
    float64 (the per-bin terms are formed in float32)."""
    ybar = attenuation_fp(mu, prj) * norm * emission_fp(img, prj) + rs
    y = prompts.to(ybar.dtype)
    ll = ybar - y * _log_where(ybar, y)
    return ll.sum(dtype=_torch().float64)


def _log_where(ybar, y):
    """log(ybar) where y > 0, 0 elsewhere (so 0 log 0 = 0)."""
    torch = _torch()
    return torch.where(y > 0, torch.log(torch.where(y > 0, ybar, torch.ones_like(ybar))),
                       torch.zeros_like(ybar))


def quadratic_prior(x):
    """R(x) = 1/2 sum over the 6-neighbourhood pairs (x_j - x_k)^2 (each pair
    once), on the (320,320,127) image grid."""
    r = (x[1:] - x[:-1]).pow(2).sum() + (x[:, 1:] - x[:, :-1]).pow(2).sum() + \
        (x[:, :, 1:] - x[:, :, :-1]).pow(2).sum()
    return 0.5 * r.to(_torch().float64)


def map_recon(scanner_params, prompts, norm, rs, mu, subsets, beta=1e-3, steps=200, step=1.0,
              img0=None, mask=None, sens=None, checkpoint=True, callback=None):
    """Gradient MAP reconstruction: minimises
    Phi(x) = sum_s NLL_s(x) + beta R(x) (``poisson_nll`` over subset s's bins,
    ``quadratic_prior`` R) by preconditioned projected gradient steps
    x <- max(0, x - step * x / sens * grad Phi(x)), inside ``mask``, with
    sens = A^T(a n) (all bins). With beta = 0 and step = 1 one step is exactly
    one MLEM iteration. Every step evaluates the full objective; each subset's
    likelihood term runs under ``torch.utils.checkpoint`` so that only one
    subset's sinograms are alive at a time in the backward pass.

    prompts / norm / rs: (Naw, 837) device tensors; mu: (320,320,127) mu-map
    (held fixed); subsets: list of transaxial-bin index arrays. Returns the
    image and the list of objective values (one per step, before the step)."""
    torch = _torch()
    from torch.utils.checkpoint import checkpoint as ckpt
    prjs = [Projector(scanner_params, s) for s in subsets]
    with torch.no_grad():
        prompts = prompts.to(torch.float32)
        att = [prj.fp(mu, att=True) for prj in prjs]
        rows = [p.subs.long() for p in prjs]
        an = [a * norm[r] for a, r in zip(att, rows)]
        ys = [prompts[r] for r in rows]
        rss = [rs[r] for r in rows]
        if sens is None:
            sens = sum(p.bp(a) for p, a in zip(prjs, an))
        if mask is None:
            from .osem import cylinder_mask
            mask = cylinder_mask(scanner_params['Cnt'], 29.).bool()
        x = torch.ones((320, 320, 127), device='cuda') if img0 is None else img0.clone()
        x = torch.where(mask, x, torch.zeros_like(x))
        pre_scale = torch.where(sens > 0, 1.0 / sens, torch.zeros_like(sens))

    def nll_s(x, s):
        ybar = an[s] * emission_fp(x, prjs[s]) + rss[s]
        ll = ybar - ys[s] * _log_where(ybar, ys[s])
        return ll.sum(dtype=torch.float64)

    hist = []
    for _ in range(steps):
        xv = x.detach().requires_grad_(True)
        phi = beta * quadratic_prior(xv) if beta else xv.sum() * 0
        for s in range(len(prjs)):
            phi = phi + (ckpt(nll_s, xv, s, use_reentrant=False) if checkpoint else nll_s(xv, s))
        g, = torch.autograd.grad(phi, xv)
        hist.append(float(phi.detach()))
        with torch.no_grad():
            x = torch.clamp_min(x - step * x * pre_scale * g, 0.0)
            x = torch.where(mask, x, torch.zeros_like(x))
        if callback is not None:
            callback(x, hist[-1])
    return x, hist
