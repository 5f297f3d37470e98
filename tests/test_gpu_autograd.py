"""Differentiable projectors (nipet_amd.prj.autograd) on the GPU:
adjoint dot-product tests through autograd, Poisson-NLL gradients with
respect to the activity image and the mu-map against FP64 central finite
differences (oracle orc_nll_fd, pinned in test_nll_fd_oracle.py) on 1000
random voxels, and gradient MAP with beta = 0 reproducing one MLEM iteration."""
import numpy as np
import pytest

from conftest import nll_case

pytestmark = pytest.mark.gpu


def central_rays(params, n=24):
    """One transaxial bin within 40 radial bins of the centre at each of n
    evenly spaced angles (chords through the phantom)."""
    ali = np.asarray(params['txLUT']['aw2ali'])
    iw, ia = ali % 344, ali // 344
    out = []
    for k, ang in enumerate(np.linspace(0, 251, n).astype(int)):
        cand = np.nonzero((ia == ang) & (np.abs(iw - 172) < 40))[0]
        out.append(cand[(7 * k) % len(cand)])
    return np.array(out, np.int32)


@pytest.fixture(scope='module')
def torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


@pytest.fixture(scope='module')
def case(params):
    return nll_case(params, central_rays(params), nvox=1000)


def test_emission_adjoint_autograd(torch, params):
    """<A x, y> = <x, A^T y> over all 68,516 bins, A^T from autograd: 1e-6 rel."""
    from nipet_amd.prj.autograd import Projector, emission_fp
    prj = Projector(params)
    g = torch.Generator(device='cuda').manual_seed(3)
    x = torch.rand((320, 320, 127), device='cuda', generator=g).requires_grad_(True)
    y = torch.rand((prj.nprj, 837), device='cuda', generator=g)
    ax = emission_fp(x, prj)
    aty, = torch.autograd.grad(ax, x, grad_outputs=y)
    lhs = float((ax.detach().double() * y.double()).sum())
    rhs = float((x.detach().double() * aty.double()).sum())
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs), (lhs, rhs)


def test_attenuation_span1_and_adjoint(torch, params):
    """Span-11 attenuation factors equal the span-1 mean of exp(-line integral)
    (1e-6 rel), and the attenuation VJP is the transpose of the linearisation
    J dmu = -(1/n_b) sum_LOR exp(-l) (A1 dmu): <J dmu, g> = <dmu, J^T g>."""
    from nipet_amd.prj.autograd import Projector, attenuation_fp
    from nipet_amd.prj.mmrsim import mumap_phantom
    prj = Projector(params, np.arange(0, 68516, 7))
    mu = torch.from_numpy(mumap_phantom()).cuda().requires_grad_(True)
    a = attenuation_fp(mu, prj)
    g1, s11, inv = prj.span1()
    t = torch.exp(-prj.fp(mu.detach(), geom=g1))
    a1 = torch.zeros_like(a).index_add_(1, s11, t * inv)
    assert float((a1 - a.detach()).abs().max()) <= 1e-6 * float(a.detach().abs().max())
    gen = torch.Generator(device='cuda').manual_seed(5)
    dmu = torch.rand(mu.shape, device='cuda', generator=gen) * 0.01
    gs = torch.rand(a.shape, device='cuda', generator=gen)
    jd = torch.zeros_like(a).index_add_(1, s11, -t * inv * prj.fp(dmu, geom=g1))
    jtg, = torch.autograd.grad(a, mu, grad_outputs=gs)
    lhs = float((jd.double() * gs.double()).sum())
    rhs = float((dmu.double() * jtg.double()).sum())
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs), (lhs, rhs)


@pytest.fixture(scope='module')
def gpu_grads(torch, params, case):
    from nipet_amd.prj.autograd import Projector, poisson_nll
    c = case
    prj = Projector(params, c['rays'])
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()  # noqa: E731
    x = dev(c['x']).requires_grad_(True)
    mu = dev(c['mu']).requires_grad_(True)
    nll = poisson_nll(x, mu, dev(c['y']), dev(c['n']), dev(c['rs']), prj)
    gx, gmu = torch.autograd.grad(nll, (x, mu))
    return float(nll.detach()), gx.cpu().numpy().ravel(), gmu.cpu().numpy().ravel()


def test_nll_value(params, case, gpu_grads):
    from oracle import oracle as O
    c = case
    a = O.fprj(params, c['mu'], c['rays'], att=True)
    p = O.fprj(params, c['x'], c['rays'])
    yb = a * c['n'] * p + c['rs']
    ref = float((yb - np.where(c['y'] > 0, c['y'] * np.log(yb), 0)).sum())
    assert abs(gpu_grads[0] - ref) <= 1e-5 * abs(ref), (gpu_grads[0], ref)


@pytest.mark.parametrize('wrt,h', [(0, 1e-3), (1, 1e-4)])
def test_grad_vs_fd64(params, case, gpu_grads, wrt, h):
    """dNLL/dx and dNLL/dmu (float32 GPU autograd) against FP64 central finite
    differences on 1000 random voxels of the rays' columns. Bound: within
    1e-5 of max |g| and 1e-5 relative L2 over the voxels (measured on MI355X:
    about 1e-7 for both)."""
    from oracle import oracle as O
    c = case
    fd = O.nll_fd(params, c['x'], c['mu'], c['rays'], c['y'], c['n'], c['rs'], c['vox'], wrt, h)
    g = gpu_grads[1 + wrt][c['vox']]
    err = np.abs(g - fd)
    rel_l2 = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    print(f'wrt={"mu" if wrt else "x"}: max|fd|={np.abs(fd).max():.4g} '
          f'max err={err.max():.3g} rel L2={rel_l2:.3g}')
    # voxels whose LORs miss the activity have a zero mu-gradient
    assert np.count_nonzero(fd) > (900 if wrt == 0 else 700)
    assert err.max() <= 1e-5 * np.abs(fd).max()
    assert rel_l2 <= 1e-5


def test_map_beta0_step1_is_mlem(torch, params):
    """One gradient-MAP step with beta = 0 and the EM preconditioner x/sens is
    one MLEM iteration x B / sens: equal to a single-subset OSEM iteration,
    with and without per-subset checkpointing. Bound 1e-4 rel L2: the gradient
    form x - x/sens (sens - B) cancels where B << sens, and the back-projected
    gradient terms a n (1 - y/ybar) have mixed signs at low counts, so the
    float atomics' summation order (run to run) shows at about 1e-5."""
    from nipet_amd.prj import mmrsim
    from nipet_amd.prj.autograd import map_recon
    from nipet_amd.prj.osem import osem_prepare, osem_run
    d = mmrsim.osem_inputs(params, counts=3e7, seed=8)
    allb = np.arange(68516, dtype=np.int32)
    st = osem_prepare(params, d['prompts'], d['attn'], d['norm'], d['randoms'], d['scatter'],
                      subsets=[allb])
    x0 = st.mask.to(torch.float32)                  # both start from ones inside the mask
    ref = osem_run(st, 1, img0=x0, check_counts=False)
    rs = d['randoms'] + d['scatter']
    mu = torch.from_numpy(d['mu']).cuda()
    halves = [allb[:34258], allb[34258:]]
    x1, h1 = map_recon(params, d['prompts'], d['norm'], rs, mu, halves, beta=0.0, steps=1)
    x2, h2 = map_recon(params, d['prompts'], d['norm'], rs, mu, halves, beta=0.0, steps=1,
                       checkpoint=False)
    def rel(a, b):
        return float(torch.linalg.vector_norm((a - b).double()) /
                     torch.linalg.vector_norm(b.double()))
    print(f'checkpointed vs plain {rel(x2, x1):.3g}, vs MLEM {rel(x1, ref):.3g}')
    assert rel(x2, x1) <= 1e-4 and abs(h1[0] - h2[0]) <= 1e-9 * abs(h1[0])
    assert rel(x1, ref) <= 1e-4


def test_map_prior_decreases_objective(torch, params):
    """Five MAP steps with beta = 1e-3 lower the penalised objective at each step."""
    from nipet_amd.prj import mmrsim
    from nipet_amd.prj.autograd import map_recon
    from nipet_amd.prj.mmrrec import get_subsets14
    d = mmrsim.osem_inputs(params, counts=3e7, seed=9)
    mu = torch.from_numpy(d['mu']).cuda()
    subs = [get_subsets14(k, params)[0] for k in range(14)]
    _, hist = map_recon(params, d['prompts'], d['norm'], d['randoms'] + d['scatter'], mu, subs,
                        beta=1e-3, steps=6)
    assert all(b < a for a, b in zip(hist[:-1], hist[1:])), hist
