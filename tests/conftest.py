import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a HIP GPU (MI355X) and the built library')


@pytest.fixture(scope='session')
def params():
    from nipet_amd import get_mmrparams
    return get_mmrparams()


@pytest.fixture(scope='session')
def subsets(params):
    from nipet_amd import get_subsets14
    return [get_subsets14(n, params)[0] for n in range(14)]


def phantom(seed=1, nsph=20):
    """Water ellipsoid (activity 1) with spheres of radius U(0.5, 2.5) cm and
    activity U(0.5, 4) inside it, device layout (320,320,127) float32."""
    rng = np.random.default_rng(seed)
    vox, voxz = 0.208626, 0.203125
    y = (160 - np.arange(320) - 0.5) * vox
    x = (np.arange(320) - 160 + 0.5) * vox
    z = (np.arange(127) - 63) * voxz
    Y, X, Z = np.meshgrid(y, x, z, indexing='ij')
    ax = (9.0, 7.0, 8.0)
    inside = (X / ax[0])**2 + (Y / ax[1])**2 + (Z / ax[2])**2 <= 1.0
    img = inside.astype(np.float32)
    for _ in range(nsph):
        while True:
            c = rng.uniform(-1, 1, 3) * np.array(ax)
            if ((c / np.array(ax))**2).sum() <= 1:
                break
        r = rng.uniform(0.5, 2.5)
        a = rng.uniform(0.5, 4.0)
        m = (X - c[0])**2 + (Y - c[1])**2 + (Z - c[2])**2 <= r * r
        img[m] = a
    return img


def rel_err(got, ref, floor=1e-6):
    """Largest per-element relative error |got - ref| / max(|ref|, floor *
    max|ref|) (projection bins or image voxels)."""
    ref = np.asarray(ref, np.float64)
    den = np.maximum(np.abs(ref), floor * np.abs(ref).max())
    return float(np.max(np.abs(np.asarray(got, np.float64) - ref) / den))


def worst_elem(got, ref, floor=1e-6):
    """(index, |ref| / max|ref|, relative error) of the element that sets
    rel_err: where a margin is decided, and at what signal level."""
    ref = np.asarray(ref, np.float64)
    den = np.maximum(np.abs(ref), floor * np.abs(ref).max())
    e = np.abs(np.asarray(got, np.float64) - ref) / den
    i = int(np.argmax(e))
    return np.unravel_index(i, ref.shape), float(abs(ref.flat[i]) / np.abs(ref).max()), float(e.flat[i])


def margin(name, err, bar):
    """Print a parity margin (measured error against its bar) and return
    whether it holds."""
    print(f'parity {name}: {err:.3g} (bar {bar:g}, margin x{bar / max(err, 1e-300):.3g})',
          flush=True)
    return err < bar


def gpu_ok():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def nll_case(params, rays, seed=11, nvox=40):
    """Poisson-likelihood test case over transaxial bins ``rays``: current
    image x (phantom 3), mu-map, norm U(0.8, 1.2), randoms+scatter, prompts
    y ~ Poisson(a n A x_true + rs) with x_true = phantom 1 (FP64 oracle
    projections), and ``nvox`` voxels on the rays' columns (flat indices)."""
    from oracle import oracle as O
    from nipet_amd.prj.mmrsim import mumap_phantom
    rng = np.random.default_rng(seed)
    rays = np.asarray(rays, np.int32)
    xt, x, mu = phantom(1), phantom(3), mumap_phantom()
    a = O.fprj(params, mu, rays, att=True)
    p = O.fprj(params, xt, rays)
    # norm and rs are float32 values (what a float32 device model sees)
    n = rng.uniform(0.8, 1.2, a.shape).astype(np.float32).astype(np.float64)
    rs = (0.3 * (a * n * p).mean() * rng.uniform(0.5, 1.5, a.shape)).astype(np.float32)
    rs = rs.astype(np.float64)
    y = rng.poisson(a * n * p + rs).astype(np.float64)
    off, col, _, _, _, _ = O.tx_lists(params, rays)
    vox = np.zeros(0, np.int64)
    while len(vox) < nvox:
        k = rng.integers(0, off[-1], nvox)
        vox = np.unique(np.concatenate([vox, col[k].astype(np.int64) * 127 +
                                        rng.integers(2, 125, nvox)]))
    return dict(rays=rays, x=x, mu=mu, n=n, rs=rs, y=y, vox=rng.permutation(vox)[:nvox])
