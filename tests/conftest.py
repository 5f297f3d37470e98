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


def gpu_ok():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
