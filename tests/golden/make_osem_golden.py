"""Writes tests/golden/osem4x14_sample.npz: the FP64 oracle's OSEM image after
4 iterations x 14 subsets on the synthetic inputs of osem_data.py (full mMR
span-11 geometry, 320x320x127 image, mask radius 29 cm), sampled at 200,000
voxels inside the mask plus global statistics. Run from the repository root:
    python tests/golden/make_osem_golden.py
(about 10 minutes on 8 cores)."""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from osem_data import checksum, make_osem_data  # noqa: E402

import nipet_amd as nip  # noqa: E402
from oracle import oracle as O  # noqa: E402

NITER, NSAMPLE, SAMPLE_SEED = 4, 200_000, 11


def main():
    t0 = time.time()
    params = nip.get_mmrparams()
    d = make_osem_data(params)
    a = d['attn'].astype(np.float64)
    n = d['norm'].astype(np.float64)
    an = a * n
    rs = d['randoms'].astype(np.float64) + d['scatter'].astype(np.float64)
    rho = np.where(np.abs(an) < 1e-6, 0.0, rs / np.where(np.abs(an) < 1e-6, 1.0, an))
    subsets = [nip.get_subsets14(k, params)[0] for k in range(14)]
    sens = np.stack([O.bprj(params, an[s], s) for s in subsets])
    print(f'sensitivity done {time.time() - t0:.0f}s', flush=True)
    mask = nip.get_cylinder(params['Cnt'], rad=29., unival=1, gpu_dim=True) > 0.9
    x = np.ones((320, 320, 127))
    for it in range(NITER):
        x = O.osem(params, d['prompts'], rho, sens, mask, subsets, x, 1)
        print(f'iteration {it + 1} done {time.time() - t0:.0f}s', flush=True)
    inside = np.flatnonzero(mask)
    idx = np.sort(np.random.default_rng(SAMPLE_SEED).choice(inside, NSAMPLE, replace=False))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'osem4x14_sample.npz')
    np.savez_compressed(out, idx=idx.astype(np.int64), val=x.ravel()[idx],
                        total=x.sum(), l2=np.sqrt((x * x).sum()),
                        checksum=np.array(checksum(d['prompts']), dtype=np.uint64),
                        niter=NITER)
    print('wrote', out, f'{time.time() - t0:.0f}s', flush=True)


if __name__ == '__main__':
    main()
