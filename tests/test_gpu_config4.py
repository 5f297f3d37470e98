"""The batched bootstrap / dynamic-frame workload at its full size: 64
Poisson(1) bootstrap replicates of a 600 s, 6e8-word list-mode acquisition
plus its 24 dynamic frames of 10-300 s, every replicate and frame
reconstructed 4 x 14 in ONE captured HIP graph (the bench's workload, built by
nipet_amd.prj.batch.replicates_and_frames), against plain sequential
reconstructions (one histogram session with nrep = 1, rep0 = r, or one frame
window, then osem_prepare / osem_run): 8 replicates and 4 frames, the
shortest and the longest among them, within 1e-5 relative L2."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ITR = 4


@pytest.fixture(scope='module')
def torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


def _rel(torch, a, b):
    return float(torch.linalg.vector_norm((a - b).double()) / torch.linalg.vector_norm(b.double()))


def test_64_replicates_and_24_frames_one_graph(torch, params):
    from nipet_amd.lm.mmrhist import hist_words
    from nipet_amd.prj import mmrsim
    from nipet_amd.prj.batch import FRAMES24, _remgaps_b, frame_prompts, replicates_and_frames
    from nipet_amd.prj.osem import osem_prepare, osem_run
    print('config-4 data', flush=True)
    d = mmrsim.osem_inputs(params, counts=3e8)
    words, rs = mmrsim.lm_workload(params, d, seconds=600, per_ms=1000, seed=5)
    assert words.numel() > 6.0e8
    seed = 2024
    print('config-4 batch: 64 replicates + 24 frames, one graph', flush=True)
    reps, frms, info = replicates_and_frames(words, params, d['attn'], d['norm'], rs, nrep=64,
                                             seed=seed, itr=ITR)
    assert len(reps) == 64 and len(frms) == 24
    print(f"timing {info['timing']}", flush=True)
    for r in (0, 1, 17, 31, 32, 45, 62, 63):
        h = hist_words(words, params, btp=2, seed=seed, nrep=1, rep0=r)
        p = _remgaps_b(h, params, 1, 0, 1)[..., 0].contiguous()
        del h
        st = osem_prepare(params, p, d['attn'], d['norm'], rs)
        ref = osem_run(st, ITR, check_counts=False)
        rel = _rel(torch, reps[r], ref)
        print(f'parity config-4 replicate {r}: {rel:.3g} rel L2 (bar 1e-5)', flush=True)
        assert rel <= 1e-5
        del st, ref, p
    assert _rel(torch, reps[0], reps[1]) > 1e-3
    dur = 600
    lengths = [t1 - t0 for t0, t1 in FRAMES24]
    picks = sorted({int(np.argmin(lengths)), int(np.argmax(lengths)), 14, 18})
    assert len(picks) == 4
    for f in picks:
        t0, t1 = FRAMES24[f]
        p, _ = frame_prompts(words, params, [FRAMES24[f]])
        st = osem_prepare(params, p[..., 0].contiguous(), d['attn'], d['norm'],
                          rs * ((t1 - t0) / dur))
        ref = osem_run(st, ITR, check_counts=False)
        rel = _rel(torch, frms[f], ref)
        print(f'parity config-4 frame {f} ({t0}-{t1} s): {rel:.3g} rel L2 (bar 1e-5)',
              flush=True)
        assert rel <= 1e-5
        del st, ref, p
