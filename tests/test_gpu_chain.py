"""mmrchain (img/pipe.py:20-373) on the GPU: frames of a list-mode file in the
reference's 'def' form, histogrammed per frame, reconstructed as one batched
graph launch, equal to frame-by-frame reconstructions (batch=False, the
osemone path) within 1e-5 relative L2 per frame."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


def test_mmrchain_batched_equals_frame_by_frame(torch, params, tmp_path):
    import nipet_amd as nip
    from nipet_amd.lm.lmsim import synth_lm
    from nipet_amd.prj import mmrsim
    from nipet_amd.prj.mmrprj import frwd_prj
    Cnt = params['Cnt']
    img = torch.from_numpy(mmrsim.ellipsoid_phantom(1)).cuda()
    mu = mmrsim.mumap_phantom(2)
    attn = frwd_prj(torch.from_numpy(mu).cuda(), params, dev_out=True, attenuation=True)
    emis = frwd_prj(img, params, dev_out=True)
    words = synth_lm(params, 12, 900, weights=attn * emis, seed=21).cpu().numpy()
    f = os.path.join(tmp_path, 'chain.bf')
    words.tofile(f)
    mu_o = {'im': nip.convert2e7(mu, Cnt)}
    mu_h = {'im': np.zeros((127, 344, 344), np.float32)}
    frames = ['def', [2, 3], [1, 6]]
    kw = dict(frames=frames, mu_h=mu_h, mu_o=mu_o, itr=2, ret_histo=True)
    print('mmrchain batched', flush=True)
    out_b = nip.mmrchain({'lm_bf': f}, params, **kw)
    print('mmrchain frame by frame', flush=True)
    out_s = nip.mmrchain({'lm_bf': f}, params, batch=False, **kw)
    assert out_b['frames'] == [[0, 3], [3, 6], [6, 12]] and out_b['#frames'] == 3
    assert out_b['recmod'] == 3 and out_b['im'].shape == (3, 127, 344, 344)
    assert set(out_b['hst']) == {'0-3', '3-6', '6-12'}
    for k in range(3):
        a, b = out_b['im'][k].astype(np.float64), out_s['im'][k].astype(np.float64)
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f'parity mmrchain frame {k}: {rel:.3g} rel L2 (bar 1e-5)', flush=True)
        assert np.linalg.norm(b) > 0 and rel < 1e-5
    # the longest frame has the most counts
    assert out_b['im'][2].sum() > out_b['im'][0].sum()
    assert out_b['fpet'].endswith('_nfrm-3_itr-2.nii.gz')
