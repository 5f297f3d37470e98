"""Synthetic Siemens mMR list-mode streams (int32 words) generated on the GPU.

Layout per millisecond k (t_first + k ms, k = 0 .. duration*1000): a time tag
word 0x80000000 | ms, then, at ms % 1000 == 500, one singles report per bucket
(0xA0000000 | bucket << 19 | value), then `per_ms` events. Events are prompts
((1 << 30) | address) with probability `prompt_frac`, else delayeds
(address). Prompt addresses follow a span-11 weight sinogram (Naw, 837): a
span-11 bin is drawn with probability proportional to its weight and one of
its span-1 sinograms uniformly; delayed addresses are uniform over the active
span-1 bins (sinogram >= 1, so that no word is 0).
"""
import numpy as np

__all__ = ['synth_lm', 'span11_members']


def span11_members(axLUT):
    """(837, 6) span-1 sinogram indices of each span-11 sinogram (-1 padded), counts."""
    s11 = np.asarray(axLUT['sn1_sn11'], np.int64)
    mem = -np.ones((837, 6), np.int64)
    cnt = np.zeros(837, np.int64)
    for si, s in enumerate(s11):
        mem[s, cnt[s]] = si
        cnt[s] += 1
    return mem, cnt


def synth_lm(scanner_params, duration_s, per_ms, weights=None, prompt_frac=0.85, seed=5,
             t_first=1_000_000, singles=True, chunk=1 << 25):
    """Device int32 word tensor of a synthetic acquisition of ``duration_s`` seconds."""
    import torch
    tx, ax = scanner_params['txLUT'], scanner_params['axLUT']
    g = torch.Generator(device='cuda').manual_seed(seed)
    ntag = duration_s * 1000 + 1
    ms = torch.arange(ntag, device='cuda', dtype=torch.int64)
    nsg = torch.where((ms % 1000 == 500) & torch.tensor(singles, device='cuda'), 224, 0)
    nev = torch.full_like(ms, per_ms)
    nev[-1] = 0
    per = 1 + nsg + nev
    off = torch.cumsum(per, 0) - per
    total = int(per.sum())
    words = torch.empty(total, dtype=torch.int32, device='cuda')
    words[off] = (torch.tensor(-(1 << 31), device='cuda', dtype=torch.int64) |
                  (t_first + ms)).to(torch.int32)
    if singles:
        sec = torch.nonzero(nsg > 0).squeeze(1)
        b = torch.arange(224, device='cuda')
        pos = (off[sec][:, None] + 1 + b[None, :]).reshape(-1)
        val = torch.randint(0, 1 << 19, (len(pos),), generator=g, device='cuda')
        val[::7] = 0                                    # some empty reports
        w = (0xA0000000 - (1 << 32)) | (b.repeat(len(sec)) << 19) | val
        words[pos] = w.to(torch.int32)
    aw2ali = torch.from_numpy(np.asarray(tx['aw2ali'], np.int64)).cuda()
    mem, cnt = span11_members(ax)
    mem, cnt = torch.from_numpy(mem).cuda(), torch.from_numpy(cnt).cuda()
    naw = len(tx['aw2ali'])
    if weights is not None:
        # integer weights (2^32 for the largest): an integer scan is exact, so
        # the stream is the same on every run (a float scan regroups its sums
        # with the device's scheduling and moves an event now and then)
        wf = weights.reshape(-1).to(torch.float64)
        wi = torch.round(wf.clamp(min=0) * (2.0**32 / float(wf.max()))).to(torch.int64)
        cdf = torch.cumsum(wi, 0).to(torch.float64)
        cdf /= float(cdf[-1])
    nevents = int(nev.sum())
    for e0 in range(0, nevents, chunk):
        e = torch.arange(e0, min(nevents, e0 + chunk), device='cuda', dtype=torch.int64)
        k = e // per_ms
        pos = off[k] + 1 + nsg[k] + (e - k * per_ms)
        prompt = torch.rand(len(e), generator=g, device='cuda') < prompt_frac
        if weights is not None:
            idx = torch.searchsorted(cdf, torch.rand(len(e), generator=g, device='cuda',
                                                     dtype=torch.float64))
            idx = idx.clamp_(max=cdf.numel() - 1)
            aw_p, s11 = idx // 837, idx % 837
        else:
            aw_p = torch.randint(0, naw, (len(e),), generator=g, device='cuda')
            s11 = torch.randint(0, 837, (len(e),), generator=g, device='cuda')
        j = (torch.rand(len(e), generator=g, device='cuda') * cnt[s11]).long()
        si_p = mem[s11, j.clamp(max=5)]
        si_d = torch.randint(1, 4084, (len(e),), generator=g, device='cuda')
        aw_d = torch.randint(0, naw, (len(e),), generator=g, device='cuda')
        addr = torch.where(prompt, si_p * 86688 + aw2ali[aw_p], si_d * 86688 + aw2ali[aw_d])
        words[pos] = (addr | (prompt.to(torch.int64) << 30)).to(torch.int32)
    return words
