"""GPU list-mode histogramming, bit-exact against the sequential CPU oracle
(oracle/lm_oracle.c) on every output: span-11 prompt/delayed sinograms,
SSRB sinogram, head curves, centre-of-mass sums, projection views, fan-sums
and singles buckets; plain, bootstrap, frame window, chunked and replicates."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OUT = ('psino', 'dsino', 'ssr', 'phc', 'dhc', 'zR', 'zM', 'pvs', 'fan', 'bck')


@pytest.fixture(scope='module')
def torch():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return torch


@pytest.fixture(scope='module')
def stream(torch, params):
    """6 s synthetic stream (1 ms tags, singles each second) with events
    before the first tag, zero words and out-of-range addresses mixed in."""
    from nipet_amd.lm.lmsim import synth_lm
    w = synth_lm(params, 6, 400, seed=7).cpu().numpy()
    rng = np.random.default_rng(1)
    pre = ((1 << 30) | rng.integers(86688, 4084 * 86688, 300)).astype(np.int32)
    junk_at = rng.choice(np.arange(10, len(w) - 10), 50, replace=False)
    w[junk_at[:25]] = 0
    w[junk_at[25:]] = (1 << 30) | 0x3ffffff0
    return np.concatenate([pre, w])


def _toff_nitag(w):
    t = np.nonzero((w >> 29) == -4)[0]
    tf, tl = int(w[t[0]] & 0x1fffffff), int(w[t[-1]] & 0x1fffffff)
    return tf, ((tl - tf) + 999) // 1000


def _gpu(torch, params, w, toff, nitag, chunks=(len,), **kw):
    from nipet_amd.lm.mmrhist import LmHistogram
    h = LmHistogram(params, toff, nitag, **kw)
    wd = torch.from_numpy(w).cuda()
    cuts = [0] + [c(w) if callable(c) else c for c in chunks]
    for a, b in zip(cuts[:-1], cuts[1:]):
        h.add(wd[a:b])
    h.finish()
    torch.cuda.synchronize()
    return h, {k: v.cpu().numpy() for k, v in h.out.items()}


def _compare(g, o, nrep):
    for k in ('psino', 'dsino', 'phc', 'dhc'):
        gv = g[k].astype(np.uint32)                       # [n][nrep]
        ov = o[k].T                                       # [nrep][n] -> [n][nrep]
        assert np.array_equal(gv, ov), k
    for k in ('ssr', 'pvs', 'fan', 'bck'):
        assert np.array_equal(g[k].astype(np.uint32).ravel(), o[k].ravel()), k
    for k in ('zR', 'zM'):
        assert np.array_equal(g[k].astype(np.uint64), o[k]), k


def test_philox_kat_gpu(torch):
    import ctypes as ct

    from nipet_amd import _lib
    kat = [([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
           ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
            [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]
    for ctr, key, exp in kat:
        c, k, o = (np.array(ctr, np.uint32), np.array(key, np.uint32), np.zeros(4, np.uint32))
        _lib.check(_lib.lib().nipet_philox4x32(c.ctypes.data, k.ctypes.data, o.ctypes.data))
        assert list(o) == exp


@pytest.mark.parametrize('btp', [0, 2])
def test_hist_bit_exact(torch, params, stream, btp):
    from oracle import oracle as O
    toff, nitag = _toff_nitag(stream)
    h, g = _gpu(torch, params, stream, toff, nitag, btp=btp, seed=99)
    o = O.lm_hist(stream, params, toff, nitag, btp=btp, seed=99)
    _compare(g, o, 1)
    n, bad, ovf = h.status()
    assert n == len(stream) and bad == 25 and ovf == 0
    # reference invariants (tests/test_amyloid_pvc.py:131-133)
    assert g['phc'].sum() == g['psino'].sum() and g['dhc'].sum() == g['dsino'].sum()


@pytest.mark.parametrize('span', [1, 0])
def test_hist_span1_and_ssrb(torch, params, stream, span):
    """Span-1 (4084 sinograms) and SSRB (127) sessions: the bucketed LDS counts
    cover 21,609 and 672 buckets of 16,384 bins; plain and bootstrap, bit-exact
    against the oracle."""
    from oracle import oracle as O
    toff, nitag = _toff_nitag(stream)
    w = stream[:900_000]
    for btp in (0, 2):
        h, g = _gpu(torch, params, w, toff, nitag, span=span, btp=btp, seed=31)
        o = O.lm_hist(w, params, toff, nitag, span=span, btp=btp, seed=31)
        for k in ('psino', 'dsino', 'phc', 'dhc'):
            assert np.array_equal(g[k].astype(np.uint32), o[k].T), (span, btp, k)
        assert g['psino'].shape[0] == (4084 if span == 1 else 127) * 86688


def test_hist_sparse_span1(torch, params, stream):
    """A short span-1 session: one tile of staged events spans all 338 coarse
    buckets (21,609 buckets, more than one LDS window of the partition pass);
    every output bit-exact against the oracle."""
    from oracle import oracle as O
    toff, nitag = _toff_nitag(stream)
    w = stream[:20_000]
    for btp in (0, 2):
        h, g = _gpu(torch, params, w, toff, nitag, span=1, btp=btp, seed=8)
        o = O.lm_hist(w, params, toff, nitag, span=1, btp=btp, seed=8)
        _compare(g, o, 1)
        assert o['psino'].sum() > 10_000


def test_hist_chunked_frame_window(torch, params, stream):
    from oracle import oracle as O
    toff, nitag = _toff_nitag(stream)
    cuts = (1234, 700_001, 700_002, 1_500_000, len)
    h, g = _gpu(torch, params, stream, toff, nitag, chunks=cuts, tstart=2, tstop=4, btp=2,
                seed=5)
    o = O.lm_hist(stream, params, toff, nitag, tstart=2, tstop=4, btp=2, seed=5)
    _compare(g, o, 1)
    assert g['phc'][:2].sum() == 0 and g['phc'][4:].sum() == 0


def test_hist_replicates(torch, params, stream):
    from oracle import oracle as O
    toff, nitag = _toff_nitag(stream)
    w = stream[:900_000]
    h, g = _gpu(torch, params, w, toff, nitag, btp=2, seed=2024, nrep=3)
    o = O.lm_hist(w, params, toff, nitag, btp=2, seed=2024, nrep=3)
    _compare(g, o, 3)
    # replicates differ from each other
    assert not np.array_equal(g['psino'][:, 0], g['psino'][:, 1])


def test_mmrhist_file(torch, params, stream, tmp_path):
    import nipet_amd as nip
    from oracle import oracle as O
    f = os.path.join(tmp_path, 'lm.bf')
    stream.tofile(f)
    toff, nitag = _toff_nitag(stream)
    ele, tt, _ = nip.lminfo(f)
    assert ele == len(stream) and tt[0] == toff
    hst = nip.mmrhist({'lm_bf': f}, params, t0=1, t1=5)
    o = O.lm_hist(stream, params, toff, nitag, tstart=1, tstop=5)
    assert hst['psino'].shape == (837, 252, 344) and hst['psino'].dtype == np.uint16
    assert np.array_equal(hst['psino'].ravel(), (o['psino'][0] & 0xffff).astype(np.uint16))
    assert np.array_equal(hst['dsino'].ravel(), (o['dsino'][0] & 0xffff).astype(np.uint16))
    assert np.array_equal(hst['phc'], o['phc'][0]) and hst['dur'] == 4
    assert hst['fansums'].shape == (64, 504) and hst['pssr'].shape == (127, 252, 344)


def test_mmrhist_sinogram_cache(torch, params, stream, tmp_path):
    """store=True writes the raw histogram of the frame as
    sino/sinos_s11_frm-t0-t1.npz under outpath; use_stored=True with
    datain['sinos'] naming that file returns the same outputs without the
    list-mode file, and a cache file of another frame is not used
    (lm/mmrhist.py:65-69, 116-125)."""
    import nipet_amd as nip
    f = os.path.join(tmp_path, 'lm.bf')
    stream.tofile(f)
    out = os.path.join(tmp_path, 'out')
    hst = nip.mmrhist({'lm_bf': f}, params, t0=1, t1=5, store=True, outpath=out)
    cache = os.path.join(out, 'sino', 'sinos_s11_frm-1-5.npz')
    assert os.path.isfile(cache)
    os.remove(f)
    again = nip.mmrhist({'sinos': cache, 'lm_bf': f}, params, t0=1, t1=5, use_stored=True)
    assert again.keys() == hst.keys()
    for k, v in hst.items():
        if isinstance(v, np.ndarray):
            # cmass is NaN wherever a second holds no prompt (zR / zM = 0 / 0, as
            # the reference computes it)
            assert v.dtype == again[k].dtype, k
            assert np.array_equal(v, again[k], equal_nan=v.dtype.kind == 'f'), k
        else:
            assert v == again[k], k
    # the cached frame is 1-5 s: another frame falls back to the (now missing) file
    with pytest.raises(ValueError):
        nip.mmrhist({'sinos': cache, 'lm_bf': f}, params, t0=0, t1=5, use_stored=True)


def test_time_index_on_device(torch, stream):
    """First / last time tag and the first tag word of every second, found on
    the device, against a NumPy scan of the same words; a stream without tags
    and seconds past the last tag give the word count."""
    from nipet_amd.lm.mmrhist import second_index, tag_range
    w = torch.from_numpy(stream).cuda()
    t = np.nonzero((stream >> 29) == -4)[0]
    tf, tl = int(stream[t[0]] & 0x1fffffff), int(stream[t[-1]] & 0x1fffffff)
    assert tag_range(w) == (tf, tl)
    nsec = ((tl - tf) + 999) // 1000
    got = second_index(w, tf, nsec + 2)
    ms = (stream[t] & 0x1fffffff).astype(np.int64) - tf
    want = [int(t[np.argmax(ms >= 1000 * s)]) if np.any(ms >= 1000 * s) else len(stream)
            for s in range(nsec + 3)]
    assert got.tolist() == want
    assert got[-1] == len(stream)
    # a slice without tags
    events = torch.from_numpy(stream[1:200].copy()).cuda()
    assert ((stream[1:200] >> 29) == -4).sum() == 0
    with pytest.raises(ValueError):
        tag_range(events)
    assert second_index(events, tf, 3).tolist() == [199] * 4


def test_tag_span_long_streams(torch):
    """First / last tag of streams longer than the two end windows (2^20 words
    each): tags in both ends, only in the middle, only at one end, and a tag
    in the middle beside end tags (never chosen)."""
    from nipet_amd.lm.mmrhist import tag_range
    n = 3_500_000
    cases = [[5, 1_700_000, n - 3], [1_500_000, 2_000_000], [17, 900_000, 1_600_000],
             [1_600_000, n - 1], [0, n - 1]]
    for pos in cases:
        w = np.full(n, (1 << 30) | 12345, np.int32)
        for k, i in enumerate(pos):
            w[i] = np.int32(-(1 << 31) | (1000 + 7 * k))
        got = tag_range(torch.from_numpy(w).cuda())
        assert got == (1000, 1000 + 7 * (len(pos) - 1)), (pos, got)


def test_rep0_selects_replicate(torch, params, stream):
    """A session with nrep = 1, rep0 = 2 histograms exactly replicate 2 of an
    nrep = 3 session (oracle), so replicate sets can be split across sessions."""
    from oracle import oracle as O
    toff, nitag = _toff_nitag(stream)
    w = stream[:900_000]
    _, g = _gpu(torch, params, w, toff, nitag, btp=2, seed=2024, nrep=1, rep0=2)
    o = O.lm_hist(w, params, toff, nitag, btp=2, seed=2024, nrep=3)
    assert np.array_equal(g['psino'][:, 0].astype(np.uint32), o['psino'][2])
    assert np.array_equal(g['phc'][:, 0].astype(np.uint32), o['phc'][2])


def test_remgaps_u16b(torch, params, stream):
    """Replicate-innermost histogram -> batched uint16 prompts (naw, 837, nb)
    for replicates r0 .. r0+nb-1: equal to the NumPy gather by aw2ali."""
    import ctypes as ct

    from nipet_amd import _lib
    toff, nitag = _toff_nitag(stream)
    h, g = _gpu(torch, params, stream[:900_000], toff, nitag, btp=2, seed=7, nrep=5)
    ali = np.asarray(params['txLUT']['aw2ali'], np.int32)
    naw = len(ali)
    out = torch.empty((naw, 837, 3), dtype=torch.uint16, device='cuda')
    ali_d = torch.from_numpy(ali).cuda()
    _lib.check(_lib.lib().nipet_remgaps_u16b(out.data_ptr(), h.out['psino'].data_ptr(),
                                             ali_d.data_ptr(), naw, 837, 86688, 5, 1, 3, 3, 0,
                                             _lib.stream_handle()), 'remgaps_u16b')
    torch.cuda.synchronize()
    ref = g['psino'].reshape(837, 86688, 5)[:, ali, 1:4].transpose(1, 0, 2)
    assert np.array_equal(out.cpu().numpy().astype(np.int64), np.minimum(ref, 65535))
    # replicate 4 alone into column 2 of a 3-wide output, columns 0-1 untouched
    out.fill_(7)
    _lib.check(_lib.lib().nipet_remgaps_u16b(out.data_ptr(), h.out['psino'].data_ptr(),
                                             ali_d.data_ptr(), naw, 837, 86688, 5, 4, 1, 3, 2,
                                             _lib.stream_handle()), 'remgaps_u16b')
    got = out.cpu().numpy().astype(np.int64)
    assert np.array_equal(got[..., 2], np.minimum(g['psino'].reshape(837, 86688, 5)[:, ali, 4].T,
                                                  65535))
    assert np.all(got[..., :2] == 7)
    bad = _lib.lib().nipet_remgaps_u16b(out.data_ptr(), h.out['psino'].data_ptr(),
                                        ali_d.data_ptr(), naw, 837, 86688, 5, 3, 3, 3, 0, None)
    assert bad != 0
    bad = _lib.lib().nipet_remgaps_u16b(out.data_ptr(), h.out['psino'].data_ptr(),
                                        ali_d.data_ptr(), naw, 837, 86688, 5, 0, 2, 3, 2, None)
    assert bad != 0


@pytest.mark.parametrize('nrep', [1, 4])
def test_hist_nonparametric_bootstrap(torch, params, stream, nrep):
    """btp = 1 (events redrawn from their 8192-word patch) bit-exact against
    the oracle, in one chunk and in chunks that start on patch boundaries."""
    from oracle import oracle as O
    toff, nitag = _toff_nitag(stream)
    w = stream[:1_000_000]
    o = O.lm_hist(w, params, toff, nitag, btp=1, seed=31, nrep=nrep)
    for cuts in ((len,), (8192 * 3, 8192 * 40, len)):
        print(f'btp=1 nrep={nrep} chunks={cuts}', flush=True)
        h, g = _gpu(torch, params, w, toff, nitag, chunks=cuts, btp=1, seed=31, nrep=nrep)
        _compare(g, o, nrep)
        assert g['phc'][:, 0].sum() + g['dhc'][:, 0].sum() == o['counted']


def test_hist_nonparametric_chunk_alignment(torch, params, stream):
    """A btp = 1 chunk that does not start on a patch boundary is refused
    with a status code (the patches are blocks of the session's stream)."""
    from nipet_amd import _lib
    from nipet_amd.lm.mmrhist import LmHistogram
    toff, nitag = _toff_nitag(stream)
    h = LmHistogram(params, toff, nitag, btp=1, seed=1)
    wd = torch.from_numpy(stream[:50_000]).cuda()
    h.add(wd[:8192 * 2])
    with pytest.raises(_lib.NipetError, match='multiple of 8192'):
        h.add(wd[:1000])
        h.add(wd[1000:2000])


@pytest.mark.parametrize('btp', [2, 1])
def test_hist_64_replicates_vs_split_sessions(torch, params, stream, btp):
    """An nrep = 64 session (the bootstrap batch's histogram) equals 64
    sessions with nrep = 1 and rep0 = r, replicate by replicate, on every
    per-replicate output (bit-exact; the single sessions are pinned against
    the oracle above)."""
    from nipet_amd.lm.mmrhist import LmHistogram
    toff, nitag = _toff_nitag(stream)
    w = torch.from_numpy(stream[:400_000]).cuda()
    big = LmHistogram(params, toff, nitag, btp=btp, seed=4242, nrep=64)
    big.add(w)
    big.finish()
    for r in range(64):
        one = LmHistogram(params, toff, nitag, btp=btp, seed=4242, nrep=1, rep0=r)
        one.add(w)
        one.finish()
        for k in ('psino', 'dsino', 'phc', 'dhc'):
            assert torch.equal(big.out[k][:, r], one.out[k][:, 0]), (r, k)
        if r == 0:
            for k in ('ssr', 'zR', 'zM', 'pvs', 'fan', 'bck'):
                assert torch.equal(big.out[k], one.out[k]), k
        del one
    tot = big.out['psino'].sum(dim=0, dtype=torch.int64)
    assert int((tot != tot[0]).sum()) > 0          # replicates differ
