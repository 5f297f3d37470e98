"""Pinning the list-mode oracle (oracle/lm_oracle.c): Philox4x32-10 known-answer
vectors of the published generator (Random123 kat_vectors), the Poisson CDF
table, and a hand-built word stream whose every output is known."""
import numpy as np

from oracle import oracle as O

TAG = -(1 << 31)


def test_philox_known_answers():
    kat = [([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
           ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
           ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
            [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]
    for ctr, key, exp in kat:
        assert list(O.philox4x32(ctr, key)) == exp


def test_poisson_cdf_table():
    from math import exp, factorial
    c = O.poisson_cdf(1.0, 40).astype(np.float64)
    p = np.cumsum([exp(-1) / factorial(k) for k in range(40)])
    assert np.all(np.abs(c[:-1] - np.floor(p[:-1] * 2**32)) <= 1)
    assert np.all(np.diff(c) >= 0) and c[-1] == 2**32 - 1


def _addr(si, a, w):
    return si * 86688 + a * 344 + w


def test_hand_built_stream(params):
    ax, tx = params['axLUT'], params['txLUT']
    toff = 5000
    P = 1 << 30
    sgl = lambda b, v: -(1 << 31) + (1 << 29) | (b << 19) | v  # noqa: E731
    words = np.array([
        P | _addr(7, 5, 9),                 # before the first tag -> second 0
        TAG | toff,                         # t = 0 s
        P | _addr(100, 0, 17),              # a = 0 -> projection view
        _addr(200, 10, 20),                 # delayed
        sgl(3, 0),                          # empty report, itagu 0: counter stays 0
        sgl(3, 0),
        TAG | (toff + 1250),                # t = 1.25 s
        P | _addr(100, 126, 17),            # a = 126 -> projection view << 8
        sgl(3, 5),                          # first non-empty report
        sgl(3, 0),                          # counter already non-zero -> itagu << 10
        sgl(250, 1),                        # bucket >= 224 ignored
        0,                                  # zero word: ignored
        P | 0x3fffffff,                     # address beyond span-1: skipped
        TAG | (toff + 3100),                # t = 3.1 s (outside the frame [0, 2))
        P | _addr(1, 1, 1),
        TAG | (toff + 3999),
    ], dtype=np.int64).astype(np.int32)
    nitag = 4
    h = O.lm_hist(words, params, toff, nitag, tstart=0, tstop=2)
    s11, ssrb, rno = (ax[k].astype(np.int64) for k in ('sn1_sn11', 'sn1_ssrb', 'sn1_rno'))
    ps = np.zeros(837 * 86688, np.uint32)
    for si, a, w in [(7, 5, 9), (100, 0, 17), (100, 126, 17)]:
        ps[s11[si] * 86688 + a * 344 + w] += 1
    assert np.array_equal(h['psino'][0], ps)
    ds = np.zeros_like(ps)
    ds[s11[200] * 86688 + 10 * 344 + 20] = 1
    assert np.array_equal(h['dsino'][0], ds)
    assert list(h['phc'][0]) == [2, 1, 0, 0] and list(h['dhc'][0]) == [1, 0, 0, 0]
    assert h['counted'] == 4
    assert list(h['zR']) == [ssrb[7] + ssrb[100], ssrb[100], 0, 0]
    assert list(h['zM']) == [2, 1, 0, 0]
    pv = h['pvs'].reshape(-1, 127, 344)
    assert pv[0, ssrb[100], 17] == 1 + (1 << 8) and pv.sum() == 1 + (1 << 8)
    c = tx['s2cF'][10 + 252 * 20].astype(np.int64)
    fan = h['fan'].reshape(64, 504)
    assert fan[rno[200, 0], c[0]] == 1 and fan[rno[200, 1], c[1]] == 1 and fan.sum() == 2
    ssr = h['ssr'].reshape(127, 86688)
    assert ssr[ssrb[100], 17] == 1 and ssr[ssrb[100], 126 * 344 + 17] == 1 and ssr.sum() == 3
    bck = h['bck'].reshape(2, nitag, 224)
    assert bck[0, 0, 3] == 0 and bck[1, 0, 3] == 0
    assert bck[0, 1, 3] == 5 << 3
    assert bck[1, 1, 3] == (1 << 30) + (250 << 10) + (250 << 10)
    assert bck.sum() == bck[0, 1, 3] + bck[1, 1, 3]


def test_bootstrap_weights_statistics(params):
    # Poisson(1) weights: mean 1 and variance 1 over many events
    n = 200_000
    words = np.full(n + 2, (1 << 30) | _addr(10, 3, 3), np.int32)
    words[0] = TAG | 0
    words[-1] = TAG | 999
    h = O.lm_hist(words, params, 0, 1, btp=2, btprt=1.0, seed=12345)
    tot = int(h['phc'][0, 0])
    assert abs(tot / n - 1.0) < 0.01
    z0 = n - h['counted']                   # N == 0 draws
    assert abs(z0 / n - np.exp(-1)) < 0.005


def test_bootstrap_counter_layout(params):
    """Every event's weight, per replicate, from the KAT-pinned generator:
    N = #{k < 39 : u >= cdf[k]} with u = output word (g mod 4) of
    Philox4x32-10(counter (g // 4, 0, replicate, 0), key (seed lo, seed hi)),
    g the global word index; counted when 0 < N < 32. Events sit in distinct
    bins, so every bin holds one event's weight."""
    s11 = params['axLUT']['sn1_sn11'].astype(np.int64)
    seed, nrep, base, nev = (7 << 32) | 4242, 3, 1001, 61
    addrs = [(si, (5 * si) % 252, (7 * si) % 344) for si in range(1, 4084, 4084 // nev)][:nev]
    words = np.array([TAG | 0] + [(1 << 30) | _addr(*a) for a in addrs] + [TAG | 999],
                     np.int64).astype(np.int32)
    h = O.lm_hist(words, params, 0, 1, btp=2, btprt=1.0, seed=seed, nrep=nrep, base=base)
    cdf = O.poisson_cdf(1.0, 40).astype(np.int64)
    key = [seed & 0xffffffff, seed >> 32]
    nonzero = 0
    for r in range(nrep):
        exp = np.zeros(837 * 86688, np.uint32)
        for i, (si, a, w) in enumerate(addrs):
            g = base + 1 + i
            u = int(O.philox4x32([(g >> 2) & 0xffffffff, g >> 34, r, 0], key)[g & 3])
            n = int(np.sum(u >= cdf[:39]))
            if 0 < n < 32:
                exp[s11[si] * 86688 + a * 344 + w] += n
        assert np.array_equal(h['psino'][r], exp)
        nonzero += int(np.count_nonzero(exp))
    assert 0 < nonzero < nrep * nev     # some events drawn 0, some counted


def _patch_stream(words_per_patch, fill, npatch):
    """Stream of npatch 8192-word patches: time tags at the first word of
    every patch (1 s apart), then `fill(k)` words (k = patch index)."""
    out = []
    for k in range(npatch):
        p = np.zeros(8192, np.int64)
        p[0] = TAG | (1000 * k)
        p[1:1 + words_per_patch] = fill(k)
        out.append(p)
    out.append(np.array([TAG | (1000 * npatch)]))
    return np.concatenate(out).astype(np.int32)


def test_nonparametric_bootstrap_stays_in_patch(params):
    """btp = 1 draws from the 8192-word patch of the event: patches whose
    events are all one word reproduce the plain histogram exactly (a draw
    from a neighbouring patch would move counts between bins / prompt and
    delayed), and each event keeps the time tag of its position."""
    bins = [(1 << 30) | _addr(5, 3, 7), _addr(900, 100, 200), (1 << 30) | _addr(3000, 250, 1)]
    w = _patch_stream(5000, lambda k: bins[k % 3], 6)
    plain = O.lm_hist(w, params, 0, 6)
    boot = O.lm_hist(w, params, 0, 6, btp=1, seed=3, nrep=2)
    for r in range(2):
        for k in ('psino', 'dsino', 'phc', 'dhc'):
            assert np.array_equal(boot[k][r], plain[k][0]), k
    for k in ('ssr', 'zR', 'zM', 'pvs', 'fan'):
        assert np.array_equal(boot[k], plain[k]), k


def test_nonparametric_bootstrap_draws(params):
    """Within a patch the drawn words are uniform over the patch's events
    (chi-square over 4 distinct words with weights 1:2:3:4), the counts per
    replicate equal the in-frame event positions, each second keeps its
    event count (prompts + delayeds), and replicates differ."""
    words4 = np.array([(1 << 30) | _addr(10, 0, 5), (1 << 30) | _addr(11, 1, 6),
                       _addr(12, 2, 7), (1 << 30) | _addr(13, 3, 8)], np.int64)
    rng = np.random.default_rng(0)

    def fill(k):
        return rng.permutation(np.repeat(words4, [800, 1600, 2400, 3200]))

    w = _patch_stream(8000, fill, 8)
    nev = 8 * 8000
    h = O.lm_hist(w, params, 0, 8, btp=1, seed=77, nrep=3)
    plain = O.lm_hist(w, params, 0, 8)
    s11 = params['axLUT']['sn1_sn11'].astype(np.int64)
    for r in range(3):
        got = []
        for x in words4:
            v = int(x) & 0x3fffffff
            si, aw = v // 86688, v % 86688
            arr = h['psino'] if (int(x) >> 30) & 1 else h['dsino']
            got.append(int(arr[r][s11[si] * 86688 + aw]))
        assert sum(got) == nev
        exp = nev * np.array([1, 2, 3, 4]) / 10
        chi2 = float(((np.array(got) - exp) ** 2 / exp).sum())
        assert chi2 < 16.3, (got, chi2)          # 3 dof, p = 0.001
        assert np.array_equal(h['phc'][r] + h['dhc'][r], plain['phc'][0] + plain['dhc'][0])
    assert not np.array_equal(h['psino'][0], h['psino'][1])
