"""Deterministic synthetic OSEM inputs on the mMR geometry, computed without
any projector (so that the oracle run that wrote the golden fixture and the
GPU test see bit-identical inputs on any machine).

Emission model per transaxial bin aw and span-11 sinogram s:
  trues(aw, s)  = chord(aw) * axial(s) with chord(aw) the analytic length of
                  the transaxial ray inside the water ellipse x^2/9^2 + y^2/7^2
                  <= 1 (cm) plus a hot disc of radius 2.5 cm at (3, -2),
                  axial(s) a smooth profile over the segment-wise axial position;
  attenuation   = E(0.096 * water chord * (1 + 0.02 segment index)) with
                  E(x) = 1/(1 + x + x^2/2 + x^3/6) ~ exp(-x);
  norm          = U(0.8, 1.2) (numpy default_rng(3));
  randoms       = flat, 12 % of the total trues;
  scatter       = 25 % of the trues, spread as a Lorentzian of the radial bin;
  prompts       = Poisson(scale * (attn * norm * trues) + randoms + scatter) as
                  uint16 (default_rng(4)), scaled to `counts` expected counts.
Only +, -, *, / and sqrt (all correctly rounded) are used before the Poisson
draw, so the inputs do not depend on the CPU's vector math library;
`checksum` lets a test confirm it sees the inputs the fixture was made from.
"""
import numpy as np


def _chords(params, cx, cy, ax, ay):
    tx = params['txLUT']
    crs = tx['crs'].astype(np.float64)
    c = tx['s2c'].astype(np.int64)
    p = 0.5 * (crs[c[:, 0], :2] + crs[c[:, 0], 2:])
    q = 0.5 * (crs[c[:, 1], :2] + crs[c[:, 1], 2:])
    d = q - p
    # ellipse ((x-cx)/ax)^2 + ((y-cy)/ay)^2 = 1 along p + t d, t in R
    px, py = (p[:, 0] - cx) / ax, (p[:, 1] - cy) / ay
    dx, dy = d[:, 0] / ax, d[:, 1] / ay
    A = dx * dx + dy * dy
    B = 2 * (px * dx + py * dy)
    C = px * px + py * py - 1
    disc = np.clip(B * B - 4 * A * C, 0, None)
    return np.sqrt(disc) / A * np.hypot(d[:, 0], d[:, 1])


def make_osem_data(params, counts=3e8):
    Cnt = params['Cnt']
    naw, nsn = params['txLUT']['Naw'], Cnt['NSN11']
    water = _chords(params, 0.0, 0.0, 9.0, 7.0)
    hot = _chords(params, 3.0, -2.0, 2.5, 2.5)
    chord = water + 3.0 * hot
    seg = np.repeat(np.arange(11), Cnt['SEG'])
    pos = np.concatenate([np.arange(n) - (n - 1) / 2 for n in Cnt['SEG']])
    axial = (1.0 - 0.03 * seg) / (1.0 + (pos / 40.0)**2)
    trues = chord[:, None] * axial[None, :]
    x = 0.096 * water[:, None] * (1 + 0.02 * seg[None, :])
    attn = 1.0 / (1.0 + x + x * x / 2 + x * x * x / 6)
    norm = np.random.default_rng(3).uniform(0.8, 1.2, size=(naw, nsn))
    t = attn * norm * trues
    iw = params['txLUT']['aw2sn'][:, 1].astype(np.float64)
    sprof = 1.0 / (1.0 + ((iw - 172) / 60.0)**2)
    sct = sprof[:, None] * axial[None, :]
    sct *= 0.25 * t.sum() / sct.sum()
    rnd = np.full((naw, nsn), 0.12 * t.sum() / (naw * nsn))
    mean = t + sct + rnd
    scale = counts / mean.sum()
    lam = scale * mean
    prompts = np.random.default_rng(4).poisson(lam).astype(np.uint16)
    return {'prompts': prompts, 'attn': attn.astype(np.float32), 'norm': norm.astype(np.float32),
            'randoms': (scale * rnd).astype(np.float32),
            'scatter': (scale * sct).astype(np.float32)}


def checksum(prompts):
    """Order-sensitive integer checksum of the prompts array."""
    p = np.asarray(prompts, dtype=np.uint64).ravel()
    w = (np.arange(p.size, dtype=np.uint64) % np.uint64(1000003)) + np.uint64(1)
    return int(p.sum()), int((p * w).sum() % np.uint64(2**61 - 1))
