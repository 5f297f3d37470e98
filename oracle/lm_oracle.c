/*
 * CPU oracle of mMR list-mode histogramming with bootstrap.
 *
 * TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py's CPU
 * leg); the product library does not link it.
 *
 * Restates, sequentially over the word stream:
 *  - word decoding and routing, lm/src/hst.cu:104-239: event if word > 0
 *    (bit 30 prompt, bits 0-29 span-1 Siemens address si*86688 + a*344 + w);
 *    time tag if word>>29 == -4 (ms in bits 0-28); singles if word>>29 == -3
 *    (bucket (word & 0x1fffffff) >> 19, value word & 0x7ffff);
 *  - itag = (ms - toff) / 1000, itagu = (ms - toff) - 1000 itag
 *    (lm/src/hst.cu:88, 211-212); an event belongs to the most recent
 *    preceding time tag (events before the first tag: itag 0); events and
 *    singles count when tstart <= itag < tstop;
 *  - per counted event with weight N (1, or bootstrap): span-11 (or span-1 /
 *    SSRB) sinogram, SSRB prompt sinogram, head curves, centre-of-mass sums,
 *    projection views (a = 0 or 126, itag < 5400), delayed fan-sums
 *    (lm/src/hst.cu:141-202); singles buckets (lm/src/hst.cu:215-235);
 *  - parametric bootstrap (btp = 2, lm/src/hst.cu:127-139): N ~ Poisson(btprt),
 *    counted when 0 < N < 32. Here N is drawn by inversion of a uint32 CDF
 *    table against output word (g mod 4) of Philox4x32-10 (Salmon, Moraes,
 *    Dror, Shaw, SC'11) with counter (g / 4 lo, hi, replicate, 0) and key
 *    (seed lo, seed hi), g the global word index (one generator call serves
 *    four consecutive words), replacing the reference's clock-seeded cuRAND;
 *  - non-parametric bootstrap (btp = 1, lm/src/hst.cu:110-126): an event
 *    word is replaced by an event word drawn uniformly from its patch, the
 *    8192-word block (global word index / 8192) that holds it: with u the
 *    first output word of Philox4x32-10 (counter (word index lo, hi,
 *    replicate, 1)), the floor(u nev / 2^32)-th of the patch's nev event words
 *    in word order. The event keeps the time tag of its position; the drawn
 *    word gives the bin and the prompt bit. (The reference draws positions of
 *    its thread's patch with cuRAND until it hits an event: also uniform over
 *    the patch's events.)
 * Prompt and delayed counts are kept in separate uint32 arrays (the reference
 * packs both into one uint32, lm/src/hst.cu:173,194).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NSBINANG 86688
#define NSBINS 344
#define NSANGLES 252
#define NCRS 504
#define NBUCKTS 224
#define SEG0 127
#define MXNITAG 5400
#define VTIME 2
#define ITIME 1000
#define LM_PATCH 8192

/* Philox4x32-10 */
static void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)M0 * c[0], p1 = (uint64_t)M1 * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k0, n1 = lo1, n2 = hi0 ^ c[3] ^ k1, n3 = lo0;
    c[0] = n0;
    c[1] = n1;
    c[2] = n2;
    c[3] = n3;
    k0 += W0;
    k1 += W1;
  }
}

void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  philox(c, key[0], key[1]);
  memcpy(out, c, sizeof(c));
}

/* uint32 CDF table of Poisson(lambda): cdf[k] = floor(P(X <= k) 2^32) for
 * k < K-1, cdf[K-1] = 2^32 - 1 (saturated). N = min{k : u < cdf[k]}. */
void orc_poisson_cdf(double lambda, uint32_t *cdf, int K) {
  double p = exp(-lambda), s = 0.0;
  for (int k = 0; k < K; k++) {
    s += p;
    double v = floor(s * 4294967296.0);
    cdf[k] = (k == K - 1 || v >= 4294967295.0) ? 0xFFFFFFFFu : (uint32_t)v;
    p = p * lambda / (double)(k + 1);
  }
}

typedef struct {
  uint32_t *psino, *dsino; /* [nrep][nsino_bins] */
  uint32_t *ssr;           /* [127*86688] */
  uint32_t *phc, *dhc;     /* [nrep][nitag] */
  uint64_t *zR, *zM;       /* [nitag] */
  uint32_t *pvs;           /* [tn][127][344] */
  uint32_t *fan;           /* [64][504] */
  uint32_t *bck;           /* [2][nitag][224] */
} orc_hist_out;

/* Histogram words[0..n) (global word index = base + i). Returns the number of
 * counted events of replicate 0. */
int64_t orc_lm_hist(const int32_t *words, int64_t n, int64_t base, int toff, int nitag, int tstart,
                    int tstop, int span, const int16_t *sn1_sn11, const int16_t *sn1_ssrb,
                    const int16_t *sn1_rno, const int16_t *s2cF, int btp, const uint32_t *cdf,
                    int ncdf, uint64_t seed, int nrep, int64_t nbins, uint32_t *psino,
                    uint32_t *dsino, uint32_t *ssr, uint32_t *phc, uint32_t *dhc, uint64_t *zR,
                    uint64_t *zM, uint32_t *pvs, uint32_t *fan, uint32_t *bck) {
  int itag = 0, itagu = 0;
  int64_t counted = 0;
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  /* btp = 1: event words of the current patch, in word order */
  int32_t pev[LM_PATCH];
  int64_t patch = -1;
  uint32_t nev = 0;
  for (int64_t i = 0; i < n; i++) {
    int32_t word = words[i];
    if (word > 0) {
      if (itag < tstart || itag >= tstop) continue;
      const int64_t g = base + i;
      if (btp == 1 && g / LM_PATCH != patch) {
        patch = g / LM_PATCH;
        int64_t p0 = patch * LM_PATCH - base, p1 = p0 + LM_PATCH;
        if (p0 < 0) p0 = 0;
        if (p1 > n) p1 = n;
        nev = 0;
        for (int64_t k = p0; k < p1; k++)
          if (words[k] > 0) pev[nev++] = words[k];
      }
      for (int r = 0; r < nrep; r++) {
        uint32_t N = 1;
        int32_t wr = word;
        if (btp == 1) {
          uint32_t c[4] = {(uint32_t)g, (uint32_t)((uint64_t)g >> 32), (uint32_t)r, 1u};
          philox(c, k0, k1);
          wr = pev[((uint64_t)c[0] * nev) >> 32];
        }
        int32_t val = wr & 0x3fffffff;
        int si = val / NSBINANG;
        int aw = val - si * NSBINANG;
        int a = aw / NSBINS, w = aw - a * NSBINS;
        if (si >= 4084) continue; /* address beyond the span-1 range: skipped */
        int64_t addr = span == 11 ? (int64_t)sn1_sn11[si] * NSBINANG + aw
                                  : (span == 1 ? (int64_t)val : (int64_t)sn1_ssrb[si] * NSBINANG + aw);
        int prompt = (wr >> 30) & 1;
        int si_ssrb = sn1_ssrb[si];
        if (btp == 2) {
          uint64_t g = (uint64_t)(base + i), q = g >> 2;
          uint32_t c[4] = {(uint32_t)q, (uint32_t)(q >> 32), (uint32_t)r, 0u};
          philox(c, k0, k1);
          uint32_t u = c[g & 3];
          N = 0;
          while ((int)N < ncdf - 1 && u >= cdf[N]) N++;
        }
        if (N == 0 || N >= 32) continue;
        if (r == 0) counted++;
        if (prompt) {
          psino[(int64_t)r * nbins + addr] += N;
          phc[(int64_t)r * nitag + itag] += N;
          if (r == 0) {
            ssr[(int64_t)si_ssrb * NSBINANG + aw] += N;
            zR[itag] += (uint64_t)si_ssrb;
            zM[itag] += N;
            if ((a == 0 || a == 126) && itag < MXNITAG)
              pvs[(int64_t)(itag >> VTIME) * SEG0 * NSBINS + si_ssrb * NSBINS + w] +=
                  N << (8 * (a == 126));
          }
        } else {
          dsino[(int64_t)r * nbins + addr] += N;
          dhc[(int64_t)r * nitag + itag] += N;
          if (r == 0) {
            int c0 = s2cF[2 * (a + NSANGLES * w)], c1 = s2cF[2 * (a + NSANGLES * w) + 1];
            fan[NCRS * sn1_rno[2 * si] + c0] += N;
            fan[NCRS * sn1_rno[2 * si + 1] + c1] += N;
          }
        }
      }
    } else if ((word >> 29) == -4) {
      int ms = (word & 0x1fffffff) - toff;
      itag = ms / ITIME;
      itagu = ms - itag * ITIME;
    } else if ((word >> 29) == -3 && itag >= tstart && itag < tstop) {
      uint32_t ib = (uint32_t)((word & 0x1fffffff) >> 19);
      if (ib < NBUCKTS) {
        uint32_t v = (uint32_t)(word & 0x0007ffff);
        bck[ib + NBUCKTS * itag] += v << 3;
        uint32_t *b1 = bck + ib + NBUCKTS * itag + (int64_t)NBUCKTS * nitag;
        *b1 += (uint32_t)(v > 0) << 30;
        if (*b1 == 0)
          *b1 += (uint32_t)itagu;
        else
          *b1 += (uint32_t)itagu << 10;
      }
    }
  }
  return counted;
}
