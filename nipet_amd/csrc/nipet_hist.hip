// List-mode histogramming of Siemens mMR 32-bit words with bootstrap on gfx950.
//
// Reference: hst kernel lm/src/hst.cu:37-244, gpu_hst :347-528, lmproc
// lm/src/lmproc.cu:12-230. Semantics (restated):
//  - word > 0: event, bit 30 prompt, bits 0-29 span-1 address si*86688 + a*344 + w;
//    word >> 29 == -4: time tag (ms in bits 0-28); word >> 29 == -3: singles;
//  - itag = (ms - toff) / 1000 of the most recent preceding time tag (0 before
//    the first tag); events / singles count for tstart <= itag < tstop;
//  - parametric bootstrap (btp = 2): weight N ~ Poisson via inversion of a
//    uint32 CDF against Philox4x32-10(counter = (word index lo, hi,
//    replicate, 0), key = (seed lo, hi)) output word 0; counted when
//    0 < N < 32;
//  - non-parametric bootstrap (btp = 1, lm/src/hst.cu:110-126): every event
//    word is replaced by an event word drawn uniformly from its patch, the
//    block of LM_BLOCK words (global word index / LM_BLOCK) that holds it:
//    draw j = floor(u nev / 2^32) with u = word 0 of Philox4x32-10(counter =
//    (word index lo, hi, replicate, 1)), nev the patch's events, and take the
//    j-th event of the patch in word order. The event keeps the time tag of
//    its own position (as the reference); the drawn word gives the bin and
//    the prompt / delayed bit.
// Layout: with nrep replicates, sinograms and head curves are replicate-
// innermost ([bin][nrep], [itag][nrep]): one event updates nrep consecutive
// counters with one coalesced wave atomic. Integer atomics make every count
// independent of the processing order (bit-exact against a sequential pass).
//
// Kernels per chunk of words:
//  k_lm_tags   last time tag of every block of LM_BLOCK words;
//  k_lm_carry  parallel "last tag before block b" scan (carry across chunks);
//  k_lm_hist   events (head curves and centre of mass privatised in LDS over a
//              window of LM_HW seconds), singles appended to a list;
//  k_lm_sgl_first / k_lm_sgl_apply  singles buckets in two order-free passes.
#include <stdlib.h>

#include <cmath>
#include <vector>

#include "nipet_common.h"

namespace nipet {

constexpr int LM_THREADS = 256;
constexpr int LM_PER_LANE = 32;
constexpr int LM_BLOCK = LM_THREADS * LM_PER_LANE;  // words per block
constexpr int LM_WAVES = LM_THREADS / 64;
constexpr int LM_HW = 16;                           // LDS head-curve window [s]
constexpr int LM_MAXREP = 64;
constexpr int LM_NCDF = 40;
constexpr int NSBINANG = 86688, NSBINS = 344, NSANGLES = 252, NCRS = 504, NBUCKTS = 224;
constexpr int NSN1 = 4084, SEG0 = 127, MXNITAG = 5400, VTIME = 2, ITIME = 1000;
constexpr int64_t NSN1_BINS_MAX = (int64_t)NSN1 * NSBINANG;
// LDS-privatised sinogram counts (one replicate): bins are counted per bucket
// of LM_BBINS consecutive bins, whose prompt and delayed counters fit in LDS
constexpr int LM_BSHIFT = 14;
constexpr int LM_BBINS = 1 << LM_BSHIFT;
// Events reach their bucket in two partition passes whose writes are runs:
// each block routes its events to coarse buckets of 2^LM_CSHIFT bins (about a
// hundred events per coarse bucket and block), then tiles of LM_TILE staged
// events are sorted in LDS by bucket. LM_MAXCB covers span 1.
constexpr int LM_CSHIFT = 20;
constexpr int LM_FPC = 1 << (LM_CSHIFT - LM_BSHIFT);  // buckets per coarse bucket
constexpr int LM_MAXCB = (NSN1_BINS_MAX + (1 << LM_CSHIFT) - 1) >> LM_CSHIFT;
constexpr int LM_TILE = 8192;
static_assert(LM_MAXCB < 0xffff, "k_lm_hist keeps a coarse bucket in 16 bits");
static_assert(((NSN1_BINS_MAX + LM_BBINS - 1) >> LM_BSHIFT) < 65536 && LM_TILE <= 65536,
              "k_lm_bin_part packs a bucket index and a rank into 16 bits each");
constexpr int32_t TMSK = 0x1fffffff;
constexpr int64_t LM_SPAN_END = 1 << 20;  // words searched at each end for the first / last tag

__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)M0 * c[0], p1 = (uint64_t)M1 * c[2];
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
    k0 += W0;
    k1 += W1;
  }
}

struct LmDev {
  const int32_t *words;
  int64_t n, base;
  int toff, nitag, tstart, tstop, span, btp, nrep, rep0;
  uint32_t k0, k1;
  int64_t nbins;
  const int32_t *lut;   // [NSN1] (sn1_sn11 << 16) | sn1_ssrb
  const short2 *rno;    // [NSN1] sn1_rno
  const short2 *s2cF;   // [86688] crystal pair of (a + 252 w)
  const uint32_t *cdf;  // [LM_NCDF]
  const int32_t *blk_carry;
  uint32_t *psino, *dsino, *ssr, *phc, *dhc, *pvs, *fan;
  unsigned long long *zR, *zM;
  unsigned long long *sgl_idx;
  int32_t *sgl_word, *sgl_ms;
  unsigned int *sgl_count;
  unsigned int sgl_cap;
  unsigned int *bad;
  // LDS-privatised sinogram counts: events staged per coarse bucket
  uint32_t *ccur;   // [ncb] next free slot of each coarse bucket
  uint32_t *stage;  // (bin within coarse bucket) | prompt << LM_CSHIFT | N << (LM_CSHIFT + 1)
  int64_t stage_cap;
  int ncb;
};

__device__ inline bool is_tag(int32_t w) { return (w >> 29) == -4; }

// Output bin of an event word's span-1 address (span 11, 1 or SSRB).
__device__ inline int64_t lm_bin(int span, int32_t val, const int32_t *lut) {
  const int si = val / NSBINANG, aw = val - si * NSBINANG;
  const int32_t lu = lut[si];
  return span == 11 ? (int64_t)(lu >> 16) * NSBINANG + aw
                    : (span == 1 ? (int64_t)val : (int64_t)(lu & 0xffff) * NSBINANG + aw);
}

// The LM_PER_LANE words of one lane of a block's wave (word w0 + 64 j + lane;
// 0 past the end); returns the last time tag (ms) among the wave's words, -1
// if none.
__device__ inline int32_t lm_load(const LmDev &p, int64_t w0, int lane,
                                  int32_t (&wd)[LM_PER_LANE]) {
#pragma unroll
  for (int j = 0; j < LM_PER_LANE; j++) {
    const int64_t i = w0 + j * 64 + lane;
    wd[j] = i < p.n ? p.words[i] : 0;
  }
  int32_t wl = -1;
#pragma unroll
  for (int j = 0; j < LM_PER_LANE; j++) {
    const unsigned long long m = __ballot(is_tag(wd[j]));
    if (m) wl = __builtin_amdgcn_readlane(wd[j] & TMSK, 63 - __clzll(m));
  }
  return wl;
}

// Time tag (ms) in effect at this lane's word of a 64-word group: the most
// recent tag at or before it, c (wave-uniform) when the group holds none
// before it; c then moves to the group's last tag. le = lanes 0 .. lane.
__device__ inline int32_t lm_tag_at(int32_t word, int32_t &c, unsigned long long le) {
  int32_t ms = c;
  const unsigned long long m = __ballot(is_tag(word));
  if (m) {
    const unsigned long long ml = m & le;
    const int32_t tv = word & TMSK;
    const int src = ml ? 63 - __clzll(ml) : (int)__lane_id();
    const int32_t got = __shfl(tv, src);
    if (ml) ms = got;
    c = __builtin_amdgcn_readlane(tv, 63 - __clzll(m));
  }
  return ms;
}

// last time tag (ms) of every block, -1 if none; also clears the chunk's
// singles-list counter (a kernel write rather than a memset, so that a
// captured graph replays it)
__global__ __launch_bounds__(LM_THREADS) void k_lm_tags(const int32_t *words, int64_t n,
                                                         int32_t *blk_last,
                                                         unsigned int *sgl_count) {
  __shared__ long long best[LM_THREADS / 64];
  if (blockIdx.x == 0 && threadIdx.x == 0) *sgl_count = 0u;
  const int64_t b0 = (int64_t)blockIdx.x * LM_BLOCK;
  long long key = -1;  // (index << 32) | ms of the last tag seen by this lane
  for (int j = 0; j < LM_PER_LANE; j++) {
    int64_t i = b0 + (int64_t)j * LM_THREADS + threadIdx.x;
    if (i < n) {
      int32_t w = words[i];
      if (is_tag(w)) key = ((long long)(i - b0) << 32) | (long long)(w & TMSK);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    long long v = __shfl_xor(key, o);
    key = v > key ? v : key;
  }
  if ((threadIdx.x & 63) == 0) best[threadIdx.x >> 6] = key;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long k = -1;
    for (int i = 0; i < LM_THREADS / 64; i++) k = best[i] > k ? best[i] : k;
    blk_last[blockIdx.x] = k < 0 ? -1 : (int32_t)(k & 0xffffffffll);
  }
}

// blk_carry[b] = last tag before block b (carry[0] from the previous chunk);
// carry[0] <- last tag of the chunk. One block of 1024 threads.
__global__ __launch_bounds__(1024) void k_lm_carry(const int32_t *blk_last, int32_t *blk_carry,
                                                   int nb, int32_t *carry) {
  __shared__ int32_t part[1024];
  const int t = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int b0 = t * per, b1 = min(nb, b0 + per);
  int32_t last = -1;
  for (int b = b0; b < b1; b++)
    if (blk_last[b] >= 0) last = blk_last[b];
  part[t] = last;
  __syncthreads();
  // inclusive "last valid" scan over the 1024 partials (Hillis-Steele)
  for (int o = 1; o < 1024; o <<= 1) {
    int32_t v = t >= o ? part[t - o] : -1;
    __syncthreads();
    if (part[t] < 0 && v >= 0) part[t] = v;
    __syncthreads();
  }
  int32_t c = t == 0 ? carry[0] : (part[t - 1] >= 0 ? part[t - 1] : carry[0]);
  for (int b = b0; b < b1; b++) {
    blk_carry[b] = c;
    if (blk_last[b] >= 0) c = blk_last[b];
  }
  __syncthreads();
  if (t == 1023 && part[1023] >= 0) carry[0] = part[1023];
}

// N = #{k < LM_NCDF - 1 : u >= cdf[k]}: the inversion count of a
// non-decreasing uint32 CDF. Counted without branches over the first
// LM_FASTK entries (held in registers, c8); the remaining entries are read only
// when some lane of the wave passes all of them (for a mean of 1, a draw
// reaches N >= 8 with probability 1e-5). A per-lane while loop here compiled
// into divergent exec-mask bookkeeping that saturated the scalar unit.
constexpr int LM_FASTK = 8;
__device__ inline uint32_t poisson_n(uint32_t u, const uint32_t (&c8)[LM_FASTK],
                                     const uint32_t *cdf) {
  uint32_t N = 0;
#pragma unroll
  for (int k = 0; k < LM_FASTK; k++) N += u >= c8[k] ? 1u : 0u;
  if (__any(N == LM_FASTK)) {
    for (int k = LM_FASTK; k < LM_NCDF - 1; k++) N += u >= cdf[k] ? 1u : 0u;
  }
  return N;
}

// Uniform draw of the parametric bootstrap for word g and replicate r: output
// word (g mod 4) of Philox4x32-10(counter (g / 4 lo, hi, r, 0), key (k0, k1)).
// One generator call serves four consecutive words of a replicate.
__device__ inline uint32_t boot_pick(const uint32_t (&c)[4], uint64_t g) {
  const int m = (int)(g & 3);
  return m == 0 ? c[0] : (m == 1 ? c[1] : (m == 2 ? c[2] : c[3]));
}
__device__ inline void boot_quad(uint32_t (&c)[4], uint64_t g, uint32_t r, uint32_t k0,
                                 uint32_t k1) {
  const uint64_t q = g >> 2;
  c[0] = (uint32_t)q;
  c[1] = (uint32_t)(q >> 32);
  c[2] = r;
  c[3] = 0u;
  philox4x32_10(c, k0, k1);
}

// Head curves and centre-of-mass sums, one replicate: each lane keeps running
// sums for one second (its words are 64 apart, so a wave's 2,048 words rarely
// span two seconds) and adds them when its second changes and after the walk.
// Per-lane atomics on one LDS word serialise 64 deep, and a butterfly per word
// row waited on the LDS crossbar; this adds once per lane and second.
struct LmHead {
  int k = INT32_MIN;  // second - window base of the sums (none yet)
  uint32_t p = 0, d = 0;
  unsigned long long z = 0;
};

__device__ inline void lm_head_flush(LmHead &h, int itag0, uint32_t *hcp, uint32_t *hcd,
                                     unsigned long long *mzr, unsigned long long *mzm,
                                     const LmDev &p) {
  if (h.p | h.d) {
    const int it = itag0 + h.k;
    if (h.k >= 0 && h.k < LM_HW) {
      if (h.p) {
        atomicAdd(&hcp[h.k], h.p);
        atomicAdd(&mzr[h.k], h.z);
        atomicAdd(&mzm[h.k], (unsigned long long)h.p);
      }
      if (h.d) atomicAdd(&hcd[h.k], h.d);
    } else {
      if (h.p) {
        atomicAdd(p.phc + it, h.p);
        atomicAdd(p.zR + it, h.z);
        atomicAdd(p.zM + it, (unsigned long long)h.p);
      }
      if (h.d) atomicAdd(p.dhc + it, h.d);
    }
  }
  h.p = h.d = 0;
  h.z = 0;
}

// one counted event: prompt weight hp with SSRB plane hz, or delayed weight hd
__device__ inline void lm_head_add(LmHead &h, uint32_t hp, uint32_t hz, uint32_t hd, int k,
                                   int itag0, uint32_t *hcp, uint32_t *hcd,
                                   unsigned long long *mzr, unsigned long long *mzm,
                                   const LmDev &p) {
  if (k != h.k) {
    lm_head_flush(h, itag0, hcp, hcd, mzr, mzm, p);
    h.k = k;
  }
  h.p += hp;
  h.d += hd;
  h.z += hp ? hz : 0u;
}

// REP = false: one event per lane (nrep == 1).
// REP = true: the wave walks its events one by one, lane r = replicate r.
// NP: non-parametric bootstrap (btp = 1): the block's event words are
// compacted in word order into LDS (the resampling patch) and every counted
// event is replaced by a draw from them.
// BIN (one replicate, no resampling): the prompt / delayed sinogram counts are
// not added here. The block counts its events per coarse bucket, reserves one
// run of slots per coarse bucket, sorts its counted events into those runs in
// LDS and writes the runs out with consecutive lanes on consecutive slots (the
// coarse slot ranges were counted by k_lm_bin_count); k_lm_bin_part and
// k_lm_bin_apply then count them in LDS.
template <bool REP, bool NP, bool BIN = false>
__global__ __launch_bounds__(LM_THREADS) void k_lm_hist(LmDev p) {
  static_assert(!BIN || (!REP && !NP), "binned counts serve one replicate without resampling");
  __shared__ int32_t lut[NSN1];
  __shared__ int32_t evw[NP ? LM_BLOCK : 1];
  __shared__ int32_t wave_nev[LM_WAVES];
  __shared__ uint32_t cdf[LM_NCDF];
  __shared__ uint32_t hcp[LM_HW * (REP ? LM_MAXREP : 1)];
  __shared__ uint32_t hcd[LM_HW * (REP ? LM_MAXREP : 1)];
  __shared__ unsigned long long mzr[LM_HW], mzm[LM_HW];
  __shared__ int32_t wave_last[LM_WAVES];
  __shared__ uint32_t crank[BIN ? LM_MAXCB : 1], cbase[BIN ? LM_MAXCB : 1];
  __shared__ uint32_t stg[BIN ? LM_BLOCK : 1];  // the block's staged events, by coarse bucket
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nr = REP ? p.nrep : 1;
  for (int i = tid; i < NSN1; i += LM_THREADS) lut[i] = p.lut[i];
  if constexpr (BIN) {
    for (int i = tid; i < p.ncb; i += LM_THREADS) crank[i] = 0;
  }
  if (tid < LM_NCDF) cdf[tid] = p.cdf[tid];
  for (int i = tid; i < LM_HW * nr; i += LM_THREADS) {
    hcp[i] = 0;
    hcd[i] = 0;
  }
  if (tid < LM_HW) {
    mzr[tid] = 0;
    mzm[tid] = 0;
  }
  const int64_t b0 = (int64_t)blockIdx.x * LM_BLOCK;
  const int64_t w0 = b0 + (int64_t)wave * (LM_PER_LANE * 64);
  const int32_t carry_blk = p.blk_carry[blockIdx.x];
  const int itag0 = carry_blk < 0 ? 0 : (carry_blk - p.toff) / ITIME;  // window base

  int32_t wd[LM_PER_LANE];
  // pass 1: last tag of each wave's words
  const int32_t wl = lm_load(p, w0, lane, wd);
  if (lane == 0) wave_last[wave] = wl;
  int nev = 0;  // events of the block (NP)
  if constexpr (NP) {
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < LM_PER_LANE; j++) cnt += __popcll(__ballot(wd[j] > 0));
    if (lane == 0) wave_nev[wave] = cnt;
  }
  __syncthreads();
  uint32_t c8[LM_FASTK];  // the first CDF entries, wave-uniform
#pragma unroll
  for (int k = 0; k < LM_FASTK; k++) c8[k] = __builtin_amdgcn_readfirstlane(cdf[k]);
  int32_t c = carry_blk;
  for (int k = 0; k < wave; k++)
    if (wave_last[k] >= 0) c = wave_last[k];
  if constexpr (NP) {
    // compact the event words in word order (wave-major, then j, then lane)
    int run = 0;
    for (int k = 0; k < LM_WAVES; k++) {
      if (k < wave) run += wave_nev[k];
      nev += wave_nev[k];
    }
    const unsigned long long lt = (1ull << lane) - 1;
#pragma unroll
    for (int j = 0; j < LM_PER_LANE; j++) {
      unsigned long long m = __ballot(wd[j] > 0);
      if (wd[j] > 0) evw[run + __popcll(m & lt)] = wd[j];
      run += __popcll(m);
    }
    __syncthreads();
  }

  const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1);
  // REP: the last generator call of the wave (its word quad, the 4 draws)
  uint64_t bq = ~0ull;
  uint32_t bc[4] = {0u, 0u, 0u, 0u};
  LmHead hs;  // one replicate: this lane's head-curve sums
  // BIN: each word's staged value and coarse bucket (16 bits each, two per
  // register; 0xffff: not counted), kept until the block's runs are laid out
  uint32_t bpay[BIN ? LM_PER_LANE : 1], bcb[BIN ? LM_PER_LANE / 2 : 1];
  if constexpr (BIN) {
#pragma unroll
    for (int j = 0; j < LM_PER_LANE / 2; j++) bcb[j] = ~0u;
  }
  for (int j = 0; j < LM_PER_LANE; j++) {
    const int32_t word = wd[j];
    const int64_t gi = w0 + j * 64 + lane;
    const int32_t ms = lm_tag_at(word, c, le);
    const int rel = ms < 0 ? 0 : ms - p.toff;
    const int itag = rel / ITIME;
    const bool in_frame = itag >= p.tstart && itag < p.tstop;
    bool ev = word > 0 && in_frame;
    if (!NP && ev && (word & 0x3fffffff) >= NSN1 * NSBINANG) {
      atomicAdd(p.bad, 1u);
      ev = false;
    }
    // singles reports: appended for the order-free bucket passes
    if ((word >> 29) == -3 && in_frame) {
      unsigned int k = atomicAdd(p.sgl_count, 1u);
      if (k < p.sgl_cap) {
        p.sgl_idx[k] = (unsigned long long)(p.base + gi);
        p.sgl_word[k] = word;
        p.sgl_ms[k] = rel;
      }
    }
    if (!REP) {
      // this lane's head-curve adds: counted prompt weight and its SSRB plane,
      // counted delayed weight
      uint32_t hp = 0, hz = 0, hd = 0;
      if (ev) {
        uint32_t N = 1;
        int32_t wv = word;
        bool drawn = true;
        if constexpr (NP) {
          uint64_t g = (uint64_t)(p.base + gi);
          uint32_t cc[4] = {(uint32_t)g, (uint32_t)(g >> 32), (uint32_t)p.rep0, 1u};
          philox4x32_10(cc, p.k0, p.k1);
          wv = evw[(int)(((uint64_t)cc[0] * (uint32_t)nev) >> 32)];
          if ((wv & 0x3fffffff) >= NSN1 * NSBINANG) {
            atomicAdd(p.bad, 1u);
            drawn = false;
          }
        } else if (p.btp == 2) {
          const uint64_t g = (uint64_t)(p.base + gi);
          uint32_t cc[4];
          boot_quad(cc, g, (uint32_t)p.rep0, p.k0, p.k1);
          N = poisson_n(boot_pick(cc, g), c8, cdf);
        }
        const bool okN = N > 0 && N < 32;
        if (drawn && (BIN || okN)) {
          const int32_t val = wv & 0x3fffffff;
          const int si = val / NSBINANG, aw = val - si * NSBINANG;
          const int a = aw / NSBINS, w = aw - a * NSBINS;
          const int32_t lu = lut[si];
          const int s11 = lu >> 16, sr = lu & 0xffff;
          const int64_t addr = p.span == 11 ? (int64_t)s11 * NSBINANG + aw
                                            : (p.span == 1 ? (int64_t)val : (int64_t)sr * NSBINANG + aw);
          if constexpr (BIN) {
            // every counted event takes a slot of its coarse run (weight 0
            // when not counted): counted here, placed after the walk
            const uint32_t cb = (uint32_t)(addr >> LM_CSHIFT);
            atomicAdd(&crank[cb], 1u);
            const int sh = 16 * (j & 1);
            bcb[j >> 1] = (bcb[j >> 1] & ~(0xffffu << sh)) | (cb << sh);
            bpay[j] = (uint32_t)(addr & ((1 << LM_CSHIFT) - 1)) |
                      ((uint32_t)((wv >> 30) & 1) << LM_CSHIFT) | ((okN ? N : 0u) << (LM_CSHIFT + 1));
          }
          if (okN) {
            if ((wv >> 30) & 1) {
              if constexpr (!BIN) atomicAdd(p.psino + addr, N);
              hp = N;
              hz = (uint32_t)sr;
              if ((a == 0 || a == 126) && itag < MXNITAG)
                atomicAdd(p.pvs + (int64_t)(itag >> VTIME) * SEG0 * NSBINS + sr * NSBINS + w,
                          N << (8 * (a == 126)));
            } else {
              if constexpr (!BIN) atomicAdd(p.dsino + addr, N);
              hd = N;
              if constexpr (!BIN) {  // binned sessions: k_lm_fan
                short2 rr = p.rno[si], cr = p.s2cF[a + NSANGLES * w];
                atomicAdd(p.fan + NCRS * rr.x + cr.x, N);
                atomicAdd(p.fan + NCRS * rr.y + cr.y, N);
              }
            }
          }
        }
      }
      if (hp | hd) lm_head_add(hs, hp, hz, hd, itag - itag0, itag0, hcp, hcd, mzr, mzm, p);
    } else if constexpr (NP) {
      // lane r draws replicate r's word: bins differ between lanes
      unsigned long long evm = __ballot(ev);
      while (evm) {
        const int l = __ffsll((long long)evm) - 1;
        evm &= evm - 1;
        const int it = __builtin_amdgcn_readlane(itag, l);
        const uint64_t g = (uint64_t)(p.base + w0 + j * 64 + l);
        bool ok = false;
        int32_t wv = 0;
        if (lane < nr) {
          uint32_t cc[4] = {(uint32_t)g, (uint32_t)(g >> 32), (uint32_t)(p.rep0 + lane), 1u};
          philox4x32_10(cc, p.k0, p.k1);
          wv = evw[(int)(((uint64_t)cc[0] * (uint32_t)nev) >> 32)];
          ok = (wv & 0x3fffffff) < NSN1 * NSBINANG;
          if (!ok && lane == 0) atomicAdd(p.bad, 1u);
        }
        if (ok) {
          const int32_t v = wv & 0x3fffffff;
          const int si = v / NSBINANG, aw = v - si * NSBINANG;
          const int a = aw / NSBINS, w = aw - a * NSBINS;
          const int32_t lu = lut[si];
          const int s11 = lu >> 16, sr = lu & 0xffff;
          const int64_t addr = p.span == 11 ? (int64_t)s11 * NSBINANG + aw
                                            : (p.span == 1 ? (int64_t)v : (int64_t)sr * NSBINANG + aw);
          const int k = it - itag0;
          const bool inwin = k >= 0 && k < LM_HW;
          const bool prompt = (wv >> 30) & 1;
          atomicAdd((prompt ? p.psino : p.dsino) + addr * nr + lane, 1u);
          uint32_t *hl = prompt ? hcp : hcd;
          uint32_t *hg = prompt ? p.phc : p.dhc;
          if (inwin) atomicAdd(&hl[k * nr + lane], 1u); else atomicAdd(hg + (int64_t)it * nr + lane, 1u);
          if (lane == 0) {
            if (prompt) {
              if (inwin) {
                atomicAdd(&mzr[k], (unsigned long long)sr);
                atomicAdd(&mzm[k], 1ull);
              } else {
                atomicAdd(p.zR + it, (unsigned long long)sr);
                atomicAdd(p.zM + it, 1ull);
              }
              if ((a == 0 || a == 126) && it < MXNITAG)
                atomicAdd(p.pvs + (int64_t)(it >> VTIME) * SEG0 * NSBINS + sr * NSBINS + w,
                          1u << (8 * (a == 126)));
            } else {
              short2 rr = p.rno[si], cr = p.s2cF[a + NSANGLES * w];
              atomicAdd(p.fan + NCRS * rr.x + cr.x, 1u);
              atomicAdd(p.fan + NCRS * rr.y + cr.y, 1u);
            }
          }
        }
      }
    } else {
      // Each lane decodes its own word once (VALU); the loop over the wave's
      // events then broadcasts the decoded fields with readlane, so the
      // scalar unit only walks the event mask.
      int32_t e_addr = 0, e_x1 = 0, e_x2 = -1;
      if (ev) {
        const int32_t v = word & 0x3fffffff;
        const int si = v / NSBINANG, aw = v - si * NSBINANG;
        const int a = aw / NSBINS, w = aw - a * NSBINS;
        const int32_t lu = lut[si];
        const int s11 = lu >> 16, sr = lu & 0xffff;
        // bin address (< 4084 * 86688 < 2^31)
        e_addr = p.span == 11 ? s11 * NSBINANG + aw : (p.span == 1 ? v : sr * NSBINANG + aw);
        if ((word >> 30) & 1) {
          // prompt: SSRB plane (centre of mass), projection-view offset with
          // bit 30 set for the second view (a = 126), -1 for none
          e_x1 = sr;
          if ((a == 0 || a == 126) && itag < MXNITAG)
            e_x2 = ((itag >> VTIME) * SEG0 * NSBINS + sr * NSBINS + w) | ((a == 126) << 30);
        } else {
          // delayed: the two fan-sum entries
          const short2 rr = p.rno[si], cr = p.s2cF[a + NSANGLES * w];
          e_x1 = NCRS * rr.x + cr.x;
          e_x2 = NCRS * rr.y + cr.y;
        }
      }
      unsigned long long evm = __ballot(ev);
      while (evm) {
        const int l = __ffsll((long long)evm) - 1;
        evm &= evm - 1;
        const int32_t wv = __builtin_amdgcn_readlane(word, l);
        const int it = __builtin_amdgcn_readlane(itag, l);
        const int32_t addr = __builtin_amdgcn_readlane(e_addr, l);
        const int32_t x1 = __builtin_amdgcn_readlane(e_x1, l);
        const int32_t x2 = __builtin_amdgcn_readlane(e_x2, l);
        const uint64_t g = (uint64_t)(p.base + w0 + j * 64 + l);
        uint32_t N = lane < nr ? 1u : 0u;
        if (p.btp == 2) {
          if ((g >> 2) != bq) {  // wave-uniform: one generator call per quad of words
            bq = g >> 2;
            boot_quad(bc, g, (uint32_t)(p.rep0 + lane), p.k0, p.k1);
          }
          const uint32_t n0 = poisson_n(boot_pick(bc, g), c8, cdf);
          N = lane < nr ? n0 : 0u;
        }
        const bool ok = N > 0 && N < 32;
        const int k = it - itag0;
        const bool inwin = k >= 0 && k < LM_HW;
        const bool prompt = (wv >> 30) & 1;
        if (ok) {
          atomicAdd((prompt ? p.psino : p.dsino) + (int64_t)addr * nr + lane, N);
          uint32_t *hl = prompt ? hcp : hcd;
          uint32_t *hg = prompt ? p.phc : p.dhc;
          if (inwin) atomicAdd(&hl[k * nr + lane], N); else atomicAdd(hg + (int64_t)it * nr + lane, N);
        }
        if (lane == 0 && ok) {
          if (prompt) {
            if (inwin) {
              atomicAdd(&mzr[k], (unsigned long long)x1);
              atomicAdd(&mzm[k], (unsigned long long)N);
            } else {
              atomicAdd(p.zR + it, (unsigned long long)x1);
              atomicAdd(p.zM + it, (unsigned long long)N);
            }
            if (x2 >= 0) atomicAdd(p.pvs + (x2 & 0x3fffffff), N << (8 * ((x2 >> 30) & 1)));
          } else {
            atomicAdd(p.fan + x1, N);
            atomicAdd(p.fan + x2, N);
          }
        }
      }
    }
  }
  if (!REP) lm_head_flush(hs, itag0, hcp, hcd, mzr, mzm, p);
  if constexpr (BIN) {
    // one run of slots per coarse bucket: crank becomes the run's first slot
    // in stg (exclusive scan by wave 0), cbase the run's global slot minus
    // that local start (one cursor atomic per coarse bucket and block)
    __syncthreads();
    if (wave == 0) {
      uint32_t run = 0;
      for (int i0 = 0; i0 < p.ncb; i0 += 64) {
        const int i = i0 + lane;
        const uint32_t nc = i < p.ncb ? crank[i] : 0u;
        uint32_t x = nc;  // inclusive scan over the wave
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t y = __shfl_up(x, o);
          if (lane >= o) x += y;
        }
        const uint32_t ex = run + x - nc;
        if (i < p.ncb) {
          cbase[i] = (nc ? atomicAdd(p.ccur + i, nc) : 0u) - ex;
          crank[i] = ex;
        }
        run += __shfl(x, 63);
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < LM_PER_LANE; j++) {
      const uint32_t cb = (bcb[j >> 1] >> (16 * (j & 1))) & 0xffffu;
      if (cb != 0xffffu) {
        const uint32_t sl = atomicAdd(&crank[cb], 1u);
        if (sl < LM_BLOCK) stg[sl] = bpay[j];  // the runs were counted for exactly these events
      }
    }
  }
  __syncthreads();
  if constexpr (BIN) {
    // crank now holds the end of every run: wave k writes runs k, k + 4, ...,
    // consecutive lanes on consecutive slots
    for (int cb = wave; cb < p.ncb; cb += LM_WAVES) {
      const uint32_t e = crank[cb], gb = cbase[cb];
      for (uint32_t sl = (cb ? crank[cb - 1] : 0u) + lane; sl < e; sl += 64) {
        const uint32_t pos = gb + sl;
        if ((int64_t)pos < p.stage_cap) p.stage[pos] = stg[sl];
      }
    }
  }
  // flush the LDS windows
  for (int i = tid; i < LM_HW * nr; i += LM_THREADS) {
    const int it = itag0 + i / nr;
    if (it >= p.nitag) continue;
    if (hcp[i]) atomicAdd(p.phc + (int64_t)it * nr + i % nr, hcp[i]);
    if (hcd[i]) atomicAdd(p.dhc + (int64_t)it * nr + i % nr, hcd[i]);
  }
  if (tid < LM_HW && itag0 + tid < p.nitag) {
    if (mzr[tid]) atomicAdd(p.zR + itag0 + tid, mzr[tid]);
    if (mzm[tid]) atomicAdd(p.zM + itag0 + tid, mzm[tid]);
  }
}

// Singles buckets (lm/src/hst.cu:215-235). The reference adds itagu to the
// second counter unshifted when that counter is still zero, shifted by 10
// otherwise; a counter is zero before report k exactly when every earlier
// report of the same (second, bucket) had value 0 and itagu 0. Pass 1 finds,
// per (second, bucket), the first report with a non-zero value or itagu.
__global__ void k_lm_sgl_first(const unsigned long long *idx, const int32_t *word,
                               const int32_t *ms, const unsigned int *count, unsigned int cap,
                               int nitag, unsigned long long *first_nz) {
  unsigned int n = min(*count, cap);
  for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int32_t w = word[i];
    unsigned int ib = (unsigned int)((w & TMSK) >> 19);
    int itag = ms[i] / ITIME, itagu = ms[i] - itag * ITIME;
    if (ib >= NBUCKTS || itag >= nitag) continue;
    if ((w & 0x7ffff) != 0 || itagu != 0) atomicMin(first_nz + (int64_t)itag * NBUCKTS + ib, idx[i]);
  }
}

__global__ void k_lm_sgl_apply(const unsigned long long *idx, const int32_t *word,
                               const int32_t *ms, const unsigned int *count, unsigned int cap,
                               int nitag, const unsigned long long *first_nz, uint32_t *bck) {
  unsigned int n = min(*count, cap);
  for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int32_t w = word[i];
    unsigned int ib = (unsigned int)((w & TMSK) >> 19);
    int itag = ms[i] / ITIME, itagu = ms[i] - itag * ITIME;
    if (ib >= NBUCKTS || itag >= nitag) continue;
    uint32_t v = (uint32_t)(w & 0x7ffff);
    int64_t k = (int64_t)itag * NBUCKTS + ib;
    bool zero = v == 0 && idx[i] <= first_nz[k];
    atomicAdd(bck + k, v << 3);
    atomicAdd(bck + k + (int64_t)NBUCKTS * nitag,
              ((uint32_t)(v > 0) << 30) + (zero ? (uint32_t)itagu : (uint32_t)itagu << 10));
  }
}

// SSRB prompt sinogram from the replicate-0 prompt sinogram: every span-11
// (or span-1) sinogram s holds LORs of a single axial position ssrb(s), so
// ssr[p][aw] = sum over s with ssrb(s) = p of psino[s][aw] (exact in integers).
__global__ void k_lm_ssrb(const uint32_t *psino, int nrep, const int32_t *s_off,
                          const int32_t *s_list, uint32_t *ssr) {
  const int64_t n = (int64_t)SEG0 * NSBINANG;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int pl = (int)(i / NSBINANG), aw = (int)(i - (int64_t)pl * NSBINANG);
    uint32_t acc = 0;
    for (int k = s_off[pl]; k < s_off[pl + 1]; k++)
      acc += psino[((int64_t)s_list[k] * NSBINANG + aw) * nrep];
    ssr[i] = acc;
  }
}

// Delayed fan-sums of one-replicate sessions (the sessions whose sinogram
// counts are LDS-privatised). In the event kernel each delayed event's two
// crystal ends were global atomics into the 64 x 504 fan-sums (about 2 ms of
// device atomics for a 600 s stream). Here each workgroup (one per CU: the
// fan-sums take 126 KB of LDS) walks LM_FAN_SUB word blocks at a time, with
// the same tag walk and tests as k_lm_hist, counts the crystal ends in LDS and
// writes its copy to part[blockIdx.x]; k_lm_fan_sum adds the copies to the
// output.
constexpr int LM_FAN = 64 * NCRS;
constexpr int LM_FAN_WG = 256;
constexpr int LM_FAN_SUB = 4;
// COUNT: the same walk also does k_lm_bin_count's counting (buckets in the
// dynamic LDS, coarse buckets and a 16-bit bin table in static LDS): one read
// of the stream instead of two, for the spans whose buckets fit beside the
// fan-sums (span 11 and SSRB; not span 1's 21,609 buckets)
constexpr int LM_FAN_NBK_MAX = 6144;

template <bool COUNT>
__global__ __launch_bounds__(LM_THREADS * LM_FAN_SUB) void k_lm_fan(LmDev p, int64_t nblk,
                                                                   uint32_t *part, int nbk,
                                                                   uint32_t *gcnt,
                                                                   uint32_t *gccnt) {
  extern __shared__ uint32_t bcnt[];  // COUNT: events per bucket
  __shared__ uint32_t fan[LM_FAN];
  __shared__ uint32_t cdf[LM_NCDF];
  __shared__ int32_t wave_last[LM_FAN_SUB][LM_WAVES];
  __shared__ uint32_t ccnt[COUNT ? LM_MAXCB : 1];
  __shared__ uint16_t lut16[COUNT ? NSN1 : 1];  // output sinogram of each span-1 sinogram
  constexpr int NT = LM_THREADS * LM_FAN_SUB;
  const int tid = threadIdx.x, lane = tid & 63, sub = tid / LM_THREADS;
  const int wave = (tid / 64) % LM_WAVES;
  for (int i = tid; i < LM_FAN; i += NT) fan[i] = 0;
  if (tid < LM_NCDF) cdf[tid] = p.cdf[tid];
  if constexpr (COUNT) {
    for (int i = tid; i < nbk; i += NT) bcnt[i] = 0;
    for (int i = tid; i < p.ncb; i += NT) ccnt[i] = 0;
    for (int i = tid; i < NSN1; i += NT)
      lut16[i] = (uint16_t)(p.span == 11 ? p.lut[i] >> 16 : p.lut[i] & 0xffff);
  }
  __syncthreads();
  uint32_t c8[LM_FASTK];
#pragma unroll
  for (int k = 0; k < LM_FASTK; k++) c8[k] = __builtin_amdgcn_readfirstlane(cdf[k]);
  const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1);
  const int64_t step = (int64_t)gridDim.x * LM_FAN_SUB;
  const int64_t rounds = (nblk + step - 1) / step;  // the same for every thread
  for (int64_t r = 0; r < rounds; r++) {
    const int64_t blk = r * step + (int64_t)blockIdx.x * LM_FAN_SUB + sub;
    const bool on = blk < nblk;  // uniform over the sub-block
    const int64_t w0 = blk * LM_BLOCK + (int64_t)wave * (LM_PER_LANE * 64);
    int32_t wd[LM_PER_LANE];
    __syncthreads();  // wave_last of the previous round read
    int32_t wl = -1;
    if (on) wl = lm_load(p, w0, lane, wd);
    if (lane == 0) wave_last[sub][wave] = wl;
    __syncthreads();
    if (!on) continue;
    int32_t c = p.blk_carry[blk];
    for (int k = 0; k < wave; k++)
      if (wave_last[sub][k] >= 0) c = wave_last[sub][k];
    for (int j = 0; j < LM_PER_LANE; j++) {
      const int32_t word = wd[j];
      const int32_t ms = lm_tag_at(word, c, le);
      const int itag = (ms < 0 ? 0 : ms - p.toff) / ITIME;
      const bool cnt = word > 0 && itag >= p.tstart && itag < p.tstop &&
                       (word & 0x3fffffff) < NSN1 * NSBINANG;
      const bool del = cnt && !((word >> 30) & 1);
      if constexpr (COUNT) {
        if (cnt) {  // the events k_lm_hist stages (k_lm_bin_count's test and bin)
          const int32_t val = word & 0x3fffffff;
          const int si = val / NSBINANG, aw = val - si * NSBINANG;
          const int64_t bin = (int64_t)lut16[si] * NSBINANG + aw;
          atomicAdd(&bcnt[bin >> LM_BSHIFT], 1u);
          atomicAdd(&ccnt[bin >> LM_CSHIFT], 1u);
        }
      }
      uint32_t N = 1;
      if (p.btp == 2) {  // the event's weight, as k_lm_hist draws it
        const uint64_t g = (uint64_t)(p.base + w0 + j * 64 + lane);
        uint32_t cc[4];
        boot_quad(cc, g, (uint32_t)p.rep0, p.k0, p.k1);
        N = poisson_n(boot_pick(cc, g), c8, cdf);
      }
      if (del && N > 0 && N < 32) {
        const int32_t val = word & 0x3fffffff;
        const int si = val / NSBINANG, aw = val - si * NSBINANG;
        const int a = aw / NSBINS, w = aw - a * NSBINS;
        const short2 rr = p.rno[si], cr = p.s2cF[a + NSANGLES * w];
        atomicAdd(&fan[NCRS * rr.x + cr.x], N);
        atomicAdd(&fan[NCRS * rr.y + cr.y], N);
      }
    }
  }
  __syncthreads();
  for (int i = tid; i < LM_FAN; i += NT) part[(int64_t)blockIdx.x * LM_FAN + i] = fan[i];
  if constexpr (COUNT) {
    for (int i = tid; i < nbk; i += NT)
      if (bcnt[i]) atomicAdd(gcnt + i, bcnt[i]);
    for (int i = tid; i < p.ncb; i += NT)
      if (ccnt[i]) atomicAdd(gccnt + i, ccnt[i]);
  }
}

__global__ void k_lm_fan_sum(const uint32_t *__restrict__ part, int nwg, uint32_t *fan) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= LM_FAN) return;
  uint32_t acc = 0;
  for (int b = 0; b < nwg; b++) acc += part[(int64_t)b * LM_FAN + i];
  fan[i] += acc;
}

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_philox_kat(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  philox4x32_10(c, key[0], key[1]);
  for (int i = 0; i < 4; i++) out[i] = c[i];
}

// ---- LDS-privatised sinogram counts (one replicate, no resampling) ----
// Pass 1: events per bucket of LM_BBINS bins and per coarse bucket, counted
// in LDS over the blocks blockIdx.x, + gridDim.x, ... (the events k_lm_hist
// counts: in the frame, address in range), added to gcnt[nbk] and gccnt[ncb]
// once per workgroup.
__global__ __launch_bounds__(LM_THREADS) void k_lm_bin_count(LmDev p, int64_t nblk, int nbk,
                                                             uint32_t *gcnt, uint32_t *gccnt) {
  extern __shared__ uint32_t bcnt[];
  __shared__ uint32_t ccnt[LM_MAXCB];
  __shared__ int32_t lut[NSN1];
  __shared__ int32_t wave_last[LM_WAVES];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int i = tid; i < nbk; i += LM_THREADS) bcnt[i] = 0;
  for (int i = tid; i < p.ncb; i += LM_THREADS) ccnt[i] = 0;
  for (int i = tid; i < NSN1; i += LM_THREADS) lut[i] = p.lut[i];
  const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1);
  for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int64_t w0 = blk * LM_BLOCK + (int64_t)wave * (LM_PER_LANE * 64);
    int32_t wd[LM_PER_LANE];
    __syncthreads();  // LDS set up; wave_last of the previous block read
    const int32_t wl = lm_load(p, w0, lane, wd);
    if (lane == 0) wave_last[wave] = wl;
    __syncthreads();
    int32_t c = p.blk_carry[blk];
    for (int k = 0; k < wave; k++)
      if (wave_last[k] >= 0) c = wave_last[k];
    for (int j = 0; j < LM_PER_LANE; j++) {
      const int32_t word = wd[j];
      const int32_t ms = lm_tag_at(word, c, le);
      const int itag = (ms < 0 ? 0 : ms - p.toff) / ITIME;
      if (word > 0 && itag >= p.tstart && itag < p.tstop &&
          (word & 0x3fffffff) < NSN1 * NSBINANG) {
        const int64_t bin = lm_bin(p.span, word & 0x3fffffff, lut);
        atomicAdd(&bcnt[bin >> LM_BSHIFT], 1u);
        atomicAdd(&ccnt[bin >> LM_CSHIFT], 1u);
      }
    }
  }
  __syncthreads();
  for (int i = tid; i < nbk; i += LM_THREADS)
    if (bcnt[i]) atomicAdd(gcnt + i, bcnt[i]);
  for (int i = tid; i < p.ncb; i += LM_THREADS)
    if (ccnt[i]) atomicAdd(gccnt + i, ccnt[i]);
}

// Slot ranges from counts: bucket b's events go to [bstart[b], bstart[b + 1]);
// bcur = bstart (the cursors the next pass takes slots from). One block.
__global__ __launch_bounds__(1024) void k_lm_bin_scan(const uint32_t *gcnt, int nbk,
                                                      uint32_t *bstart, uint32_t *bcur) {
  __shared__ uint32_t t[1024];
  __shared__ uint32_t carry;
  const int tid = threadIdx.x;
  if (tid == 0) carry = 0;
  for (int b0 = 0; b0 < nbk; b0 += 1024) {
    __syncthreads();
    const int i = b0 + tid;
    const uint32_t v = i < nbk ? gcnt[i] : 0u;
    t[tid] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {  // inclusive scan
      const uint32_t x = tid >= o ? t[tid - o] : 0u;
      __syncthreads();
      t[tid] += x;
      __syncthreads();
    }
    const uint32_t ex = carry + t[tid] - v;
    if (i < nbk) {
      bstart[i] = ex;
      bcur[i] = ex;
    }
    __syncthreads();
    if (tid == 1023) carry += t[1023];
  }
  __syncthreads();
  if (tid == 0) bstart[nbk] = carry;
}

// Pass 3: tiles of LM_TILE events of the coarse stage (s1, coarse buckets
// [cstart[c], cstart[c + 1])) sorted by bucket in LDS and written to the
// buckets' ranges of s2 (cursors fcur) in runs, as (bin within bucket) |
// prompt << LM_BSHIFT | N << (LM_BSHIFT + 1). A thread keeps its 32 events
// and their buckets in registers; per window of at most LM_PART_W buckets
// (one window unless the tile spans many sparse coarse buckets) the events
// are ranked with LDS atomics, one cursor atomic per bucket reserves each
// run, the events are placed in LDS by run and wave k writes runs k, k + 4,
// ... with consecutive lanes on consecutive slots.
constexpr int LM_PART_W = 2048;

__global__ __launch_bounds__(LM_THREADS) void k_lm_bin_part(const uint32_t *__restrict__ s1,
                                                            const uint32_t *__restrict__ cstart,
                                                            int ncb, int nbk, uint32_t *fcur,
                                                            uint32_t *s2) {
  __shared__ uint32_t cnt[LM_PART_W + 1];  // counts, then run starts (cnt[nw] = total)
  __shared__ uint32_t gb[LM_PART_W];       // global slot of the run's local slot 0
  __shared__ uint32_t stg[LM_TILE];
  __shared__ uint32_t cs[LM_MAXCB + 1];
  __shared__ int win[2];  // first and last coarse bucket of the tile
  constexpr int PER = LM_TILE / LM_THREADS;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t total = cstart[ncb];
  const int64_t t0 = (int64_t)blockIdx.x * LM_TILE;
  if (t0 >= (int64_t)total) return;  // block-uniform, before any barrier
  const int nt = (int)((int64_t)total - t0 < LM_TILE ? (int64_t)total - t0 : LM_TILE);
  // the whole tile is loaded first (one memory latency per tile)
  uint32_t v[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int i = k * LM_THREADS + tid;
    v[k] = i < nt ? s1[t0 + i] : 0u;
  }
  for (int i = tid; i <= ncb; i += LM_THREADS) cs[i] = cstart[i];
  __syncthreads();
  // coarse bucket of stage position e: the largest c with cs[c] <= e
  auto coarse = [&](uint32_t e, int lo, int hi) {
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (cs[mid] <= e) lo = mid; else hi = mid - 1;
    }
    return lo;
  };
  if (tid == 0) {
    win[0] = coarse((uint32_t)t0, 0, ncb - 1);
    win[1] = coarse((uint32_t)(t0 + nt - 1), 0, ncb - 1);
  }
  __syncthreads();
  const int c0 = win[0], c1 = win[1], f0 = c0 * LM_FPC;
  const int nf = ((c1 + 1) * LM_FPC < nbk ? (c1 + 1) * LM_FPC : nbk) - f0;
  // bucket within the tile's window << 16 | rank within the bucket
  uint32_t rf[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int i = k * LM_THREADS + tid;
    const int cb = i < nt ? coarse((uint32_t)(t0 + i), c0, c1) : c0;
    rf[k] = (uint32_t)(cb * LM_FPC + (int)((v[k] & ((1u << LM_CSHIFT) - 1)) >> LM_BSHIFT) - f0)
            << 16;
  }
  for (int w0 = 0; w0 < nf; w0 += LM_PART_W) {  // block-uniform
    const int nw = nf - w0 < LM_PART_W ? nf - w0 : LM_PART_W;
    __syncthreads();  // the previous window's runs written out
    for (int i = tid; i < nw; i += LM_THREADS) cnt[i] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int f = (int)(rf[k] >> 16) - w0;
      if (k * LM_THREADS + tid < nt && f >= 0 && f < nw)
        rf[k] = (rf[k] & 0xffff0000u) | atomicAdd(&cnt[f], 1u);
    }
    __syncthreads();
    if (wave == 0) {  // run starts (exclusive scan) and one cursor atomic per run
      uint32_t run = 0;
      for (int i0 = 0; i0 < nw; i0 += 64) {
        const int i = i0 + lane;
        const uint32_t n = i < nw ? cnt[i] : 0u;
        uint32_t x = n;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t y = __shfl_up(x, o);
          if (lane >= o) x += y;
        }
        const uint32_t ex = run + x - n;
        if (i < nw) {
          gb[i] = (n ? atomicAdd(fcur + f0 + w0 + i, n) : 0u) - ex;
          cnt[i] = ex;
        }
        run += __shfl(x, 63);
      }
      if (lane == 0) cnt[nw] = run;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int f = (int)(rf[k] >> 16) - w0;
      if (k * LM_THREADS + tid < nt && f >= 0 && f < nw)
        stg[cnt[f] + (rf[k] & 0xffffu)] =
            (v[k] & (LM_BBINS - 1)) | ((v[k] >> LM_CSHIFT) << LM_BSHIFT);
    }
    __syncthreads();
    for (int f = wave; f < nw; f += LM_WAVES) {
      const uint32_t e = cnt[f + 1], base = gb[f];
      for (uint32_t sl = cnt[f] + lane; sl < e; sl += 64) {
        const uint32_t pos = base + sl;
        if (pos < total) s2[pos] = stg[sl];  // the ranges were counted for exactly these events
      }
    }
  }
}

// Pass 4: one block per bucket counts its staged events in LDS and adds the
// bucket's bins to the sinograms (each bin belongs to one block: no atomics
// on the outputs, and buckets without events are not touched).
__global__ __launch_bounds__(1024) void k_lm_bin_apply(const uint32_t *__restrict__ stage,
                                                       const uint32_t *__restrict__ bstart,
                                                       int64_t nbins, uint32_t *psino,
                                                       uint32_t *dsino) {
  __shared__ uint32_t pc[LM_BBINS], dc[LM_BBINS];
  const int tid = threadIdx.x;
  const uint32_t s0 = bstart[blockIdx.x], s1 = bstart[blockIdx.x + 1];
  if (s0 == s1) return;  // block-uniform, before any barrier
  for (int i = tid; i < LM_BBINS; i += 1024) {
    pc[i] = 0;
    dc[i] = 0;
  }
  __syncthreads();
  // eight loads in flight per thread before their counts
  constexpr int PER = 8;
  for (uint32_t e0 = s0; e0 < s1; e0 += PER * 1024) {
    uint32_t v[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const uint32_t e = e0 + k * 1024 + tid;
      v[k] = e < s1 ? stage[e] : 0u;  // weight 0: not counted
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const uint32_t n = v[k] >> (LM_BSHIFT + 1);
      if (n) atomicAdd(((v[k] >> LM_BSHIFT) & 1u ? pc : dc) + (v[k] & (LM_BBINS - 1)), n);
    }
  }
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x << LM_BSHIFT;
  const int lim = (int)(nbins - base < LM_BBINS ? nbins - base : LM_BBINS);
  for (int i = tid; i < lim; i += 1024) {
    if (pc[i]) psino[base + i] += pc[i];
    if (dc[i]) dsino[base + i] += dc[i];
  }
}

__global__ void k_lm_zero_u32(uint32_t *p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

}  // namespace nipet

using namespace nipet;

struct NipetLm {
  NipetLmConfig cfg;
  int device = 0;
  int32_t *lut = nullptr;
  short2 *rno = nullptr, *s2cF = nullptr;
  uint32_t *cdf = nullptr;
  int32_t *carry = nullptr;  // device: last tag (ms) seen so far
  int32_t *blk_last = nullptr, *blk_carry = nullptr;
  int64_t blk_cap = 0;
  unsigned long long *first_nz = nullptr;
  unsigned long long *sgl_idx = nullptr;
  int32_t *sgl_word = nullptr, *sgl_ms = nullptr;
  unsigned int *sgl_count = nullptr, *bad = nullptr;
  unsigned int sgl_cap = 0;
  int32_t *ssrb_off = nullptr, *ssrb_list = nullptr;  // sinograms of each SSRB plane
  int64_t base = 0;  // global index of the next word
  // LDS-privatised sinogram counts: buckets of LM_BBINS output bins (nbk),
  // coarse buckets of 2^LM_CSHIFT bins (ncb): counts, slot ranges, cursors
  int64_t nbins = 0;
  int nbk = 0, ncb = 0;
  uint32_t *gcnt = nullptr, *bstart = nullptr, *bcur = nullptr;
  uint32_t *gccnt = nullptr, *cstart = nullptr, *ccur = nullptr;
  uint32_t *stage = nullptr, *stage2 = nullptr;  // coarse order, bucket order
  int64_t stage_cap = 0;
  uint32_t *fan_part = nullptr;  // [LM_FAN_WG][64 x 504] fan-sum copies
  // set when a chunk failed after its first kernels had advanced the
  // session's time-tag carry: the session can no longer be continued
  bool broken = false;
};

static void lm_free(NipetLm *h) {
  void *ps[] = {h->lut, h->rno, h->s2cF, h->cdf, h->carry, h->blk_last, h->blk_carry,
                h->first_nz, h->sgl_idx, h->sgl_word, h->sgl_ms, h->sgl_count, h->bad,
                h->ssrb_off, h->ssrb_list, h->gcnt, h->bstart, h->bcur, h->gccnt,
                h->cstart, h->ccur, h->stage, h->stage2, h->fan_part};
  for (void *q : ps) (void)hipFree(q);
}

// uint32 Poisson CDF (same formula and order as the CPU definition)
static void poisson_cdf(double lambda, uint32_t *cdf, int K) {
  double p = std::exp(-lambda), s = 0.0;
  for (int k = 0; k < K; k++) {
    s += p;
    double v = std::floor(s * 4294967296.0);
    cdf[k] = (k == K - 1 || v >= 4294967295.0) ? 0xFFFFFFFFu : (uint32_t)v;
    p = p * lambda / (double)(k + 1);
  }
}

static int lm_build(const NipetLmConfig *c, NipetLm *h) {
  NIPET_HIP(hipGetDevice(&h->device));
  std::vector<int32_t> lut(NSN1);
  for (int i = 0; i < NSN1; i++) {
    if (c->sn1_sn11[i] < 0 || c->sn1_ssrb[i] < 0 || c->sn1_ssrb[i] >= SEG0) {
      set_error("nipet_lm_create: sn1_sn11/sn1_ssrb[%d] out of range", i);
      return NIPET_EINVAL;
    }
    lut[i] = ((int32_t)c->sn1_sn11[i] << 16) | (int32_t)c->sn1_ssrb[i];
  }
  for (int i = 0; i < 2 * NSN1; i++)
    if (c->sn1_rno[i] < 0 || c->sn1_rno[i] >= 64) {
      set_error("nipet_lm_create: sn1_rno out of range");
      return NIPET_EINVAL;
    }
  for (int i = 0; i < 2 * NSBINANG; i++)
    if (c->s2cF[i] < 0 || c->s2cF[i] >= NCRS) {
      set_error("nipet_lm_create: s2cF out of range");
      return NIPET_EINVAL;
    }
  uint32_t cdf[LM_NCDF];
  poisson_cdf(c->btprt, cdf, LM_NCDF);
  // sinograms of the output span that fall on each SSRB plane
  const int nsino = c->span == 11 ? 837 : (c->span == 1 ? NSN1 : SEG0);
  std::vector<int> plane_of(nsino, -1);
  for (int si = 0; si < NSN1; si++) {
    int s = c->span == 11 ? c->sn1_sn11[si] : (c->span == 1 ? si : c->sn1_ssrb[si]);
    if (s < 0 || s >= nsino) {
      set_error("nipet_lm_create: span map out of range at %d", si);
      return NIPET_EINVAL;
    }
    if (plane_of[s] >= 0 && plane_of[s] != c->sn1_ssrb[si]) {
      set_error("nipet_lm_create: output sinogram %d mixes axial positions", s);
      return NIPET_EINVAL;
    }
    plane_of[s] = c->sn1_ssrb[si];
  }
  std::vector<int32_t> soff(SEG0 + 1, 0), slist;
  for (int pl = 0; pl < SEG0; pl++) {
    for (int s = 0; s < nsino; s++)
      if (plane_of[s] == pl) slist.push_back(s);
    soff[pl + 1] = (int32_t)slist.size();
  }
  NIPET_HIP(hipMalloc(&h->ssrb_off, (SEG0 + 1) * sizeof(int32_t)));
  NIPET_HIP(hipMalloc(&h->ssrb_list, std::max<size_t>(slist.size(), 1) * sizeof(int32_t)));
  NIPET_HIP(hipMemcpy(h->ssrb_off, soff.data(), (SEG0 + 1) * sizeof(int32_t),
                      hipMemcpyHostToDevice));
  if (!slist.empty())
    NIPET_HIP(hipMemcpy(h->ssrb_list, slist.data(), slist.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice));
  NIPET_HIP(hipMalloc(&h->lut, NSN1 * sizeof(int32_t)));
  NIPET_HIP(hipMalloc(&h->rno, NSN1 * sizeof(short2)));
  NIPET_HIP(hipMalloc(&h->s2cF, NSBINANG * sizeof(short2)));
  NIPET_HIP(hipMalloc(&h->cdf, LM_NCDF * sizeof(uint32_t)));
  NIPET_HIP(hipMalloc(&h->carry, sizeof(int32_t)));
  NIPET_HIP(hipMalloc(&h->sgl_count, sizeof(unsigned int)));
  NIPET_HIP(hipMalloc(&h->bad, sizeof(unsigned int)));
  NIPET_HIP(hipMalloc(&h->first_nz, (size_t)c->nitag * NBUCKTS * sizeof(unsigned long long)));
  h->nbins = (int64_t)nsino * NSBINANG;
  h->nbk = (int)((h->nbins + LM_BBINS - 1) >> LM_BSHIFT);
  h->ncb = (int)((h->nbins + (1 << LM_CSHIFT) - 1) >> LM_CSHIFT);
  if (c->nrep == 1 && c->btp != 1) {  // LDS-privatised sinogram counts
    NIPET_HIP(hipMalloc(&h->gcnt, (size_t)h->nbk * sizeof(uint32_t)));
    NIPET_HIP(hipMalloc(&h->bstart, (size_t)(h->nbk + 1) * sizeof(uint32_t)));
    NIPET_HIP(hipMalloc(&h->bcur, (size_t)h->nbk * sizeof(uint32_t)));
    NIPET_HIP(hipMalloc(&h->gccnt, (size_t)h->ncb * sizeof(uint32_t)));
    NIPET_HIP(hipMalloc(&h->cstart, (size_t)(h->ncb + 1) * sizeof(uint32_t)));
    NIPET_HIP(hipMalloc(&h->ccur, (size_t)h->ncb * sizeof(uint32_t)));
    NIPET_HIP(hipMalloc(&h->fan_part, (size_t)LM_FAN_WG * LM_FAN * sizeof(uint32_t)));
  }
  NIPET_HIP(hipMemcpy(h->lut, lut.data(), NSN1 * sizeof(int32_t), hipMemcpyHostToDevice));
  NIPET_HIP(hipMemcpy(h->rno, c->sn1_rno, NSN1 * sizeof(short2), hipMemcpyHostToDevice));
  NIPET_HIP(hipMemcpy(h->s2cF, c->s2cF, NSBINANG * sizeof(short2), hipMemcpyHostToDevice));
  NIPET_HIP(hipMemcpy(h->cdf, cdf, sizeof(cdf), hipMemcpyHostToDevice));
  int32_t none = -1;
  NIPET_HIP(hipMemcpy(h->carry, &none, sizeof(int32_t), hipMemcpyHostToDevice));
  NIPET_HIP(hipMemset(h->bad, 0, sizeof(unsigned int)));
  int64_t nf = (int64_t)c->nitag * NBUCKTS;
  k_fill_u64<<<(int)std::min<int64_t>((nf + 255) / 256, 4096), 256>>>(h->first_nz, ~0ull, nf);
  NIPET_HIP(hipGetLastError());
  NIPET_HIP(hipDeviceSynchronize());
  return NIPET_OK;
}

// ---- time-tag index of a device word stream: what getLMinfo finds by scanning
// the file (first / last time tag, word positions of given times,
// lm/src/lmaux.cu:17-200), for a stream already in device memory ----
// span[0] = smallest index of a time-tag word, span[1] = largest index + 1,
// over the indices [a0, a1) and [b0, b1). With skip set (a scan of the middle
// [a0, a1) after the two ends), a launch returns at once when the first tag lay
// before a0 and the last at or after a1.
__global__ void k_lm_tag_span(const int32_t *__restrict__ w, int64_t n, int64_t a0, int64_t a1,
                              int64_t b0, int64_t b1, int skip, unsigned long long *span) {
  if (skip && span[0] < (unsigned long long)a0 && span[1] > (unsigned long long)a1)
    return;  // grid-uniform
  unsigned long long lo = (unsigned long long)n, hi = 0ull;
  const int64_t na = a1 - a0, nt = na + (b1 - b0);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t < na ? a0 + t : b0 + (t - na);
    if (is_tag(w[i])) {
      lo = lo < (unsigned long long)i ? lo : (unsigned long long)i;
      hi = hi > (unsigned long long)(i + 1) ? hi : (unsigned long long)(i + 1);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long l2 = __shfl_xor(lo, o), h2 = __shfl_xor(hi, o);
    lo = l2 < lo ? l2 : lo;
    hi = h2 > hi ? h2 : hi;
  }
  if ((threadIdx.x & 63) == 0) {
    if (lo < (unsigned long long)n) atomicMin(span, lo);
    if (hi > 0ull) atomicMax(span + 1, hi);
  }
}

// first[b] = smallest index of a tag word whose second (ms - toff) / 1000 is b
// (tags past the last second count for nsec, tags before toff for none)
__global__ void k_lm_sec_first(const int32_t *__restrict__ w, int64_t n, int toff, int nsec,
                               unsigned long long *first) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t x = w[i];
    if (!is_tag(x)) continue;
    const int rel = (x & TMSK) - toff;
    if (rel < 0) continue;
    const int b = rel / ITIME < nsec ? rel / ITIME : nsec;
    atomicMin(first + b, (unsigned long long)i);
  }
}

// suffix minimum over first[0 .. m): first[s] = min_{b >= s} first[b]; one block
__global__ __launch_bounds__(1024) void k_lm_suffix_min(unsigned long long *first, int m) {
  __shared__ unsigned long long t[1024];
  __shared__ unsigned long long carry;
  const int tid = threadIdx.x;
  if (tid == 0) carry = ~0ull;
  for (int end = m; end > 0; end -= 1024) {
    const int i = end - 1024 + tid;  // tile [end - 1024, end)
    __syncthreads();
    t[tid] = i >= 0 ? first[i] : ~0ull;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const unsigned long long v = tid + o < 1024 ? t[tid + o] : ~0ull;
      __syncthreads();
      if (v < t[tid]) t[tid] = v;
      __syncthreads();
    }
    const unsigned long long c = carry;
    const unsigned long long r = t[tid] < c ? t[tid] : c;
    if (i >= 0) first[i] = r;
    __syncthreads();
    if (tid == 0) carry = r;  // the tile's first element holds the tile minimum
  }
}

extern "C" {

int nipet_lm_create(const NipetLmConfig *c, NipetLm **out) {
  if (!c || !out || !c->sn1_sn11 || !c->sn1_ssrb || !c->sn1_rno || !c->s2cF) {
    set_error("nipet_lm_create: NULL argument");
    return NIPET_EINVAL;
  }
  if (c->nitag <= 0 || c->tstart < 0 || c->tstop < c->tstart || c->tstop > c->nitag ||
      (c->span != 11 && c->span != 1 && c->span != 0) || c->btp < 0 || c->btp > 2 ||
      c->nrep < 1 || c->nrep > LM_MAXREP || c->rep0 < 0 || c->rep0 > (1 << 30) ||
      !(c->btprt > 0.0 && c->btprt < 20.0)) {
    set_error("nipet_lm_create: bad configuration (nitag=%d tstart=%d tstop=%d span=%d btp=%d "
              "nrep=%d)", c->nitag, c->tstart, c->tstop, c->span, c->btp, c->nrep);
    return NIPET_EINVAL;
  }
  NipetLm *h = new NipetLm();
  h->cfg = *c;
  int rc = lm_build(c, h);
  if (rc) {
    lm_free(h);
    delete h;
    return rc;
  }
  *out = h;
  return NIPET_OK;
}

void nipet_lm_destroy(NipetLm *h) {
  if (!h) return;
  lm_free(h);
  delete h;
}

// Singles list and the binned path's staging slots (one per word of the
// chunk). Allocated after the chunk's first kernels are queued, so that the
// host-side cost of these large allocations overlaps their run time.
static int grow_late(NipetLm *h, unsigned int sgl, int64_t stage) {
  if (sgl > h->sgl_cap) {
    (void)hipFree(h->sgl_idx);
    (void)hipFree(h->sgl_word);
    (void)hipFree(h->sgl_ms);
    h->sgl_idx = nullptr;
    h->sgl_word = h->sgl_ms = nullptr;
    h->sgl_cap = 0;
    NIPET_HIP(hipMalloc(&h->sgl_idx, (size_t)sgl * sizeof(unsigned long long)));
    NIPET_HIP(hipMalloc(&h->sgl_word, (size_t)sgl * sizeof(int32_t)));
    NIPET_HIP(hipMalloc(&h->sgl_ms, (size_t)sgl * sizeof(int32_t)));
    h->sgl_cap = sgl;
  }
  if (stage > h->stage_cap) {
    (void)hipFree(h->stage);
    (void)hipFree(h->stage2);
    h->stage = h->stage2 = nullptr;
    h->stage_cap = 0;
    NIPET_HIP(hipMalloc(&h->stage, (size_t)stage * sizeof(uint32_t)));
    NIPET_HIP(hipMalloc(&h->stage2, (size_t)stage * sizeof(uint32_t)));
    h->stage_cap = stage;
  }
  return NIPET_OK;
}

static int grow(NipetLm *h, int64_t nblk) {
  if (nblk > h->blk_cap) {
    (void)hipFree(h->blk_last);
    (void)hipFree(h->blk_carry);
    h->blk_last = h->blk_carry = nullptr;
    h->blk_cap = 0;
    NIPET_HIP(hipMalloc(&h->blk_last, nblk * sizeof(int32_t)));
    NIPET_HIP(hipMalloc(&h->blk_carry, nblk * sizeof(int32_t)));
    h->blk_cap = nblk;
  }
  return NIPET_OK;
}

int nipet_lm_chunk(NipetLm *h, const int32_t *words, int64_t n, const NipetLmOut *o,
                   void *stream) {
  if (!h || !o || (!words && n > 0) || n < 0) {
    set_error("nipet_lm_chunk: bad argument");
    return NIPET_EINVAL;
  }
  if (!o->psino || !o->dsino || !o->ssr || !o->phc || !o->dhc || !o->zR || !o->zM || !o->pvs ||
      !o->fan || !o->bck) {
    set_error("nipet_lm_chunk: every output buffer is required");
    return NIPET_EINVAL;
  }
  if (h->broken) {
    set_error("nipet_lm_chunk: an earlier chunk of this session failed part-way");
    return NIPET_EINVAL;
  }
  if (n == 0) return NIPET_OK;
  const NipetLmConfig &c = h->cfg;
  if (c.btp == 1 && h->base % LM_BLOCK != 0) {
    // the resampling patches are the LM_BLOCK-word blocks of the session's
    // word stream: a chunk must start on a patch boundary
    set_error("nipet_lm_chunk: non-parametric bootstrap needs every chunk but the last to hold "
              "a multiple of %d words (this chunk starts at word %lld)", LM_BLOCK,
              (long long)h->base);
    return NIPET_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  int64_t nblk = (n + LM_BLOCK - 1) / LM_BLOCK;
  if (nblk > INT32_MAX) {
    set_error("nipet_lm_chunk: chunk too large");
    return NIPET_EINVAL;
  }
  // singles list: at most one entry per word; size it generously for the
  // chunk, entries beyond the capacity are reported as an error
  unsigned int cap = (unsigned int)std::min<int64_t>(std::max<int64_t>(n / 32, 1 << 20), 1 << 28);
  // one replicate without resampling: sinogram counts privatised in LDS per
  // bucket (staged events need a slot per word of the chunk)
  const bool bin = h->gcnt && n <= (int64_t)UINT32_MAX;
  int rc = grow(h, nblk);
  if (rc) return rc;
  prof_begin(PK_HIST, st);
  k_lm_tags<<<(unsigned)nblk, LM_THREADS, 0, st>>>(words, n, h->blk_last, h->sgl_count);
  k_lm_carry<<<1, 1024, 0, st>>>(h->blk_last, h->blk_carry, (int)nblk, h->carry);
  LmDev p = {};  // singles list and staging slots are set once allocated
  p.words = words;
  p.n = n;
  p.base = h->base;
  p.toff = c.toff;
  p.nitag = c.nitag;
  p.tstart = c.tstart;
  p.tstop = c.tstop;
  p.span = c.span;
  p.btp = c.btp;
  p.nrep = c.nrep;
  p.rep0 = c.rep0;
  p.k0 = (uint32_t)c.seed;
  p.k1 = (uint32_t)(c.seed >> 32);
  p.nbins = 0;
  p.lut = h->lut;
  p.rno = h->rno;
  p.s2cF = h->s2cF;
  p.cdf = h->cdf;
  p.blk_carry = h->blk_carry;
  p.psino = o->psino;
  p.dsino = o->dsino;
  p.ssr = o->ssr;
  p.phc = o->phc;
  p.dhc = o->dhc;
  p.pvs = o->pvs;
  p.fan = o->fan;
  p.zR = (unsigned long long *)o->zR;
  p.zM = (unsigned long long *)o->zM;
  p.sgl_count = h->sgl_count;
  p.bad = h->bad;
  p.ccur = h->ccur;
  p.ncb = h->ncb;
  // binned path: bucket counts (with the fan-sums where they fit) and their
  // scans go first; they read the words and the tag carry only
  const int nwg = (int)std::min<int64_t>((nblk + LM_FAN_SUB - 1) / LM_FAN_SUB, LM_FAN_WG);
  const bool merged = h->nbk <= LM_FAN_NBK_MAX;  // counts in the fan-sum walk
  if (bin) {
    k_lm_zero_u32<<<(h->nbk + 255) / 256, 256, 0, st>>>(h->gcnt, h->nbk);
    k_lm_zero_u32<<<(h->ncb + 255) / 256, 256, 0, st>>>(h->gccnt, h->ncb);
    if (merged) {
      k_lm_fan<true><<<(unsigned)nwg, LM_THREADS * LM_FAN_SUB, (size_t)h->nbk * sizeof(uint32_t),
                       st>>>(p, nblk, h->fan_part, h->nbk, h->gcnt, h->gccnt);
    } else {
      k_lm_bin_count<<<(unsigned)std::min<int64_t>(nblk, 2048), LM_THREADS,
                       (size_t)h->nbk * sizeof(uint32_t), st>>>(p, nblk, h->nbk, h->gcnt,
                                                                 h->gccnt);
    }
    k_lm_bin_scan<<<1, 1024, 0, st>>>(h->gcnt, h->nbk, h->bstart, h->bcur);
    k_lm_bin_scan<<<1, 1024, 0, st>>>(h->gccnt, h->ncb, h->cstart, h->ccur);
  }
  // the tag carry has advanced: a failure from here on ends the session
  rc = grow_late(h, cap, bin ? n : 0);
  if (rc) {
    h->broken = true;
    return rc;
  }
  p.sgl_idx = h->sgl_idx;
  p.sgl_word = h->sgl_word;
  p.sgl_ms = h->sgl_ms;
  p.sgl_cap = h->sgl_cap;
  p.stage = bin ? h->stage : nullptr;
  p.stage_cap = bin ? n : 0;
  if (c.btp == 1) {
    if (c.nrep == 1)
      k_lm_hist<false, true><<<(unsigned)nblk, LM_THREADS, 0, st>>>(p);
    else
      k_lm_hist<true, true><<<(unsigned)nblk, LM_THREADS, 0, st>>>(p);
  } else if (bin) {
    k_lm_hist<false, false, true><<<(unsigned)nblk, LM_THREADS, 0, st>>>(p);
    k_lm_bin_part<<<(unsigned)((n + LM_TILE - 1) / LM_TILE), LM_THREADS, 0, st>>>(
        h->stage, h->cstart, h->ncb, h->nbk, h->bcur, h->stage2);
    k_lm_bin_apply<<<(unsigned)h->nbk, 1024, 0, st>>>(h->stage2, h->bstart, h->nbins, o->psino,
                                                       o->dsino);
    if (!merged)
      k_lm_fan<false><<<(unsigned)nwg, LM_THREADS * LM_FAN_SUB, 0, st>>>(p, nblk, h->fan_part, 0,
                                                                         nullptr, nullptr);
    k_lm_fan_sum<<<(LM_FAN + 255) / 256, 256, 0, st>>>(h->fan_part, nwg, o->fan);
  } else if (c.nrep == 1) {
    k_lm_hist<false, false><<<(unsigned)nblk, LM_THREADS, 0, st>>>(p);
  } else {
    k_lm_hist<true, false><<<(unsigned)nblk, LM_THREADS, 0, st>>>(p);
  }
  k_lm_sgl_first<<<256, 256, 0, st>>>(h->sgl_idx, h->sgl_word, h->sgl_ms, h->sgl_count,
                                      h->sgl_cap, c.nitag, h->first_nz);
  k_lm_sgl_apply<<<256, 256, 0, st>>>(h->sgl_idx, h->sgl_word, h->sgl_ms, h->sgl_count,
                                      h->sgl_cap, c.nitag, h->first_nz, o->bck);
  prof_end(PK_HIST, st);
  NIPET_HIP(hipGetLastError());
  h->base += n;
  return NIPET_OK;
}

int nipet_lm_finish(NipetLm *h, const NipetLmOut *o, void *stream) {
  if (!h || !o || !o->psino || !o->ssr) {
    set_error("nipet_lm_finish: bad argument");
    return NIPET_EINVAL;
  }
  if (h->broken) {
    set_error("nipet_lm_finish: a chunk of this session failed part-way");
    return NIPET_EINVAL;
  }
  k_lm_ssrb<<<4096, 256, 0, (hipStream_t)stream>>>(o->psino, h->cfg.nrep, h->ssrb_off,
                                                     h->ssrb_list, o->ssr);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_lm_status(NipetLm *h, int64_t *words_done, unsigned int *bad_words,
                    unsigned int *singles_overflow) {
  if (!h) {
    set_error("nipet_lm_status: NULL handle");
    return NIPET_EINVAL;
  }
  unsigned int bad = 0, cnt = 0;
  NIPET_HIP(hipMemcpy(&bad, h->bad, sizeof(bad), hipMemcpyDeviceToHost));
  NIPET_HIP(hipMemcpy(&cnt, h->sgl_count, sizeof(cnt), hipMemcpyDeviceToHost));
  if (words_done) *words_done = h->base;
  if (bad_words) *bad_words = bad;
  if (singles_overflow) *singles_overflow = cnt > h->sgl_cap ? cnt - h->sgl_cap : 0;
  return NIPET_OK;
}

int nipet_lm_tag_span(const int32_t *words, int64_t n, int64_t *span, void *stream) {
  if (!span || n < 0 || (n > 0 && !words)) {
    set_error("nipet_lm_tag_span: bad argument");
    return NIPET_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  unsigned long long *sp = (unsigned long long *)span;
  k_fill_u64<<<1, 64, 0, st>>>(sp, (unsigned long long)n, 1);
  k_fill_u64<<<1, 64, 0, st>>>(sp + 1, 0ull, 1);
  if (n > 0) {
    // the first and last LM_SPAN_END words (a stream's first and last tags
    // lie there), then the rest only when an end held no tag
    const int64_t e = std::min<int64_t>(LM_SPAN_END, n), b = std::max<int64_t>(e, n - LM_SPAN_END);
    const int64_t ne = e + (n - b);
    k_lm_tag_span<<<(unsigned)std::min<int64_t>((ne + 255) / 256, 2048), 256, 0, st>>>(
        words, n, 0, e, b, n, 0, sp);
    if (b > e)
      k_lm_tag_span<<<(unsigned)std::min<int64_t>((b - e + 255) / 256, 8192), 256, 0, st>>>(
          words, n, e, b, b, b, 1, sp);
  }
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_lm_second_index(const int32_t *words, int64_t n, int toff, int nsec, int64_t *first,
                          void *stream) {
  if (!first || n < 0 || (n > 0 && !words) || nsec < 0) {
    set_error("nipet_lm_second_index: bad argument");
    return NIPET_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  unsigned long long *f = (unsigned long long *)first;
  const int64_t m = (int64_t)nsec + 1;
  k_fill_u64<<<(unsigned)std::min<int64_t>((m + 255) / 256, 1024), 256, 0, st>>>(
      f, (unsigned long long)n, m);
  if (n > 0) {
    const int64_t nb = std::min<int64_t>((n + 255) / 256, 8192);
    k_lm_sec_first<<<(unsigned)nb, 256, 0, st>>>(words, n, toff, nsec, f);
  }
  k_lm_suffix_min<<<1, 1024, 0, st>>>(f, (int)m);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_philox4x32(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
  if (!ctr || !key || !out) {
    set_error("nipet_philox4x32: NULL argument");
    return NIPET_EINVAL;
  }
  uint32_t *d = nullptr;
  NIPET_HIP(hipMalloc(&d, 10 * sizeof(uint32_t)));
  NIPET_HIP(hipMemcpy(d, ctr, 4 * sizeof(uint32_t), hipMemcpyHostToDevice));
  NIPET_HIP(hipMemcpy(d + 4, key, 2 * sizeof(uint32_t), hipMemcpyHostToDevice));
  k_philox_kat<<<1, 1>>>(d, d + 4, d + 6);
  NIPET_HIP(hipGetLastError());
  NIPET_HIP(hipMemcpy(out, d + 6, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost));
  (void)hipFree(d);
  return NIPET_OK;
}

}  // extern "C"
