// Randoms estimation from delayed fan-sums for gfx950 (mmr_lmproc.rand,
// gpu_randoms lm/src/rnd.cu:180-420).
//
// Crystal singles rates s[c][r] (448 active crystals x 64 rings) are found from
// the delayed fan-sums f by the reference's fixed-point iteration
//   s <- s / 2 + f / (2 F s),   s0 = sqrt(f / F 1),   10 iterations,
// where (F s)[c][r] sums s over the crystals in coincidence with c and the
// rings r' with |r' - r| <= MRD. As in the reference (lm/src/rnd.cu:120-150,
// 205-226), the crystals in coincidence with c are the circular run of ones of
// row c of cij, walked from its first crystal for at most 320 crystals.
// One kernel per iteration: a block per crystal c sums the run for all 64
// rings (4 interleaved partial sums per ring, combined in a fixed order), then
// the ring band, then updates s (double-buffered: other blocks still read the
// previous iterate). The randoms sinogram of LOR (c0, r0; c1, r1) is
// s[c0][r0] s[c1][r1]; span-11 bins sum their span-1 LORs in a fixed order
// (the reference uses float atomics).
#include <vector>

#include "nipet_common.h"
#include "nipet_pairs.h"

namespace nipet {

constexpr int RND_NR = 64;     // rings
constexpr int RND_NC = 448;    // active crystals
constexpr int RND_NCRS = 504;  // crystals with gaps
constexpr int RND_WALK = 320;  // crystals walked per fan (16 x 20 threads, lm/src/rnd.cu:126-131)

struct RndDev {
  int mrd;
  const short2 *run;       // [448] first crystal and length of the coincidence run
  const int16_t *crsri;    // [504] reduced crystal or -1
};

// fansums [64][504] (gaps included) -> f[c][r] over active crystals
__global__ void k_rnd_fsum(float *f, const uint32_t *fan, const int16_t *crsri) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= RND_NR * RND_NCRS) return;
  const int r = i / RND_NCRS, c = i - r * RND_NCRS;
  const int cr = crsri[c];
  if (cr >= 0) f[cr * RND_NR + r] = (float)fan[i];
}

// INIT: so = sqrt(f / F 1) (F 1 = run length x ring-band length);
// else: so = si / 2 + (F si > 0 ? f / F si : 0) / 2. Block = crystal c, 256
// threads = 4 partial sums x 64 rings.
template <bool INIT>
__global__ __launch_bounds__(256) void k_rnd_step(float *__restrict__ so,
                                                  const float *__restrict__ si,
                                                  const float *__restrict__ f, RndDev d) {
  __shared__ float part[4][RND_NR];
  __shared__ float u[RND_NR];
  const int c = blockIdx.x, g = threadIdx.x >> 6, r = threadIdx.x & 63;
  const short2 run = d.run[c];
  const int lo = r - d.mrd < 0 ? 0 : r - d.mrd;
  const int hi = r + d.mrd > RND_NR - 1 ? RND_NR - 1 : r + d.mrd;
  const int i = c * RND_NR + r;
  if (INIT) {
    if (g == 0) so[i] = sqrtf(f[i] / (float)(run.y * (hi - lo + 1)));
    return;
  }
  float acc = 0.f;
  for (int k = g; k < run.y; k += 4) {
    int q = run.x + k;
    if (q >= RND_NC) q -= RND_NC;
    acc += si[q * RND_NR + r];
  }
  part[g][r] = acc;
  __syncthreads();
  if (g == 0) u[r] = (part[0][r] + part[1][r]) + (part[2][r] + part[3][r]);
  __syncthreads();
  if (g == 0) {
    float fs = 0.f;
    for (int q = lo; q <= hi; q++) fs += u[q];
    const float qv = fs > 0.f ? f[i] / fs : 0.f;
    so[i] = 0.5f * si[i] + 0.5f * qv;
  }
}

// crystal map with gaps, cmap[c][r] (504 x 64)
__global__ void k_rnd_cmap(float *cmap, const float *s, const int16_t *crsri) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= RND_NCRS * RND_NR) return;
  const int c = i / RND_NR, r = i - c * RND_NR;
  const int cr = crsri[c];
  cmap[i] = cr >= 0 ? s[cr * RND_NR + r] : 0.f;
}

}  // namespace nipet

using namespace nipet;

struct NipetRnd {
  RndDev d;
  PairSino ps;
  float *f = nullptr, *s[2] = {nullptr, nullptr};
  std::vector<void *> alloc;
};

static void rnd_free(NipetRnd *h) {
  for (void *p : h->alloc) (void)hipFree(p);
  h->alloc.clear();
}

template <typename T>
static int rnd_dev(NipetRnd *h, T **dst, const T *src, size_t n) {
  NIPET_HIP(hipMalloc((void **)dst, n * sizeof(T)));
  h->alloc.push_back(*dst);
  if (src) NIPET_HIP(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return NIPET_OK;
}

extern "C" {

int nipet_rnd_create(const NipetRndDesc *desc, NipetRnd **out) {
  if (!desc || !out || !desc->cij || !desc->crsri || !desc->s2cr || !desc->sn1_rno ||
      desc->naw <= 0 || desc->nsn1 != 4084 ||
      (desc->nsn != 4084 && (desc->nsn != 837 || !desc->sn1_sn11)) || desc->mrd < 0 ||
      desc->mrd > 63) {
    set_error("nipet_rnd_create: bad description (span 1 or 11 of the 64-ring mMR)");
    return NIPET_EINVAL;
  }
  {
    int n = 0;
    for (int c = 0; c < RND_NCRS; c++)
      if (desc->crsri[c] >= 0 && desc->crsri[c] != n++) n = -RND_NCRS - 1;
    if (n != RND_NC) {
      set_error("nipet_rnd_create: crsri must number the 448 active crystals 0..447 in order");
      return NIPET_EINVAL;
    }
  }
  // coincidence run of each crystal: first crystal after a 0 -> 1 step of row
  // c of cij (circularly), its length, walked for at most 320 crystals
  std::vector<short2> run(RND_NC);
  for (int c = 0; c < RND_NC; c++) {
    const int8_t *row = desc->cij + (size_t)c * RND_NC;
    int nup = 0, start = 0, len = 0;
    for (int q = 0; q < RND_NC; q++) {
      const int prv = row[(q + RND_NC - 1) % RND_NC] != 0, cur = row[q] != 0;
      if (cur && !prv) {
        nup++;
        start = q;
      }
      len += cur;
    }
    if (nup != 1) {
      set_error("nipet_rnd_create: row %d of cij is not one circular run of crystals", c);
      return NIPET_EINVAL;
    }
    run[c].x = (short)start;
    run[c].y = (short)(len < RND_WALK ? len : RND_WALK);
  }
  for (int i = 0; i < 2 * desc->naw; i++)
    if (desc->s2cr[i] < 0 || desc->s2cr[i] >= RND_NC) {
      set_error("nipet_rnd_create: s2cr entry %d out of range", i);
      return NIPET_EINVAL;
    }
  for (int i = 0; i < 2 * 4084; i++)
    if (desc->sn1_rno[i] < 0 || desc->sn1_rno[i] >= RND_NR) {
      set_error("nipet_rnd_create: sn1_rno entry %d out of range", i);
      return NIPET_EINVAL;
    }
  // span-1 members of each output sinogram, ascending
  std::vector<int32_t> off(desc->nsn + 1, 0);
  std::vector<short2> mem(4084);
  for (int si = 0; si < 4084; si++) {
    const int s = desc->nsn == 4084 ? si : desc->sn1_sn11[si];
    if (s < 0 || s >= desc->nsn) {
      set_error("nipet_rnd_create: sn1_sn11 entry %d out of range", si);
      return NIPET_EINVAL;
    }
    off[s + 1]++;
  }
  for (int s = 0; s < desc->nsn; s++) off[s + 1] += off[s];
  {
    std::vector<int32_t> fill(off.begin(), off.end() - 1);
    for (int si = 0; si < 4084; si++) {
      const int s = desc->nsn == 4084 ? si : desc->sn1_sn11[si];
      mem[fill[s]++] = make_short2(desc->sn1_rno[2 * si], desc->sn1_rno[2 * si + 1]);
    }
  }
  NipetRnd *h = new NipetRnd();
  short2 *drun = nullptr, *s2cr = nullptr, *mrr = nullptr;
  int16_t *crsri = nullptr;
  int32_t *doff = nullptr;
  const size_t ns = (size_t)RND_NC * RND_NR;
  int rc;
  if ((rc = rnd_dev(h, &drun, run.data(), run.size())) ||
      (rc = rnd_dev(h, &crsri, desc->crsri, (size_t)RND_NCRS)) ||
      (rc = rnd_dev(h, &s2cr, (const short2 *)desc->s2cr, (size_t)desc->naw)) ||
      (rc = rnd_dev(h, &doff, off.data(), off.size())) ||
      (rc = rnd_dev(h, &mrr, mem.data(), mem.size())) ||
      (rc = rnd_dev(h, &h->f, (const float *)nullptr, ns)) ||
      (rc = rnd_dev(h, &h->s[0], (const float *)nullptr, ns)) ||
      (rc = rnd_dev(h, &h->s[1], (const float *)nullptr, ns))) {
    rnd_free(h);
    delete h;
    return rc;
  }
  h->d.mrd = desc->mrd;
  h->d.run = drun;
  h->d.crsri = crsri;
  h->ps = PairSino{nullptr, RND_NR, 1, s2cr, nullptr, doff, mrr, nullptr, desc->naw, desc->nsn};
  *out = h;
  return NIPET_OK;
}

int nipet_rnd_estimate(NipetRnd *h, const uint32_t *fansums, int niter, float *rsino,
                       float *cmap, void *stream) {
  if (!h || !fansums || niter < 0 || niter > 1000) {
    set_error("nipet_rnd_estimate: bad arguments");
    return NIPET_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  k_rnd_fsum<<<(RND_NR * RND_NCRS + 255) / 256, 256, 0, st>>>(h->f, fansums, h->d.crsri);
  k_rnd_step<true><<<RND_NC, 256, 0, st>>>(h->s[0], h->s[1], h->f, h->d);
  int cur = 0;
  for (int k = 0; k < niter; k++, cur ^= 1)
    k_rnd_step<false><<<RND_NC, 256, 0, st>>>(h->s[cur ^ 1], h->s[cur], h->f, h->d);
  if (rsino) {
    PairSino p = h->ps;
    p.tab = h->s[cur];
    k_pair_sino<<<pair_blocks(p.naw), PAIR_THREADS, 0, st>>>(rsino, p);
  }
  if (cmap)
    k_rnd_cmap<<<(RND_NCRS * RND_NR + 255) / 256, 256, 0, st>>>(cmap, h->s[cur], h->d.crsri);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

void nipet_rnd_destroy(NipetRnd *h) {
  if (!h) return;
  rnd_free(h);
  delete h;
}

}  // extern "C"
