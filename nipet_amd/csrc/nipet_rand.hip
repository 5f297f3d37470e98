// Randoms estimation from delayed fan-sums for gfx950 (mmr_lmproc.rand,
// gpu_randoms lm/src/rnd.cu:180-420).
//
// Crystal singles rates s[c][r] (448 active crystals x 64 rings) are found from
// the delayed fan-sums f by the reference's fixed-point iteration
//   s <- s / 2 + f / (2 F s),   s0 = sqrt(f / F 1),   10 iterations,
// where (F s)[c][r] sums s over the crystals c' in coincidence with c (cij)
// and the rings r' with |r' - r| <= MRD. F runs as a ring-band prefix sum
// followed by a 448 x 448 crystal product (cij row in LDS). The randoms
// sinogram is s[c0][r0] s[c1][r1] per span-1 LOR, and span-11 bins sum their
// span-1 LORs in a fixed order (the reference uses float atomics).
#include <vector>

#include "nipet_common.h"

namespace nipet {

constexpr int RND_NR = 64;    // rings
constexpr int RND_NC = 448;   // active crystals
constexpr int RND_NCRS = 504; // crystals with gaps

struct RndDev {
  int naw, nsn, mrd;
  const int8_t *cij;       // [448][448]
  const int16_t *crsri;    // [504] reduced crystal or -1
  const short2 *s2cr;      // [naw] reduced crystal pair
  const short2 *sn1_rno;   // [4084] ring pair of each span-1 sinogram
  const int32_t *sn_off;   // [nsn + 1] span-1 sinograms of each output sinogram
  const int16_t *sn_mem;
};

// fansums [64][504] (gaps included) -> f[c][r] over active crystals
__global__ void k_rnd_fsum(float *f, const uint32_t *fan, const int16_t *crsri) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= RND_NR * RND_NCRS) return;
  const int r = i / RND_NCRS, c = i - r * RND_NCRS;
  const int cr = crsri[c];
  if (cr >= 0) f[cr * RND_NR + r] = (float)fan[i];
}

// t[c][r] = sum over rings r' with |r' - r| <= mrd of s[c][r'] (block per c)
__global__ void k_rnd_rings(float *t, const float *s, int mrd) {
  __shared__ float v[RND_NR];
  const int c = blockIdx.x, r = threadIdx.x;
  v[r] = s[c * RND_NR + r];
  __syncthreads();
  const int a = r - mrd < 0 ? 0 : r - mrd, b = r + mrd > RND_NR - 1 ? RND_NR - 1 : r + mrd;
  float acc = 0.f;
  for (int q = a; q <= b; q++) acc += v[q];
  t[c * RND_NR + r] = acc;
}

// out[c][r] = sum over crystals c' with cij[c][c'] of t[c'][r] (block per c,
// thread per ring; crystals summed in ascending order)
__global__ void k_rnd_crys(float *out, const float *t, const int8_t *cij) {
  __shared__ int8_t row[RND_NC];
  const int c = blockIdx.x, r = threadIdx.x;
  for (int i = r; i < RND_NC; i += RND_NR) row[i] = cij[c * RND_NC + i];
  __syncthreads();
  float acc = 0.f;
  for (int q = 0; q < RND_NC; q++)
    if (row[q]) acc += t[q * RND_NR + r];
  out[c * RND_NR + r] = acc;
}

__global__ void k_rnd_ones(float *s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < RND_NC * RND_NR) s[i] = 1.f;
}

// init: s = sqrt(f / F1); step: s = s / 2 + (Fs > 0 ? f / Fs : 0) / 2
__global__ void k_rnd_update(float *s, const float *f, const float *fs, int init) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= RND_NC * RND_NR) return;
  if (init) {
    s[i] = fs[i] > 0.f ? sqrtf(f[i] / fs[i]) : 0.f;
  } else {
    const float q = fs[i] > 0.f ? f[i] / fs[i] : 0.f;
    s[i] = 0.5f * s[i] + 0.5f * q;
  }
}

// randoms sinogram, device layout out[aw][sn]
__global__ void k_rnd_sino(float *out, const float *s, RndDev d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)d.naw * d.nsn) return;
  const int aw = (int)(i / d.nsn), sn = (int)(i - (int64_t)aw * d.nsn);
  const short2 cc = d.s2cr[aw];
  float acc = 0.f;
  for (int m = d.sn_off[sn]; m < d.sn_off[sn + 1]; m++) {
    const short2 rr = d.sn1_rno[d.sn_mem[m]];
    acc += s[cc.x * RND_NR + rr.x] * s[cc.y * RND_NR + rr.y];
  }
  out[i] = acc;
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
  float *f = nullptr, *s = nullptr, *t = nullptr, *fs = nullptr;
  void *alloc[8] = {};
};

static void rnd_free(NipetRnd *h) {
  for (void *p : h->alloc) (void)hipFree(p);
  (void)hipFree(h->f);
  (void)hipFree(h->s);
  (void)hipFree(h->t);
  (void)hipFree(h->fs);
}

template <typename T>
static int rnd_upload(NipetRnd *h, int slot, T **dst, const T *src, size_t n) {
  NIPET_HIP(hipMalloc((void **)dst, n * sizeof(T)));
  h->alloc[slot] = *dst;
  NIPET_HIP(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
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
  // span-1 members of each output sinogram
  std::vector<int32_t> off(desc->nsn + 1, 0);
  std::vector<int16_t> mem(4084);
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
      mem[fill[s]++] = (int16_t)si;
    }
  }
  NipetRnd *h = new NipetRnd();
  h->d.naw = desc->naw;
  h->d.nsn = desc->nsn;
  h->d.mrd = desc->mrd;
  int8_t *cij = nullptr;
  int16_t *crsri = nullptr, *snm = nullptr;
  short2 *s2cr = nullptr, *rno = nullptr;
  int32_t *sno = nullptr;
  int rc = NIPET_OK;
  if ((rc = rnd_upload(h, 0, &cij, desc->cij, (size_t)RND_NC * RND_NC)) ||
      (rc = rnd_upload(h, 1, &crsri, desc->crsri, (size_t)RND_NCRS)) ||
      (rc = rnd_upload(h, 2, &s2cr, (const short2 *)desc->s2cr, (size_t)desc->naw)) ||
      (rc = rnd_upload(h, 3, &rno, (const short2 *)desc->sn1_rno, (size_t)4084)) ||
      (rc = rnd_upload(h, 4, &sno, off.data(), off.size())) ||
      (rc = rnd_upload(h, 5, &snm, mem.data(), mem.size()))) {
    rnd_free(h);
    delete h;
    return rc;
  }
  h->d.cij = cij;
  h->d.crsri = crsri;
  h->d.s2cr = s2cr;
  h->d.sn1_rno = rno;
  h->d.sn_off = sno;
  h->d.sn_mem = snm;
  const size_t n = (size_t)RND_NC * RND_NR * sizeof(float);
  hipError_t e = hipMalloc(&h->f, n);
  if (e == hipSuccess) e = hipMalloc(&h->s, n);
  if (e == hipSuccess) e = hipMalloc(&h->t, n);
  if (e == hipSuccess) e = hipMalloc(&h->fs, n);
  if (e != hipSuccess) {
    rnd_free(h);
    delete h;
    return hip_fail(e, "nipet_rnd_create");
  }
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
  const int n = RND_NC * RND_NR;
  const unsigned nbl = (n + 255) / 256;
  k_rnd_fsum<<<(RND_NR * RND_NCRS + 255) / 256, 256, 0, st>>>(h->f, fansums, h->d.crsri);
  // F 1 -> fs, init s
  k_rnd_ones<<<nbl, 256, 0, st>>>(h->s);
  k_rnd_rings<<<RND_NC, RND_NR, 0, st>>>(h->t, h->s, h->d.mrd);
  k_rnd_crys<<<RND_NC, RND_NR, 0, st>>>(h->fs, h->t, h->d.cij);
  k_rnd_update<<<nbl, 256, 0, st>>>(h->s, h->f, h->fs, 1);
  for (int k = 0; k < niter; k++) {
    k_rnd_rings<<<RND_NC, RND_NR, 0, st>>>(h->t, h->s, h->d.mrd);
    k_rnd_crys<<<RND_NC, RND_NR, 0, st>>>(h->fs, h->t, h->d.cij);
    k_rnd_update<<<nbl, 256, 0, st>>>(h->s, h->f, h->fs, 0);
  }
  if (rsino) {
    const int64_t m = (int64_t)h->d.naw * h->d.nsn;
    k_rnd_sino<<<(unsigned)((m + 255) / 256), 256, 0, st>>>(rsino, h->s, h->d);
  }
  if (cmap)
    k_rnd_cmap<<<(RND_NCRS * RND_NR + 255) / 256, 256, 0, st>>>(cmap, h->s, h->d.crsri);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

void nipet_rnd_destroy(NipetRnd *h) {
  if (!h) return;
  rnd_free(h);
  delete h;
}

}  // extern "C"
