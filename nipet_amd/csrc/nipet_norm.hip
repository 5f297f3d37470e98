// Normalisation sinogram from components (mmr_auxe.norm: dev_norm
// src/norm.cu:13-51, norm_from_components src/norm.cu:54-220) and the online
// mean / variance update (mmr_auxe.varon: var src/auxmath.cu:15-22) for gfx950.
//
// The reference's normalisation factor of span-1 LOR (bin aw, rings r0, r1,
// crystals c1, c2 with gaps) is
//   geo[w] cinf[w][a % 9] E(c1, r0) E(c2, r1),
//   E(c, r) = ceff[r][c] exp(q p_r / (1 + q n_r)) / (1 + q n_r),  q = B_b / 2,
// B_b the singles rate of bucket b = c / 18 + 28 (r / 8), p_r / n_r the
// paralysing / non-paralysing ring dead-time parameters. Span 1 scales by
// axf1[si] / axe1[s11(si)], span 11 averages its span-1 LORs and divides by
// axe1[s11]. Here E is formed once per call (504 x 64 values, L2 resident),
// and each output element sums its span-1 members in a fixed order (the
// reference loops over the 4084 span-1 sinograms per bin with float atomics).
#include <cmath>
#include <vector>

#include "nipet_common.h"
#include "nipet_pairs.h"

namespace nipet {

constexpr int NRM_NRNG = 64;
constexpr int NRM_NCRS = 504;
constexpr int NRM_NSN1 = 4084;
constexpr int NRM_W = 344;
constexpr int NRM_NBCK = 224;

struct NrmDev {
  int naw, nsn, span;
  const int32_t *aw2ali;   // [naw] w + 344 a
  const int16_t *mem_si;   // [4084] span-1 sinogram of each member (CSR order)
  const int16_t *sn1_sn11; // [4084]
  const int8_t *nos;       // [4084] span-1 sinograms in the span-11 sinogram of si
  // per call
  const float *geo, *cinf, *ceff, *axe1, *axf1, *dtp, *dtnp;
  const int32_t *bck;
  float *E;                // [64][504] crystal factor E(c, r) at r * 504 + c
  float *wmem;             // [4084] axial weight of each member (CSR order)
  float *tw;               // [naw] geo[w] cinf[w][a % 9]
};

__global__ void k_nrm_eff(NrmDev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NRM_NRNG * NRM_NCRS) return;
  const int r = i / NRM_NCRS, c = i - r * NRM_NCRS;
  const float q = 0.5f * (float)d.bck[c / 18 + 28 * (r / 8)];
  const float den = 1.f + q * d.dtnp[r];
  d.E[i] = d.ceff[i] * (expf(q * d.dtp[r] / den) / den);
}

__global__ void k_nrm_axial(NrmDev d) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= NRM_NSN1) return;
  const int si = d.mem_si[m], s11 = d.sn1_sn11[si];
  d.wmem[m] = d.span == 1 ? d.axf1[si] / d.axe1[s11] : 1.f / (d.axe1[s11] * (float)d.nos[si]);
}

__global__ void k_nrm_tx(NrmDev d) {
  const int aw = blockIdx.x * blockDim.x + threadIdx.x;
  if (aw >= d.naw) return;
  const int ali = d.aw2ali[aw];
  const int w = ali % NRM_W, a = ali / NRM_W;
  d.tw[aw] = d.geo[w] * d.cinf[w * 9 + a % 9];
}

// Welford update with instance b (0-based): M1 += (x - M1) / (b + 1),
// M2 += (x - M1_old)(x - M1_new)
__global__ void k_varon(float *m1, float *m2, const float *x, int b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float xi = x[i], delta = xi - m1[i];
  const float mn = m1[i] + delta / (float)(b + 1);
  m1[i] = mn;
  m2[i] += delta * (xi - mn);
}

}  // namespace nipet

using namespace nipet;

struct NipetNorm {
  NrmDev d;
  PairSino ps;
  std::vector<void *> alloc;
  float *comp = nullptr;  // geo | cinf | ceff | axe1 | axf1 | dtp | dtnp
  int32_t *bck = nullptr;
};

static void nrm_free(NipetNorm *h) {
  for (void *p : h->alloc) (void)hipFree(p);
  h->alloc.clear();
}

template <typename T>
static int nrm_dev(NipetNorm *h, T **dst, const T *src, size_t n) {
  NIPET_HIP(hipMalloc((void **)dst, n * sizeof(T)));
  h->alloc.push_back(*dst);
  if (src) NIPET_HIP(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return NIPET_OK;
}

// component offsets in NipetNorm::comp
static constexpr size_t NRM_O_GEO = 0, NRM_O_CINF = NRM_W, NRM_O_CEFF = NRM_O_CINF + 9 * NRM_W,
                        NRM_O_AXE1 = NRM_O_CEFF + NRM_NRNG * NRM_NCRS,
                        NRM_O_AXF1 = NRM_O_AXE1 + 837, NRM_O_DTP = NRM_O_AXF1 + NRM_NSN1,
                        NRM_O_DTNP = NRM_O_DTP + NRM_NRNG, NRM_O_END = NRM_O_DTNP + NRM_NRNG;

extern "C" {

int nipet_norm_create(const NipetNormDesc *desc, NipetNorm **out) {
  if (!desc || !out || !desc->aw2ali || !desc->sn1_rno || !desc->sn1_sn11 || !desc->sn1_sn11no ||
      desc->naw <= 0 || (desc->span != 1 && desc->span != 11)) {
    set_error("nipet_norm_create: bad description (aw2ali, sn1_rno, sn1_sn11, sn1_sn11no; span 1 "
              "or 11)");
    return NIPET_EINVAL;
  }
  const int nsn = desc->span == 1 ? NRM_NSN1 : 837;
  std::vector<short2> cp(desc->naw);
  for (int aw = 0; aw < desc->naw; aw++) {
    const int ali = desc->aw2ali[aw];
    if (ali < 0 || ali >= NRM_W * 252) {
      set_error("nipet_norm_create: aw2ali entry %d out of range", aw);
      return NIPET_EINVAL;
    }
    // crystals of dev_norm (src/norm.cu:25-26); all terms are multiples of 1/2
    const int w = ali % NRM_W, a = ali / NRM_W;
    const double c1 = std::fmod(a + 0.5 * (NRM_NCRS - 2 + NRM_W / 2 - w), (double)NRM_NCRS);
    const double c2 = std::fmod(a + 0.5 * (2 * NRM_NCRS - 2 - NRM_W / 2 + w), (double)NRM_NCRS);
    cp[aw].x = (short)std::floor(c1);
    cp[aw].y = (short)std::floor(c2);
  }
  std::vector<int32_t> off(nsn + 1, 0);
  std::vector<int16_t> mem(NRM_NSN1);
  std::vector<short2> mrr(NRM_NSN1);
  for (int si = 0; si < NRM_NSN1; si++) {
    const int r0 = desc->sn1_rno[2 * si], r1 = desc->sn1_rno[2 * si + 1];
    const int s11 = desc->sn1_sn11[si];
    if (r0 < 0 || r0 >= NRM_NRNG || r1 < 0 || r1 >= NRM_NRNG || s11 < 0 || s11 >= 837 ||
        desc->sn1_sn11no[si] < 1) {
      set_error("nipet_norm_create: axial LUT entry %d out of range", si);
      return NIPET_EINVAL;
    }
    off[(desc->span == 1 ? si : s11) + 1]++;
  }
  for (int s = 0; s < nsn; s++) off[s + 1] += off[s];
  {
    std::vector<int32_t> fill(off.begin(), off.end() - 1);
    for (int si = 0; si < NRM_NSN1; si++) {
      const int m = fill[desc->span == 1 ? si : desc->sn1_sn11[si]]++;
      mem[m] = (int16_t)si;
      mrr[m] = make_short2(desc->sn1_rno[2 * si], desc->sn1_rno[2 * si + 1]);
    }
  }
  NipetNorm *h = new NipetNorm();
  NrmDev &d = h->d;
  d.naw = desc->naw;
  d.nsn = nsn;
  d.span = desc->span;
  short2 *cpair = nullptr, *rno = nullptr;
  int32_t *ali = nullptr, *sno = nullptr;
  int16_t *s11 = nullptr, *snm = nullptr;
  int8_t *nos = nullptr;
  float *E = nullptr, *wmem = nullptr, *tw = nullptr;
  int rc;
  if ((rc = nrm_dev(h, &cpair, cp.data(), cp.size())) ||
      (rc = nrm_dev(h, &ali, desc->aw2ali, (size_t)desc->naw)) ||
      (rc = nrm_dev(h, &rno, mrr.data(), mrr.size())) ||
      (rc = nrm_dev(h, &s11, desc->sn1_sn11, (size_t)NRM_NSN1)) ||
      (rc = nrm_dev(h, &nos, desc->sn1_sn11no, (size_t)NRM_NSN1)) ||
      (rc = nrm_dev(h, &sno, off.data(), off.size())) ||
      (rc = nrm_dev(h, &snm, mem.data(), mem.size())) ||
      (rc = nrm_dev(h, &E, (const float *)nullptr, (size_t)NRM_NRNG * NRM_NCRS)) ||
      (rc = nrm_dev(h, &wmem, (const float *)nullptr, (size_t)NRM_NSN1)) ||
      (rc = nrm_dev(h, &tw, (const float *)nullptr, (size_t)desc->naw)) ||
      (rc = nrm_dev(h, &h->comp, (const float *)nullptr, NRM_O_END)) ||
      (rc = nrm_dev(h, &h->bck, (const int32_t *)nullptr, (size_t)NRM_NBCK))) {
    nrm_free(h);
    delete h;
    return rc;
  }
  d.aw2ali = ali;
  d.mem_si = snm;
  d.sn1_sn11 = s11;
  d.nos = nos;
  d.E = E;
  d.wmem = wmem;
  d.tw = tw;
  d.geo = h->comp + NRM_O_GEO;
  d.cinf = h->comp + NRM_O_CINF;
  d.ceff = h->comp + NRM_O_CEFF;
  d.axe1 = h->comp + NRM_O_AXE1;
  d.axf1 = h->comp + NRM_O_AXF1;
  d.dtp = h->comp + NRM_O_DTP;
  d.dtnp = h->comp + NRM_O_DTNP;
  d.bck = h->bck;
  // out[aw][sn] = tw[aw] sum_m wmem[m] E[r0][c1] E[r1][c2]
  h->ps = PairSino{E, 1, NRM_NCRS, cpair, tw, sno, rno, wmem, desc->naw, nsn};
  *out = h;
  return NIPET_OK;
}

int nipet_norm_sino(NipetNorm *h, const NipetNormComp *c, const int32_t *buckets, float *out,
                    void *stream) {
  if (!h || !c || !c->geo || !c->cinf || !c->ceff || !c->axe1 || !c->dtp || !c->dtnp ||
      !buckets || !out || (h->d.span == 1 && !c->axf1)) {
    set_error("nipet_norm_sino: bad arguments (host components geo, cinf, ceff, axe1, dtp, dtnp"
              "%s, host buckets[224], device output)", h && h->d.span == 1 ? ", axf1" : "");
    return NIPET_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  // host sources may be pageable and are released by the caller after return:
  // the copies complete (stream synchronised) before the kernels are queued
  const struct {
    const float *src;
    size_t off, n;
  } parts[] = {{c->geo, NRM_O_GEO, NRM_W},
               {c->cinf, NRM_O_CINF, 9 * NRM_W},
               {c->ceff, NRM_O_CEFF, (size_t)NRM_NRNG * NRM_NCRS},
               {c->axe1, NRM_O_AXE1, 837},
               {c->axf1, NRM_O_AXF1, NRM_NSN1},
               {c->dtp, NRM_O_DTP, NRM_NRNG},
               {c->dtnp, NRM_O_DTNP, NRM_NRNG}};
  for (const auto &p : parts)
    if (p.src)
      NIPET_HIP(hipMemcpyAsync(h->comp + p.off, p.src, p.n * sizeof(float), hipMemcpyHostToDevice,
                               st));
  NIPET_HIP(hipMemcpyAsync(h->bck, buckets, NRM_NBCK * sizeof(int32_t), hipMemcpyHostToDevice, st));
  NIPET_HIP(hipStreamSynchronize(st));
  k_nrm_eff<<<(NRM_NRNG * NRM_NCRS + 255) / 256, 256, 0, st>>>(h->d);
  k_nrm_axial<<<(NRM_NSN1 + 255) / 256, 256, 0, st>>>(h->d);
  k_nrm_tx<<<(h->d.naw + 255) / 256, 256, 0, st>>>(h->d);
  k_pair_sino<<<pair_blocks(h->d.naw), PAIR_THREADS, 0, st>>>(out, h->ps);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

void nipet_norm_destroy(NipetNorm *h) {
  if (!h) return;
  nrm_free(h);
  delete h;
}

int nipet_varon(float *m1, float *m2, const float *x, int b, int64_t n, void *stream) {
  if (!m1 || !m2 || !x || b < 0 || n < 0) {
    set_error("nipet_varon: bad arguments (device m1, m2, x; b >= 0)");
    return NIPET_EINVAL;
  }
  if (n == 0) return NIPET_OK;
  k_varon<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(m1, m2, x, b, n);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

}  // extern "C"
