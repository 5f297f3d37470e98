// Geometry: FP64 transaxial intersection lists of the mMR, built once per
// session on the device, plus the axial (Michelogram) tables.
//
// Semantics of a transaxial list (reference: sddn_tx prj/src/tprj.cu:13-172
// and the segment walk of prj/src/prjf.cu:59-86):
//  - the ray joins the centres of the faces of crystals s2c[aw][0] and
//    s2c[aw][1]; t is the transaxial distance from the first crystal;
//  - the ray is clipped to the FOV circle x^2+y^2 = tfov2: t in (t1, t2);
//  - every crossing of a grid line x = i*vox or y = j*vox strictly inside
//    (t1, t2) is a boundary; the pieces before the first and after the last
//    crossing are not part of the list; a list entry is the voxel between two
//    consecutive crossings;
//  - crossings are merged in increasing t (x-line first on an exact tie); the
//    voxel index is stepped by +-1 at each crossing starting from the voxel
//    holding the midpoint of [t1, first crossing];
//  - u = 160 + floor(x/vox), v = 160 - ceil(y/vox) (prj/src/tprj.cu:62,82).
// Everything is computed in double without contraction, so the lists are
// bit-identical to a straightforward double-precision CPU tracer.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "nipet_common.h"

namespace nipet {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char *what) {
  set_error("HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
  return (e == hipErrorOutOfMemory) ? NIPET_ENOMEM : NIPET_EHIP;
}

// ---- kernel timing ----
struct ProfRec {
  int kernel;
  hipEvent_t e0, e1;
};
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::vector<ProfRec> g_prof_pending;
static std::vector<hipEvent_t> g_prof_pool;
static double g_prof_ms[PK_COUNT];
static int64_t g_prof_n[PK_COUNT];
static thread_local hipEvent_t g_prof_open[PK_COUNT];

static hipEvent_t prof_event() {
  hipEvent_t e = nullptr;
  if (!g_prof_pool.empty()) {
    e = g_prof_pool.back();
    g_prof_pool.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    e = nullptr;
  }
  return e;
}

void prof_begin(int k, hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (!g_prof_on) return;
  hipEvent_t e = prof_event();
  if (e && hipEventRecord(e, st) == hipSuccess) g_prof_open[k] = e;
}

void prof_end(int k, hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (!g_prof_on || !g_prof_open[k]) return;
  hipEvent_t e = prof_event();
  if (e && hipEventRecord(e, st) == hipSuccess)
    g_prof_pending.push_back(ProfRec{k, g_prof_open[k], e});
  g_prof_open[k] = nullptr;
}

static int prof_resolve() {
  for (auto &r : g_prof_pending) {
    NIPET_HIP(hipEventSynchronize(r.e1));
    float ms = 0.f;
    NIPET_HIP(hipEventElapsedTime(&ms, r.e0, r.e1));
    g_prof_ms[r.kernel] += ms;
    g_prof_n[r.kernel] += 1;
    g_prof_pool.push_back(r.e0);
    g_prof_pool.push_back(r.e1);
  }
  g_prof_pending.clear();
  return NIPET_OK;
}

#pragma clang fp contract(off)

struct RayD {
  double px, py, ax, ay, atn, t1, t2;
};

__device__ static void ray_setup(const float *crs, int c1, int c2, double tfov2, RayD &r) {
  double x1 = 0.5 * ((double)crs[4 * c1 + 0] + (double)crs[4 * c1 + 2]);
  double y1 = 0.5 * ((double)crs[4 * c1 + 1] + (double)crs[4 * c1 + 3]);
  double x2 = 0.5 * ((double)crs[4 * c2 + 0] + (double)crs[4 * c2 + 2]);
  double y2 = 0.5 * ((double)crs[4 * c2 + 1] + (double)crs[4 * c2 + 3]);
  double dx = x2 - x1, dy = y2 - y1;
  double atn = sqrt(dx * dx + dy * dy);
  r.px = x1;
  r.py = y1;
  r.ax = dx / atn;
  r.ay = dy / atn;
  r.atn = atn;
  double b = x1 * r.ax + y1 * r.ay;
  double c = x1 * x1 + y1 * y1 - tfov2;
  double disc = b * b - c;
  double sq = disc > 0 ? sqrt(disc) : 0.0;
  r.t1 = -b - sq;
  r.t2 = -b + sq;
}

// Crossings of the lines coord = i*vox in increasing t.
struct LineWalk {
  double p, a, vox, t1, t2, t;
  long i, iend;
  int dir;
  bool live;

  __device__ void init(double p_, double a_, double vox_, double t1_, double t2_) {
    p = p_;
    a = a_;
    vox = vox_;
    t1 = t1_;
    t2 = t2_;
    live = (a != 0.0);
    if (!live) return;
    double c1 = p + a * t1, c2 = p + a * t2;
    double lo = c1 < c2 ? c1 : c2, hi = c1 < c2 ? c2 : c1;
    long i0 = (long)floor(lo / vox) - 1, i1 = (long)ceil(hi / vox) + 1;
    if (a > 0) {
      i = i0;
      iend = i1;
      dir = 1;
    } else {
      i = i1;
      iend = i0;
      dir = -1;
    }
    // skip to the first crossing with t > t1
    for (;;) {
      t = ((double)i * vox - p) / a;
      if (t > t1) break;
      if (i == iend) {
        live = false;
        return;
      }
      i += dir;
    }
    live = (t < t2);
  }
  __device__ void advance() {
    if (i == iend) {
      live = false;
      return;
    }
    i += dir;
    t = ((double)i * vox - p) / a;
    live = (t > t1 && t < t2);
  }
};

// One thread per transaxial bin. Pass 1 (col == nullptr): counts. Pass 2:
// writes the segment records at off[aw].
__global__ void k_tx_lists(const float *crs, const short2 *s2c, int naw, double vox,
                           double tfov2, int nz, const int32_t *off, int32_t *count,
                           SegRec *seg, double *atn_out, float *lmax_out, int *err) {
  int aw = blockIdx.x * blockDim.x + threadIdx.x;
  if (aw >= naw) return;
  short2 cc = s2c[aw];
  RayD r;
  ray_setup(crs, cc.x, cc.y, tfov2, r);
  LineWalk wx, wy;
  wx.init(r.px, r.ax, vox, r.t1, r.t2);
  wy.init(r.py, r.ay, vox, r.t1, r.t2);
  int n = 0;
  if (wx.live || wy.live) {
    bool xfirst = wx.live && (!wy.live || wx.t <= wy.t);
    double tf = xfirst ? wx.t : wy.t;
    double tm = 0.5 * (r.t1 + tf);
    long u = NX / 2 + (long)floor((r.px + r.ax * tm) / vox);
    long v = NY / 2 - (long)ceil((r.py + r.ay * tm) / vox);
    int su = r.ax > 0 ? 1 : -1;
    int sv = r.ay > 0 ? 1 : -1;
    double tprev = 0;
    bool first = true;
    int32_t base = seg ? off[aw] : 0;
    int32_t cap = seg ? off[aw + 1] - base : 0;
    float lmax = 0.f;
    while (wx.live || wy.live) {
      bool takex = wx.live && (!wy.live || wx.t <= wy.t);
      double t = takex ? wx.t : wy.t;
      if (!first) {
        if (seg && n < cap) {
          SegRec rec;
          rec.s = tprev;
          rec.len = (float)(t - tprev);
          rec.colz = (int32_t)(nz * (u + NX * v));
          seg[base + n] = rec;
          lmax = fmaxf(lmax, rec.len);
        }
        n++;
      }
      if (takex) {
        u += su;
        wx.advance();
      } else {
        v -= sv;
        wy.advance();
      }
      if (u < 0 || u >= NX || v < 0 || v >= NY) {
        atomicOr(err, 1);
        return;
      }
      tprev = t;
      first = false;
    }
    if (seg) lmax_out[aw] = lmax;
  }
  if (!seg) {
    count[aw] = n;
    atn_out[aw] = r.atn;
  }
}

// Bins of a subset ordered by decreasing segment count (bins outside
// 0 .. naw-1 last), one workgroup: a counting sort by count with LDS buckets.
// Bins of equal count keep no particular order: the order only decides which
// bins are dispatched first (long rays first: a shorter tail).
__global__ __launch_bounds__(1024) void k_bins_longest_first(const int32_t *__restrict__ off,
                                                             int naw,
                                                             const int32_t *__restrict__ subs,
                                                             int n, int32_t *__restrict__ out) {
  constexpr int NK = MAX_SEG + 2;  // bucket MAX_SEG - count; MAX_SEG + 1: invalid bins
  __shared__ int cnt[NK];
  const int tid = threadIdx.x;
  for (int i = tid; i < NK; i += blockDim.x) cnt[i] = 0;
  __syncthreads();
  auto bucket = [&](int aw) {
    if (aw < 0 || aw >= naw) return MAX_SEG + 1;
    const int c = off[aw + 1] - off[aw];
    return MAX_SEG - min(max(c, 0), MAX_SEG);
  };
  for (int i = tid; i < n; i += blockDim.x) atomicAdd(&cnt[bucket(subs[i])], 1);
  __syncthreads();
  if (tid == 0) {  // exclusive scan (643 entries)
    int acc = 0;
    for (int k = 0; k < NK; k++) {
      const int c = cnt[k];
      cnt[k] = acc;
      acc += c;
    }
  }
  __syncthreads();
  for (int i = tid; i < n; i += blockDim.x) {
    const int aw = subs[i];
    out[atomicAdd(&cnt[bucket(aw)], 1)] = aw;
  }
}

}  // namespace nipet

using namespace nipet;

extern "C" {

const char *nipet_last_error(void) { return g_err; }

const char *nipet_version(void) { return "nipet_amd 0.1 (gfx950)"; }

int nipet_prof_enable(int on) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_on = on != 0;
  return NIPET_OK;
}

int nipet_prof_reset(void) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  int rc = prof_resolve();
  for (int k = 0; k < PK_COUNT; k++) {
    g_prof_ms[k] = 0.0;
    g_prof_n[k] = 0;
  }
  return rc;
}

int nipet_prof_read(int kernel, double *total_ms, int64_t *count) {
  if (kernel < 0 || kernel >= PK_COUNT || !total_ms || !count) {
    set_error("nipet_prof_read: bad argument");
    return NIPET_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_prof_mu);
  int rc = prof_resolve();
  *total_ms = g_prof_ms[kernel];
  *count = g_prof_n[kernel];
  return rc;
}

static void geom_free(Geom *g) {
  // teardown: a failing free leaves nothing to undo
  void *bufs[] = {g->off,   g->seg,    g->atn,    g->lmax,   g->li_z,   g->li_sn,
                  g->li_nosi, g->dr_li, g->sn_off, g->sn_mem, g->fp_slot, g->bp_tab};
  for (void *p : bufs) (void)hipFree(p);
}

static int geom_build(const NipetGeomDesc *d, NipetGeom *g) {
  NIPET_HIP(hipGetDevice(&g->device));
  g->naw = d->naw;
  g->nli = d->nli;
  g->nsn = d->nsn;
  g->ncrs = d->ncrs;
  g->vox_z = d->vox_z;

  // axial tables: zeta (z in voxel units, 0 at the bottom face of plane 0)
  // of the ring centres of LOR a (r0 -> r1); the mirror b runs r1 -> r0.
  std::vector<float2> liz(d->nli);
  std::vector<short2> lisn(d->nli);
  std::vector<float> linos(d->nli);
  int ndirect = 0;
  for (int li = 0; li < d->nli; li++) {
    double za = (double)d->li2rng[2 * li] + 0.5 * d->ring_pitch;
    double zb = (double)d->li2rng[2 * li + 1] + 0.5 * d->ring_pitch;
    double zeta_a = za / d->vox_z + 0.5 * NZ;
    double zeta_b = zb / d->vox_z + 0.5 * NZ;
    if (zeta_a < 0 || zeta_a >= NZ || zeta_b < 0 || zeta_b >= NZ || zeta_b < zeta_a) {
      set_error("li2rng[%d] outside the image or with r1 < r0", li);
      return NIPET_EINVAL;
    }
    liz[li] = make_float2((float)zeta_a, (float)(zeta_b - zeta_a));
    int sa = d->li2sn[2 * li], sb = d->li2sn[2 * li + 1];
    if (sa < 0 || sa >= d->nsn || sb < 0 || sb >= d->nsn || d->li2nos[li] <= 0) {
      set_error("li2sn/li2nos[%d] out of range", li);
      return NIPET_EINVAL;
    }
    lisn[li] = make_short2((short)sa, (short)sb);
    linos[li] = 1.0f / (float)d->li2nos[li];
    if (zeta_b == zeta_a) {
      if (li != ndirect) {
        set_error("direct planes must come first in the li tables (li=%d)", li);
        return NIPET_EINVAL;
      }
      ndirect++;
    }
  }
  g->ndirect = ndirect;

  // ring-difference tables: ring centres sit at zeta = 2 r + 0.5 (ring pitch =
  // two z-voxels), so LOR a of li starts on ring r0 = (zeta_a - 0.5) / 2 and
  // spans d = (zeta_b - zeta_a) / 2 rings
  std::vector<int16_t> drli(64 * 64, -1);
  std::vector<int32_t> snoff(d->nsn + 1, 0);
  std::vector<int16_t> snmem;
  if (d->nli > LI_MAX) {
    set_error("more than %d Michelogram elements", LI_MAX);
    return NIPET_EINVAL;
  }
  g->dmax = 0;
  for (int li = 0; li < d->nli; li++) {
    const float2 z = liz[li];
    const int r0 = (int)floorf(z.x) >> 1, dd = (int)z.y >> 1;
    if (z.x != 2.f * r0 + 0.5f || z.y != 2.f * dd || r0 < 0 || r0 + dd > 63) {
      set_error("li2rng[%d]: rings must sit at z-voxel 2r + 0.5 of a 64-ring scanner", li);
      return NIPET_EINVAL;
    }
    if (drli[dd * 64 + r0] >= 0) {
      set_error("li2rng[%d]: ring pair listed twice", li);
      return NIPET_EINVAL;
    }
    drli[dd * 64 + r0] = (int16_t)li;
    g->dmax = dd > g->dmax ? dd : g->dmax;
  }
  for (int li = 0; li < d->nli; li++) {
    snoff[lisn[li].x + 1]++;
    if (liz[li].y > 0.f) snoff[lisn[li].y + 1]++;
  }
  for (int s = 0; s < d->nsn; s++) snoff[s + 1] += snoff[s];
  snmem.assign(snoff[d->nsn] > 0 ? snoff[d->nsn] : 1, 0);
  {
    std::vector<int32_t> fill(snoff.begin(), snoff.end() - 1);
    for (int li = 0; li < d->nli; li++) {
      snmem[fill[lisn[li].x]++] = (int16_t)(2 * li);
      if (liz[li].y > 0.f) snmem[fill[lisn[li].y]++] = (int16_t)(2 * li + 1);
    }
  }
  // forward-projector slots: the LORs of one ring difference d in groups of
  // FP_G consecutive r0 (ordered by d, then r0); LOR (d, r0 + k) lies exactly
  // 2k planes above (d, r0) along the whole ray
  {
    std::vector<int16_t> slots;
    for (int dd = 0; dd <= g->dmax; dd++)
      for (int r0 = 0; r0 + dd <= 63; r0 += FP_G) {
        bool any = false;
        for (int k = 0; k < FP_G && r0 + k + dd <= 63; k++) any |= drli[dd * 64 + r0 + k] >= 0;
        if (any) slots.push_back((int16_t)(dd * 64 + r0));
      }
    g->fp_nslot = (int)slots.size();
    if (g->fp_nslot > FP_G_SLOTS) {
      set_error("%d forward-projector slots exceed the kernel's %d lanes", g->fp_nslot,
                FP_G_SLOTS);
      return NIPET_EINVAL;
    }
    NIPET_HIP(hipMalloc(&g->fp_slot, slots.size() * sizeof(int16_t)));
    NIPET_HIP(hipMemcpy(g->fp_slot, slots.data(), slots.size() * sizeof(int16_t),
                        hipMemcpyHostToDevice));
  }
  {
    // back-projector LOR tables, scaled for 1, 2 and 4 images: 16-bit byte
    // offsets into a row of nsn + 1 sinograms of 4 images
    if (d->nsn * 4 * 4 > 0xffff) {
      set_error("%d sinograms: the back projector's 16-bit row offsets reach 4095", d->nsn);
      return NIPET_EINVAL;
    }
    std::vector<uint16_t> tab(6 * (size_t)BP_TABN);
    const int nbs[3] = {1, 2, 4};
    for (int w = 0; w < 3; w++) {
      uint16_t *t = tab.data() + (size_t)w * 2 * BP_TABN;
      const uint16_t zr = (uint16_t)(d->nsn * nbs[w] * 4);
      for (int i = 0; i < 2 * BP_TABN; i++) t[i] = zr;
      for (int li = 0; li < d->nli; li++) {
        const int r0 = (int)floorf(liz[li].x) >> 1, dd = (int)liz[li].y >> 1;
        const int k = bp_row(dd) + dd + 1 + r0;
        t[k] = (uint16_t)(lisn[li].x * nbs[w] * 4);
        if (dd > 0) t[BP_TABN + k] = (uint16_t)(lisn[li].y * nbs[w] * 4);
      }
    }
    NIPET_HIP(hipMalloc(&g->bp_tab, tab.size() * sizeof(uint16_t)));
    NIPET_HIP(hipMemcpy(g->bp_tab, tab.data(), tab.size() * sizeof(uint16_t),
                        hipMemcpyHostToDevice));
  }
  NIPET_HIP(hipMalloc(&g->dr_li, drli.size() * sizeof(int16_t)));
  NIPET_HIP(hipMalloc(&g->sn_off, snoff.size() * sizeof(int32_t)));
  NIPET_HIP(hipMalloc(&g->sn_mem, snmem.size() * sizeof(int16_t)));
  NIPET_HIP(hipMemcpy(g->dr_li, drli.data(), drli.size() * sizeof(int16_t),
                      hipMemcpyHostToDevice));
  NIPET_HIP(hipMemcpy(g->sn_off, snoff.data(), snoff.size() * sizeof(int32_t),
                      hipMemcpyHostToDevice));
  NIPET_HIP(hipMemcpy(g->sn_mem, snmem.data(), snmem.size() * sizeof(int16_t),
                      hipMemcpyHostToDevice));

  NIPET_HIP(hipMalloc(&g->li_z, d->nli * sizeof(float2)));
  NIPET_HIP(hipMalloc(&g->li_sn, d->nli * sizeof(short2)));
  NIPET_HIP(hipMalloc(&g->li_nosi, d->nli * sizeof(float)));
  NIPET_HIP(hipMemcpy(g->li_z, liz.data(), d->nli * sizeof(float2), hipMemcpyHostToDevice));
  NIPET_HIP(hipMemcpy(g->li_sn, lisn.data(), d->nli * sizeof(short2), hipMemcpyHostToDevice));
  NIPET_HIP(hipMemcpy(g->li_nosi, linos.data(), d->nli * sizeof(float), hipMemcpyHostToDevice));

  // transaxial lists
  float *d_crs = nullptr;
  short2 *d_s2c = nullptr;
  int32_t *d_cnt = nullptr;
  int *d_err = nullptr;
  // the build's temporaries are freed on every return path
  struct Tmp {
    void **p[4];
    ~Tmp() {
      for (void **q : p) (void)hipFree(*q);
    }
  } tmp{{(void **)&d_crs, (void **)&d_s2c, (void **)&d_cnt, (void **)&d_err}};
  NIPET_HIP(hipMalloc(&d_crs, d->ncrs * 4 * sizeof(float)));
  NIPET_HIP(hipMalloc(&d_s2c, d->naw * sizeof(short2)));
  NIPET_HIP(hipMalloc(&d_cnt, d->naw * sizeof(int32_t)));
  NIPET_HIP(hipMalloc(&d_err, sizeof(int)));
  NIPET_HIP(hipMalloc(&g->atn, d->naw * sizeof(double)));
  NIPET_HIP(hipMalloc(&g->lmax, d->naw * sizeof(float)));
  NIPET_HIP(hipMalloc(&g->off, (d->naw + 1) * sizeof(int32_t)));
  NIPET_HIP(hipMemcpy(d_crs, d->crs, d->ncrs * 4 * sizeof(float), hipMemcpyHostToDevice));
  NIPET_HIP(hipMemcpy(d_s2c, d->s2c, d->naw * sizeof(short2), hipMemcpyHostToDevice));
  NIPET_HIP(hipMemset(d_err, 0, sizeof(int)));
  NIPET_HIP(hipMemset(g->lmax, 0, d->naw * sizeof(float)));
  int nb = (d->naw + 255) / 256;
  k_tx_lists<<<nb, 256>>>(d_crs, d_s2c, d->naw, (double)d->vox_xy, (double)d->tfov2, NZ,
                          nullptr, d_cnt, nullptr, g->atn, g->lmax, d_err);
  NIPET_HIP(hipGetLastError());
  std::vector<int32_t> cnt(d->naw);
  int herr = 0;
  NIPET_HIP(hipMemcpy(cnt.data(), d_cnt, d->naw * sizeof(int32_t), hipMemcpyDeviceToHost));
  NIPET_HIP(hipMemcpy(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost));
  if (herr) {
    set_error("transaxial ray left the image grid (inconsistent geometry)");
    return NIPET_EINVAL;
  }
  std::vector<int32_t> off(d->naw + 1);
  int64_t tot = 0;
  for (int i = 0; i < d->naw; i++) {
    if (cnt[i] > MAX_SEG) {  // the projectors stage one ray's list in LDS
      set_error("transaxial ray %d has %d segments (at most %d)", i, cnt[i], MAX_SEG);
      return NIPET_EINVAL;
    }
    off[i] = (int32_t)tot;
    tot += cnt[i];
  }
  off[d->naw] = (int32_t)tot;
  if (tot >= INT32_MAX) {
    set_error("too many transaxial segments");
    return NIPET_EINVAL;
  }
  g->nseg = tot;
  NIPET_HIP(hipMemcpy(g->off, off.data(), (d->naw + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
  NIPET_HIP(hipMalloc(&g->seg, (tot + 1) * sizeof(SegRec)));
  k_tx_lists<<<nb, 256>>>(d_crs, d_s2c, d->naw, (double)d->vox_xy, (double)d->tfov2, NZ, g->off,
                          nullptr, g->seg, nullptr, g->lmax, d_err);
  NIPET_HIP(hipGetLastError());
  NIPET_HIP(hipDeviceSynchronize());
  // the projectors split a segment into at most three z pieces: no segment may
  // span two z-boundary spacings atn / (2 dmax) of the steepest LOR
  {
    std::vector<float> lm(d->naw);
    std::vector<double> at(d->naw);
    NIPET_HIP(hipMemcpy(lm.data(), g->lmax, d->naw * sizeof(float), hipMemcpyDeviceToHost));
    NIPET_HIP(hipMemcpy(at.data(), g->atn, d->naw * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < d->naw; i++)
      if ((double)lm[i] * 2.0 * g->dmax >= 1.99 * at[i]) {
        set_error("transaxial bin %d: segments cross more than two z-planes", i);
        return NIPET_EINVAL;
      }
  }
  return NIPET_OK;
}

int nipet_geom_create(const NipetGeomDesc *d, NipetGeom **out) {
  if (!d || !out || !d->crs || !d->s2c || !d->li2rng || !d->li2sn || !d->li2nos) {
    set_error("nipet_geom_create: NULL argument");
    return NIPET_EINVAL;
  }
  if (d->naw <= 0 || d->nli <= 0 || d->nsn <= 0 || d->nsn > MAXNSN || d->ncrs <= 0 ||
      d->vox_xy <= 0 || d->vox_z <= 0 || d->tfov2 <= 0) {
    set_error("nipet_geom_create: bad sizes (naw=%d nli=%d nsn=%d)", d->naw, d->nli, d->nsn);
    return NIPET_EINVAL;
  }
  for (int i = 0; i < 2 * d->naw; i++)
    if (d->s2c[i] < 0 || d->s2c[i] >= d->ncrs) {
      set_error("nipet_geom_create: s2c entry %d out of range", i);
      return NIPET_EINVAL;
    }
  NipetGeom *g = new NipetGeom();
  int rc = geom_build(d, g);
  if (rc != NIPET_OK) {
    geom_free(g);
    delete g;
    return rc;
  }
  *out = g;
  return NIPET_OK;
}

void nipet_geom_destroy(NipetGeom *g) {
  if (!g) return;
  geom_free(g);
  delete g;
}

int nipet_geom_info(const NipetGeom *g, int *naw, int64_t *nseg_total, int *nli, int *nsn) {
  if (!g) {
    set_error("nipet_geom_info: NULL geometry");
    return NIPET_EINVAL;
  }
  if (naw) *naw = g->naw;
  if (nseg_total) *nseg_total = g->nseg;
  if (nli) *nli = g->nli;
  if (nsn) *nsn = g->nsn;
  return NIPET_OK;
}

int nipet_bins_longest_first(const NipetGeom *g, const int32_t *subs, int n, int32_t *out,
                             void *stream) {
  if (!g || n < 0 || (n > 0 && (!subs || !out)) || subs == out) {
    set_error("nipet_bins_longest_first: bad argument (geometry, n >= 0, distinct subs/out)");
    return NIPET_EINVAL;
  }
  if (n == 0) return NIPET_OK;
  k_bins_longest_first<<<1, 1024, 0, (hipStream_t)stream>>>(g->off, g->naw, subs, n, out);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_tx_lists(const NipetGeom *g, int64_t *offsets, int32_t *col, float *t, float *len) {
  const bool lists = col || t || len;
  if (!g || !offsets || (lists && (!col || !t || !len))) {
    set_error("nipet_tx_lists: NULL argument (offsets required; col, t, len all or none)");
    return NIPET_EINVAL;
  }
  std::vector<int32_t> off(g->naw + 1);
  NIPET_HIP(hipMemcpy(off.data(), g->off, (g->naw + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int i = 0; i <= g->naw; i++) offsets[i] = off[i];
  if (!lists) return NIPET_OK;
  std::vector<SegRec> sg(g->nseg);
  NIPET_HIP(hipMemcpy(sg.data(), g->seg, g->nseg * sizeof(SegRec), hipMemcpyDeviceToHost));
  for (int64_t k = 0; k < g->nseg; k++) {
    col[k] = sg[k].colz / NZ;
    t[k] = (float)sg[k].s;
    len[k] = sg[k].len;
  }
  return NIPET_OK;
}

}  // extern "C"
