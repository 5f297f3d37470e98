/*
 * Float32 C++/OpenMP CPU baseline of the reference's projection tracer,
 * OSEM subset update and list-mode histogram loop.
 *
 * TEST INFRASTRUCTURE ONLY: timed by bench.py's cpu_baseline leg and checked
 * by tests/; the product library never links it.
 *
 * It runs the reference's own algorithm (not the exact FP64 crossing walk of
 * nipet_oracle.c), so that its time is what the reference's tracer costs on
 * the host cores:
 *  - cb_tx: per transaxial bin, the float32 row / column crossing
 *    parameters and the row / column step sequence (prj/src/tprj.cu:13-172,
 *    sddn_tx): rec[] = {tr1, tc1, dtr, dtc, t1, min(tr1, tc1), t2, atn,
 *    u + (v << 9), k} and the step bytes (two direction bytes, then one byte
 *    per step: 1 = row crossing, 0 = column crossing);
 *  - cb_fprj: every LOR of a bin walks the bin's steps from the third on;
 *    direct LORs read one z plane, oblique LORs change plane at most once per
 *    transaxial step, at z crossings spaced (tz2 - tz1) / nz, and read the
 *    mirror LOR's plane alongside (prj/src/prjf.cu:33-206, fprj_drct /
 *    fprj_oblq). The step lengths and voxel columns of a bin are the same for
 *    all its LORs, so they are decoded once per bin (same float values as a
 *    per-LOR decode). One OpenMP task per transaxial bin;
 *  - cb_bprj: the same walk scattering into one image accumulator per
 *    thread, then a reduction over the threads (prj/src/prjb.cu:33-187);
 *  - cb_osem_subset: FP, + rho, prompts / e, BP, / sens, mask
 *    (prj/src/recon.cu:459-497 with the element-wise kernels of :218-311);
 *  - cb_hst: the event loop of lm/src/hst.cu:104-239 without bootstrap.
 *    Words are cut into patches of CB_PATCH words; each patch starts from the
 *    time of its first time tag (the reference's thread patches, :74-94) and
 *    counts with atomic adds; delayeds go into the prompt sinogram shifted by
 *    16 bits (:173, :194).
 * cb_time_pair / cb_time_hst time those calls with std::chrono: one warm-up
 * call, then `runs` timed calls, each time written out.
 */
#include <omp.h>
#include <stdint.h>
#include <string.h>

#include <chrono>
#include <cmath>
#include <vector>

namespace {

constexpr int NTV = 1807;  // step bytes per bin (2 direction bytes + steps)
constexpr int NTT = 10;    // floats per bin record
constexpr int IMX = 320, IMY = 320, IMZ = 127;
constexpr int64_t NVOX = (int64_t)IMX * IMY * IMZ;
constexpr float VOXY = 0.208626f, VOXZ = 0.203125f, VOXZI = 4.923077f;
constexpr float RING = 0.40625f, TFOV2 = 890.0f;
constexpr float L21 = 0.001f, TA1 = 0.7885139f, TA2 = -0.7822831f;
constexpr int NDIRECT = 64;
constexpr int NSBINANG = 86688, NSBINS = 344, NSANGLES = 252, NCRS = 504, NBUCKTS = 224;
constexpr int SEG0 = 127, MXNITAG = 5400, VTIME = 2, ITIME = 1000;
constexpr int CB_PATCH = 4096;
constexpr int MAXSTEP = NTV;

inline float sgnbit(float x) { return std::signbit(x) ? 1.f : 0.f; }

// axial voxel coordinate of z [cm]: the float of 63.5 + z / vox_z
inline float zcoord(float z) { return (float)(0.5 * IMZ + (double)(VOXZI * z)); }

// Crossing record and step bytes of one bin; returns the end index k of the
// step bytes, or -1 when they would overflow NTV.
int trace_bin(const float *crs, int c1, int c2, float *rec, uint8_t *tv) {
  const float *a = crs + 4 * c1, *b = crs + 4 * c2;
  const float ax = 0.5f * (a[0] + a[2]), ay = 0.5f * (a[1] + a[3]);
  const float bx = 0.5f * (b[0] + b[2]), by = 0.5f * (b[1] + b[3]);
  const float px = a[0] + 0.5f * (a[2] - a[0]), py = a[1] + 0.5f * (a[3] - a[1]);
  float dx = bx - ax, dy = by - ay;
  const float atn = sqrtf(dx * dx + dy * dy);
  dx /= atn;
  dy /= atn;
  // chord of the FOV circle
  const float B = 2 * (px * dx + py * dy), C = 4 * (-TFOV2 + px * px + py * py);
  const float t1 = 0.5f * (-B - sqrtf(B * B - C)), t2 = 0.5f * (-B + sqrtf(B * B - C));
  // rows (y grid lines)
  const float y1 = py + dy * t1, y2 = py + dy * t2;
  const float lr1 = VOXY * (ceilf(y1 / VOXY) - sgnbit(dy));
  const int v = (int)(0.5 * IMY - std::ceil((double)(y1 / VOXY)));
  const float lr2 = VOXY * (floorf(y2 / VOXY) + sgnbit(dy));
  const float tr1 = (lr1 - py) / dy, tr2 = (lr2 - py) / dy;
  const bool y21 = fabsf(y2 - y1) >= VOXY, lr21 = fabsf(lr1 - lr2) < L21;
  const int nr = (int)(y21 * roundf(fabsf(lr2 - lr1) / VOXY)) + lr21;
  const float dtr = nr > 0 ? (tr2 - tr1) / nr + lr21 * t2 : t2;
  // columns (x grid lines)
  const double x1 = (double)(px + dx * t1);
  const float lc1 = (float)(VOXY * (std::ceil(x1 / VOXY) - sgnbit(dx)));
  const int u = (int)(0.5 * IMX + std::floor(x1 / VOXY));
  const float x2 = px + dx * t2;
  const float lc2 = VOXY * (floorf(x2 / VOXY) + sgnbit(dx));
  const float tc1 = (lc1 - px) / dx, tc2 = (lc2 - px) / dx;
  const bool x21 = fabsf(x2 - (float)x1) >= VOXY, lc21 = fabsf(lc1 - lc2) < L21;
  const int nc = (int)(x21 * roundf(fabsf(lc2 - lc1) / VOXY)) + lc21;
  const float dtc = nc > 0 ? (tc2 - tc1) / nc + lc21 * t2 : t2;

  tv[0] = dx >= 0 ? 2 : 0;
  tv[1] = dy >= 0 ? 2 : 0;
  int k = 2;
  // the faster-changing crossing family drives the loop; a crossing of the
  // other family is inserted before it whenever that one comes first
  if (atanf(dy / dx) < TA1 && atanf(dy / dx) > TA2) {
    float f = tc1, s = tr1;
    for (int i = 0; i <= nc; i++) {
      if (k + 4 > NTV) return -1;
      if (f - s > 0) {
        tv[k++] = 1;
        s += dtr;
      }
      tv[k++] = 0;
      f += dtc;
    }
    if (tr2 > tc2) tv[k++] = 1;
  } else {
    float f = tr1, s = tc1;
    for (int i = 0; i <= nr; i++) {
      if (k + 4 > NTV) return -1;
      if (f - s > 0) {
        tv[k++] = 0;
        s += dtc;
      }
      tv[k++] = 1;
      f += dtr;
    }
    if (tc2 > tr2) tv[k++] = 0;
  }
  rec[0] = tr1;
  rec[1] = tc1;
  rec[2] = dtr;
  rec[3] = dtc;
  rec[4] = t1;
  rec[5] = fminf(tr1, tc1);
  rec[6] = t2;
  rec[7] = atn;
  rec[8] = (float)(u + (v << 9));
  rec[9] = (float)k;
  return k;
}

// The steps of one bin: segment [tp, tn] in voxel column uv (z stride 1).
struct Steps {
  int n = 0;
  float tp[MAXSTEP], tn[MAXSTEP];
  int uv[MAXSTEP];
};

void decode(const float *rec, const uint8_t *tv, Steps &s) {
  const int sa0 = tv[0] - 1, sa1 = tv[1] - 1;
  int rbit = tv[2] & 1;
  const int u = (int)rec[8], v = u >> 9;
  int uv = IMZ * ((u & 0x1ff) + IMX * v);
  // the first (fractional) voxel is skipped
  uv += !rbit * sa0 * IMZ;
  uv -= rbit * sa1 * IMZ * IMX;
  const float dtr = rec[2], dtc = rec[3];
  float trc = rec[0] + rbit * dtr, tcc = rec[1] + dtc * !rbit;
  rbit = tv[3] & 1;
  float tn = trc * rbit + tcc * !rbit, tp = rec[5];
  const int kend = (int)rec[9];
  int n = 0;
  for (int k = 3; k < kend; k++) {
    s.tp[n] = tp;
    s.tn[n] = tn;
    s.uv[n] = uv;
    n++;
    trc += dtr * rbit;
    tcc += dtc * !rbit;
    uv += !rbit * sa0 * IMZ;
    uv -= rbit * sa1 * IMZ * IMX;
    tp = tn;
    rbit = tv[k + 1] & 1;
    tn = trc * rbit + tcc * !rbit;
  }
  s.n = n;
}

// Axial set-up of an oblique LOR (prj/src/prjf.cu:128-156).
struct Axial {
  float s;       // sqrt(1 + (az / atn)^2)
  int sgn;       // +1 / -1 plane step of the LOR (the mirror steps the other way)
  int w, w_;     // planes of the LOR and of its mirror at the first step
  float tz, dtz; // first z crossing and the spacing of the crossings
};

inline Axial axial(const float *rng, float atn, float tp0, float t2) {
  Axial x;
  const float az = rng[1] - rng[0];
  const float az_atn = az / atn;
  x.s = sqrtf(az_atn * az_atn + 1);
  x.sgn = az >= 0 ? 1 : -1;
  const float pz = (float)((double)rng[0] + 0.5 * RING);
  float z = pz + az_atn * tp0;
  x.w = (int)floorf(zcoord(z));
  const float lz1 = (float)(ceilf(zcoord(z)) * VOXZ - 0.5 * IMZ * VOXZ);
  z = (float)((double)rng[1] + 0.5 * RING - (double)(az_atn * tp0));
  x.w_ = (int)floorf(zcoord(z));
  z = pz + az_atn * t2;
  const float lz2 = (float)(floorf(zcoord(z)) * VOXZ - 0.5 * IMZ * VOXZ);
  const int nz = (int)(fabsf(lz2 - lz1) / VOXZ);
  const float tz1 = (lz1 - pz) / az_atn, tz2 = (lz2 - pz) / az_atn;
  x.dtz = (tz2 - tz1) / nz;
  x.tz = tz1;
  return x;
}

struct Tables {
  const float *rec;
  const uint8_t *tv;
  const float *li2rng;   // [nli][2]
  const int16_t *li2sn;  // [nli][2]
  const int8_t *li2nos;  // [nli]
  int nli, nsn;
};

// Forward projection of one bin into srow[nsn] (zeroed by the caller).
void fp_bin(const Tables &T, const Steps &st, const float *rec, const float *img, float *srow,
            int att) {
  for (int ixz = 0; ixz < T.nli; ixz++) {
    const float *rng = T.li2rng + 2 * ixz;
    const int sn0 = T.li2sn[2 * ixz], sn1 = T.li2sn[2 * ixz + 1];
    if (ixz < NDIRECT) {
      const int w = (int)floorf(zcoord((float)((double)rng[0] + 0.5 * RING)));
      float acc = 0;
      for (int k = 0; k < st.n; k++) acc += (st.tn[k] - st.tp[k]) * img[w + st.uv[k]];
      if (att)
        srow[sn0] += expf(-acc) / (float)T.li2nos[ixz];
      else
        srow[sn0] += acc;
      continue;
    }
    Axial x = axial(rng, rec[7], rec[5], rec[6]);
    int w = x.w, w_ = x.w_;
    float tzc = x.tz, acc = 0, acc_ = 0;
    for (int k = 0; k < st.n; k++) {
      const float tp = st.tp[k], tn = st.tn[k], lt = tn - tp;
      const float *col = img + st.uv[k];
      if (tn - tzc > 0) {
        const float fr = (tzc - tp) / lt;
        acc += fr * lt * x.s * col[w];
        acc_ += fr * lt * x.s * col[w_];
        w += x.sgn;
        w_ -= x.sgn;
        acc += (1 - fr) * lt * x.s * col[w];
        acc_ += (1 - fr) * lt * x.s * col[w_];
        tzc += x.dtz;
      } else {
        acc += lt * x.s * col[w];
        acc_ += lt * x.s * col[w_];
      }
    }
    if (att) {
      srow[sn0] += expf(-acc) / (float)T.li2nos[ixz];
      srow[sn1] += expf(-acc_) / (float)T.li2nos[ixz];
    } else {
      srow[sn0] += acc;
      srow[sn1] += acc_;
    }
  }
}

// Back projection of one bin's row srow[nsn] into the accumulator img.
void bp_bin(const Tables &T, const Steps &st, const float *rec, const float *srow, float *img) {
  for (int ixz = 0; ixz < T.nli; ixz++) {
    const float *rng = T.li2rng + 2 * ixz;
    const float bin = srow[T.li2sn[2 * ixz]];
    if (ixz < NDIRECT) {
      const int w = (int)floorf(zcoord((float)((double)rng[0] + 0.5 * RING)));
      for (int k = 0; k < st.n; k++) img[w + st.uv[k]] += (st.tn[k] - st.tp[k]) * bin;
      continue;
    }
    const float bin_ = srow[T.li2sn[2 * ixz + 1]];
    Axial x = axial(rng, rec[7], rec[5], rec[6]);
    int w = x.w, w_ = x.w_;
    float tzc = x.tz;
    for (int k = 0; k < st.n; k++) {
      const float tp = st.tp[k], tn = st.tn[k], lt = tn - tp;
      float *col = img + st.uv[k];
      if (tn - tzc > 0) {
        const float fr = (tzc - tp) / lt;
        col[w] += fr * lt * x.s * bin;
        col[w_] += fr * lt * x.s * bin_;
        w += x.sgn;
        w_ -= x.sgn;
        col[w] += (1 - fr) * lt * x.s * bin;
        col[w_] += (1 - fr) * lt * x.s * bin_;
        tzc += x.dtz;
      } else {
        col[w] += lt * x.s * bin;
        col[w_] += lt * x.s * bin_;
      }
    }
  }
}

// Every voxel index the walks of these bins would touch, checked against the
// image; returns the number of out-of-range accesses (none may happen before
// the unchecked walks above run on these bins).
int64_t check_bin(const Tables &T, const Steps &st, const float *rec) {
  int64_t bad = 0;
  auto in = [&](int64_t i) { bad += (i < 0 || i >= NVOX); };
  for (int ixz = 0; ixz < T.nli; ixz++) {
    const float *rng = T.li2rng + 2 * ixz;
    if (ixz < NDIRECT) {
      const int w = (int)floorf(zcoord((float)((double)rng[0] + 0.5 * RING)));
      for (int k = 0; k < st.n; k++) {
        in(w + (int64_t)st.uv[k]);
        bad += (w < 0 || w >= IMZ);
      }
      continue;
    }
    Axial x = axial(rng, rec[7], rec[5], rec[6]);
    int w = x.w, w_ = x.w_;
    float tzc = x.tz;
    for (int k = 0; k < st.n; k++) {
      const int64_t c = st.uv[k];
      in(c + w);
      in(c + w_);
      bad += (w < 0 || w >= IMZ || w_ < 0 || w_ >= IMZ);
      if (st.tn[k] - tzc > 0) {
        w += x.sgn;
        w_ -= x.sgn;
        in(c + w);
        in(c + w_);
        bad += (w < 0 || w >= IMZ || w_ < 0 || w_ >= IMZ);
        tzc += x.dtz;
      }
    }
  }
  return bad;
}

inline void add_u32(uint32_t *p, uint32_t v) { __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }
inline void add_u64(uint64_t *p, uint64_t v) { __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }

}  // namespace

extern "C" {

int cb_threads(void) { return omp_get_max_threads(); }

/* Crossing records rec[naw][10] and step bytes tv[naw][1807] of all bins.
 * Returns the number of bins whose steps overflow (0 expected). */
int cb_tx(const float *crs, const int16_t *s2c, int naw, float *rec, uint8_t *tv) {
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
  for (int i = 0; i < naw; i++) {
    uint8_t *t = tv + (int64_t)i * NTV;
    memset(t, 0, NTV);
    bad += trace_bin(crs, s2c[2 * i], s2c[2 * i + 1], rec + (int64_t)i * NTT, t) < 0;
  }
  return bad;
}

/* Out-of-image voxel accesses the FP / BP walks of bins subs[nprj] would make. */
int64_t cb_check(const float *rec, const uint8_t *tv, const int32_t *subs, int nprj,
                 const float *li2rng, const int16_t *li2sn, const int8_t *li2nos, int nli,
                 int nsn) {
  const Tables T{rec, tv, li2rng, li2sn, li2nos, nli, nsn};
  int64_t bad = 0;
#pragma omp parallel reduction(+ : bad)
  {
    std::vector<Steps> sv(1);
#pragma omp for schedule(dynamic, 8)
    for (int i = 0; i < nprj; i++) {
      const int b = subs[i];
      decode(rec + (int64_t)b * NTT, tv + (int64_t)b * NTV, sv[0]);
      bad += check_bin(T, sv[0], rec + (int64_t)b * NTT);
    }
  }
  return bad;
}

/* sino[nprj][nsn] = FP of img over bins subs[nprj] (emission, or attenuation
 * factors averaged per span-11 bin when att). */
int cb_fprj(float *sino, const float *img, const float *rec, const uint8_t *tv,
            const int32_t *subs, int nprj, const float *li2rng, const int16_t *li2sn,
            const int8_t *li2nos, int nli, int nsn, int att) {
  const Tables T{rec, tv, li2rng, li2sn, li2nos, nli, nsn};
#pragma omp parallel
  {
    std::vector<Steps> sv(1);
#pragma omp for schedule(dynamic, 8)
    for (int i = 0; i < nprj; i++) {
      const int b = subs[i];
      float *srow = sino + (int64_t)i * nsn;
      memset(srow, 0, sizeof(float) * nsn);
      decode(rec + (int64_t)b * NTT, tv + (int64_t)b * NTV, sv[0]);
      fp_bin(T, sv[0], rec + (int64_t)b * NTT, img, srow, att);
    }
  }
  return 0;
}

/* img = BP of sino[nprj][nsn] over bins subs[nprj]. work: cb_threads() images
 * (the per-thread accumulators). */
int cb_bprj(float *img, const float *sino, const float *rec, const uint8_t *tv,
            const int32_t *subs, int nprj, const float *li2rng, const int16_t *li2sn,
            const int8_t *li2nos, int nli, int nsn, float *work) {
  const Tables T{rec, tv, li2rng, li2sn, li2nos, nli, nsn};
  const int nt = omp_get_max_threads();
#pragma omp parallel num_threads(nt)
  {
    float *acc = work + (int64_t)omp_get_thread_num() * NVOX;
    memset(acc, 0, sizeof(float) * NVOX);
    std::vector<Steps> sv(1);
#pragma omp for schedule(dynamic, 8)
    for (int i = 0; i < nprj; i++) {
      const int b = subs[i];
      decode(rec + (int64_t)b * NTT, tv + (int64_t)b * NTV, sv[0]);
      bp_bin(T, sv[0], rec + (int64_t)b * NTT, sino + (int64_t)i * nsn, acc);
    }
#pragma omp for schedule(static)
    for (int64_t v = 0; v < NVOX; v++) {
      float s = 0;
      for (int t = 0; t < nt; t++) s += work[(int64_t)t * NVOX + v];
      img[v] = s;
    }
  }
  return 0;
}

/* One OSEM subset update (prj/src/recon.cu:459-497): e = FP(img) + rho[subs],
 * e = prompts[subs] / e (0 where |e| < 1e-6), b = BP(e), b = b / sens (0
 * where |sens| < 1e-6), img = mask ? img * b : 0. prompts uint16 and rho
 * float are full [naw][nsn]; esino [nprj][nsn] and bimg are work buffers. */
int cb_osem_subset(float *img, const uint16_t *prompts, const float *rho, const float *sens,
                   const uint8_t *mask, const float *rec, const uint8_t *tv,
                   const int32_t *subs, int nprj, const float *li2rng, const int16_t *li2sn,
                   const int8_t *li2nos, int nli, int nsn, float *esino, float *bimg,
                   float *work) {
  cb_fprj(esino, img, rec, tv, subs, nprj, li2rng, li2sn, li2nos, nli, nsn, 0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nprj; i++) {
    const int64_t r = (int64_t)subs[i] * nsn;
    for (int s = 0; s < nsn; s++) {
      float e = esino[(int64_t)i * nsn + s] + rho[r + s];
      esino[(int64_t)i * nsn + s] = fabsf(e) < 1e-6f ? 0.f : (float)prompts[r + s] / e;
    }
  }
  cb_bprj(bimg, esino, rec, tv, subs, nprj, li2rng, li2sn, li2nos, nli, nsn, work);
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < NVOX; v++) {
    const float b = fabsf(sens[v]) < 1e-6f ? 0.f : bimg[v] / sens[v];
    img[v] = mask[v] ? img[v] * b : 0.f;
  }
  return 0;
}

/* Timed FP + BP pairs: one warm-up pair, then `runs` pairs; fp_s[runs] and
 * bp_s[runs] receive the wall times [s]. */
int cb_time_pair(const float *img, const float *rec, const uint8_t *tv, const int32_t *subs,
                 int nprj, const float *li2rng, const int16_t *li2sn, const int8_t *li2nos,
                 int nli, int nsn, float *sino, float *bimg, float *work, int runs, double *fp_s,
                 double *bp_s) {
  using clk = std::chrono::steady_clock;
  for (int r = -1; r < runs; r++) {
    auto t0 = clk::now();
    cb_fprj(sino, img, rec, tv, subs, nprj, li2rng, li2sn, li2nos, nli, nsn, 0);
    auto t1 = clk::now();
    cb_bprj(bimg, sino, rec, tv, subs, nprj, li2rng, li2sn, li2nos, nli, nsn, work);
    auto t2 = clk::now();
    if (r >= 0) {
      fp_s[r] = std::chrono::duration<double>(t1 - t0).count();
      bp_s[r] = std::chrono::duration<double>(t2 - t1).count();
    }
  }
  return 0;
}

/* Histogram of words[n] (no bootstrap): span-11 sinogram psino[837][86688]
 * (prompts + delayeds << 16), SSRB prompts ssrb[127][86688], head curves
 * rprmt / rdlyd [nitag], centre of mass zR / zM [nitag], projection views
 * snview [(min(nitag, 5400) + 3) / 4][127][344], delayed fan-sums
 * fansums[64][504], singles buckets bucks[2][nitag][224]. Outputs are
 * accumulated (the caller zeroes them). Returns the counted events. */
int64_t cb_hst(const int32_t *words, int64_t n, int toff, int nitag, int tstart, int tstop,
               const int16_t *sn1_sn11, const int16_t *sn1_ssrb, const int16_t *sn1_rno,
               const int16_t *s2cF, uint32_t *psino, uint32_t *ssrb, uint32_t *rprmt,
               uint32_t *rdlyd, uint64_t *zR, uint64_t *zM, uint32_t *snview, uint32_t *fansums,
               uint32_t *bucks) {
  const int64_t npatch = (n + CB_PATCH - 1) / CB_PATCH;
  int64_t counted = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : counted)
  for (int64_t pt = 0; pt < npatch; pt++) {
    const int64_t i0 = pt * CB_PATCH, i1 = i0 + CB_PATCH < n ? i0 + CB_PATCH : n;
    // time of the patch's first time tag (none: the patch counts nothing)
    int itag = nitag, itagu = 0;
    for (int64_t i = i0; i < i1; i++)
      if ((words[i] >> 29) == -4) {
        const int ms = (words[i] & 0x1fffffff) - toff;
        itag = ms / ITIME;
        itagu = ms - itag * ITIME;
        break;
      }
    for (int64_t i = i0; i < i1; i++) {
      const int32_t word = words[i];
      const int32_t val = word & 0x3fffffff;
      // as the reference: outside [tstart, tstop) nothing is decoded, time
      // tags included (lm/src/hst.cu:145, :207-213)
      if (!(itag >= tstart && itag < tstop)) continue;
      if (word > 0) {
        const int si = val / NSBINANG, aw = val - si * NSBINANG;
        const int a = aw / NSBINS, w = aw - a * NSBINS;
        const int s11 = sn1_sn11[si], sr = sn1_ssrb[si];
        const int64_t addr = (int64_t)s11 * NSBINANG + aw;
        counted++;
        if ((word >> 30) & 1) {
          add_u32(rprmt + itag, 1);
          add_u32(ssrb + (int64_t)sr * NSBINANG + aw, 1);
          add_u32(psino + addr, 1);
          add_u64(zR + itag, (uint64_t)sr);
          add_u64(zM + itag, 1);
          if ((a == 0 || a == 126) && itag < MXNITAG)
            add_u32(snview + (int64_t)(itag >> VTIME) * SEG0 * NSBINS + sr * NSBINS + w,
                    1u << (8 * (a == 126)));
        } else {
          add_u32(psino + addr, 1u << 16);
          add_u32(rdlyd + itag, 1);
          add_u32(fansums + NCRS * sn1_rno[2 * si] + s2cF[2 * (a + NSANGLES * w)], 1);
          add_u32(fansums + NCRS * sn1_rno[2 * si + 1] + s2cF[2 * (a + NSANGLES * w) + 1], 1);
        }
      } else if ((word >> 29) == -4) {
        const int ms = val - toff;
        itag = ms / ITIME;
        itagu = ms - itag * ITIME;
      } else if ((word >> 29) == -3) {
        const int ibck = (word & 0x1fffffff) >> 19;
        if (ibck < NBUCKTS) {
          uint32_t *b0 = bucks + ibck + (int64_t)NBUCKTS * itag;
          uint32_t *b1 = b0 + (int64_t)NBUCKTS * nitag;
          add_u32(b0, (uint32_t)(word & 0x7ffff) << 3);
          add_u32(b1, (uint32_t)((word & 0x7ffff) > 0) << 30);
          add_u32(b1, __atomic_load_n(b1, __ATOMIC_RELAXED) == 0 ? (uint32_t)itagu
                                                                 : (uint32_t)itagu << 10);
        }
      }
    }
  }
  return counted;
}

/* Timed histograms: one warm-up, then `runs` calls, each zeroing the outputs
 * first (inside the timed region, as the reference clears its device arrays
 * for every histogram); secs[runs] receive the wall times. */
int64_t cb_time_hst(const int32_t *words, int64_t n, int toff, int nitag, const int16_t *sn1_sn11,
                    const int16_t *sn1_ssrb, const int16_t *sn1_rno, const int16_t *s2cF,
                    uint32_t *psino, uint32_t *ssrb, uint32_t *rprmt, uint32_t *rdlyd,
                    uint64_t *zR, uint64_t *zM, uint32_t *snview, uint32_t *fansums,
                    uint32_t *bucks, int runs, double *secs) {
  using clk = std::chrono::steady_clock;
  const int tn = ((nitag < MXNITAG ? nitag : MXNITAG) + 3) / 4;
  const int64_t sizes[] = {(int64_t)837 * NSBINANG * 4, (int64_t)SEG0 * NSBINANG * 4,
                           (int64_t)nitag * 4, (int64_t)nitag * 4, (int64_t)nitag * 8,
                           (int64_t)nitag * 8, (int64_t)tn * SEG0 * NSBINS * 4,
                           (int64_t)64 * NCRS * 4, (int64_t)2 * nitag * NBUCKTS * 4};
  void *ptrs[] = {psino, ssrb, rprmt, rdlyd, zR, zM, snview, fansums, bucks};
  int64_t counted = 0;
  for (int r = -1; r < runs; r++) {
    auto t0 = clk::now();
    for (int k = 0; k < 9; k++) {
      char *p = (char *)ptrs[k];
      const int64_t sz = sizes[k];
#pragma omp parallel for schedule(static)
      for (int64_t c = 0; c < (sz + (1 << 20) - 1) >> 20; c++) {
        const int64_t o = c << 20, e = o + (1 << 20) < sz ? o + (1 << 20) : sz;
        memset(p + o, 0, e - o);
      }
    }
    counted = cb_hst(words, n, toff, nitag, 0, nitag, sn1_sn11, sn1_ssrb, sn1_rno, s2cF, psino,
                     ssrb, rprmt, rdlyd, zR, zM, snview, fansums, bucks);
    auto t1 = clk::now();
    if (r >= 0) secs[r] = std::chrono::duration<double>(t1 - t0).count();
  }
  return counted;
}

}  // extern "C"
