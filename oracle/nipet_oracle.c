/*
 * FP64 CPU oracle for the NIPET projection / reconstruction / histogramming
 * path on the Siemens mMR geometry.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product library links or calls
 * this file: it is used by tests/ (as the checker), by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg.
 *
 * What it restates (reference file:line):
 *  - transaxial ray tracing: prj/src/tprj.cu:13-172 (sddn_tx). Rays join the
 *    centres of the two crystal faces; the chord is clipped to the FOV circle
 *    r^2 = TFOV2; every x- and y-grid-line crossing strictly inside the chord
 *    is collected; the partial segments before the first and after the last
 *    crossing are skipped (prj/src/prjf.cu:59-76 starts at min(tr1,tc1) and
 *    stops at the last crossing).
 *    Here: all crossing parameters are computed in double, sorted by a merge of
 *    the x- and y-crossing sequences (x first on an exact tie), and the voxel
 *    is stepped by +-1 per crossing from the voxel that holds the midpoint of
 *    [t1, first crossing].
 *  - axial tracing: prj/src/prjf.cu:33-206 (fprj_drct/fprj_oblq). A span-1 LOR
 *    between rings r0 (at crystal s2c[.][0]) and r1 has z-voxel coordinate
 *    zeta(t) = zeta0 + (zeta1-zeta0) t/atn (zeta in voxel units, 2r+0.5 at
 *    ring centres); each transaxial segment is split at EVERY z-plane it
 *    crosses (the reference splits at most once per segment and spaces the
 *    splits by (tz2-tz1)/nz with nz truncated; this oracle uses the exact
 *    intersection, which is what those formulas approximate).
 *    Weight of a piece = transaxial length * sqrt(1+(az/atn)^2).
 *    The mirror LOR (rings swapped) uses plane 2(r0+r1) - w for every piece.
 *  - span-11: emission bins sum their span-1 LORs; attenuation bins average
 *    exp(-line integral) over their span-1 LORs (prj/src/prjf.cu:88-94,193-204).
 *  - back projection: exact adjoint of the emission projector
 *    (prj/src/prjb.cu:33-187).
 *
 * Image layout (y=320, x=320, z=127), index w + 127*(u + 320*v),
 * u = 160 + floor(x/vox), v = 160 - ceil(y/vox) (prj/src/tprj.cu:62,82).
 * Sinogram layout [row][sino], sino fastest.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NX 320
#define NY 320
#define NZ 127
#define NSB 344
#define NSA 252
#define MAXCROSS 2048

static const double VOX = (double)0.208626f;   /* transaxial voxel [cm] (float constant) */
static const double VOXZ = 0.203125;           /* axial voxel [cm] */
static const double AXR = 0.40625;             /* ring pitch [cm] */
static const double TFOV2 = 890.0;             /* FOV radius^2 [cm^2] */

typedef struct {
  double px, py, ax, ay, atn, t1, t2;
} ray_t;

static void ray_setup(const float *crs, int c1, int c2, ray_t *r) {
  double x1 = 0.5 * ((double)crs[4 * c1 + 0] + (double)crs[4 * c1 + 2]);
  double y1 = 0.5 * ((double)crs[4 * c1 + 1] + (double)crs[4 * c1 + 3]);
  double x2 = 0.5 * ((double)crs[4 * c2 + 0] + (double)crs[4 * c2 + 2]);
  double y2 = 0.5 * ((double)crs[4 * c2 + 1] + (double)crs[4 * c2 + 3]);
  double dx = x2 - x1, dy = y2 - y1;
  double atn = sqrt(dx * dx + dy * dy);
  r->px = x1;
  r->py = y1;
  r->ax = dx / atn;
  r->ay = dy / atn;
  r->atn = atn;
  double b = x1 * r->ax + y1 * r->ay;
  double c = x1 * x1 + y1 * y1 - TFOV2;
  double disc = b * b - c;
  double sq = disc > 0 ? sqrt(disc) : 0.0;
  r->t1 = -b - sq;
  r->t2 = -b + sq;
}

/* crossing parameters of the lines  coord = i*VOX  strictly inside (t1,t2),
 * in increasing t. Returns the count. */
static int line_crossings(double p, double a, double t1, double t2, double *out) {
  int n = 0;
  if (a == 0.0) return 0;
  double c1 = p + a * t1, c2 = p + a * t2;
  double lo = c1 < c2 ? c1 : c2, hi = c1 < c2 ? c2 : c1;
  long i0 = (long)floor(lo / VOX) - 1, i1 = (long)ceil(hi / VOX) + 1;
  if (a > 0) {
    for (long i = i0; i <= i1; i++) {
      double t = ((double)i * VOX - p) / a;
      if (t > t1 && t < t2 && n < MAXCROSS) out[n++] = t;
    }
  } else {
    for (long i = i1; i >= i0; i--) {
      double t = ((double)i * VOX - p) / a;
      if (t > t1 && t < t2 && n < MAXCROSS) out[n++] = t;
    }
  }
  return n;
}

/* Segment list of one transaxial bin: nseg segments, each with its voxel
 * column index (u + 320 v), start parameter t and length. Returns nseg or -1. */
static int trace_tx(const float *crs, int c1, int c2, ray_t *r, int32_t *col, double *ts,
                    double *len) {
  double tx[MAXCROSS], ty[MAXCROSS];
  ray_setup(crs, c1, c2, r);
  int nx = line_crossings(r->px, r->ax, r->t1, r->t2, tx);
  int ny = line_crossings(r->py, r->ay, r->t1, r->t2, ty);
  int ncr = nx + ny;
  if (ncr < 2) return 0;
  /* first crossing decides the entry voxel: midpoint of [t1, first crossing] */
  double tf = (nx > 0 && (ny == 0 || tx[0] <= ty[0])) ? tx[0] : ty[0];
  double tm = 0.5 * (r->t1 + tf);
  long u = NX / 2 + (long)floor((r->px + r->ax * tm) / VOX);
  long v = NY / 2 - (long)ceil((r->py + r->ay * tm) / VOX);
  int su = r->ax > 0 ? 1 : -1;
  int sv = r->ay > 0 ? 1 : -1;
  int ix = 0, iy = 0, nseg = 0;
  double tprev = 0;
  for (int k = 0; k < ncr; k++) {
    double t;
    if (ix < nx && (iy >= ny || tx[ix] <= ty[iy])) {
      t = tx[ix++];
      if (k > 0) {
        col[nseg] = (int32_t)(u + NX * v);
        ts[nseg] = tprev;
        len[nseg] = t - tprev;
        nseg++;
      }
      u += su;
    } else {
      t = ty[iy++];
      if (k > 0) {
        col[nseg] = (int32_t)(u + NX * v);
        ts[nseg] = tprev;
        len[nseg] = t - tprev;
        nseg++;
      }
      v -= sv;
    }
    if (u < 0 || u >= NX || v < 0 || v >= NY) return -1;
    tprev = t;
  }
  return nseg;
}

/* Transaxial lists for a set of transaxial bins (rays[i] indexes s2c).
 * Pass col == NULL to count only: offsets[i+1]-offsets[i] = number of
 * segments of ray i. atn (optional) receives the crystal distance. */
int orc_tx_lists(const float *crs, const int16_t *s2c, const int32_t *rays, int nray,
                 int64_t *offsets, int32_t *col, double *ts, double *len, double *atn,
                 double *t12) {
  int32_t bc[MAXCROSS];
  double bt[MAXCROSS], bl[MAXCROSS];
  offsets[0] = 0;
  for (int i = 0; i < nray; i++) {
    ray_t r;
    int aw = rays[i];
    int n = trace_tx(crs, s2c[2 * aw], s2c[2 * aw + 1], &r, bc, bt, bl);
    if (n < 0) return -1;
    if (col) {
      memcpy(col + offsets[i], bc, n * sizeof(int32_t));
      memcpy(ts + offsets[i], bt, n * sizeof(double));
      memcpy(len + offsets[i], bl, n * sizeof(double));
    }
    if (atn) atn[i] = r.atn;
    if (t12) {
      t12[2 * i] = r.t1;
      t12[2 * i + 1] = r.t2;
    }
    offsets[i + 1] = offsets[i] + n;
  }
  return 0;
}

/* Axial description of one Michelogram element li: LOR a runs from ring
 * r0 (at the first crystal) to ring r1, its mirror b from r1 to r0. */
typedef struct {
  double z0, slope, sfac; /* zeta at t=0, d zeta/dt, 3D length factor */
  int wsum;               /* w_a + w_b for every piece (2 r0 + 2 r1) */
  int direct;
} lor_t;

static void lor_setup(const float *li2rng, int li, double atn, lor_t *L) {
  double za = (double)li2rng[2 * li] + 0.5 * AXR;     /* ring centre z [cm] */
  double zb = (double)li2rng[2 * li + 1] + 0.5 * AXR;
  double zeta_a = za / VOXZ + 0.5 * NZ;
  double zeta_b = zb / VOXZ + 0.5 * NZ;
  double az = zb - za;
  L->z0 = zeta_a;
  L->slope = (zeta_b - zeta_a) / atn;
  L->sfac = sqrt(1.0 + (az / atn) * (az / atn));
  L->wsum = (int)floor(zeta_a) + (int)floor(zeta_b);
  L->direct = (zb == za);
}

/* Walk the pieces of LOR li over a ray's segments, calling ACT(w, col, weight)
 * for LOR a and ACTB for its mirror. */
#define WALK_LOR(L, nseg, col, ts, len, ACT, NPIECE)                                        \
  do {                                                                                     \
    for (int k = 0; k < (nseg); k++) {                                                     \
      double s = ts[k], e = ts[k] + len[k];                                                \
      double zs = (L).z0 + (L).slope * s;                                                  \
      long w = (long)floor(zs);                                                            \
      if ((L).slope <= 0.0) {                                                              \
        ACT(w, col[k], len[k] * (L).sfac);                                                 \
        NPIECE++;                                                                          \
        continue;                                                                          \
      }                                                                                    \
      double pos = s;                                                                      \
      for (;;) {                                                                           \
        double tb = ((double)(w + 1) - (L).z0) / (L).slope;                                \
        if (tb < e) {                                                                      \
          ACT(w, col[k], (tb - pos) * (L).sfac);                                           \
          NPIECE++;                                                                        \
          pos = tb;                                                                        \
          w++;                                                                             \
        } else {                                                                           \
          ACT(w, col[k], (e - pos) * (L).sfac);                                            \
          NPIECE++;                                                                        \
          break;                                                                           \
        }                                                                                  \
      }                                                                                    \
    }                                                                                      \
  } while (0)

/* Forward projection of rays[0..nray) (rows of sino). att=0: emission (span
 * bins sum their LORs); att=1: exp(-integral), bins average their LORs.
 * li2sn (nli,2) maps LOR a / mirror b to the output sinogram; li2nos (nli)
 * is the number of span-1 LORs per output bin. nterms (optional) receives
 * the number of (LOR, voxel) pieces. */
static int fprj_impl(const float *imgf, const double *imgd, double *sino, const int32_t *rays,
                     int nray, const float *crs, const int16_t *s2c, const float *li2rng,
                     const int16_t *li2sn, const int8_t *li2nos, int nli, int nsn, int att,
                     int64_t *nterms) {
  int64_t tot = 0;
  int err = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : tot)
  for (int i = 0; i < nray; i++) {
    int32_t col[MAXCROSS];
    double ts[MAXCROSS], len[MAXCROSS];
    ray_t r;
    int aw = rays[i];
    int nseg = trace_tx(crs, s2c[2 * aw], s2c[2 * aw + 1], &r, col, ts, len);
    double *out = sino + (size_t)i * nsn;
    memset(out, 0, nsn * sizeof(double));
    if (nseg < 0) {
      err = 1;
      continue;
    }
    for (int li = 0; li < nli; li++) {
      lor_t L;
      lor_setup(li2rng, li, r.atn, &L);
      double acc = 0, accb = 0;
      int64_t np = 0;
#define IMG(k) (imgd ? imgd[k] : (double)imgf[k])
#define ACC_A(w, c, wt) (acc += (wt) * IMG((size_t)(c) * NZ + (w)),                          \
                         accb += (wt) * IMG((size_t)(c) * NZ + (L.wsum - (w))))
      WALK_LOR(L, nseg, col, ts, len, ACC_A, np);
#undef ACC_A
#undef IMG
      if (L.direct) {
        tot += np;
        if (att)
          out[li2sn[2 * li]] += exp(-acc) / (double)li2nos[li];
        else
          out[li2sn[2 * li]] += acc;
      } else {
        tot += 2 * np;
        if (att) {
          out[li2sn[2 * li]] += exp(-acc) / (double)li2nos[li];
          out[li2sn[2 * li + 1]] += exp(-accb) / (double)li2nos[li];
        } else {
          out[li2sn[2 * li]] += acc;
          out[li2sn[2 * li + 1]] += accb;
        }
      }
    }
  }
  if (nterms) *nterms = tot;
  return err ? -1 : 0;
}

int orc_fprj(const float *img, double *sino, const int32_t *rays, int nray, const float *crs,
             const int16_t *s2c, const float *li2rng, const int16_t *li2sn, const int8_t *li2nos,
             int nli, int nsn, int att, int64_t *nterms) {
  return fprj_impl(img, NULL, sino, rays, nray, crs, s2c, li2rng, li2sn, li2nos, nli, nsn, att,
                   nterms);
}

/* Same, for a double-precision image. */
int orc_fprj_d(const double *img, double *sino, const int32_t *rays, int nray, const float *crs,
               const int16_t *s2c, const float *li2rng, const int16_t *li2sn,
               const int8_t *li2nos, int nli, int nsn, int att, int64_t *nterms) {
  return fprj_impl(NULL, img, sino, rays, nray, crs, s2c, li2rng, li2sn, li2nos, nli, nsn, att,
                   nterms);
}

/* Back projection (adjoint of the emission forward projection) of
 * sino[row][sn] over rays[row] into img (accumulated; caller zeroes it). */
int orc_bprj(double *img, const double *sino, const int32_t *rays, int nray, const float *crs,
             const int16_t *s2c, const float *li2rng, const int16_t *li2sn, int nli, int nsn,
             int nthreads) {
  size_t nvox = (size_t)NX * NY * NZ;
  int nt = 1;
#ifdef _OPENMP
  nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
  double **priv = (double **)calloc(nt, sizeof(double *));
  int err = 0;
  for (int t = 0; t < nt; t++) {
    priv[t] = (double *)calloc(nvox, sizeof(double));
    if (!priv[t]) err = 1;
  }
  if (!err) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
    for (int i = 0; i < nray; i++) {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      double *im = priv[tid];
      int32_t col[MAXCROSS];
      double ts[MAXCROSS], len[MAXCROSS];
      ray_t r;
      int aw = rays[i];
      int nseg = trace_tx(crs, s2c[2 * aw], s2c[2 * aw + 1], &r, col, ts, len);
      if (nseg < 0) {
        err = 1;
        continue;
      }
      const double *in = sino + (size_t)i * nsn;
      for (int li = 0; li < nli; li++) {
        lor_t L;
        lor_setup(li2rng, li, r.atn, &L);
        double va = in[li2sn[2 * li]];
        double vb = L.direct ? 0.0 : in[li2sn[2 * li + 1]];
        int64_t np = 0;
#define ADD_A(w, c, wt) (im[(size_t)(c) * NZ + (w)] += (wt) * va,                            \
                         im[(size_t)(c) * NZ + (L.wsum - (w))] += (wt) * vb)
        WALK_LOR(L, nseg, col, ts, len, ADD_A, np);
#undef ADD_A
        (void)np;
      }
    }
    for (int t = 0; t < nt; t++)
      for (size_t j = 0; j < nvox; j++) img[j] += priv[t][j];
  }
  for (int t = 0; t < nt; t++) free(priv[t]);
  free(priv);
  return err ? -1 : 0;
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---------------------------------------------------------------------------
 * Finite-difference checker of the Poisson log-likelihood gradient.
 *
 * L(x, mu) = sum over rays[] and bins b of  ybar_b - y_b log ybar_b  (the
 * y_b log term only where y_b > 0), ybar_b = a_b(mu) n_b p_b(x) + rs_b, with p
 * the emission projection of x and a the attenuation projection of mu (both
 * as in fprj_impl, FP64). grad[i] = (L(v + h) - L(v - h)) / (2 h) for voxel
 * vox[i] of x (wrt = 0) or of mu (wrt = 1); the other voxels and image are
 * held. Only the rays whose transaxial list holds the voxel's column are
 * re-evaluated: the terms of the other rays cancel exactly in the difference.
 */

/* projection of one ray into out[nsn] (zeroed here) with img[ov] replaced by
 * ovv (ov < 0: no replacement) */
static void ray_proj(const double *img, int64_t ov, double ovv, const ray_t *r, int nseg,
                     const int32_t *col, const double *ts, const double *len,
                     const float *li2rng, const int16_t *li2sn, const int8_t *li2nos, int nli,
                     int nsn, int att, double *out) {
  memset(out, 0, nsn * sizeof(double));
  for (int li = 0; li < nli; li++) {
    lor_t L;
    lor_setup(li2rng, li, r->atn, &L);
    double acc = 0, accb = 0;
    int64_t np = 0;
#define IMGV(k) ((int64_t)(k) == ov ? ovv : img[k])
#define ACC_A(w, c, wt) (acc += (wt) * IMGV((size_t)(c) * NZ + (w)),                          \
                         accb += (wt) * IMGV((size_t)(c) * NZ + (L.wsum - (w))))
    WALK_LOR(L, nseg, col, ts, len, ACC_A, np);
#undef ACC_A
#undef IMGV
    (void)np;
    double fa = att ? exp(-acc) / (double)li2nos[li] : acc;
    out[li2sn[2 * li]] += fa;
    if (!L.direct) out[li2sn[2 * li + 1]] += att ? exp(-accb) / (double)li2nos[li] : accb;
  }
}

static double ray_nll(const double *p, const double *a, const double *y, const double *n,
                      const double *rs, int nsn) {
  double s = 0;
  for (int b = 0; b < nsn; b++) {
    double yb = a[b] * n[b] * p[b] + rs[b];
    s += yb;
    if (y[b] > 0) s -= y[b] * log(yb);
  }
  return s;
}

int orc_nll_fd(const double *x, const double *mu, const int32_t *rays, int nray,
               const double *y, const double *nrm, const double *rs, const float *crs,
               const int16_t *s2c, const float *li2rng, const int16_t *li2sn,
               const int8_t *li2nos, int nli, int nsn, const int64_t *vox, int nv, int wrt,
               double h, double *grad) {
  /* per ray: transaxial list, column mask, base projections */
  int32_t **cols = (int32_t **)calloc(nray, sizeof(void *));
  double **tss = (double **)calloc(nray, sizeof(void *));
  double **lns = (double **)calloc(nray, sizeof(void *));
  double *pbase = (double *)calloc((size_t)nray * nsn, sizeof(double));
  double *abase = (double *)calloc((size_t)nray * nsn, sizeof(double));
  uint8_t *cmask = (uint8_t *)calloc((size_t)nray * NX * NY, 1);
  ray_t *rr = (ray_t *)calloc(nray, sizeof(ray_t));
  int *nsg = (int *)calloc(nray, sizeof(int));
  int err = !cols || !tss || !lns || !pbase || !abase || !cmask || !rr || !nsg;
  for (int i = 0; i < nray && !err; i++) {
    cols[i] = (int32_t *)malloc(MAXCROSS * sizeof(int32_t));
    tss[i] = (double *)malloc(MAXCROSS * sizeof(double));
    lns[i] = (double *)malloc(MAXCROSS * sizeof(double));
    if (!cols[i] || !tss[i] || !lns[i]) {
      err = 1;
      break;
    }
    int aw = rays[i];
    nsg[i] = trace_tx(crs, s2c[2 * aw], s2c[2 * aw + 1], &rr[i], cols[i], tss[i], lns[i]);
    if (nsg[i] < 0) {
      err = 1;
      break;
    }
    for (int k = 0; k < nsg[i]; k++) cmask[(size_t)i * NX * NY + cols[i][k]] = 1;
  }
  if (!err) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < nray; i++) {
      ray_proj(x, -1, 0.0, &rr[i], nsg[i], cols[i], tss[i], lns[i], li2rng, li2sn, li2nos, nli,
               nsn, 0, pbase + (size_t)i * nsn);
      ray_proj(mu, -1, 0.0, &rr[i], nsg[i], cols[i], tss[i], lns[i], li2rng, li2sn, li2nos, nli,
               nsn, 1, abase + (size_t)i * nsn);
    }
#pragma omp parallel
    {
      double *q = (double *)malloc(nsn * sizeof(double));
#pragma omp for schedule(dynamic, 1)
      for (int k = 0; k < nv; k++) {
        int64_t v = vox[k];
        int c = (int)(v / NZ);
        const double *img = wrt ? mu : x;
        double d = 0;
        for (int i = 0; i < nray; i++) {
          if (!cmask[(size_t)i * NX * NY + c]) continue;
          size_t o = (size_t)i * nsn;
          double lp, lm;
          ray_proj(img, v, img[v] + h, &rr[i], nsg[i], cols[i], tss[i], lns[i], li2rng, li2sn,
                   li2nos, nli, nsn, wrt, q);
          lp = wrt ? ray_nll(pbase + o, q, y + o, nrm + o, rs + o, nsn)
                   : ray_nll(q, abase + o, y + o, nrm + o, rs + o, nsn);
          ray_proj(img, v, img[v] - h, &rr[i], nsg[i], cols[i], tss[i], lns[i], li2rng, li2sn,
                   li2nos, nli, nsn, wrt, q);
          lm = wrt ? ray_nll(pbase + o, q, y + o, nrm + o, rs + o, nsn)
                   : ray_nll(q, abase + o, y + o, nrm + o, rs + o, nsn);
          d += lp - lm;
        }
        grad[k] = d / (2.0 * h);
      }
      free(q);
    }
  }
  for (int i = 0; i < nray && cols; i++) {
    free(cols[i]);
    free(tss[i]);
    free(lns[i]);
  }
  free(cols);
  free(tss);
  free(lns);
  free(pbase);
  free(abase);
  free(cmask);
  free(rr);
  free(nsg);
  return err ? -1 : 0;
}
