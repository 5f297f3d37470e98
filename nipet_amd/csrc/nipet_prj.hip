// Siddon forward / back projectors and the fused OSEM subset update for gfx950.
//
// Work decomposition (one workgroup = one transaxial bin of the subset):
//  - the bin's transaxial segment list (FP64-built, device resident) is
//    walked in chunks of CHUNK segments; for each chunk the z-columns
//    img[col][0:127] of its voxels are staged in LDS (127 contiguous floats
//    each: coalesced 508-byte reads) -- the FP gathers then never touch HBM;
//  - each lane owns Michelogram elements li (LOR a = ring r0 -> r1 and its
//    mirror b = r1 -> r0, which share their z-boundary parameters, reference
//    prj/src/prjf.cu:137-190), li = lane + 704*r for r < 3 (2074 elements);
//  - per segment a LOR needs its z-voxel w and the t of the next z-plane
//    boundary tb; a segment contributes p1 = clamp(tb - s, 0, L) to w and the
//    rest to w+1 (mirror: w_b and w_b-1), and a second boundary inside the
//    segment (possible only when atn/dz < the longest segment of the bin) is
//    handled by a wave-uniform slow path;
//  - FP: span-1 -> span-11 reduction in LDS, then one coalesced store per
//    sinogram bin (no global atomics; prj/src/prjf.cu:88-94 uses atomicAdd);
//  - BP: a gather over ring differences into registers (see k_bprj), one
//    coalesced global float atomic per (segment, plane) instead of one per
//    (LOR, piece) (prj/src/prjb.cu:70-82, 153-185).
#include "nipet_common.h"

namespace nipet {

constexpr int PRJ_THREADS = 704;  // 11 waves of 64
constexpr int LI_ROUNDS = 3;      // 3*704 = 2112 >= 2074 Michelogram elements
constexpr int CHUNK = 32;         // transaxial segments per LDS stage
constexpr int SUBSUM = 8;         // FP: float partial sums over 8 segments, folded in double
constexpr int ZSTRIDE = 128;      // LDS floats per staged z-column (127 + zero pad)
constexpr int ZGUARD = 8;         // zero guard floats on both sides of the stage
constexpr double TFAR = 1e300;    // boundary "never" (direct planes)
constexpr double SPAD = -1e30;    // start of a padding segment (contributes nothing)

struct GeomView {
  const int32_t *off;
  const SegRec *seg;
  const double *atn;
  const float *lmax;
  const float2 *li_z;
  const short2 *li_sn;
  const float *li_nosi;
  int naw, nli, ndirect, nsn;
  float vox_z;
};

static GeomView view(const Geom *g) {
  return GeomView{g->off, g->seg, g->atn, g->lmax, g->li_z, g->li_sn, g->li_nosi,
                  g->naw, g->nli, g->ndirect, g->nsn, g->vox_z};
}

// Per-lane state of one Michelogram element along a transaxial bin. The
// boundary position tb is kept in double: positions reach 67 cm along the ray
// where float32 resolves only 4e-6 cm, which would move weight between
// adjacent z-planes at object edges.
struct Lor {
  double tb;   // t of the next z boundary of LOR a (same t for its mirror)
  double inv;  // t between consecutive z boundaries (0: direct plane)
  int ia, ib;  // z-voxel of LOR a and of its mirror b (ia + ib = 2 r0 + 2 r1)
};

__device__ inline void lor_init(Lor &L, float2 z, double atn, double s0) {
  const double zeta0 = z.x, dz = z.y;  // exact (ring centres sit at 2r + 0.5)
  const int wsum = (int)floorf(z.x) + (int)floorf(z.x + z.y);
  if (z.y > 0.f) {
    L.inv = atn / dz;
    double w = floor(zeta0 + s0 / L.inv);
    L.tb = (w + 1.0 - zeta0) * L.inv;
    L.ia = (int)w;
  } else {
    L.inv = 0.0;
    L.tb = TFAR;
    L.ia = (int)floorf(z.x);
  }
  L.ib = wsum - L.ia;
}

// FP accumulation: pieces go into float partial sums folded into double
// every SUBSUM segments (repeated additions of nearly equal segment lengths
// into one float sum round in the same direction and bias the result).
struct FpAcc {
  float pa, pb;
  double aa, ab;
  __device__ void piece(const float *col, int ia, int ib, float p, int k) {
    pa = fmaf(p, col[ia + k], pa);
    pb = fmaf(p, col[ib - k], pb);
  }
  __device__ void direct(const float *col, int ia, float p) { pa = fmaf(p, col[ia], pa); }
  __device__ void fold() {
    aa += (double)pa;
    ab += (double)pb;
    pa = 0.f;
    pb = 0.f;
  }
};

// One segment with at most one z boundary inside it.
template <class Op, class F>
__device__ inline void step1(Lor &l, const SegRec &sg, F *col, Op &op) {
  const float q = (float)(l.tb - sg.s);
  const float p1 = __builtin_amdgcn_fmed3f(q, 0.f, sg.len);
  op.piece(col, l.ia, l.ib, p1, 0);
  op.piece(col, l.ia, l.ib, sg.len - p1, 1);
  const bool cr = q <= sg.len;
  l.ia += cr;
  l.ib -= cr;
  l.tb += cr ? l.inv : 0.0;
}

// One segment with up to two z boundaries inside it.
template <class Op, class F>
__device__ inline void step2(Lor &l, const SegRec &sg, F *col, Op &op) {
  const float q1 = (float)(l.tb - sg.s);
  const float q2 = (float)(l.tb + l.inv - sg.s);
  const float p1 = __builtin_amdgcn_fmed3f(q1, 0.f, sg.len);
  const float c2 = __builtin_amdgcn_fmed3f(q2, 0.f, sg.len);
  op.piece(col, l.ia, l.ib, p1, 0);
  op.piece(col, l.ia, l.ib, c2 - p1, 1);
  op.piece(col, l.ia, l.ib, sg.len - c2, 2);
  const int n = (q1 <= sg.len) + (q2 <= sg.len);
  l.ia += n;
  l.ib -= n;
  l.tb += (double)n * l.inv;
}

// Walk the CHUNK staged segments for this lane's element of round r.
template <class Op, class F>
__device__ inline void walk_chunk(Lor &l, Op &op, const SegRec *sbuf, F *zs, bool direct_wave,
                                  bool two_wave) {
  if (direct_wave) {
    F *col = zs + l.ia;
#pragma unroll
    for (int j = 0; j < CHUNK; j++) op.direct(col + j * ZSTRIDE, 0, sbuf[j].len);
    op.fold();
  } else if (two_wave) {
#pragma unroll 8
    for (int j = 0; j < CHUNK; j++) {
      step2(l, sbuf[j], zs + j * ZSTRIDE, op);
      if ((j + 1) % SUBSUM == 0) op.fold();
    }
  } else {
#pragma unroll
    for (int j = 0; j < CHUNK; j++) {
      step1(l, sbuf[j], zs + j * ZSTRIDE, op);
      if ((j + 1) % SUBSUM == 0) op.fold();
    }
  }
}

// Stage the CHUNK segment records of chunk c (padding past the end) and, for
// FP, their z-columns (BP: zero them).
__device__ inline void stage_chunk(float *zs, SegRec *sbuf, const GeomView &g, const float *img,
                                   int beg, int nseg, int c) {
  const int tid = threadIdx.x;
  if (tid < CHUNK) {
    int k = c + tid;
    SegRec pad;
    pad.s = SPAD;
    pad.len = 0.f;
    pad.colz = 0;
    sbuf[tid] = (k < nseg) ? g.seg[beg + k] : pad;
  }
  for (int idx = tid; idx < CHUNK * ZSTRIDE; idx += PRJ_THREADS) {
    int j = idx >> 7, z = idx & (ZSTRIDE - 1), k = c + j;
    float v = 0.f;
    if (img && k < nseg && z < NZ) v = img[g.seg[beg + k].colz + z];
    zs[idx] = v;
  }
}

__device__ inline float lor_sfac(float dz, float vox_z, double atn) {
  float az = dz * vox_z;
  float r = az / (float)atn;
  return sqrtf(1.f + r * r);
}

// ---------------------------------------------------------------------------
// Forward projection. MODE 0: emission; 1: attenuation factors;
// 2: OSEM ratio y / (FP + rho) with the 1e-6 guards of prj/src/recon.cu:252-283.
template <int MODE>
__global__ __launch_bounds__(PRJ_THREADS) void k_fprj(GeomView g, const float *__restrict__ img,
                                                      float *__restrict__ sino,
                                                      const int32_t *__restrict__ subs,
                                                      const uint16_t *__restrict__ prompts,
                                                      const float *__restrict__ rho) {
  __shared__ float zbuf[ZGUARD + CHUNK * ZSTRIDE + ZGUARD];
  __shared__ SegRec sbuf[CHUNK];
  __shared__ float bins[MAXNSN];
  float *zs = zbuf + ZGUARD;
  const int tid = threadIdx.x, wave = tid >> 6;
  const int row = blockIdx.x;
  const int aw = subs ? subs[row] : row;
  const int nsn = g.nsn;
  float *out = sino + (size_t)row * nsn;
  if (aw < 0 || aw >= g.naw) {  // invalid bin index: zero row
    for (int i = tid; i < nsn; i += PRJ_THREADS) out[i] = 0.f;
    return;
  }
  const int beg = g.off[aw], nseg = g.off[aw + 1] - beg;
  const double atn = g.atn[aw];
  const float lmax = g.lmax[aw];
  for (int i = tid; i < nsn; i += PRJ_THREADS) bins[i] = 0.f;
  if (tid < ZGUARD) {
    zbuf[tid] = 0.f;
    zbuf[ZGUARD + CHUNK * ZSTRIDE + tid] = 0.f;
  }
  const double s0 = nseg > 0 ? g.seg[beg].s : 0.0;

  Lor L[LI_ROUNDS];
  FpAcc A[LI_ROUNDS];
  bool two[LI_ROUNDS];
#pragma unroll
  for (int r = 0; r < LI_ROUNDS; r++) {
    int li = r * PRJ_THREADS + tid;
    float2 z = li < g.nli ? g.li_z[li] : make_float2(0.5f, 0.f);
    lor_init(L[r], z, atn, s0);
    A[r] = FpAcc{0.f, 0.f, 0.0, 0.0};
    two[r] = __any(L[r].inv > 0.0 && L[r].inv < (double)lmax);
  }

  for (int c = 0; c < nseg; c += CHUNK) {
    __syncthreads();
    stage_chunk(zs, sbuf, g, img, beg, nseg, c);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < LI_ROUNDS; r++) {
      const int li0 = r * PRJ_THREADS + wave * 64;  // wave-uniform
      if (li0 >= g.nli) continue;
      walk_chunk(L[r], A[r], sbuf, (const float *)zs, li0 + 64 <= g.ndirect, two[r]);
    }
  }

  // span-1 -> span-N reduction in LDS
#pragma unroll
  for (int r = 0; r < LI_ROUNDS; r++) {
    int li = r * PRJ_THREADS + tid;
    if (li < g.nli) {
      float2 z = g.li_z[li];
      short2 sn = g.li_sn[li];
      float sf = lor_sfac(z.y, g.vox_z, atn);
      float va = (float)A[r].aa * sf, vb = (float)A[r].ab * sf;
      if (MODE == 1) {
        float nosi = g.li_nosi[li];
        va = expf(-va) * nosi;
        vb = expf(-vb) * nosi;
      }
      atomicAdd(&bins[sn.x], va);
      if (z.y > 0.f) atomicAdd(&bins[sn.y], vb);
    }
  }
  __syncthreads();
  for (int i = tid; i < nsn; i += PRJ_THREADS) {
    float v = bins[i];
    if (MODE == 2) {
      size_t k = (size_t)aw * nsn + i;
      float e = v + rho[k];
      v = (e > -EPS && e < EPS) ? 0.f : (float)prompts[k] / e;
    }
    out[i] = v;
  }
}

// ---------------------------------------------------------------------------
// Back projection (adjoint of the emission FP) as a gather, accumulated into img.
//
// For one transaxial segment [s, s+L] and ring difference d, all LORs
// (r0, d) share the fractional z position 0.5 + 2 d t / atn (ring centres sit
// at zeta = 2 r + 0.5), so they share the base plane w0 = floor(0.5 + 2 d s/atn)
// and the pieces p1, p2, p3 between z boundaries; LOR (r0, d) puts piece k on
// plane w0 + 2 r0 + k and its mirror on plane 2 r0 + 2 d - w0 - k. A lane that
// owns the plane pair (2q, 2q+1) therefore receives, from ring difference d,
//   w0 even: 2q   <- p1 (X + Xb) + p3 (Y + Yb),   2q+1 <- p2 (X + Yb)
//   w0 odd:  2q   <- p2 (Y + Xb),   2q+1 <- p1 (X + Xb) + p3 (Y + Yb)
// with X = va[q - h], Y = va[q - h - 1], Xb = vb[q - e], Yb = vb[q - e + 1],
// h = w0 >> 1, e = (2d - w0) >> 1 (va/vb: the bin values of the LORs and of
// their mirrors times the 3D length factor, by r0). The pieces of (segment, d)
// are computed by lane d and broadcast with readlane. No atomics inside a ray;
// one coalesced global float atomic per (segment, plane) merges rays.
constexpr int BP_ROW = 72;          // LDS row of va / vb: r0 = -4 .. 67
constexpr int BP_OFF = 4;           // r0 = 0 sits at BP_OFF
constexpr int BP_MAXD = 64;         // ring differences handled (0 .. 63)

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ inline float rdlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__global__ __launch_bounds__(PRJ_THREADS) void k_bprj(GeomView g, float *__restrict__ img,
                                                      const float *__restrict__ sino,
                                                      const int32_t *__restrict__ subs) {
  __shared__ float va[BP_MAXD * BP_ROW];
  __shared__ float vb[BP_MAXD * BP_ROW];
  __shared__ float zs[CHUNK * ZSTRIDE];
  __shared__ SegRec sbuf[CHUNK];
  __shared__ int dmax_s;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int row = blockIdx.x;
  const int aw = subs ? subs[row] : row;
  if (aw < 0 || aw >= g.naw) return;
  const int nsn = g.nsn;
  const float *in = sino + (size_t)row * nsn;
  const int beg = g.off[aw], nseg = g.off[aw + 1] - beg;
  const double atn = g.atn[aw];

  // va / vb tables: zero, then the LOR values (bin value * 3D length factor)
  for (int i = tid; i < BP_MAXD * BP_ROW; i += PRJ_THREADS) {
    va[i] = 0.f;
    vb[i] = 0.f;
  }
  if (tid == 0) dmax_s = 0;
  __syncthreads();
  for (int li = tid; li < g.nli; li += PRJ_THREADS) {
    float2 z = g.li_z[li];
    short2 sn = g.li_sn[li];
    int r0 = (int)floorf(z.x) >> 1, d = (int)z.y >> 1;
    float sf = lor_sfac(z.y, g.vox_z, atn);
    int k = d * BP_ROW + BP_OFF + r0;
    va[k] = in[sn.x] * sf;
    if (d > 0) vb[k] = in[sn.y] * sf;
    atomicMax(&dmax_s, d);
  }
  __syncthreads();
  const int dmax = dmax_s;
  // per-lane constant: d = lane, slope 2d/atn of zeta along t
  const double m_lane = 2.0 * lane / atn;
  const double inv_lane = lane > 0 ? atn / (2.0 * lane) : 0.0;
  const int q = lane;

  for (int c = 0; c < nseg; c += CHUNK) {
    __syncthreads();
    if (tid < CHUNK) {
      int k = c + tid;
      SegRec pad;
      pad.s = 0.0;
      pad.len = 0.f;
      pad.colz = 0;
      sbuf[tid] = (k < nseg) ? g.seg[beg + k] : pad;
    }
    __syncthreads();
    for (int j = wave; j < CHUNK; j += PRJ_THREADS / 64) {
      const SegRec sg = sbuf[j];
      // lane d: base plane and pieces of ring difference d on this segment
      float P1 = 0.f, P2 = 0.f, P3 = 0.f;
      int W0 = 0;
      if (lane >= 1 && lane <= dmax) {
        double w0 = floor(0.5 + m_lane * sg.s);
        double tb = (w0 + 0.5) * inv_lane;
        float q1 = (float)(tb - sg.s), q2 = (float)(tb + inv_lane - sg.s);
        float p1 = __builtin_amdgcn_fmed3f(q1, 0.f, sg.len);
        float c2 = __builtin_amdgcn_fmed3f(q2, 0.f, sg.len);
        P1 = p1;
        P2 = c2 - p1;
        P3 = sg.len - c2;
        W0 = (int)w0;
      }
      // direct planes: plane 2q <- L * va0[q]
      float acc0 = sg.len * va[BP_OFF + q], acc1 = 0.f;
      for (int d = 1; d <= dmax; d++) {
        const float p1 = rdlane(P1, d), p2 = rdlane(P2, d), p3 = rdlane(P3, d);
        const int w0 = __builtin_amdgcn_readlane(W0, d);
        const int h = w0 >> 1, e = (2 * d - w0) >> 1;
        const float *ra = va + d * BP_ROW + BP_OFF;
        const float *rb = vb + d * BP_ROW + BP_OFF;
        // clamped indices always land on zero entries (r0 < 0 or r0 > 63 - d)
        const int ia = clampi(q - h, -3, 64);
        const int ib = clampi(q - e, -3, 64);
        const float X = ra[ia], Y = ra[ia - 1], Xb = rb[ib], Yb = rb[ib + 1];
        if (w0 & 1) {
          acc0 = fmaf(p2, Y + Xb, acc0);
          acc1 = fmaf(p1, X + Xb, acc1);
          acc1 = fmaf(p3, Y + Yb, acc1);
        } else {
          acc0 = fmaf(p1, X + Xb, acc0);
          acc0 = fmaf(p3, Y + Yb, acc0);
          acc1 = fmaf(p2, X + Yb, acc1);
        }
      }
      zs[j * ZSTRIDE + 2 * q] = acc0;
      zs[j * ZSTRIDE + 2 * q + 1] = acc1;
    }
    __syncthreads();
    // flush the chunk: one coalesced global atomic per (segment, plane)
    const int nj = min(CHUNK, nseg - c);
    for (int idx = tid; idx < nj * ZSTRIDE; idx += PRJ_THREADS) {
      int j = idx >> 7, z = idx & (ZSTRIDE - 1);
      float v = zs[idx];
      if (z < NZ && v != 0.f) unsafeAtomicAdd(img + sbuf[j].colz + z, v);
    }
  }
}

// ---------------------------------------------------------------------------
// Element-wise kernels.
__global__ void k_fill(float *p, float v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// rho = (r + s) / (a n) (0 where |a n| < eps), an = a n (prj/src/recon.cu:381-398)
__global__ void k_osem_corr(float *rho, float *an, const float *a, const float *n,
                            const float *r, const float *s, int64_t nel) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < nel; i += stride) {
    float x = a[i] * n[i];
    float rs = (r ? r[i] : 0.f) + (s ? s[i] : 0.f);
    an[i] = x;
    rho[i] = (x > -EPS && x < EPS) ? 0.f : rs / x;
  }
}

// img = mask ? img * (b / sens) : 0 with b/sens = 0 where |sens| < eps
// (prj/src/recon.cu:234-247, 300-315).
__global__ void k_osem_update(float *img, const float *b, const float *sens, const uint8_t *mask,
                              int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    float sv = sens[i];
    float q = (sv > -EPS && sv < EPS) ? 0.f : b[i] / sv;
    img[i] = mask[i] ? img[i] * q : 0.f;
  }
}

// Siemens (nsn, 252*344) -> device (naw, nsn) gather through aw2ali
// (d_remgaps, src/scanner_0.cu:94-105); input float32 (dtype 0) or uint16 (1).
template <typename T>
__global__ void k_remgaps(float *out, const T *in, const int32_t *aw2ali, int naw, int nsn,
                          int nbinang) {
  int64_t n = (int64_t)naw * nsn;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    int aw = (int)(i / nsn), s = (int)(i - (int64_t)aw * nsn);
    out[i] = (float)in[(int64_t)s * nbinang + aw2ali[aw]];
  }
}

// device (naw, nsn) -> Siemens (nsn, 252*344) scatter, gaps stay zero
// (d_putgaps, src/scanner_0.cu:158-166, plus the transpose of mmraux.py:1141).
__global__ void k_putgaps(float *out, const float *in, const int32_t *aw2ali, int naw, int nsn,
                          int nbinang) {
  int64_t n = (int64_t)naw * nsn;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    int aw = (int)(i / nsn), s = (int)(i - (int64_t)aw * nsn);
    out[(int64_t)s * nbinang + aw2ali[aw]] = in[i];
  }
}

static int ew_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (int)(b < 8192 ? b : 8192);
}

}  // namespace nipet

using namespace nipet;

static int check_geom(const NipetGeom *g, const char *fn) {
  if (!g) {
    set_error("%s: NULL geometry", fn);
    return NIPET_EINVAL;
  }
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != g->device) {
    set_error("%s: geometry lives on device %d, current device is %d", fn, g->device, dev);
    return NIPET_ESTATE;
  }
  return NIPET_OK;
}

static int check_nprj(const NipetGeom *g, const int32_t *subs, int nprj, const char *fn) {
  if (nprj <= 0 || (!subs && nprj != g->naw)) {
    set_error("%s: nprj=%d (must be > 0, and = naw=%d when subs is NULL)", fn, nprj, g->naw);
    return NIPET_EINVAL;
  }
  return NIPET_OK;
}

extern "C" {

int nipet_fprj(const NipetGeom *g, float *sino, const float *img, const int32_t *subs, int nprj,
               int att, void *stream) {
  int rc = check_geom(g, "nipet_fprj");
  if (rc) return rc;
  if ((rc = check_nprj(g, subs, nprj, "nipet_fprj"))) return rc;
  if (!sino || !img || (att != 0 && att != 1)) {
    set_error("nipet_fprj: NULL buffer or att not in {0,1}");
    return NIPET_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  prof_begin(PK_FPRJ, st);
  if (att)
    k_fprj<1><<<nprj, PRJ_THREADS, 0, st>>>(view(g), img, sino, subs, nullptr, nullptr);
  else
    k_fprj<0><<<nprj, PRJ_THREADS, 0, st>>>(view(g), img, sino, subs, nullptr, nullptr);
  prof_end(PK_FPRJ, st);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_bprj(const NipetGeom *g, float *img, const float *sino, const int32_t *subs, int nprj,
               void *stream) {
  int rc = check_geom(g, "nipet_bprj");
  if (rc) return rc;
  if ((rc = check_nprj(g, subs, nprj, "nipet_bprj"))) return rc;
  if (!sino || !img) {
    set_error("nipet_bprj: NULL buffer");
    return NIPET_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  prof_begin(PK_FILL, st);
  k_fill<<<ew_blocks(NVOX), 256, 0, st>>>(img, 0.f, NVOX);
  prof_end(PK_FILL, st);
  prof_begin(PK_BPRJ, st);
  k_bprj<<<nprj, PRJ_THREADS, 0, st>>>(view(g), img, sino, subs);
  prof_end(PK_BPRJ, st);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_remgaps(float *out, const void *in, int in_dtype, const int32_t *aw2ali, int naw,
                  int nsn, int nbinang, void *stream) {
  if (!out || !in || !aw2ali || naw <= 0 || nsn <= 0 || nbinang < naw ||
      (in_dtype != 0 && in_dtype != 1)) {
    set_error("nipet_remgaps: bad argument");
    return NIPET_EINVAL;
  }
  int64_t n = (int64_t)naw * nsn;
  hipStream_t st = (hipStream_t)stream;
  if (in_dtype == 0)
    k_remgaps<float><<<ew_blocks(n), 256, 0, st>>>(out, (const float *)in, aw2ali, naw, nsn,
                                                   nbinang);
  else
    k_remgaps<uint16_t><<<ew_blocks(n), 256, 0, st>>>(out, (const uint16_t *)in, aw2ali, naw,
                                                      nsn, nbinang);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_putgaps(float *out, const float *in, const int32_t *aw2ali, int naw, int nsn,
                  int nbinang, void *stream) {
  if (!out || !in || !aw2ali || naw <= 0 || nsn <= 0 || nbinang < naw) {
    set_error("nipet_putgaps: bad argument");
    return NIPET_EINVAL;
  }
  int64_t n = (int64_t)naw * nsn, nout = (int64_t)nsn * nbinang;
  hipStream_t st = (hipStream_t)stream;
  k_fill<<<ew_blocks(nout), 256, 0, st>>>(out, 0.f, nout);
  k_putgaps<<<ew_blocks(n), 256, 0, st>>>(out, in, aw2ali, naw, nsn, nbinang);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_osem_corr(float *rho, float *an, const float *a, const float *n, const float *r,
                    const float *s, int64_t nelem, void *stream) {
  if (!rho || !an || !a || !n || nelem <= 0) {
    set_error("nipet_osem_corr: NULL buffer or nelem <= 0");
    return NIPET_EINVAL;
  }
  k_osem_corr<<<ew_blocks(nelem), 256, 0, (hipStream_t)stream>>>(rho, an, a, n, r, s, nelem);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_osem_subset(const NipetGeom *g, float *img, const uint16_t *prompts, const float *rho,
                      const float *sens, const uint8_t *mask, const int32_t *subs, int nprj,
                      float *work_sino, float *work_img, void *stream) {
  int rc = check_geom(g, "nipet_osem_subset");
  if (rc) return rc;
  if (!subs || nprj <= 0 || nprj > g->naw) {
    set_error("nipet_osem_subset: subs must be given with 0 < nprj <= naw");
    return NIPET_EINVAL;
  }
  if (!img || !prompts || !rho || !sens || !mask || !work_sino || !work_img) {
    set_error("nipet_osem_subset: NULL buffer");
    return NIPET_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  prof_begin(PK_FPRJ, st);
  k_fprj<2><<<nprj, PRJ_THREADS, 0, st>>>(view(g), img, work_sino, subs, prompts, rho);
  prof_end(PK_FPRJ, st);
  prof_begin(PK_FILL, st);
  k_fill<<<ew_blocks(NVOX), 256, 0, st>>>(work_img, 0.f, NVOX);
  prof_end(PK_FILL, st);
  prof_begin(PK_BPRJ, st);
  k_bprj<<<nprj, PRJ_THREADS, 0, st>>>(view(g), work_img, work_sino, subs);
  prof_end(PK_BPRJ, st);
  prof_begin(PK_OSEM_UPDATE, st);
  k_osem_update<<<ew_blocks(NVOX), 256, 0, st>>>(img, work_img, sens, mask, NVOX);
  prof_end(PK_OSEM_UPDATE, st);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

}  // extern "C"
