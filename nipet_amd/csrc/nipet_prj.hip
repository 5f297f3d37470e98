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

constexpr int PRJ_THREADS = 704;  // FP: 11 waves of 64
constexpr int LI_ROUNDS = 3;      // 3*704 = 2112 >= 2074 Michelogram elements
constexpr int BP_THREADS = 512;   // BP: 8 waves
constexpr int SUBSUM = 8;         // FP: float partial sums over 8 segments
constexpr int ZSTRIDE = 128;      // LDS z-column (127 + zero pad)
constexpr int ZGUARD = 8;         // zero guard z-slots on both sides of the stage
constexpr double TFAR = 1e300;    // boundary "never" (direct planes)
constexpr double SPAD = -1e30;    // start of a padding segment (contributes nothing)
constexpr int MAXNB = 4;          // replicates per kernel instance (larger batches: sub-batches)
constexpr int BP_MAXD = 63;       // largest ring difference of the BP tables

template <int NB>
constexpr int fp_chunk() { return NB >= 4 ? 16 : 32; }
template <int NB>
constexpr int bp_chunk() { return NB == 1 ? 32 : (NB == 2 ? 16 : 8); }
template <int NB>
constexpr int fp_bins() { return NB == 1 ? MAXNSN : 837 * NB; }
// BP table row d holds r0 = -2 .. 65-d (two zero pads on both sides); row offset:
__host__ __device__ constexpr int bp_row(int d) { return 68 * d - d * (d - 1) / 2; }
constexpr int BP_TABN = bp_row(BP_MAXD + 1);

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

// A sub-batch of NB replicates starting at replicate b0 of a batch of nbt:
// images are [voxel][nbt], sinograms [row][nsn][nbt] (replicate innermost).
struct Batch {
  int nbt, b0;
};

template <int NB>
__device__ inline void ldv(float (&r)[NB], const float *p) {
  if constexpr (NB == 1) {
    r[0] = p[0];
  } else if constexpr (NB == 2) {
    float2 t = *reinterpret_cast<const float2 *>(p);
    r[0] = t.x;
    r[1] = t.y;
  } else {
#pragma unroll
    for (int k = 0; k < NB; k += 4) {
      float4 t = *reinterpret_cast<const float4 *>(p + k);
      r[k] = t.x;
      r[k + 1] = t.y;
      r[k + 2] = t.z;
      r[k + 3] = t.w;
    }
  }
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

// FP accumulation: pieces go into float partial sums over SUBSUM segments,
// folded into the running totals (one long float sum of nearly equal segment
// lengths would round in the same direction at every step).
template <int NB>
struct FpAcc {
  float pa[NB], pb[NB], aa[NB], ab[NB];
  __device__ void zero() {
#pragma unroll
    for (int b = 0; b < NB; b++) pa[b] = pb[b] = aa[b] = ab[b] = 0.f;
  }
  __device__ void piece(const float *col, int ia, int ib, float p, int k) {
    float x[NB];
    ldv<NB>(x, col + (ia + k) * NB);
#pragma unroll
    for (int b = 0; b < NB; b++) pa[b] = fmaf(p, x[b], pa[b]);
    ldv<NB>(x, col + (ib - k) * NB);
#pragma unroll
    for (int b = 0; b < NB; b++) pb[b] = fmaf(p, x[b], pb[b]);
  }
  __device__ void direct(const float *col, int ia, float p) {
    float x[NB];
    ldv<NB>(x, col + ia * NB);
#pragma unroll
    for (int b = 0; b < NB; b++) pa[b] = fmaf(p, x[b], pa[b]);
  }
  __device__ void fold() {
#pragma unroll
    for (int b = 0; b < NB; b++) {
      aa[b] += pa[b];
      ab[b] += pb[b];
      pa[b] = 0.f;
      pb[b] = 0.f;
    }
  }
};

// One segment with at most one z boundary inside it.
template <int NB>
__device__ inline void step1(Lor &l, const SegRec &sg, const float *col, FpAcc<NB> &op) {
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
template <int NB>
__device__ inline void step2(Lor &l, const SegRec &sg, const float *col, FpAcc<NB> &op) {
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

// Walk the CH staged segments for this lane's element of one round.
template <int NB, int CH>
__device__ inline void walk_chunk(Lor &l, FpAcc<NB> &op, const SegRec *sbuf, const float *zs,
                                  bool direct_wave, bool two_wave) {
  if (direct_wave) {
    const float *col = zs + l.ia * NB;
#pragma unroll
    for (int j = 0; j < CH; j++) op.direct(col + j * ZSTRIDE * NB, 0, sbuf[j].len);
    op.fold();
  } else if (two_wave) {
#pragma unroll 4
    for (int j = 0; j < CH; j++) {
      step2<NB>(l, sbuf[j], zs + j * ZSTRIDE * NB, op);
      if ((j + 1) % SUBSUM == 0) op.fold();
    }
  } else {
#pragma unroll
    for (int j = 0; j < CH; j++) {
      step1<NB>(l, sbuf[j], zs + j * ZSTRIDE * NB, op);
      if ((j + 1) % SUBSUM == 0) op.fold();
    }
  }
}

__device__ inline float lor_sfac(float dz, float vox_z, double atn) {
  float az = dz * vox_z;
  float r = az / (float)atn;
  return sqrtf(1.f + r * r);
}

__device__ inline SegRec pad_seg() {
  SegRec p;
  p.s = SPAD;
  p.len = 0.f;
  p.colz = 0;
  return p;
}

// ---------------------------------------------------------------------------
// Forward projection. MODE 0: emission; 1: attenuation factors;
// 2: OSEM ratio y / (FP + rho) with the 1e-6 guards of prj/src/recon.cu:252-283.
// rho_nb: 0 = one rho shared by all replicates, else rho has the batch layout.
template <int MODE, int NB>
__global__ __launch_bounds__(PRJ_THREADS) void k_fprj(GeomView g, const float *__restrict__ img,
                                                      float *__restrict__ sino,
                                                      const int32_t *__restrict__ subs,
                                                      const uint16_t *__restrict__ prompts,
                                                      const float *__restrict__ rho, int rho_nb,
                                                      Batch bt) {
  constexpr int CH = fp_chunk<NB>();
  __shared__ __align__(16) float zbuf[(ZGUARD + CH * ZSTRIDE + ZGUARD) * NB];
  __shared__ SegRec sbuf[CH];
  __shared__ float bins[fp_bins<NB>()];
  float *zs = zbuf + ZGUARD * NB;
  const int tid = threadIdx.x, wave = tid >> 6;
  const int row = blockIdx.x;
  const int aw = subs ? subs[row] : row;
  const int nsn = g.nsn;
  const int nbt = bt.nbt, b0 = bt.b0;
  float *out = sino + (size_t)row * nsn * nbt + b0;
  if (aw < 0 || aw >= g.naw) {  // invalid bin index: zero row
    for (int i = tid; i < nsn * NB; i += PRJ_THREADS) out[(i / NB) * nbt + i % NB] = 0.f;
    return;
  }
  const int beg = g.off[aw], nseg = g.off[aw + 1] - beg;
  const double atn = g.atn[aw];
  const float lmax = g.lmax[aw];
  for (int i = tid; i < nsn * NB; i += PRJ_THREADS) bins[i] = 0.f;
  if (tid < ZGUARD * NB) {
    zbuf[tid] = 0.f;
    zbuf[(ZGUARD + CH * ZSTRIDE) * NB + tid] = 0.f;
  }
  const double s0 = nseg > 0 ? g.seg[beg].s : 0.0;

  Lor L[LI_ROUNDS];
  FpAcc<NB> A[LI_ROUNDS];
  bool two[LI_ROUNDS];
#pragma unroll
  for (int r = 0; r < LI_ROUNDS; r++) {
    int li = r * PRJ_THREADS + tid;
    float2 z = li < g.nli ? g.li_z[li] : make_float2(0.5f, 0.f);
    lor_init(L[r], z, atn, s0);
    A[r].zero();
    two[r] = __any(L[r].inv > 0.0 && L[r].inv < (double)lmax);
  }

  for (int c = 0; c < nseg; c += CH) {
    __syncthreads();
    if (tid < CH) sbuf[tid] = (c + tid < nseg) ? g.seg[beg + c + tid] : pad_seg();
    for (int idx = tid; idx < CH * ZSTRIDE * NB; idx += PRJ_THREADS) {
      const int j = idx / (ZSTRIDE * NB), rem = idx - j * (ZSTRIDE * NB);
      const int z = rem / NB, b = rem - z * NB, k = c + j;
      float v = 0.f;
      if (k < nseg && z < NZ) v = img[((int64_t)g.seg[beg + k].colz + z) * nbt + b0 + b];
      zs[idx] = v;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < LI_ROUNDS; r++) {
      const int li0 = r * PRJ_THREADS + wave * 64;  // wave-uniform
      if (li0 >= g.nli) continue;
      walk_chunk<NB, CH>(L[r], A[r], sbuf, zs, li0 + 64 <= g.ndirect, two[r]);
    }
  }

  // span-1 -> span-N reduction in LDS
#pragma unroll
  for (int r = 0; r < LI_ROUNDS; r++) {
    int li = r * PRJ_THREADS + tid;
    if (li < g.nli) {
      A[r].fold();
      float2 z = g.li_z[li];
      short2 sn = g.li_sn[li];
      float sf = lor_sfac(z.y, g.vox_z, atn);
      float nosi = g.li_nosi[li];
#pragma unroll
      for (int b = 0; b < NB; b++) {
        float va = A[r].aa[b] * sf, vb = A[r].ab[b] * sf;
        if (MODE == 1) {
          va = expf(-va) * nosi;
          vb = expf(-vb) * nosi;
        }
        atomicAdd(&bins[sn.x * NB + b], va);
        if (z.y > 0.f) atomicAdd(&bins[sn.y * NB + b], vb);
      }
    }
  }
  __syncthreads();
  for (int i = tid; i < nsn * NB; i += PRJ_THREADS) {
    const int s = i / NB, b = i - s * NB;
    float v = bins[i];
    if (MODE == 2) {
      const size_t k = (size_t)aw * nsn + s;
      const float e = v + (rho_nb ? rho[k * nbt + b0 + b] : rho[k]);
      v = (e > -EPS && e < EPS) ? 0.f : (float)prompts[k * nbt + b0 + b] / e;
    }
    out[(size_t)s * nbt + b] = v;
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
// h = w0 >> 1, e = (2d - w0) >> 1 (va / vb: sinogram values of the LORs
// (r0, d) and of their mirrors). Each wave takes runs of SPW consecutive
// segments and walks all ring differences; for a given d, w0 changes between
// consecutive segments only when a z boundary lies between their starts, so
// the table reads and sums are kept in registers and reloaded only on change.
// The pieces of (segment, d) are computed by lane d (times the 3D length
// factor of d) and broadcast with readlane. va / vb are read from an LDS copy
// of the sinogram row through per-(d, r0) index tables whose padding entries
// point at a zero row, so clamped indices read zeros. No atomics inside a ray:
// each wave flushes its segments through a private LDS buffer with one
// coalesced global float atomic per (segment, plane, replicate).
__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ inline float rdlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

template <int NB>
constexpr int bp_spw() { return NB == 1 ? 8 : 4; }

template <int NB>
__global__ __launch_bounds__(BP_THREADS) void k_bprj(GeomView g, float *__restrict__ img,
                                                     const float *__restrict__ sino,
                                                     const int32_t *__restrict__ subs, Batch bt) {
  constexpr int SPW = bp_spw<NB>();
  constexpr int NWAVE = BP_THREADS / 64;
  __shared__ __align__(16) float srow[(MAXNSN / NB + 1) * NB];   // sinogram row + zero row
  __shared__ int32_t pa[BP_TABN];  // (idx_a[k-1] << 16) | idx_a[k]
  __shared__ int32_t pb[BP_TABN];  // (idx_b[k+1] << 16) | idx_b[k]
  __shared__ __align__(16) float wbuf[NWAVE][ZSTRIDE * NB];
  __shared__ int dmax_s;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int row = blockIdx.x;
  const int aw = subs ? subs[row] : row;
  if (aw < 0 || aw >= g.naw) return;
  const int nsn = g.nsn, nbt = bt.nbt, b0 = bt.b0;
  const float *in = sino + (size_t)row * nsn * nbt + b0;
  const int beg = g.off[aw], nseg = g.off[aw + 1] - beg;
  const double atn = g.atn[aw];

  // LDS: sinogram row, zero row at index nsn, LOR -> sinogram index tables
  for (int i = tid; i < (nsn + 1) * NB; i += BP_THREADS) {
    const int s = i / NB, b = i - s * NB;
    srow[i] = s < nsn ? in[(size_t)s * nbt + b] : 0.f;
  }
  __shared__ int16_t ia_tab[BP_TABN], ib_tab[BP_TABN];
  for (int i = tid; i < BP_TABN; i += BP_THREADS) {
    ia_tab[i] = (int16_t)nsn;
    ib_tab[i] = (int16_t)nsn;
  }
  if (tid == 0) dmax_s = 0;
  __syncthreads();
  for (int li = tid; li < g.nli; li += BP_THREADS) {
    float2 z = g.li_z[li];
    short2 sn = g.li_sn[li];
    int r0 = (int)floorf(z.x) >> 1, d = (int)z.y >> 1;
    int k = bp_row(d) + 2 + r0;
    ia_tab[k] = sn.x;
    if (d > 0) ib_tab[k] = sn.y;
    atomicMax(&dmax_s, d);
  }
  __syncthreads();
  for (int i = tid; i < BP_TABN; i += BP_THREADS) {
    const int prev = i > 0 ? ia_tab[i - 1] : nsn, next = i + 1 < BP_TABN ? ib_tab[i + 1] : nsn;
    pa[i] = (prev << 16) | (ia_tab[i] & 0xffff);
    pb[i] = (next << 16) | (ib_tab[i] & 0xffff);
  }
  __syncthreads();
  const int dmax = dmax_s;
  const double m_lane = 2.0 * lane / atn;
  const double inv_lane = lane > 0 ? atn / (2.0 * lane) : 0.0;
  const float sf_lane = lor_sfac(2.f * lane, g.vox_z, atn);
  const int q = lane;
  // direct planes: plane 2q <- L * va0[q]
  float X0[NB];
  ldv<NB>(X0, srow + (ia_tab[bp_row(0) + 2 + q] & 0xffff) * NB);
  float *wb = wbuf[wave];

  for (int g0 = wave * SPW; g0 < nseg; g0 += NWAVE * SPW) {
    // lane d: base plane and pieces (times the 3D factor) of each segment
    float P1[SPW], P2[SPW], P3[SPW], L[SPW];
    int W0[SPW], C[SPW];
#pragma unroll
    for (int s = 0; s < SPW; s++) {
      const int k = g0 + s;
      SegRec sg;
      if (k < nseg) {
        sg = g.seg[beg + k];
      } else {
        sg.s = 0.0;
        sg.len = 0.f;
        sg.colz = 0;
      }
      L[s] = sg.len;
      C[s] = sg.colz;
      P1[s] = P2[s] = P3[s] = 0.f;
      W0[s] = 0;
      if (lane >= 1 && lane <= dmax) {
        double w0 = floor(0.5 + m_lane * sg.s);
        double tb = (w0 + 0.5) * inv_lane;
        float q1 = (float)(tb - sg.s), q2 = (float)(tb + inv_lane - sg.s);
        float p1 = __builtin_amdgcn_fmed3f(q1, 0.f, sg.len);
        float c2 = __builtin_amdgcn_fmed3f(q2, 0.f, sg.len);
        P1[s] = p1 * sf_lane;
        P2[s] = (c2 - p1) * sf_lane;
        P3[s] = (sg.len - c2) * sf_lane;
        W0[s] = (int)w0;
      }
    }
    float acc0[SPW][NB], acc1[SPW][NB];
#pragma unroll
    for (int s = 0; s < SPW; s++)
#pragma unroll
      for (int b = 0; b < NB; b++) {
        acc0[s][b] = L[s] * X0[b];
        acc1[s][b] = 0.f;
      }
    bool any3 = false;
#pragma unroll
    for (int s = 0; s < SPW; s++) any3 |= __any(P3[s] != 0.f);

    for (int d = 1; d <= dmax; d++) {
      const int rb = bp_row(d) + 2;
      float S1[NB], S3[NB], Te[NB], To[NB];
      int wprev = -1000;
#pragma unroll
      for (int s = 0; s < SPW; s++) {
        const int w0 = __builtin_amdgcn_readlane(W0[s], d);
        if (w0 != wprev) {  // wave-uniform: new base plane -> reload the LOR values
          wprev = w0;
          const int h = w0 >> 1, e = (2 * d - w0) >> 1;
          const int ka = rb + clampi(q - h, -1, 65 - d);
          const int kb = rb + clampi(q - e, -2, 64 - d);
          const int ta = pa[ka], tbb = pb[kb];
          float X[NB], Y[NB], Xb[NB], Yb[NB];
          ldv<NB>(X, srow + (ta & 0xffff) * NB);
          ldv<NB>(Y, srow + (ta >> 16) * NB);
          ldv<NB>(Xb, srow + (tbb & 0xffff) * NB);
          ldv<NB>(Yb, srow + (tbb >> 16) * NB);
#pragma unroll
          for (int b = 0; b < NB; b++) {
            S1[b] = X[b] + Xb[b];
            S3[b] = Y[b] + Yb[b];
            Te[b] = X[b] + Yb[b];
            To[b] = Y[b] + Xb[b];
          }
        }
        const float p1 = rdlane(P1[s], d), p2 = rdlane(P2[s], d);
        if (w0 & 1) {
#pragma unroll
          for (int b = 0; b < NB; b++) {
            acc0[s][b] = fmaf(p2, To[b], acc0[s][b]);
            acc1[s][b] = fmaf(p1, S1[b], acc1[s][b]);
          }
        } else {
#pragma unroll
          for (int b = 0; b < NB; b++) {
            acc0[s][b] = fmaf(p1, S1[b], acc0[s][b]);
            acc1[s][b] = fmaf(p2, Te[b], acc1[s][b]);
          }
        }
        if (any3) {
          const float p3 = rdlane(P3[s], d);
          if (w0 & 1) {
#pragma unroll
            for (int b = 0; b < NB; b++) acc1[s][b] = fmaf(p3, S3[b], acc1[s][b]);
          } else {
#pragma unroll
            for (int b = 0; b < NB; b++) acc0[s][b] = fmaf(p3, S3[b], acc0[s][b]);
          }
        }
      }
    }
    // flush the wave's segments through its LDS buffer (wave-local, no barrier)
#pragma unroll
    for (int s = 0; s < SPW; s++) {
      if (g0 + s >= nseg) break;
#pragma unroll
      for (int b = 0; b < NB; b++) {
        wb[(2 * q) * NB + b] = acc0[s][b];
        wb[(2 * q + 1) * NB + b] = acc1[s][b];
      }
      __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the wave's LDS writes landed
      __builtin_amdgcn_wave_barrier();
      const int colz = __builtin_amdgcn_readfirstlane(C[s]);
      for (int i = lane; i < NZ * NB; i += 64) {
        const int z = i / NB, b = i - z * NB;
        const float v = wb[i];
        if (v != 0.f) unsafeAtomicAdd(img + ((int64_t)colz + z) * nbt + b0 + b, v);
      }
      __builtin_amdgcn_wave_barrier();
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
// (prj/src/recon.cu:234-247, 300-315); images [voxel][nb], sens shared
// (sens_nb = 0) or [voxel][nb].
__global__ void k_osem_update(float *img, const float *b, const float *sens, const uint8_t *mask,
                              int64_t nvox, int nb, int sens_nb) {
  const int64_t n = nvox * nb;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const int64_t v = i / nb;
    float sv = sens_nb ? sens[i] : sens[v];
    float q = (sv > -EPS && sv < EPS) ? 0.f : b[i] / sv;
    img[i] = mask[v] ? img[i] * q : 0.f;
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

// Launch a projector over nb replicates as sub-batches of 4, 2 and 1.
template <int MODE>
static void launch_fprj(const Geom *g, float *sino, const float *img, const int32_t *subs,
                        int nprj, const uint16_t *prompts, const float *rho, int rho_nb, int nb,
                        hipStream_t st) {
  for (int b0 = 0; b0 < nb;) {
    const int rem = nb - b0;
    const Batch bt{nb, b0};
    if (rem >= 4) {
      k_fprj<MODE, 4><<<nprj, PRJ_THREADS, 0, st>>>(view(g), img, sino, subs, prompts, rho,
                                                   rho_nb, bt);
      b0 += 4;
    } else if (rem >= 2) {
      k_fprj<MODE, 2><<<nprj, PRJ_THREADS, 0, st>>>(view(g), img, sino, subs, prompts, rho,
                                                   rho_nb, bt);
      b0 += 2;
    } else {
      k_fprj<MODE, 1><<<nprj, PRJ_THREADS, 0, st>>>(view(g), img, sino, subs, prompts, rho,
                                                   rho_nb, bt);
      b0 += 1;
    }
  }
}

static void launch_bprj(const Geom *g, float *img, const float *sino, const int32_t *subs,
                        int nprj, int nb, hipStream_t st) {
  for (int b0 = 0; b0 < nb;) {
    const int rem = nb - b0;
    const Batch bt{nb, b0};
    if (rem >= 4) {
      k_bprj<4><<<nprj, BP_THREADS, 0, st>>>(view(g), img, sino, subs, bt);
      b0 += 4;
    } else if (rem >= 2) {
      k_bprj<2><<<nprj, BP_THREADS, 0, st>>>(view(g), img, sino, subs, bt);
      b0 += 2;
    } else {
      k_bprj<1><<<nprj, BP_THREADS, 0, st>>>(view(g), img, sino, subs, bt);
      b0 += 1;
    }
  }
}

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

static int check_nprj(const NipetGeom *g, const int32_t *subs, int nprj, int nb, const char *fn) {
  if (nprj <= 0 || (!subs && nprj != g->naw)) {
    set_error("%s: nprj=%d (must be > 0, and = naw=%d when subs is NULL)", fn, nprj, g->naw);
    return NIPET_EINVAL;
  }
  if (nb < 1 || nb > NIPET_MAXBATCH) {
    set_error("%s: batch %d outside 1..%d", fn, nb, NIPET_MAXBATCH);
    return NIPET_EINVAL;
  }
  if (nb > 1 && g->nsn > 837) {
    set_error("%s: batches need at most 837 sinograms (span 11)", fn);
    return NIPET_EINVAL;
  }
  return NIPET_OK;
}

extern "C" {

int nipet_fprj_b(const NipetGeom *g, float *sino, const float *img, const int32_t *subs,
                 int nprj, int nb, int att, void *stream) {
  int rc = check_geom(g, "nipet_fprj");
  if (rc) return rc;
  if ((rc = check_nprj(g, subs, nprj, nb, "nipet_fprj"))) return rc;
  if (!sino || !img || (att != 0 && att != 1)) {
    set_error("nipet_fprj: NULL buffer or att not in {0,1}");
    return NIPET_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  prof_begin(PK_FPRJ, st);
  if (att)
    launch_fprj<1>(g, sino, img, subs, nprj, nullptr, nullptr, 0, nb, st);
  else
    launch_fprj<0>(g, sino, img, subs, nprj, nullptr, nullptr, 0, nb, st);
  prof_end(PK_FPRJ, st);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_fprj(const NipetGeom *g, float *sino, const float *img, const int32_t *subs, int nprj,
               int att, void *stream) {
  return nipet_fprj_b(g, sino, img, subs, nprj, 1, att, stream);
}

int nipet_bprj_b(const NipetGeom *g, float *img, const float *sino, const int32_t *subs,
                 int nprj, int nb, void *stream) {
  int rc = check_geom(g, "nipet_bprj");
  if (rc) return rc;
  if ((rc = check_nprj(g, subs, nprj, nb, "nipet_bprj"))) return rc;
  if (!sino || !img) {
    set_error("nipet_bprj: NULL buffer");
    return NIPET_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  prof_begin(PK_FILL, st);
  k_fill<<<ew_blocks(NVOX * nb), 256, 0, st>>>(img, 0.f, NVOX * nb);
  prof_end(PK_FILL, st);
  prof_begin(PK_BPRJ, st);
  launch_bprj(g, img, sino, subs, nprj, nb, st);
  prof_end(PK_BPRJ, st);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_bprj(const NipetGeom *g, float *img, const float *sino, const int32_t *subs, int nprj,
               void *stream) {
  return nipet_bprj_b(g, img, sino, subs, nprj, 1, stream);
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

int nipet_osem_subset_b(const NipetGeom *g, float *img, const uint16_t *prompts,
                        const float *rho, int rho_per_rep, const float *sens, int sens_per_rep,
                        const uint8_t *mask, const int32_t *subs, int nprj, int nb,
                        float *work_sino, float *work_img, void *stream) {
  int rc = check_geom(g, "nipet_osem_subset");
  if (rc) return rc;
  if (!subs || nprj <= 0 || nprj > g->naw) {
    set_error("nipet_osem_subset: subs must be given with 0 < nprj <= naw");
    return NIPET_EINVAL;
  }
  if ((rc = check_nprj(g, subs, nprj, nb, "nipet_osem_subset"))) return rc;
  if (!img || !prompts || !rho || !sens || !mask || !work_sino || !work_img) {
    set_error("nipet_osem_subset: NULL buffer");
    return NIPET_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  prof_begin(PK_FPRJ, st);
  launch_fprj<2>(g, work_sino, img, subs, nprj, prompts, rho, rho_per_rep ? 1 : 0, nb, st);
  prof_end(PK_FPRJ, st);
  prof_begin(PK_FILL, st);
  k_fill<<<ew_blocks(NVOX * nb), 256, 0, st>>>(work_img, 0.f, NVOX * nb);
  prof_end(PK_FILL, st);
  prof_begin(PK_BPRJ, st);
  launch_bprj(g, work_img, work_sino, subs, nprj, nb, st);
  prof_end(PK_BPRJ, st);
  prof_begin(PK_OSEM_UPDATE, st);
  k_osem_update<<<ew_blocks(NVOX * nb), 256, 0, st>>>(img, work_img, sens, mask, NVOX, nb,
                                                       sens_per_rep ? 1 : 0);
  prof_end(PK_OSEM_UPDATE, st);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_osem_subset(const NipetGeom *g, float *img, const uint16_t *prompts, const float *rho,
                      const float *sens, const uint8_t *mask, const int32_t *subs, int nprj,
                      float *work_sino, float *work_img, void *stream) {
  return nipet_osem_subset_b(g, img, prompts, rho, 0, sens, 0, mask, subs, nprj, 1, work_sino,
                             work_img, stream);
}

}  // extern "C"
