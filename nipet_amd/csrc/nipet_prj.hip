// Siddon forward / back projectors and the fused OSEM subset update for gfx950.
//
// Work decomposition (one workgroup = one transaxial bin of the subset):
//  - the bin's transaxial segment list (FP64-built, device resident) is
//    walked in chunks; for each chunk the z-columns img[col][0:127] of its
//    voxels are staged in LDS (127 contiguous floats each: coalesced 508-byte
//    reads) -- the projections then never gather from HBM;
//  - FP (k_fprj): a lane owns Michelogram elements (a LOR and its mirror,
//    which share their z boundaries) and walks the staged segments, splitting
//    each at the next z boundary; span-1 -> span-11 sums gathered in a fixed
//    order, one coalesced store per sinogram bin (prj/src/prjf.cu:88-94 uses
//    atomicAdd);
//  - BP (k_bprj): all LORs of one ring difference d share their z-boundary
//    positions along the bin (ring centres at z-voxel 2r + 0.5), so the pieces
//    of a segment are computed once per (segment, d); lanes own z-plane pairs
//    and gather over d into registers, one coalesced global float atomic per
//    (segment, plane) instead of one per (LOR, piece) (prj/src/prjb.cu:70-82,
//    153-185).
#include <atomic>
#include <mutex>
#include <utility>

#include "nipet_common.h"

namespace nipet {

constexpr int BP_THREADS = 512;   // BP: 8 waves
constexpr int SUBSUM = 8;         // float partial sums over 8 segments
constexpr int ZSTRIDE = 128;      // LDS z-column (127 + zero pad)
constexpr int MAXNB = 4;          // replicates per kernel instance (larger batches: sub-batches)

// BP table row d holds r0 = -2 .. 65-d (two zero pads on both sides); row offset:
// Row d of the BP index tables covers r0 = -d-1 .. 64 (d + 66 entries), every
// index the gather can form for ring difference d, so no clamping is needed.

struct GeomView {
  const int32_t *off;
  const SegRec *seg;
  const double *atn;
  const float *lmax;
  const float2 *li_z;
  const short2 *li_sn;
  const float *li_nosi;
  const int16_t *dr_li;
  const int32_t *sn_off;
  const int16_t *sn_mem;
  const int16_t *fp_slot;
  const uint16_t *bp_tab;
  int fp_nslot;
  int naw, nli, ndirect, nsn, dmax;
  float vox_z;
};

static GeomView view(const Geom *g) {
  GeomView v{};
  v.off = g->off;
  v.seg = g->seg;
  v.atn = g->atn;
  v.lmax = g->lmax;
  v.li_z = g->li_z;
  v.li_sn = g->li_sn;
  v.li_nosi = g->li_nosi;
  v.dr_li = g->dr_li;
  v.sn_off = g->sn_off;
  v.sn_mem = g->sn_mem;
  v.fp_slot = g->fp_slot;
  v.bp_tab = g->bp_tab;
  v.fp_nslot = g->fp_nslot;
  v.naw = g->naw;
  v.nli = g->nli;
  v.ndirect = g->ndirect;
  v.nsn = g->nsn;
  v.dmax = g->dmax;
  v.vox_z = g->vox_z;
  return v;
}

// A sub-batch of NB replicates starting at replicate b0 of a batch of nbt:
// images are [voxel][nbt], sinograms [row][nsn][nbt] (replicate innermost).
struct Batch {
  int nbt, b0;
  // optional per-bin segment window [first, first + count) of the bin's list
  // (nipet_osem_window): the segments whose columns meet the reconstruction
  // mask; the rest carry zero image values (FP) or feed voxels the masked
  // update discards (BP)
  const int32_t *win = nullptr;
};

// The bin's segment list, or its window. A window is clamped into the list
// (nipet_osem_window's always lies inside it), so that the kernels' LDS copy
// of a bin's records (at most MAX_SEG) holds whatever a caller passes.
__device__ inline void seg_range(const GeomView &g, const Batch &bt, int aw, int &beg,
                                 int &nseg) {
  beg = g.off[aw];
  const int n = g.off[aw + 1] - beg;
  nseg = n;
  if (bt.win) {
    const int f = min(max(bt.win[2 * aw], 0), n);
    beg += f;
    nseg = min(max(bt.win[2 * aw + 1], 0), n - f);
  }
}

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

template <int NB>
__device__ inline void stv(float *p, const float (&r)[NB]) {
  if constexpr (NB == 1) {
    p[0] = r[0];
  } else if constexpr (NB == 2) {
    *reinterpret_cast<float2 *>(p) = make_float2(r[0], r[1]);
  } else {
#pragma unroll
    for (int k = 0; k < NB; k += 4)
      *reinterpret_cast<float4 *>(p + k) = make_float4(r[k], r[k + 1], r[k + 2], r[k + 3]);
  }
}

// Compile-time loop: f(std::integral_constant<int, i>) for i = 0 .. N-1, so
// that register arrays are only ever indexed by constants.
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

__device__ inline float lor_sfac(float dz, float vox_z, double atn) {
  float az = dz * vox_z;
  float r = az / (float)atn;
  return sqrtf(1.f + r * r);
}

__device__ inline float rdlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// ---------------------------------------------------------------------------
// Forward projection. MODE 0: emission; 1: attenuation factors;
// 2: OSEM ratio y / (FP + rho) with the 1e-6 guards of prj/src/recon.cu:252-283.
// rho_nb: 0 = one rho shared by all replicates, else rho has the batch layout.
//
// Lane = Michelogram element li = lane + FP_THREADS r (r < FP_ROUNDS): LOR a
// (ring r0 -> r1) and its mirror b (r1 -> r0) share their z-boundary
// positions (reference prj/src/prjf.cu:137-190). Per segment [s, s + L] of
// the staged chunk, LOR a is on plane ia and the next z boundary lies at
// t = tb: it adds p1 = clamp(tb - s, 0, L) times plane ia and L - p1 times
// plane ia + 1 (mirror: ib and ib - 1), and steps to the next plane when
// tb <= s + L. tb is tracked in float relative to the start of the chunk's
// middle segment and recomputed exactly (FP64, from the plane index) at every
// chunk, so positions along 67 cm rays keep 1e-7 cm resolution without FP64
// in the inner loop. Bins whose segments can hold two z boundaries (segment longer
// than the boundary spacing) take a three-piece step. Pieces go into float
// partial sums over SUBSUM segments folded into float totals (one long float
// sum of nearly equal segment lengths would round in the same direction at
// every step). Span-1 -> span-N: each output bin sums its LORs in a fixed
// order from LDS (no atomics, bit-reproducible).
constexpr int FP_THREADS = 704;  // 11 waves
constexpr int FP_ROUNDS = 3;     // 3 * 704 = 2112 >= LI_MAX elements
constexpr int FP_GUARD = 2;      // zero planes below 0 and above 127 (pieces reach ia + 2, ib - 2)
constexpr int FP_PLANES = ZSTRIDE + 2 * FP_GUARD;

// Staged chunk layout: value of (plane p, segment j, replicate b) at
// ((p + FP_GUARD) * PS + j) * NB + b with PS = CH + 1: a lane's planes for
// consecutive segments are adjacent (small immediate LDS offsets), and lanes
// on different planes fall in different banks.
// One chunk length for every batch width: the exact FP64 boundary rebase then
// happens at the same segments, so a replicate's arithmetic (and result) is the
// same whether it is projected alone or in a batch.
template <int NB>
constexpr int fp_chunk() { return 16; }
template <int NB>
constexpr int fp_ps() { return fp_chunk<NB>() + 1; }

template <int NB>
__device__ inline void fma_nb(float (&acc)[NB], float p, const float (&x)[NB]) {
#pragma unroll
  for (int b = 0; b < NB; b++) acc[b] = fmaf(p, x[b], acc[b]);
}

// One bin's segment records [n] copied to LDS with 16-byte loads.
__device__ inline void load_segs(SegRec *dst, const SegRec *__restrict__ src, int n, int tid,
                                 int nthr) {
  const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
  uint4 *d4 = reinterpret_cast<uint4 *>(dst);
  for (int i = tid; i < n; i += nthr) d4[i] = s4[i];
}

struct FpLor {
  float tbr;   // next z boundary of LOR a relative to the chunk's origin (float)
  float invf;  // boundary spacing (float)
  int ao;      // staged-chunk offset of plane ia (LOR a), segment 0
  int bo;      // staged-chunk offset of plane ib - 2 (mirror b: ib = 2 r0 + 2 r1 - ia)
};

template <int NB>
struct FpSum {
  float pa[NB], pb[NB], ta[NB], tb[NB];
  __device__ void fold() {
#pragma unroll
    for (int b = 0; b < NB; b++) {
      ta[b] += pa[b];
      tb[b] += pb[b];
      pa[b] = 0.f;
      pb[b] = 0.f;
    }
  }
};

// Walk the nch staged segments (SR: start - chunk origin, length; padding
// entries (-1e30, 0) contribute nothing and never cross). TWO: segments may
// hold two z boundaries; DIRECT: direct planes only.
template <int NB, bool DIRECT, bool TWO>
__device__ inline void fp_walk(FpLor &l, FpSum<NB> &acc, const float2 *__restrict__ sr,
                               const float *__restrict__ zs, int nch) {
  constexpr int PS = fp_ps<NB>() * NB;  // plane stride (floats)
  constexpr int UNROLL = NB >= 4 ? 2 : 4;
  for (int j8 = 0; j8 < nch; j8 += SUBSUM) {
    const float *za = zs + l.ao + j8 * NB;
    const float *zb = zs + l.bo + j8 * NB;
#pragma unroll UNROLL
    for (int jj = 0; jj < SUBSUM; jj++) {
      const float2 sl = sr[j8 + jj];
      float xa0[NB], xa1[NB];
      if constexpr (DIRECT) {
        ldv<NB>(xa0, za + jj * NB);
        fma_nb<NB>(acc.pa, sl.y, xa0);
      } else if constexpr (!TWO) {
        const float q = l.tbr - sl.x;
        const float p1 = __builtin_amdgcn_fmed3f(q, 0.f, sl.y);
        const bool cr = q <= sl.y;
        float xb0[NB], xb1[NB];
        ldv<NB>(xa0, za + jj * NB);
        ldv<NB>(xb0, zb + 2 * PS + jj * NB);
        fma_nb<NB>(acc.pa, p1, xa0);
        fma_nb<NB>(acc.pb, p1, xb0);
        // the second piece L - p1 is non-zero only when the boundary lies inside
        // the segment (about one step in four): lanes without a crossing skip
        // its plane reads, which halves the LDS traffic of the walk
        if (cr) {
          const float p2 = sl.y - p1;
          ldv<NB>(xa1, za + PS + jj * NB);
          ldv<NB>(xb1, zb + PS + jj * NB);
          fma_nb<NB>(acc.pa, p2, xa1);
          fma_nb<NB>(acc.pb, p2, xb1);
        }
        const int step = cr ? PS : 0;
        za += step;
        zb -= step;
        l.tbr += cr ? l.invf : 0.f;
      } else {
        const float q1 = l.tbr - sl.x, q2 = q1 + l.invf;
        const float p1 = __builtin_amdgcn_fmed3f(q1, 0.f, sl.y);
        const float c2 = __builtin_amdgcn_fmed3f(q2, 0.f, sl.y);
        const float p2 = c2 - p1, p3 = sl.y - c2;
        float xa2[NB], xb0[NB], xb1[NB], xb2[NB];
        ldv<NB>(xa0, za + jj * NB);
        ldv<NB>(xa1, za + PS + jj * NB);
        ldv<NB>(xa2, za + 2 * PS + jj * NB);
        ldv<NB>(xb0, zb + 2 * PS + jj * NB);
        ldv<NB>(xb1, zb + PS + jj * NB);
        ldv<NB>(xb2, zb + jj * NB);
        fma_nb<NB>(acc.pa, p1, xa0);
        fma_nb<NB>(acc.pa, p2, xa1);
        fma_nb<NB>(acc.pa, p3, xa2);
        fma_nb<NB>(acc.pb, p1, xb0);
        fma_nb<NB>(acc.pb, p2, xb1);
        fma_nb<NB>(acc.pb, p3, xb2);
        const int n = (int)(q1 <= sl.y) + (int)(q2 <= sl.y);
        za += n * PS;
        zb -= n * PS;
        l.tbr += (float)n * l.invf;
      }
    }
    // za / zb moved by whole planes only (segments are immediate offsets)
    l.ao = (int)(za - zs) - j8 * NB;
    l.bo = (int)(zb - zs) - j8 * NB;
    acc.fold();
  }
}

// The output sinograms' LOR lists (CSR sn_off / sn_mem, at most 2 LI_MAX
// members < 65536) copied into LDS at the start of a bin, so that the final
// per-sinogram sums read LDS instead of chained global loads.
struct SnCsr {
  uint16_t off[MAXNSN + 1];
  uint16_t mem[2 * LI_MAX];
};

__device__ inline void load_sn_csr(SnCsr &c, const GeomView &g, int tid, int nthreads) {
  for (int i = tid; i <= g.nsn; i += nthreads) c.off[i] = (uint16_t)g.sn_off[i];
  const int nm = g.sn_off[g.nsn];
  for (int i = tid; i < nm; i += nthreads) c.mem[i] = (uint16_t)g.sn_mem[i];
}

// waves per SIMD: width 4 holds 12 column values in flight per thread (about
// 154 VGPRs, 3 waves: one 11-wave workgroup per CU either way)
template <int MODE, int NB>
__global__ __launch_bounds__(FP_THREADS, NB >= 4 ? 3 : 6) void k_fprj(GeomView g, const float *__restrict__ img,
                                                     float *__restrict__ sino,
                                                     const int32_t *__restrict__ subs,
                                                     const uint16_t *__restrict__ prompts,
                                                     const float *__restrict__ rho, int rho_nb,
                                                     Batch bt) {
  constexpr int CH = fp_chunk<NB>(), PS = fp_ps<NB>();
  // staged chunk (plane-major, see fp_ps), then reused for the per-LOR values
  constexpr int ZN = FP_PLANES * PS * NB, LN = 2 * LI_MAX;
  __shared__ __align__(16) float lds[ZN > LN ? ZN : LN];
  __shared__ float2 sr[CH];
  __shared__ double inv[FP_ROUNDS][FP_THREADS];  // exact boundary spacing (0: direct)
  __shared__ float z0[FP_ROUNDS][FP_THREADS];    // zeta of LOR a at t = 0
  const int tid = threadIdx.x, wave = tid >> 6;
  const int row = blockIdx.x;
  const int aw = subs ? subs[row] : row;
  const int nsn = g.nsn, nbt = bt.nbt, b0 = bt.b0;
  float *out = sino + (size_t)row * nsn * nbt + b0;
  if (aw < 0 || aw >= g.naw) {  // invalid bin index: zero row
    for (int i = tid; i < nsn * NB; i += FP_THREADS) out[(i / NB) * nbt + i % NB] = 0.f;
    return;
  }
  int beg, nseg;
  seg_range(g, bt, aw, beg, nseg);
  const double atn = g.atn[aw];
  const float lmax = g.lmax[aw];
  float *zs = lds;
  __shared__ SnCsr csr;
  load_sn_csr(csr, g, tid, FP_THREADS);
  // guard planes (never written by the staging loop)
  for (int i = tid; i < FP_GUARD * PS * NB; i += FP_THREADS) {
    zs[i] = 0.f;
    zs[(FP_GUARD + ZSTRIDE) * PS * NB + i] = 0.f;
  }

  FpLor L[FP_ROUNDS];
  FpSum<NB> A[FP_ROUNDS];
  bool two[FP_ROUNDS], direct[FP_ROUNDS], live[FP_ROUNDS];
  const double s0 = nseg > 0 ? g.seg[beg].s : 0.0;
#pragma unroll
  for (int r = 0; r < FP_ROUNDS; r++) {
    const int li = r * FP_THREADS + tid;
    const float2 z = li < g.nli ? g.li_z[li] : make_float2(0.5f, 0.f);
    const int wsum = (int)floorf(z.x) + (int)floorf(z.x + z.y);
    int ia;
    double iv = 0.0;
    if (z.y > 0.f) {
      iv = atn / (double)z.y;
      ia = (int)floor((double)z.x + s0 / iv);
    } else {
      ia = (int)floorf(z.x);
    }
    z0[r][tid] = z.x;
    inv[r][tid] = iv;
    L[r].ao = (ia + FP_GUARD) * PS * NB;
    L[r].bo = (wsum - ia - 2 + FP_GUARD) * PS * NB;
    L[r].invf = (float)iv;
#pragma unroll
    for (int b = 0; b < NB; b++) A[r].pa[b] = A[r].pb[b] = A[r].ta[b] = A[r].tb[b] = 0.f;
    const int li0 = r * FP_THREADS + wave * 64;  // wave-uniform
    live[r] = li0 < g.nli;
    direct[r] = li0 + 64 <= g.ndirect;
    two[r] = __any(iv > 0.0 && iv < (double)lmax);
  }

  // Software pipeline over chunks of CH segments: while chunk c is walked,
  // each thread has its NST z-column values of chunk c + CH in flight. The
  // bin's segment records sit in LDS (copied once), so neither the column
  // addresses nor a chunk's positions wait on a global load.
  // A thread stages whole (segment, z) groups of NB values: one vector load
  // per group where the rows' replicates are aligned for it (every width of
  // an aligned batch), NB scalar loads otherwise.
  constexpr int NG = CH * ZSTRIDE, NSG = (NG + FP_THREADS - 1) / FP_THREADS;
  __shared__ SegRec segl[MAX_SEG];
  load_segs(segl, g.seg + beg, nseg, tid, FP_THREADS);
  const bool vec = nbt % NB == 0 && b0 % NB == 0 &&
                   (uintptr_t)img % (sizeof(float) * NB) == 0;  // grid-uniform
  float pre[NSG][NB];
  auto fetch = [&](int c) {
#pragma unroll
    for (int i = 0; i < NSG; i++) {
      const int gi = tid + i * FP_THREADS;
      const int j = gi / ZSTRIDE, z = gi - j * ZSTRIDE;
#pragma unroll
      for (int b = 0; b < NB; b++) pre[i][b] = 0.f;
      if (gi < NG && c + j < nseg && z < NZ) {
        const float *q = img + ((int64_t)segl[c + j].colz + z) * nbt + b0;
        if (vec) {
          ldv<NB>(pre[i], q);
        } else {
#pragma unroll
          for (int b = 0; b < NB; b++) pre[i][b] = q[b];
        }
      }
    }
  };
  __syncthreads();
  if (nseg > 0) fetch(0);
  for (int c = 0; c < nseg; c += CH) {
    const int nch = nseg - c < CH ? nseg - c : CH;
    __syncthreads();
    // positions are relative to the chunk's middle segment: float positions
    // within +-8 segments carry half the rounding of ones within 16
    const double sb = segl[c + (nch >> 1)].s;
    if (tid < CH)
      sr[tid] = c + tid < nseg ? make_float2((float)(segl[c + tid].s - sb), segl[c + tid].len)
                               : make_float2(-1e30f, 0.f);
#pragma unroll
    for (int i = 0; i < NSG; i++) {
      const int gi = tid + i * FP_THREADS;
      if (gi < NG) {
        const int j = gi / ZSTRIDE, z = gi - j * ZSTRIDE;
        stv<NB>(zs + ((z + FP_GUARD) * PS + j) * NB, pre[i]);
      }
    }
    // exact next-boundary positions relative to the chunk's middle segment
#pragma unroll
    for (int r = 0; r < FP_ROUNDS; r++) {
      const int ia = L[r].ao / (PS * NB) - FP_GUARD;
      const double iv = inv[r][tid];
      L[r].tbr = iv > 0.0 ? (float)(((double)(ia + 1) - (double)z0[r][tid]) * iv - sb) : 3e38f;
    }
    __syncthreads();
    if (c + CH < nseg) fetch(c + CH);
#pragma unroll
    for (int r = 0; r < FP_ROUNDS; r++) {
      if (!live[r]) continue;
      if (direct[r])
        fp_walk<NB, true, false>(L[r], A[r], sr, zs, nch);
      else if (two[r])
        fp_walk<NB, false, true>(L[r], A[r], sr, zs, nch);
      else
        fp_walk<NB, false, false>(L[r], A[r], sr, zs, nch);
    }
  }

  // per-LOR values -> LDS; each output bin then sums its LORs in CSR order
  float *lorv = lds;
  float sf[FP_ROUNDS];
  bool have[FP_ROUNDS], mir[FP_ROUNDS];
#pragma unroll
  for (int r = 0; r < FP_ROUNDS; r++) {
    const int li = r * FP_THREADS + tid;
    have[r] = li < g.nli;
    const float2 z = have[r] ? g.li_z[li] : make_float2(0.5f, 0.f);
    mir[r] = z.y > 0.f;
    sf[r] = lor_sfac(z.y, g.vox_z, atn);
    A[r].fold();
  }
  static_for<NB>([&](auto bc) {
    constexpr int b = decltype(bc)::value;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < FP_ROUNDS; r++) {
      const int lv = r * FP_THREADS + tid;  // li
      if (have[r]) {
        float va = A[r].ta[b] * sf[r], vb = A[r].tb[b] * sf[r];
        if (MODE == 1) {
          const float nosi = g.li_nosi[lv];
          va = expf(-va) * nosi;
          vb = expf(-vb) * nosi;
        }
        lorv[2 * lv] = va;
        if (mir[r]) lorv[2 * lv + 1] = vb;
      }
    }
    __syncthreads();
    for (int s = tid; s < nsn; s += FP_THREADS) {
      float v = 0.f;
      const int m1 = csr.off[s + 1];
      for (int m = csr.off[s]; m < m1; m++) v += lorv[csr.mem[m]];
      if (MODE == 2) {
        const size_t kk = (size_t)aw * nsn + s;
        const float ee = v + (rho_nb ? rho[kk * nbt + b0 + b] : rho[kk]);
        v = (ee > -EPS && ee < EPS) ? 0.f : (float)prompts[kk * nbt + b0 + b] / ee;
      }
      out[(size_t)s * nbt + b] = v;
    }
  });
}

// ---------------------------------------------------------------------------
// Forward projection with lanes over groups of G LORs of one ring difference.
//
// LOR (d, r0 + k) lies exactly 2k planes above LOR (d, r0) along the whole ray
// (ring centres at zeta = 2 r + 0.5), so the G LORs of a slot (and their
// mirrors, 2k planes above the mirror of (d, r0)) share the boundary
// positions, the pieces p1 / p2 and the plane steps: per segment a lane
// computes those once and adds 2 (4 across a boundary) products per LOR pair,
// reading the G planes at immediate offsets from one pointer. Per LOR the
// arithmetic equals k_fprj's, so both kernels give the same values. Slots (d, r0) come from the
// geometry (fp_slot); lanes whose group runs past ring 63 walk phantom LORs
// whose planes lie in zero guard planes and are never written out.
// G = 3 (odd): lanes of one ring difference then sit 6 planes (3 plane pairs)
// apart; in the pair layout (fpg_addr) that is 3 banks, so 32 lanes of one d
// reach 32 distinct banks. Measured on MI355X with a linear plane stride of 17
// floats: G = 3 1.19 ms per subset, G = 5 1.31 ms, G = 4 3.4 ms (bank
// conflicts), one LOR per lane 1.71 ms. With two replicates (plane stride 34
// floats) every G conflicts and k_fprj<., 2> is faster, so only single images
// use this kernel.
template <int G>
constexpr int fpg_threads() { return FP_G_SLOTS; }  // >= 712 slots of 3 LORs
template <int G>
constexpr int fpg_rounds() { return 1; }
template <int G>
constexpr int fpg_top() { return 2 + 2 * (G - 1); }    // zero planes above 127
template <int G>
constexpr int fpg_planes() { return FP_GUARD + ZSTRIDE + fpg_top<G>(); }

template <int NB, int G>
struct FpSumG {
  float pa[G][NB], pb[G][NB], ta[G][NB], tb[G][NB];
  __device__ void fold() {
    static_for<G>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
#pragma unroll
      for (int b = 0; b < NB; b++) {
        ta[k][b] += pa[k][b];
        tb[k][b] += pb[k][b];
        pa[k][b] = 0.f;
        pb[k][b] = 0.f;
      }
    });
  }
};

// Staged chunk of the grouped kernel: planes in pairs, plane p at
// fpg_addr(p) = (p >> 1) * FPG_PP + (p & 1) * CH floats (segment j at + j),
// pair stride FPG_PP = 2 CH + 1. Planes 2k above p stay at the immediate
// offset FPG_PP * k; the next plane up is CH or CH + 1 floats on, by the
// parity of p (kept per lane as oa / ob). Lanes of one ring difference sit 6
// planes (3 pairs) apart at every segment, so their reads step 3 banks and
// reach all 32 banks; a linear plane stride maps 6 planes to an even number of
// banks, 16 of the 32.
constexpr int FPG_CH = 16;
constexpr int FPG_PP = 2 * FPG_CH + 1;
__host__ __device__ constexpr int fpg_addr(int p) { return (p >> 1) * FPG_PP + (p & 1) * FPG_CH; }

struct FpLorG {
  float tbr;   // next z boundary of LOR a relative to the chunk's origin (float)
  float invf;  // boundary spacing (float)
  int ao;      // staged-chunk offset of plane pa = ia + FP_GUARD (LOR a), segment 0
  int bo;      // staged-chunk offset of plane pb = ib - 2 + FP_GUARD (mirror b)
  // fpg_addr(p + 1) - fpg_addr(p) = CH + (p & 1) for p = pa and p = pb:
  // pa + pb = 4 r0 + 2 d + 2 is even, so both planes share their parity (and
  // keep sharing it, since a boundary moves pa up and pb down by one)
  int o;
};

// srx / srl: position and length of the chunk's segment j in lane j (every
// wave holds the chunk's records), broadcast to the scalar unit per segment
template <int G, bool TWO>
__device__ inline void fpg_walk(FpLorG &l, FpSumG<1, G> &acc, float srx, float srl,
                                const float *__restrict__ zs, int nch) {
  constexpr int PP = FPG_PP;
  for (int j8 = 0; j8 < nch; j8 += SUBSUM) {
    const float *za = zs + l.ao + j8;
    const float *zb = zs + l.bo + j8;
    int o = l.o;
#pragma unroll 4
    for (int jj = 0; jj < SUBSUM; jj++) {
      const float2 sl = make_float2(rdlane(srx, j8 + jj), rdlane(srl, j8 + jj));
      if constexpr (!TWO) {
        const float q = l.tbr - sl.x;
        const float p1 = __builtin_amdgcn_fmed3f(q, 0.f, sl.y);
        const bool cr = q <= sl.y;
        static_for<G>([&](auto kc) {
          constexpr int k = decltype(kc)::value;
          acc.pa[k][0] = fmaf(p1, za[k * PP + jj], acc.pa[k][0]);       // plane pa + 2k
          acc.pb[k][0] = fmaf(p1, zb[(k + 1) * PP + jj], acc.pb[k][0]);  // plane pb + 2k + 2
        });
        // second piece only where the segment holds a boundary: the same
        // segment for every lane of a ring difference, so whole lane groups
        // skip these reads (and the plane steps) together
        if (cr) {
          const float p2 = sl.y - p1;
          const float *za1 = za + o, *zb1 = zb + o;  // planes pa + 1, pb + 1
          static_for<G>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            acc.pa[k][0] = fmaf(p2, za1[k * PP + jj], acc.pa[k][0]);
            acc.pb[k][0] = fmaf(p2, zb1[k * PP + jj], acc.pb[k][0]);
          });
          za = za1;       // pa + 1
          zb = zb1 - PP;  // pb - 1
          o = PP - o;
          l.tbr += l.invf;
        }
      } else {
        const float q1 = l.tbr - sl.x, q2 = q1 + l.invf;
        const float p1 = __builtin_amdgcn_fmed3f(q1, 0.f, sl.y);
        const float c2 = __builtin_amdgcn_fmed3f(q2, 0.f, sl.y);
        const float p2 = c2 - p1, p3 = sl.y - c2;
        const float *za1 = za + o, *zb1 = zb + o;
        static_for<G>([&](auto kc) {
          constexpr int k = decltype(kc)::value;
          const float xa0 = za[k * PP + jj], xa1 = za1[k * PP + jj], xa2 = za[(k + 1) * PP + jj];
          const float xb0 = zb[(k + 1) * PP + jj], xb1 = zb1[k * PP + jj], xb2 = zb[k * PP + jj];
          acc.pa[k][0] = fmaf(p1, xa0, acc.pa[k][0]);
          acc.pa[k][0] = fmaf(p2, xa1, acc.pa[k][0]);
          acc.pa[k][0] = fmaf(p3, xa2, acc.pa[k][0]);
          acc.pb[k][0] = fmaf(p1, xb0, acc.pb[k][0]);
          acc.pb[k][0] = fmaf(p2, xb1, acc.pb[k][0]);
          acc.pb[k][0] = fmaf(p3, xb2, acc.pb[k][0]);
        });
        const int n = (int)(q1 <= sl.y) + (int)(q2 <= sl.y);
        const bool n1 = n == 1, n2 = n == 2;
        za = n1 ? za1 : (n2 ? za + PP : za);
        zb = n1 ? zb1 - PP : (n2 ? zb - PP : zb);
        o = n1 ? PP - o : o;
        l.tbr += (float)n * l.invf;
      }
    }
    l.ao = (int)(za - zs) - j8;
    l.bo = (int)(zb - zs) - j8;
    l.o = o;
    acc.fold();
  }
}

// 6 waves per SIMD (two 12-wave workgroups per CU) need <= 80 VGPRs; the
// compiler's own choice for this kernel is 82
template <int MODE, int NB, int G>
__global__ __launch_bounds__(fpg_threads<G>()) __attribute__((amdgpu_waves_per_eu(6))) void k_fprj_g(GeomView g, const float *__restrict__ img,
                                                           float *__restrict__ sino,
                                                           const int32_t *__restrict__ subs,
                                                           const uint16_t *__restrict__ prompts,
                                                           const float *__restrict__ rho,
                                                           int rho_nb, Batch bt) {
  constexpr int CH = fp_chunk<NB>();
  constexpr int R = fpg_rounds<G>(), NP = fpg_planes<G>(), FPG_THREADS = fpg_threads<G>();
  static_assert(G == FP_G, "slot tables are built for FP_G LORs per lane");
  static_assert(NB == 1 && CH == FPG_CH, "pair layout of one image");
  constexpr int ZN = fpg_addr(NP), LN = 2 * LI_MAX;  // planes 0 .. NP - 1
  __shared__ __align__(16) float lds[ZN > LN ? ZN : LN];
  __shared__ double inv[R][FPG_THREADS];  // exact boundary spacing (0: direct)
  __shared__ float z0[R][FPG_THREADS];    // zeta of the slot's first LOR at t = 0
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int row = blockIdx.x;
  const int aw = subs ? subs[row] : row;
  const int nsn = g.nsn, nbt = bt.nbt, b0 = bt.b0;
  const int nslot = g.fp_nslot;
  const int16_t *slots = g.fp_slot;
  float *out = sino + (size_t)row * nsn * nbt + b0;
  if (aw < 0 || aw >= g.naw) {  // invalid bin index: zero row
    for (int i = tid; i < nsn * NB; i += FPG_THREADS) out[(i / NB) * nbt + i % NB] = 0.f;
    return;
  }
  int beg, nseg;
  seg_range(g, bt, aw, beg, nseg);
  const double atn = g.atn[aw];
  const float lmax = g.lmax[aw];
  float *zs = lds;
  __shared__ SnCsr csr;
  load_sn_csr(csr, g, tid, FPG_THREADS);
  // guard planes (never written by the staging loop)
  for (int i = tid; i < fpg_addr(FP_GUARD); i += FPG_THREADS) zs[i] = 0.f;
  for (int i = fpg_addr(FP_GUARD + ZSTRIDE) + tid; i < ZN; i += FPG_THREADS) zs[i] = 0.f;

  FpLorG L[R];
  FpSumG<NB, G> A[R];
  bool two[R], live[R];
  int dr[R];  // d * 64 + r0 of the slot
  const double s0 = nseg > 0 ? g.seg[beg].s : 0.0;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int sl = r * FPG_THREADS + tid;
    dr[r] = sl < nslot ? slots[sl] : 0;
    const int d = dr[r] >> 6, r0 = dr[r] & 63;
    const float2 z = make_float2(2.f * r0 + 0.5f, 2.f * d);
    const int wsum = (int)floorf(z.x) + (int)floorf(z.x + z.y);
    int ia;
    double iv = 0.0;
    if (z.y > 0.f) {
      iv = atn / (double)z.y;
      ia = (int)floor((double)z.x + s0 / iv);
    } else {
      ia = (int)floorf(z.x);
    }
    z0[r][tid] = z.x;
    inv[r][tid] = iv;
    const int pa = ia + FP_GUARD, pb = wsum - ia - 2 + FP_GUARD;
    L[r].ao = fpg_addr(pa);
    L[r].bo = fpg_addr(pb);
    L[r].o = CH + (pa & 1);
    L[r].invf = (float)iv;
    static_for<G>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
#pragma unroll
      for (int b = 0; b < NB; b++) A[r].pa[k][b] = A[r].pb[k][b] = A[r].ta[k][b] = A[r].tb[k][b] = 0.f;
    });
    live[r] = r * FPG_THREADS + wave * 64 < nslot;  // wave-uniform
    two[r] = __any(iv > 0.0 && iv < (double)lmax);
  }

  // Software pipeline over chunks of CH segments: while chunk c is walked,
  // each thread has its NST z-column values of chunk c + CH in flight. The
  // bin's segment records sit in LDS (copied once), so neither the column
  // addresses nor a chunk's positions wait on a global load.
  constexpr int NE = CH * ZSTRIDE, NST = (NE + FPG_THREADS - 1) / FPG_THREADS;
  __shared__ SegRec segl[MAX_SEG];
  load_segs(segl, g.seg + beg, nseg, tid, FPG_THREADS);
  float pre[NST];
  auto fetch = [&](int c) {
#pragma unroll
    for (int i = 0; i < NST; i++) {
      const int idx = tid + i * FPG_THREADS;
      const int j = idx / ZSTRIDE, z = idx - j * ZSTRIDE;
      float v = 0.f;
      if (idx < NE && c + j < nseg && z < NZ) v = img[((int64_t)segl[c + j].colz + z) * nbt + b0];
      pre[i] = v;
    }
  };
  __syncthreads();
  if (nseg > 0) fetch(0);
  for (int c = 0; c < nseg; c += CH) {
    const int nch = nseg - c < CH ? nseg - c : CH;
    __syncthreads();
    // positions are relative to the chunk's middle segment: float positions
    // within +-8 segments carry half the rounding of ones within 16
    const double sb = segl[c + (nch >> 1)].s;
    // lanes 0 .. CH-1 of every wave: the chunk's segment positions / lengths
    const int jl = lane < CH ? lane : 0;
    const float srx = c + jl < nseg ? (float)(segl[c + jl].s - sb) : -1e30f;
    const float srl = c + jl < nseg ? segl[c + jl].len : 0.f;
#pragma unroll
    for (int i = 0; i < NST; i++) {
      const int idx = tid + i * FPG_THREADS;
      if (idx < NE) {
        const int j = idx / ZSTRIDE, z = idx - j * ZSTRIDE;
        zs[fpg_addr(z + FP_GUARD) + j] = pre[i];
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      // plane of LOR a: pa = 2 (ao / PP) + (odd plane of its pair)
      const int pq = L[r].ao / FPG_PP;
      const int ia = 2 * pq + (L[r].ao - pq * FPG_PP >= CH ? 1 : 0) - FP_GUARD;
      const double iv = inv[r][tid];
      L[r].tbr = iv > 0.0 ? (float)(((double)(ia + 1) - (double)z0[r][tid]) * iv - sb) : 3e38f;
    }
    __syncthreads();
    if (c + CH < nseg) fetch(c + CH);
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (!live[r]) continue;
      if (two[r])
        fpg_walk<G, true>(L[r], A[r], srx, srl, zs, nch);
      else
        fpg_walk<G, false>(L[r], A[r], srx, srl, zs, nch);
    }
  }

  // per-LOR values -> LDS; each output bin then sums its LORs in CSR order
  float *lorv = lds;
  float sf[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int d = dr[r] >> 6;
    sf[r] = lor_sfac(2.f * d, g.vox_z, atn);
    A[r].fold();
  }
  static_for<NB>([&](auto bc) {
    constexpr int b = decltype(bc)::value;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (r * FPG_THREADS + tid >= nslot) continue;
      const int d = dr[r] >> 6, r0 = dr[r] & 63;
      static_for<G>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const int li = r0 + k + d <= 63 ? g.dr_li[dr[r] + k] : -1;
        if (li >= 0) {
          float va = A[r].ta[k][b] * sf[r], vb = A[r].tb[k][b] * sf[r];
          if (MODE == 1) {
            const float nosi = g.li_nosi[li];
            va = expf(-va) * nosi;
            vb = expf(-vb) * nosi;
          }
          lorv[2 * li] = va;
          if (d > 0) lorv[2 * li + 1] = vb;
        }
      });
    }
    __syncthreads();
    for (int s = tid; s < nsn; s += FPG_THREADS) {
      float v = 0.f;
      const int m1 = csr.off[s + 1];
      for (int m = csr.off[s]; m < m1; m++) v += lorv[csr.mem[m]];
      if (MODE == 2) {
        const size_t kk = (size_t)aw * nsn + s;
        const float ee = v + (rho_nb ? rho[kk * nbt + b0 + b] : rho[kk]);
        v = (ee > -EPS && ee < EPS) ? 0.f : (float)prompts[kk * nbt + b0 + b] / ee;
      }
      out[(size_t)s * nbt + b] = v;
    }
  });
}

// Segment windows of the reconstruction mask (nipet_osem_window): m2[c] = 1
// when column c = u + NX v of the mask holds any voxel, optionally dilated by
// a box of half-width r (the resolution model's 17-tap reach); then, one wave
// per bin, the first and last segment of the bin's list whose column is set.
__global__ void k_mask_cols(const uint8_t *__restrict__ mask, uint8_t *__restrict__ m2) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= NX * NY) return;
  const uint8_t *p = mask + (int64_t)c * NZ;
  int any = 0;
  for (int z = 0; z < NZ; z++) any |= p[z];
  m2[c] = any != 0;
}

__global__ void k_mask_dilate(const uint8_t *__restrict__ m2, uint8_t *__restrict__ m2d, int r) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= NX * NY) return;
  const int u = c % NX, v = c / NX;
  int any = 0;
  for (int vv = v - r; vv <= v + r; vv++) {
    if (vv < 0 || vv >= NY) continue;
    for (int uu = u - r; uu <= u + r; uu++)
      if (uu >= 0 && uu < NX) any |= m2[uu + NX * vv];
  }
  m2d[c] = any != 0;
}

__global__ __launch_bounds__(256) void k_seg_window(GeomView g, const uint8_t *__restrict__ m2,
                                                    int32_t *__restrict__ win) {
  const int lane = threadIdx.x & 63;
  const int aw = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (aw >= g.naw) return;  // wave-uniform
  const int beg = g.off[aw], n = g.off[aw + 1] - beg;
  int first = -1, last = -1;
  for (int k0 = 0; k0 < n; k0 += 64) {
    const int k = k0 + lane;
    const bool in = k < n && m2[g.seg[beg + k].colz / NZ] != 0;
    const unsigned long long m = __ballot(in);
    if (m) {
      if (first < 0) first = k0 + __ffsll((long long)m) - 1;
      last = k0 + 63 - __clzll(m);
    }
  }
  if (lane == 0) {
    win[2 * aw] = first < 0 ? 0 : first;
    win[2 * aw + 1] = first < 0 ? 0 : last - first + 1;
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
// Lane d computes the pieces of each segment of the run (times the 3D length
// factor sf of d) and one run word: w0 of the first segment, which segments
// hold a z boundary, and the w0 step after each segment. Without a boundary a
// segment has the single piece L sf (L: its length, known to every lane), so
// the d loop broadcasts one word per run and the pieces only for boundary
// segments (about one in four). va / vb are read from an LDS copy of the
// sinogram row through per-(d, r0) tables of byte offsets into that row; a
// table row spans every r0 the gather can form (-d-1 .. 64), with the entries
// outside 0 .. 63-d pointing at a zero row, so a reload is one add per table
// and no clamp. No atomics inside a ray: each wave flushes its segments
// through a private LDS buffer with one coalesced global float atomic per
// (segment, plane, replicate).

// Segments per run (SPW): compiled for two run lengths per width; the
// autotuner picks one (nipet_tune_set_run). Defaults, measured on MI355X:
// 8 for one image, 4 for two, 6 for four.
template <int NB>
constexpr int bp_spw_a() { return NB == 1 ? 8 : (NB == 2 ? 4 : 6); }
template <int NB>
constexpr int bp_spw_b() { return NB == 1 ? 6 : (NB == 2 ? 6 : 4); }

template <int NB, int SPW>
__global__ __launch_bounds__(BP_THREADS, NB == 1 ? 6 : 4) void k_bprj(GeomView g, float *__restrict__ img,
                                                     const float *__restrict__ sino,
                                                     const int32_t *__restrict__ subs, Batch bt,
                                                     int by_bin) {
  static_assert(SPW <= 8, "the run word holds the boundary bits and steps of 8 segments");
  constexpr int NWAVE = BP_THREADS / 64;
  __shared__ __align__(16) float srow[(MAXNSN / NB + 1) * NB];   // sinogram row + zero row
  __shared__ __align__(16) float wbuf[NWAVE][ZSTRIDE * NB];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int row = blockIdx.x;
  const int aw = subs ? subs[row] : row;
  if (aw < 0 || aw >= g.naw) return;
  const int nsn = g.nsn, nbt = bt.nbt, b0 = bt.b0;
  // sinogram rows in subset order, or (by_bin) the full [naw][nsn] sinogram
  // read at the bin's own row
  const float *in = sino + (size_t)(by_bin ? aw : row) * nsn * nbt + b0;
  int beg, nseg;
  seg_range(g, bt, aw, beg, nseg);
  const double atn = g.atn[aw];

  // LDS: sinogram row, zero row at index nsn, and the LOR -> sinogram tables
  // (ta: LOR (r0, d), tb: its mirror) as byte offsets into the row; entries
  // outside 0 <= r0 <= 63 - d point at the zero row
  for (int i = tid; i < (nsn + 1) * NB; i += BP_THREADS) {
    const int s = i / NB, b = i - s * NB;
    srow[i] = s < nsn ? in[(size_t)s * nbt + b] : 0.f;
  }
  // (built once with the geometry for this width; copied in 16-byte pieces)
  static_assert((2 * BP_TABN * sizeof(uint16_t)) % 16 == 0, "tables copy as uint4");
  __shared__ __align__(16) uint16_t tab[2 * BP_TABN];
  const uint16_t *ta = tab, *tb = tab + BP_TABN;
  {
    const uint4 *src = (const uint4 *)(g.bp_tab + (size_t)(NB == 1 ? 0 : (NB == 2 ? 1 : 2)) * 2 *
                                                      BP_TABN);
    uint4 *dst = (uint4 *)tab;
    for (int i = tid; i < 2 * BP_TABN * (int)sizeof(uint16_t) / 16; i += BP_THREADS)
      dst[i] = src[i];
  }
  __syncthreads();
  const char *srb = (const char *)srow;
  const int dmax = g.dmax;
  const double m_lane = 2.0 * lane / atn;
  const double inv_lane = lane > 0 ? atn / (2.0 * lane) : 0.0;
  const float sf_lane = lor_sfac(2.f * lane, g.vox_z, atn);
  const int q = lane;
  const char *tbyte = (const char *)tab;
  const int qb = 2 * q;  // lane's byte offset in a table row
  // direct planes: plane 2q <- L * va0[q]
  float X0[NB];
  ldv<NB>(X0, (const float *)(srb + ta[bp_row(0) + 1 + q]));
  float *wb = wbuf[wave];

  for (int g0 = wave * SPW; g0 < nseg; g0 += NWAVE * SPW) {
    // lane d: base plane and pieces (times the 3D factor) of each segment
    float P1[SPW], P2[SPW], P3[SPW], L[SPW];
    int RW = 0;  // the run word of lane d
    int wlast = -1;
    static_for<SPW>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
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
      P1[s] = P2[s] = P3[s] = 0.f;
      if (lane >= 1 && lane <= dmax) {
        // base plane, kept in 0 .. 2d so the table indices stay inside row d
        double w0 = fmin(fmax(floor(0.5 + m_lane * sg.s), 0.0), 2.0 * lane);
        double tb = (w0 + 0.5) * inv_lane;
        float q1 = (float)(tb - sg.s), q2 = (float)(tb + inv_lane - sg.s);
        float p1 = __builtin_amdgcn_fmed3f(q1, 0.f, sg.len);
        float c2 = __builtin_amdgcn_fmed3f(q2, 0.f, sg.len);
        P1[s] = p1 * sf_lane;
        P2[s] = (c2 - p1) * sf_lane;
        P3[s] = (sg.len - c2) * sf_lane;
        // base plane (bits 0-7), pieces 2 / 3 non-zero (bits 8, 9), base plane
        // differs from the previous segment's: the LOR values change (bit 10)
        // run word: base plane of the first segment (bits 0-7), a z boundary
        // inside segment s (bit 8 + s), base-plane step from segment s to
        // s + 1 (bits 16 + 2s, 2 bits; at most 2 by the geometry check)
        if (s == 0)
          RW = (int)w0;
        else if (k < nseg)
          RW |= min((int)w0 - wlast, 3) << (16 + 2 * (s - 1));
        RW |= (int)(P2[s] != 0.f || P3[s] != 0.f) << (8 + s);
        wlast = (int)w0;
      }
    });
    float acc0[SPW][NB], acc1[SPW][NB];
    static_for<SPW>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
#pragma unroll
      for (int b = 0; b < NB; b++) {
        acc0[s][b] = L[s] * X0[b];
        acc1[s][b] = 0.f;
      }
    });
    // The run word of d gives the base plane, the segments that hold a z
    // boundary and the base-plane steps. Every segment first takes the single
    // piece L * sf on plane 2q (even base plane) or 2q + 1: two FMAs per image
    // with the parity folded into the coefficients (sf * S1 or 0), no
    // broadcast. A segment with a boundary then swaps that piece for its
    // three broadcast pieces, the parity as SGPR bit masks.
    // rb: entry of r0 = 0 in row d, bp_row(d) + d + 1, carried from d to d + 1
    for (int d = 1, rb = bp_row(1) + 2; d <= dmax; rb += 67 + d, d++) {
      const int rw = __builtin_amdgcn_readlane(RW, d);
      const float sf = rdlane(sf_lane, d);
      int w0 = rw & 0xff;
      float S1[NB], S3[NB], Te[NB], To[NB], Ae[NB], Ao[NB];
      auto reload = [&]() {
        // 0 <= w0 <= 2d (the run word's start and steps keep it there) puts
        // q - h and q - e in -d .. 63, inside row d
        const int h = w0 >> 1, e = (2 * d - w0) >> 1;
        float X[NB], Y[NB], Xb[NB], Yb[NB];
        if constexpr (NB >= 4) {
          // ta[ka - 1], ta[ka] and tb[kb], tb[kb + 1]: one per-lane byte
          // offset (2 q) plus a wave-uniform one per pair, the pair's second
          // entry at an immediate offset (2 VALU instead of 6; the scalar
          // side grows, which one and two images cannot spare: measured
          // 1.352 -> 1.387 ms for one image, 0.822 -> 0.792 ms per image at 4)
          const uint16_t *pa = (const uint16_t *)(tbyte + qb + 2 * (rb - h - 1));
          const uint16_t *pb = (const uint16_t *)(tbyte + qb + 2 * (BP_TABN + rb - e));
          ldv<NB>(Y, (const float *)(srb + pa[0]));
          ldv<NB>(X, (const float *)(srb + pa[1]));
          ldv<NB>(Xb, (const float *)(srb + pb[0]));
          ldv<NB>(Yb, (const float *)(srb + pb[1]));
        } else {
          const int ka = rb - h + q, kb = rb - e + q;
          ldv<NB>(X, (const float *)(srb + ta[ka]));
          ldv<NB>(Y, (const float *)(srb + ta[ka - 1]));
          ldv<NB>(Xb, (const float *)(srb + tb[kb]));
          ldv<NB>(Yb, (const float *)(srb + tb[kb + 1]));
        }
#pragma unroll
        for (int b = 0; b < NB; b++) {
          S1[b] = X[b] + Xb[b];
          S3[b] = Y[b] + Yb[b];
          Te[b] = X[b] + Yb[b];
          To[b] = Y[b] + Xb[b];
          const float a = sf * S1[b];
          Ae[b] = (w0 & 1) ? 0.f : a;
          Ao[b] = (w0 & 1) ? a : 0.f;
        }
      };
      reload();
      static_for<SPW>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
#pragma unroll
        for (int b = 0; b < NB; b++) {
          acc0[s][b] = fmaf(L[s], Ae[b], acc0[s][b]);
          acc1[s][b] = fmaf(L[s], Ao[b], acc1[s][b]);
        }
        // one scalar test for the common step (no boundary, same base plane)
        constexpr int EV = (1 << (8 + s)) | (s + 1 < SPW ? 3 << (16 + 2 * s) : 0);
        if (rw & EV) {
          if (rw & (1 << (8 + s))) {  // wave-uniform, about one segment in four
            // the single piece is taken back; then three pieces: planes w0
            // (p1), w0 + 1 (p2) and w0 + 2 (p3) of the LORs, which by the
            // (wave-uniform) parity of w0 fall on the lane's even plane 2q
            // (acc0) or odd plane 2q + 1 (acc1)
            if constexpr (NB == 1) {
              // one image: parity as SGPR bit masks, no branch (measured
              // faster here than the branch below: 1.355 against 1.371 ms)
              const int mo = (w0 << 31) >> 31;  // all ones for an odd base plane
              const int p1 = __builtin_amdgcn_readlane(__float_as_int(P1[s]), d);
              const int p2 = __builtin_amdgcn_readlane(__float_as_int(P2[s]), d);
              const int p3 = __builtin_amdgcn_readlane(__float_as_int(P3[s]), d);
              const float a1 = __int_as_float(p1 & ~mo), c1 = __int_as_float(p1 & mo);
              const float a2 = __int_as_float(p2 & mo), c2 = __int_as_float(p2 & ~mo);
              const float a3 = __int_as_float(p3 & ~mo), c3 = __int_as_float(p3 & mo);
              acc0[s][0] = fmaf(-L[s], Ae[0], acc0[s][0]);
              acc1[s][0] = fmaf(-L[s], Ao[0], acc1[s][0]);
              acc0[s][0] = fmaf(a1, S1[0], acc0[s][0]);
              acc0[s][0] = fmaf(a2, To[0], acc0[s][0]);
              acc0[s][0] = fmaf(a3, S3[0], acc0[s][0]);
              acc1[s][0] = fmaf(c1, S1[0], acc1[s][0]);
              acc1[s][0] = fmaf(c2, Te[0], acc1[s][0]);
              acc1[s][0] = fmaf(c3, S3[0], acc1[s][0]);
            } else {
              // several images: a scalar branch on the parity skips the
              // zero-coefficient FMAs (3 + 1 per image instead of 8)
              const float p1 = rdlane(P1[s], d), p2 = rdlane(P2[s], d), p3 = rdlane(P3[s], d);
              if (w0 & 1) {
#pragma unroll
                for (int b = 0; b < NB; b++) {
                  acc1[s][b] = fmaf(-L[s], Ao[b], acc1[s][b]);
                  acc1[s][b] = fmaf(p1, S1[b], acc1[s][b]);
                  acc0[s][b] = fmaf(p2, To[b], acc0[s][b]);
                  acc1[s][b] = fmaf(p3, S3[b], acc1[s][b]);
                }
                // distinct markers keep the two branches apart: merged, they
                // select the accumulator by parity and move the accumulators
                // to scratch memory
                asm volatile("; bp odd base plane");
              } else {
#pragma unroll
                for (int b = 0; b < NB; b++) {
                  acc0[s][b] = fmaf(-L[s], Ae[b], acc0[s][b]);
                  acc0[s][b] = fmaf(p1, S1[b], acc0[s][b]);
                  acc1[s][b] = fmaf(p2, Te[b], acc1[s][b]);
                  acc0[s][b] = fmaf(p3, S3[b], acc0[s][b]);
                }
                asm volatile("; bp even base plane");
              }
            }
          }
          if constexpr (s + 1 < SPW) {
            const int dl = (rw >> (16 + 2 * s)) & 3;
            if (dl) {  // wave-uniform: the next segment starts on a new base plane
              w0 += dl;
              reload();
            }
          }
        }
      });
    }
    // flush the wave's segments through its LDS buffer (wave-local, no barrier)
    static_for<SPW>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      if (g0 + s >= nseg) return;
#pragma unroll
      for (int b = 0; b < NB; b++) {
        wb[(2 * q) * NB + b] = acc0[s][b];
        wb[(2 * q + 1) * NB + b] = acc1[s][b];
      }
      __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the wave's LDS writes landed
      __builtin_amdgcn_wave_barrier();
      const int colz = g.seg[beg + g0 + s].colz;
      for (int i = lane; i < NZ * NB; i += 64) {
        const int z = i / NB, b = i - z * NB;
        const float v = wb[i];
        if (v != 0.f) unsafeAtomicAdd(img + ((int64_t)colz + z) * nbt + b0 + b, v);
      }
      __builtin_amdgcn_wave_barrier();
    });
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

// k_osem_update without the per-element index division: one image, four
// voxels per thread (16-byte loads; mask as 4 bytes), or nb = 2 / 4 images of
// one voxel per thread. The same operations per element.
__device__ inline float osem_q(float b, float sv) {
  return (sv > -EPS && sv < EPS) ? 0.f : b / sv;
}
__global__ void k_osem_update_v4(float4 *img, const float4 *b, const float4 *sens,
                                 const uchar4 *mask, int64_t n4) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n4; i += stride) {
    const float4 x = img[i], bb = b[i], sv = sens[i];
    const uchar4 m = mask[i];
    img[i] = make_float4(m.x ? x.x * osem_q(bb.x, sv.x) : 0.f, m.y ? x.y * osem_q(bb.y, sv.y) : 0.f,
                         m.z ? x.z * osem_q(bb.z, sv.z) : 0.f, m.w ? x.w * osem_q(bb.w, sv.w) : 0.f);
  }
}
template <int NB>
__global__ void k_osem_update_w(float *img, const float *b, const float *sens,
                                const uint8_t *mask, int64_t nvox, int sens_nb) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; v < nvox; v += stride) {
    float x[NB], bb[NB], sv[NB], out[NB];
    ldv<NB>(x, img + v * NB);
    ldv<NB>(bb, b + v * NB);
    if (sens_nb) {
      ldv<NB>(sv, sens + v * NB);
    } else {
#pragma unroll
      for (int k = 0; k < NB; k++) sv[k] = sens[v];
    }
    const bool m = mask[v] != 0;
#pragma unroll
    for (int k = 0; k < NB; k++) out[k] = m ? x[k] * osem_q(bb[k], sv[k]) : 0.f;
    stv<NB>(img + v * NB, out);
  }
}

// The first-iteration pair: e (the subset's ratio sinogram, [nprj][nsn]) and
// the a n rows of the subset's bins, interleaved [nprj][nsn][2] for one
// two-image back projection (zero rows for bins outside 0..naw-1).
__global__ void k_pack_pair(float2 *out, const float *e, const float *an, const int32_t *subs,
                            int nsn, int naw, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const int64_t row = i / nsn;
    const int aw = subs[row];
    const float a = (aw >= 0 && aw < naw) ? an[(int64_t)aw * nsn + (i - row * nsn)] : 0.f;
    out[i] = make_float2(e[i], a);
  }
}

// k_osem_update (nb = 1) from the two-image back projection b2[v] = (BP(e),
// BP(a n)): the second image is the subset's sensitivity, stored to sens.
__global__ void k_osem_update_pair(float *img, float *sens, const float2 *b2,
                                   const uint8_t *mask, int64_t nvox) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < nvox; i += stride) {
    const float2 bs = b2[i];
    const float sv = bs.y;
    sens[i] = sv;
    float q = (sv > -EPS && sv < EPS) ? 0.f : bs.x / sv;
    img[i] = mask[i] ? img[i] * q : 0.f;
  }
}

// Siemens histogram with replicates innermost, in[(s * nbinang + ali) * nrep + r]
// (uint32 counts) -> device prompts out[(aw * nsn + s) * out_nb + out_b0 + b]
// (uint16, saturating at 65535) for replicates r = r0 + b, b < nb.
__global__ void k_remgaps_u16b(uint16_t *out, const uint32_t *in, const int32_t *aw2ali, int naw,
                               int nsn, int nbinang, int nrep, int r0, int nb, int out_nb,
                               int out_b0) {
  const int64_t n = (int64_t)naw * nsn * nb;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t t = i / nb;
    const int b = (int)(i - t * nb);
    const int aw = (int)(t / nsn), s = (int)(t - (int64_t)aw * nsn);
    const uint32_t v = in[((int64_t)s * nbinang + aw2ali[aw]) * nrep + r0 + b];
    out[t * out_nb + out_b0 + b] = (uint16_t)(v > 65535u ? 65535u : v);
  }
}

// Gap removal / insertion (d_remgaps / d_putgaps, src/scanner_0.cu:94-105,
// 158-166, plus the transpose of mmraux.py:1141): tiled transposes between the
// Siemens layout (nsn, nbinang) and the device layout (naw, nsn): a 64 x 64 tile of (Siemens bin, sinogram) passes through
// LDS, so Siemens rows are read / written along the bin index and device rows
// along the sinogram index, both coalesced. inv[ali] = active bin of Siemens
// bin ali, or -1 in the gaps (zeros on output).
constexpr int GT = 64;

__global__ void k_gap_inv(int32_t *inv, int nbinang) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbinang) inv[i] = -1;
}

__global__ void k_gap_inv_set(int32_t *inv, const int32_t *aw2ali, int naw, int nbinang) {
  const int aw = blockIdx.x * blockDim.x + threadIdx.x;
  if (aw < naw) {
    const int a = aw2ali[aw];
    if (a >= 0 && a < nbinang) inv[a] = aw;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_remgaps_t(float *__restrict__ out,
                                                   const T *__restrict__ in,
                                                   const int32_t *__restrict__ inv, int nsn,
                                                   int nbinang) {
  __shared__ float tile[GT][GT + 1];
  const int a0 = blockIdx.x * GT, s0 = blockIdx.y * GT;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < GT; r += 4) {  // r: sinogram, tx: Siemens bin
    const int s = s0 + r, a = a0 + tx;
    tile[r][tx] = (s < nsn && a < nbinang) ? (float)in[(int64_t)s * nbinang + a] : 0.f;
  }
  __syncthreads();
  for (int r = ty; r < GT; r += 4) {  // r: Siemens bin, tx: sinogram
    const int a = a0 + r, s = s0 + tx;
    if (a >= nbinang || s >= nsn) continue;
    const int aw = inv[a];
    if (aw >= 0) out[(int64_t)aw * nsn + s] = tile[tx][r];
  }
}

__global__ __launch_bounds__(256) void k_putgaps_t(float *__restrict__ out,
                                                   const float *__restrict__ in,
                                                   const int32_t *__restrict__ inv, int nsn,
                                                   int nbinang) {
  __shared__ float tile[GT][GT + 1];
  const int a0 = blockIdx.x * GT, s0 = blockIdx.y * GT;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < GT; r += 4) {  // r: Siemens bin, tx: sinogram
    const int a = a0 + r, s = s0 + tx;
    float v = 0.f;
    if (a < nbinang && s < nsn) {
      const int aw = inv[a];
      if (aw >= 0) v = in[(int64_t)aw * nsn + s];
    }
    tile[tx][r] = v;
  }
  __syncthreads();
  for (int r = ty; r < GT; r += 4) {  // r: sinogram, tx: Siemens bin (gaps get 0)
    const int s = s0 + r, a = a0 + tx;
    if (s < nsn && a < nbinang) out[(int64_t)s * nbinang + a] = tile[r][tx];
  }
}

static int ew_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (int)(b < 8192 ? b : 8192);
}

static void launch_osem_update(float *img, const float *b, const float *sens, const uint8_t *mask,
                               int nb, int sens_nb, hipStream_t st) {
  auto al = [](const void *p, size_t a) { return (uintptr_t)p % a == 0; };
  const size_t a = 4 * (size_t)nb;
  if (nb == 1 && NVOX % 4 == 0 && al(img, 16) && al(b, 16) && al(sens, 16) && al(mask, 4))
    k_osem_update_v4<<<ew_blocks(NVOX / 4), 256, 0, st>>>(
        (float4 *)img, (const float4 *)b, (const float4 *)sens, (const uchar4 *)mask, NVOX / 4);
  else if ((nb == 2 || nb == 4) && al(img, a) && al(b, a) && (!sens_nb || al(sens, a)))
    (nb == 2 ? k_osem_update_w<2> : k_osem_update_w<4>)<<<ew_blocks(NVOX), 256, 0, st>>>(
        img, b, sens, mask, NVOX, sens_nb);
  else
    k_osem_update<<<ew_blocks(NVOX * nb), 256, 0, st>>>(img, b, sens, mask, NVOX, nb, sens_nb);
}

}  // namespace nipet

using namespace nipet;

// Launch a projector over nb replicates as sub-batches of 4, 2 and 1.
// Largest replicate sub-batch per kernel instance (4, 2 or 1), set by the
// autotuner through nipet_tune_set; process-wide.
static std::atomic<int> g_width_fp{4}, g_width_bp{4};
// BP segments per run for widths 1, 2, 4 (one of bp_spw_a / bp_spw_b)
static std::atomic<int> g_run_bp[3] = {{bp_spw_a<1>()}, {bp_spw_a<2>()}, {bp_spw_a<4>()}};

// One BP launch of width NB with the tuned run length.
template <int NB>
static void launch_bprj_w(const Geom *g, float *img, const float *sino, const int32_t *subs,
                          int nprj, Batch bt, int by_bin, hipStream_t st) {
  // bt.win: the bin's segment window, or every segment when NULL
  const int run = g_run_bp[NB == 1 ? 0 : (NB == 2 ? 1 : 2)].load();
  if (run == bp_spw_b<NB>())
    k_bprj<NB, bp_spw_b<NB>()><<<nprj, BP_THREADS, 0, st>>>(view(g), img, sino, subs, bt, by_bin);
  else
    k_bprj<NB, bp_spw_a<NB>()><<<nprj, BP_THREADS, 0, st>>>(view(g), img, sino, subs, bt, by_bin);
}

template <int MODE>
static void launch_fprj(const Geom *g, float *sino, const float *img, const int32_t *subs,
                        int nprj, const uint16_t *prompts, const float *rho, int rho_nb, int nb,
                        hipStream_t st, const int32_t *win = nullptr) {
  const int wmax = g_width_fp.load();
  for (int b0 = 0; b0 < nb;) {
    const int rem = nb - b0;
    const Batch bt{nb, b0, win};
    if (rem >= 4 && wmax >= 4) {
      k_fprj<MODE, 4><<<nprj, FP_THREADS, 0, st>>>(view(g), img, sino, subs, prompts, rho,
                                                   rho_nb, bt);
      b0 += 4;
    } else if (rem >= 2 && wmax >= 2) {
      k_fprj<MODE, 2><<<nprj, FP_THREADS, 0, st>>>(view(g), img, sino, subs, prompts, rho,
                                                   rho_nb, bt);
      b0 += 2;
    } else {
      k_fprj_g<MODE, 1, FP_G><<<nprj, fpg_threads<FP_G>(), 0, st>>>(view(g), img, sino, subs,
                                                                    prompts, rho, rho_nb, bt);
      b0 += 1;
    }
  }
}

static void launch_bprj(const Geom *g, float *img, const float *sino, const int32_t *subs,
                        int nprj, int nb, hipStream_t st, int by_bin = 0,
                        const int32_t *win = nullptr) {
  const int wmax = g_width_bp.load();
  for (int b0 = 0; b0 < nb;) {
    const int rem = nb - b0;
    const Batch bt{nb, b0, win};
    if (rem >= 4 && wmax >= 4) {
      launch_bprj_w<4>(g, img, sino, subs, nprj, bt, by_bin, st);
      b0 += 4;
    } else if (rem >= 2 && wmax >= 2) {
      launch_bprj_w<2>(g, img, sino, subs, nprj, bt, by_bin, st);
      b0 += 2;
    } else {
      launch_bprj_w<1>(g, img, sino, subs, nprj, bt, by_bin, st);
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

int nipet_osem_sens(const NipetGeom *g, float *sens, const float *an, const int32_t *subs,
                    int nprj, void *stream) {
  int rc = check_geom(g, "nipet_osem_sens");
  if (rc) return rc;
  if ((rc = check_nprj(g, subs, nprj, 1, "nipet_osem_sens"))) return rc;
  if (!sens || !an || !subs) {
    set_error("nipet_osem_sens: NULL buffer");
    return NIPET_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  prof_begin(PK_FILL, st);
  k_fill<<<ew_blocks(NVOX), 256, 0, st>>>(sens, 0.f, NVOX);
  prof_end(PK_FILL, st);
  prof_begin(PK_BPRJ, st);
  launch_bprj(g, sens, an, subs, nprj, 1, st, 1);
  prof_end(PK_BPRJ, st);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_osem_window(const NipetGeom *g, const uint8_t *mask, int dilate, int32_t *win,
                      void *stream) {
  int rc = check_geom(g, "nipet_osem_window");
  if (rc) return rc;
  if (!mask || !win || dilate < 0 || dilate > 64) {
    set_error("nipet_osem_window: NULL buffer or dilate outside 0 .. 64");
    return NIPET_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  uint8_t *m2 = nullptr, *m2d = nullptr;
  hipError_t e = hipMallocAsync((void **)&m2, (size_t)NX * NY, st);
  if (e == hipSuccess && dilate > 0) e = hipMallocAsync((void **)&m2d, (size_t)NX * NY, st);
  if (e != hipSuccess) {
    if (m2) (void)hipFreeAsync(m2, st);
    return hip_fail(e, "nipet_osem_window: hipMallocAsync");
  }
  k_mask_cols<<<(NX * NY + 255) / 256, 256, 0, st>>>(mask, m2);
  if (dilate > 0) k_mask_dilate<<<(NX * NY + 255) / 256, 256, 0, st>>>(m2, m2d, dilate);
  k_seg_window<<<(g->naw + 3) / 4, 256, 0, st>>>(view(g), dilate > 0 ? m2d : m2, win);
  (void)hipFreeAsync(m2, st);
  if (m2d) (void)hipFreeAsync(m2d, st);
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
  hipStream_t st = (hipStream_t)stream;
  int32_t *inv = nullptr;
  NIPET_HIP(hipMallocAsync((void **)&inv, (size_t)nbinang * sizeof(int32_t), st));
  k_gap_inv<<<(nbinang + 255) / 256, 256, 0, st>>>(inv, nbinang);
  k_gap_inv_set<<<(naw + 255) / 256, 256, 0, st>>>(inv, aw2ali, naw, nbinang);
  const dim3 grid((nbinang + GT - 1) / GT, (nsn + GT - 1) / GT);
  if (in_dtype == 0)
    k_remgaps_t<float><<<grid, 256, 0, st>>>(out, (const float *)in, inv, nsn, nbinang);
  else
    k_remgaps_t<uint16_t><<<grid, 256, 0, st>>>(out, (const uint16_t *)in, inv, nsn, nbinang);
  NIPET_HIP(hipGetLastError());
  NIPET_HIP(hipFreeAsync(inv, st));
  return NIPET_OK;
}

// ---- allocation-free gap removal ----
// The inverse bin map of nipet_remgaps lives in a file-scope device array
// instead of stream-ordered scratch. Host code never passes `scratch` itself
// to a kernel (in host code that name is a host-side stand-in, not a device
// address): the address is looked up once per device with
// hipGetSymbolAddress, with that device current, and kept in g_scratch_dev.
static __device__ __align__(256) int32_t scratch[NIPET_NOALLOC_MAXBINANG];

static std::mutex g_scratch_mu;
static int32_t *g_scratch_dev[NIPET_MAXDEVICES] = {};

static int scratch_ptr(int32_t **out, const char *fn) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, fn);
  if (dev < 0 || dev >= NIPET_MAXDEVICES) {
    set_error("%s: device %d beyond the %d devices this library supports", fn, dev,
              NIPET_MAXDEVICES);
    return NIPET_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("%s: hipGetSymbolAddress(scratch) failed on device %d: %s", fn, dev,
                hipGetErrorString(e));
      return NIPET_EHIP;
    }
    g_scratch_dev[dev] = (int32_t *)p;
  }
  *out = g_scratch_dev[dev];
  return NIPET_OK;
}

int nipet_remgaps_noalloc(float *out, const void *in, int in_dtype, const int32_t *aw2ali,
                          int naw, int nsn, int nbinang, void *stream) {
  if (!out || !in || !aw2ali || naw <= 0 || nsn <= 0 || nbinang < naw ||
      (in_dtype != 0 && in_dtype != 1)) {
    set_error("nipet_remgaps_noalloc: bad argument");
    return NIPET_EINVAL;
  }
  if (nbinang > NIPET_NOALLOC_MAXBINANG) {
    set_error("nipet_remgaps_noalloc: nbinang = %d exceeds the static scratch of %d bins",
              nbinang, NIPET_NOALLOC_MAXBINANG);
    return NIPET_EINVAL;
  }
  int32_t *inv = nullptr;
  int rc = scratch_ptr(&inv, "nipet_remgaps_noalloc");
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  // the whole map is rewritten by a kernel in every call (no memset node, so
  // the call may be captured into a graph)
  k_gap_inv<<<(nbinang + 255) / 256, 256, 0, st>>>(inv, nbinang);
  k_gap_inv_set<<<(naw + 255) / 256, 256, 0, st>>>(inv, aw2ali, naw, nbinang);
  const dim3 grid((nbinang + GT - 1) / GT, (nsn + GT - 1) / GT);
  if (in_dtype == 0)
    k_remgaps_t<float><<<grid, 256, 0, st>>>(out, (const float *)in, inv, nsn, nbinang);
  else
    k_remgaps_t<uint16_t><<<grid, 256, 0, st>>>(out, (const uint16_t *)in, inv, nsn, nbinang);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_remgaps_u16b(uint16_t *out, const uint32_t *in, const int32_t *aw2ali, int naw,
                       int nsn, int nbinang, int nrep, int r0, int nb, int out_nb, int out_b0,
                       void *stream) {
  if (!out || !in || !aw2ali || naw <= 0 || nsn <= 0 || nbinang <= 0 || nrep < 1 || r0 < 0 ||
      nb < 1 || r0 + nb > nrep || out_b0 < 0 || out_b0 + nb > out_nb) {
    set_error("nipet_remgaps_u16b: bad arguments (naw=%d nsn=%d nrep=%d r0=%d nb=%d out_nb=%d "
              "out_b0=%d)", naw, nsn, nrep, r0, nb, out_nb, out_b0);
    return NIPET_EINVAL;
  }
  const int64_t n = (int64_t)naw * nsn * nb;
  k_remgaps_u16b<<<ew_blocks(n), 256, 0, (hipStream_t)stream>>>(out, in, aw2ali, naw, nsn,
                                                               nbinang, nrep, r0, nb, out_nb,
                                                               out_b0);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_tune_set(int kernel, int width) {
  if ((kernel != NIPET_PROF_FPRJ && kernel != NIPET_PROF_BPRJ) ||
      (width != 1 && width != 2 && width != 4)) {
    set_error("nipet_tune_set: kernel must be FPRJ or BPRJ and width 1, 2 or 4");
    return NIPET_EINVAL;
  }
  (kernel == NIPET_PROF_FPRJ ? g_width_fp : g_width_bp).store(width);
  return NIPET_OK;
}

int nipet_tune_set_run(int width, int spw) {
  const int k = width == 1 ? 0 : (width == 2 ? 1 : (width == 4 ? 2 : -1));
  const int a = k == 0 ? bp_spw_a<1>() : (k == 1 ? bp_spw_a<2>() : bp_spw_a<4>());
  const int b = k == 0 ? bp_spw_b<1>() : (k == 1 ? bp_spw_b<2>() : bp_spw_b<4>());
  if (k < 0 || (spw != a && spw != b)) {
    set_error("nipet_tune_set_run: width must be 1, 2 or 4 and the run length one of the "
              "compiled ones (%d or %d)", a, b);
    return NIPET_EINVAL;
  }
  g_run_bp[k].store(spw);
  return NIPET_OK;
}

int nipet_tune_get_run(int width, int *spw, int *alt) {
  const int k = width == 1 ? 0 : (width == 2 ? 1 : (width == 4 ? 2 : -1));
  if (k < 0 || !spw) {
    set_error("nipet_tune_get_run: width must be 1, 2 or 4");
    return NIPET_EINVAL;
  }
  *spw = g_run_bp[k].load();
  if (alt) {
    const int a = k == 0 ? bp_spw_a<1>() : (k == 1 ? bp_spw_a<2>() : bp_spw_a<4>());
    const int b = k == 0 ? bp_spw_b<1>() : (k == 1 ? bp_spw_b<2>() : bp_spw_b<4>());
    *alt = *spw == a ? b : a;
  }
  return NIPET_OK;
}

int nipet_tune_get(int kernel, int *width) {
  if ((kernel != NIPET_PROF_FPRJ && kernel != NIPET_PROF_BPRJ) || !width) {
    set_error("nipet_tune_get: kernel must be FPRJ or BPRJ");
    return NIPET_EINVAL;
  }
  *width = (kernel == NIPET_PROF_FPRJ ? g_width_fp : g_width_bp).load();
  return NIPET_OK;
}

int nipet_putgaps(float *out, const float *in, const int32_t *aw2ali, int naw, int nsn,
                  int nbinang, void *stream) {
  if (!out || !in || !aw2ali || naw <= 0 || nsn <= 0 || nbinang < naw) {
    set_error("nipet_putgaps: bad argument");
    return NIPET_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  int32_t *inv = nullptr;
  NIPET_HIP(hipMallocAsync((void **)&inv, (size_t)nbinang * sizeof(int32_t), st));
  k_gap_inv<<<(nbinang + 255) / 256, 256, 0, st>>>(inv, nbinang);
  k_gap_inv_set<<<(naw + 255) / 256, 256, 0, st>>>(inv, aw2ali, naw, nbinang);
  const dim3 grid((nbinang + GT - 1) / GT, (nsn + GT - 1) / GT);
  k_putgaps_t<<<grid, 256, 0, st>>>(out, in, inv, nsn, nbinang);
  NIPET_HIP(hipGetLastError());
  NIPET_HIP(hipFreeAsync(inv, st));
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
                        const uint8_t *mask, const int32_t *win, const int32_t *subs, int nprj,
                        int nb, float *work_sino, float *work_img, void *stream) {
  int rc = check_geom(g, "nipet_osem_subset");
  if (rc) return rc;
  if (!subs || nprj <= 0 || nprj > g->naw) {
    set_error("nipet_osem_subset: subs must be given with 0 < nprj <= naw");
    return NIPET_EINVAL;
  }
  if ((rc = check_nprj(g, subs, nprj, nb, "nipet_osem_subset"))) return rc;
  if (!img || !prompts || !rho || !sens || !mask || !work_sino || !work_img || work_img == img ||
      work_img == sens) {
    set_error("nipet_osem_subset: NULL or aliased buffer");
    return NIPET_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  prof_begin(PK_FPRJ, st);
  launch_fprj<2>(g, work_sino, img, subs, nprj, prompts, rho, rho_per_rep ? 1 : 0, nb, st, win);
  prof_end(PK_FPRJ, st);
  prof_begin(PK_FILL, st);
  k_fill<<<ew_blocks(NVOX * nb), 256, 0, st>>>(work_img, 0.f, NVOX * nb);
  prof_end(PK_FILL, st);
  prof_begin(PK_BPRJ, st);
  launch_bprj(g, work_img, work_sino, subs, nprj, nb, st, 0, win);
  prof_end(PK_BPRJ, st);
  prof_begin(PK_OSEM_UPDATE, st);
  launch_osem_update(img, work_img, sens, mask, nb, sens_per_rep ? 1 : 0, st);
  prof_end(PK_OSEM_UPDATE, st);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_osem_subset_rm_b(const NipetGeom *g, float *img, const uint16_t *prompts,
                           const float *rho, int rho_per_rep, const float *sens,
                           int sens_per_rep, const uint8_t *mask, const int32_t *win,
                           const int32_t *subs, int nprj, int nb, const float *krnl,
                           float *work_sino, float *work_img, float *work_img2, void *stream) {
  int rc = check_geom(g, "nipet_osem_subset_rm");
  if (rc) return rc;
  if (!subs || nprj <= 0 || nprj > g->naw) {
    set_error("nipet_osem_subset_rm: subs must be given with 0 < nprj <= naw");
    return NIPET_EINVAL;
  }
  if ((rc = check_nprj(g, subs, nprj, nb, "nipet_osem_subset_rm"))) return rc;
  if (!img || !prompts || !rho || !sens || !mask || !krnl || !work_sino || !work_img ||
      !work_img2 || work_img == work_img2 || work_img == img || work_img2 == img ||
      work_img == sens || work_img2 == sens) {
    set_error("nipet_osem_subset_rm: NULL or aliased buffer");
    return NIPET_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  // resolution-modelled image H x (work_img is the convolution scratch)
  if ((rc = psf_conv(work_img2, img, work_img, krnl, nb, st))) return rc;
  prof_begin(PK_FPRJ, st);
  launch_fprj<2>(g, work_sino, work_img2, subs, nprj, prompts, rho, rho_per_rep ? 1 : 0, nb, st,
                 win);
  prof_end(PK_FPRJ, st);
  prof_begin(PK_FILL, st);
  k_fill<<<ew_blocks(NVOX * nb), 256, 0, st>>>(work_img, 0.f, NVOX * nb);
  prof_end(PK_FILL, st);
  prof_begin(PK_BPRJ, st);
  launch_bprj(g, work_img, work_sino, subs, nprj, nb, st, 0, win);
  prof_end(PK_BPRJ, st);
  // H^T of the back projection (symmetric kernels: H), into work_img2
  if ((rc = psf_conv(work_img2, work_img, work_img, krnl, nb, st))) return rc;
  prof_begin(PK_OSEM_UPDATE, st);
  launch_osem_update(img, work_img2, sens, mask, nb, sens_per_rep ? 1 : 0, st);
  prof_end(PK_OSEM_UPDATE, st);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_osem_subset(const NipetGeom *g, float *img, const uint16_t *prompts, const float *rho,
                      const float *sens, const uint8_t *mask, const int32_t *win,
                      const int32_t *subs, int nprj, float *work_sino, float *work_img,
                      void *stream) {
  return nipet_osem_subset_b(g, img, prompts, rho, 0, sens, 0, mask, win, subs, nprj, 1,
                             work_sino, work_img, stream);
}

int nipet_osem_subset_sens(const NipetGeom *g, float *img, const uint16_t *prompts,
                           const float *rho, const float *an, float *sens, const uint8_t *mask,
                           const int32_t *win, const int32_t *subs, int nprj, float *work_sino,
                           float *work_sino2, float *work_img2, void *stream) {
  int rc = check_geom(g, "nipet_osem_subset_sens");
  if (rc) return rc;
  if (!subs || nprj <= 0 || nprj > g->naw) {
    set_error("nipet_osem_subset_sens: subs must be given with 0 < nprj <= naw");
    return NIPET_EINVAL;
  }
  if ((rc = check_nprj(g, subs, nprj, 2, "nipet_osem_subset_sens"))) return rc;
  if (!img || !prompts || !rho || !an || !sens || !mask || !work_sino || !work_sino2 ||
      !work_img2 || sens == img || work_sino == work_sino2 || work_img2 == img ||
      work_img2 == sens) {
    set_error("nipet_osem_subset_sens: NULL or aliased buffer");
    return NIPET_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  const int64_t nel = (int64_t)nprj * g->nsn;
  prof_begin(PK_FPRJ, st);
  launch_fprj<2>(g, work_sino, img, subs, nprj, prompts, rho, 0, 1, st, win);
  prof_end(PK_FPRJ, st);
  prof_begin(PK_OTHER, st);
  k_pack_pair<<<ew_blocks(nel), 256, 0, st>>>((float2 *)work_sino2, work_sino, an, subs, g->nsn,
                                               g->naw, nel);
  prof_end(PK_OTHER, st);
  prof_begin(PK_FILL, st);
  k_fill<<<ew_blocks(NVOX * 2), 256, 0, st>>>(work_img2, 0.f, NVOX * 2);
  prof_end(PK_FILL, st);
  // both images through one traversal of the rays (independent of the
  // batch autotuner's width)
  prof_begin(PK_BPRJ_PAIR, st);
  launch_bprj_w<2>(g, work_img2, work_sino2, subs, nprj, Batch{2, 0, win}, 0, st);
  prof_end(PK_BPRJ_PAIR, st);
  prof_begin(PK_OSEM_UPDATE, st);
  k_osem_update_pair<<<ew_blocks(NVOX), 256, 0, st>>>(img, sens, (const float2 *)work_img2, mask,
                                                      NVOX);
  prof_end(PK_OSEM_UPDATE, st);
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

int nipet_osem_iter(const NipetGeom *g, float *img, const uint16_t *prompts, const float *r,
                    const float *s, const float *n, const float *a, const int32_t *subs,
                    int nsub, int stride, const float *sens, const uint8_t *mask,
                    const float *krnl, void *stream) {
  int rc = check_geom(g, "nipet_osem_iter");
  if (rc) return rc;
  if (!img || !prompts || !n || !a || !subs || !sens || !mask || nsub < 1 || stride < 2) {
    set_error("nipet_osem_iter: NULL buffer, nsub < 1 or stride < 2");
    return NIPET_EINVAL;
  }
  int maxc = 0;
  for (int i = 0; i < nsub; i++) {
    const int c = subs[(int64_t)i * stride];
    if (c < 1 || c > stride - 1 || c > g->naw) {
      set_error("nipet_osem_iter: subset %d holds %d bins (must be 1 .. min(stride - 1 = %d, "
                "naw = %d))", i, c, stride - 1, g->naw);
      return NIPET_EINVAL;
    }
    maxc = c > maxc ? c : maxc;
  }
  const bool rm = krnl && krnl[0] >= 0.f;
  hipStream_t st = (hipStream_t)stream;
  const int64_t nel = (int64_t)g->naw * g->nsn;
  // per-call temporaries, stream ordered (the reference allocates and frees
  // its device copies in every call, prj/src/recon.cu:333-419, 511-520)
  float *rho = nullptr, *an = nullptr, *wsino = nullptr, *wimg = nullptr, *wimg2 = nullptr,
        *srm = nullptr;
  int32_t *dsubs = nullptr, *win = nullptr;
  auto release = [&]() {
    void *ps[] = {rho, an, wsino, wimg, wimg2, srm, dsubs, win};
    for (void *q : ps)
      if (q) (void)hipFreeAsync(q, st);
  };
  auto get = [&](void **p, size_t bytes) -> int {
    hipError_t e = hipMallocAsync(p, bytes, st);
    if (e != hipSuccess) {
      *p = nullptr;
      release();
      return hip_fail(e, "nipet_osem_iter: hipMallocAsync");
    }
    return NIPET_OK;
  };
  if ((rc = get((void **)&rho, nel * sizeof(float))) || (rc = get((void **)&an, nel * sizeof(float))) ||
      (rc = get((void **)&wsino, (size_t)maxc * g->nsn * sizeof(float))) ||
      (rc = get((void **)&wimg, NVOX * sizeof(float))) ||
      (rc = get((void **)&dsubs, (size_t)nsub * stride * sizeof(int32_t))) ||
      (rc = get((void **)&win, (size_t)g->naw * 2 * sizeof(int32_t))))
    return rc;
  if (rm && ((rc = get((void **)&wimg2, NVOX * sizeof(float))) ||
             (rc = get((void **)&srm, NVOX * sizeof(float)))))
    return rc;
  hipError_t e = hipMemcpyAsync(dsubs, subs, (size_t)nsub * stride * sizeof(int32_t),
                                hipMemcpyHostToDevice, st);
  if (e != hipSuccess) {
    release();
    return hip_fail(e, "nipet_osem_iter: subset table upload");
  }
  // rho = (r + s) / (a n) once per call (prj/src/recon.cu:381-398)
  k_osem_corr<<<ew_blocks(nel), 256, 0, st>>>(rho, an, a, n, r, s, nel);
  // the mask's segment windows (dilated by the resolution model's reach):
  // after the first update img is zero outside the mask, so the later subsets
  // walk only the segments whose columns meet it
  if ((rc = nipet_osem_window(g, mask, rm ? PSF_R : 0, win, stream))) {
    release();
    return rc;
  }
  for (int i = 0; i < nsub && rc == NIPET_OK; i++) {
    const int c = subs[(int64_t)i * stride];
    const int32_t *rows = dsubs + (int64_t)i * stride + 1;
    const float *si = sens + (int64_t)i * NVOX;
    const int32_t *wi = i > 0 ? win : nullptr;
    if (rm) {
      // the reference convolves the sensitivity images in every call
      // (prj/src/recon.cu:441-445); here one at a time, before its subset
      if ((rc = psf_conv(srm, si, wimg, krnl, 1, st))) break;
      rc = nipet_osem_subset_rm_b(g, img, prompts, rho, 0, srm, 0, mask, wi, rows, c, 1, krnl,
                                  wsino, wimg, wimg2, stream);
    } else {
      rc = nipet_osem_subset_b(g, img, prompts, rho, 0, si, 0, mask, wi, rows, c, 1, wsino, wimg,
                               stream);
    }
  }
  release();
  if (rc) return rc;
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

}  // extern "C"
