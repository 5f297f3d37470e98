// Sinograms that are products of per-(crystal, ring) factors, for gfx950:
//   out[aw][sn] = tw[aw] * sum_{m in members(sn)} w[m] A(aw, r0(m)) B(aw, r1(m)),
//   A(aw, r) = tab[c0(aw) cs + r rs],  B(aw, r) = tab[c1(aw) cs + r rs],
// where members(sn) are the span-1 sinograms (ring pairs r0, r1) of output
// sinogram sn in a fixed order (CSR). Used by the randoms sinogram (singles
// s[c][r], unit weights) and the normalisation sinogram (crystal factors
// E[r][c], axial weights w and transaxial factors tw).
//
// A block takes PAIR_AWB consecutive bins: their 2 x 64 ring factors are
// staged in LDS once, then each thread walks the members of one output
// sinogram for all PAIR_AWB bins (members read once per block) and writes
// PAIR_AWB coalesced rows. Each translation unit that includes this header
// gets its own copy of the kernel (internal linkage).
#pragma once

#include "nipet_common.h"

namespace nipet {
namespace {

constexpr int PAIR_AWB = 8;       // bins per block
constexpr int PAIR_THREADS = 256;

struct PairSino {
  const float *tab;      // factor table
  int cs, rs;            // crystal and ring strides of tab
  const short2 *cpair;   // [naw] crystals (c0, c1) of each bin
  const float *tw;       // [naw] per-bin factor, or nullptr (1)
  const int32_t *off;    // [nsn + 1] member ranges
  const short2 *mem_rr;  // [members] ring pair (r0, r1) of each member
  const float *mem_w;    // [members] member weight, or nullptr (1)
  int naw, nsn;
};

__global__ __launch_bounds__(PAIR_THREADS) void k_pair_sino(float *__restrict__ out, PairSino p) {
  __shared__ float va[PAIR_AWB][64], vb[PAIR_AWB][64];
  __shared__ float sw[PAIR_AWB];
  const int aw0 = blockIdx.x * PAIR_AWB;
  for (int i = threadIdx.x; i < PAIR_AWB * 64; i += PAIR_THREADS) {
    const int a = i >> 6, r = i & 63, aw = aw0 + a;
    float x = 0.f, y = 0.f;
    if (aw < p.naw) {
      const short2 cc = p.cpair[aw];
      x = p.tab[cc.x * p.cs + r * p.rs];
      y = p.tab[cc.y * p.cs + r * p.rs];
    }
    va[a][r] = x;
    vb[a][r] = y;
  }
  if (threadIdx.x < PAIR_AWB) {
    const int aw = aw0 + threadIdx.x;
    sw[threadIdx.x] = (aw < p.naw && p.tw) ? p.tw[aw] : 1.f;
  }
  __syncthreads();
  const int na = min(PAIR_AWB, p.naw - aw0);
  for (int sn = threadIdx.x; sn < p.nsn; sn += PAIR_THREADS) {
    float acc[PAIR_AWB];
#pragma unroll
    for (int a = 0; a < PAIR_AWB; a++) acc[a] = 0.f;
    const int m1 = p.off[sn + 1];
    for (int m = p.off[sn]; m < m1; m++) {
      const short2 rr = p.mem_rr[m];
      const float w = p.mem_w ? p.mem_w[m] : 1.f;
#pragma unroll
      for (int a = 0; a < PAIR_AWB; a++) acc[a] += w * (va[a][rr.x] * vb[a][rr.y]);
    }
#pragma unroll
    for (int a = 0; a < PAIR_AWB; a++)
      if (a < na) out[(int64_t)(aw0 + a) * p.nsn + sn] = sw[a] * acc[a];
  }
}

inline unsigned pair_blocks(int naw) { return (unsigned)((naw + PAIR_AWB - 1) / PAIR_AWB); }

}  // namespace
}  // namespace nipet
