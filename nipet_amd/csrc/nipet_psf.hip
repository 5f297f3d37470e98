// Separable resolution-modelling (PSF) convolution of device images for gfx950.
//
// The reference convolves padded single images with row / column kernels per
// z-slab and per y-row (d_conv, prj/src/recon.cu:49-215). Here one pass per
// axis covers the whole image, for one image or a replicate-innermost batch
// [voxel][nb]:
//  - z (contiguous axis): each workgroup stages whole z-lines (127 values + 8
//    zero planes on each side) in LDS and every thread sums its 17 taps;
//  - x and y (strided axes): one thread per line along the axis, consecutive
//    threads on consecutive inner elements (coalesced), a 24-value register
//    window slides along the line, 8 outputs per load of 8 inputs: one read
//    and one write per voxel and pass.
// Kernel taps follow the reference's order: out[i] = sum_{j=-R..R} k[R-j] in[i+j]
// with zero outside the image (prj/src/recon.cu:117-121, 171-176); rows of
// krnl[3][17] act on z, x and y in that order (psf_config, prj/mmrrec.py:127-130).
#include "nipet_common.h"

namespace nipet {

constexpr int PSF_TAPS = 2 * PSF_R + 1;

struct PsfTaps {
  float k[PSF_TAPS];
};

constexpr int PSF_SEG = 80;  // line piece per thread along x and y (multiple of 8)

// z-lines: line l = (column a, replicate b) at base a * NZ * nb + b, stride nb.
constexpr int PSF_ZLINES = 2;  // lines per 256-thread block (128 threads each)

__global__ __launch_bounds__(256) void k_psf_z(float *__restrict__ out,
                                               const float *__restrict__ in, PsfTaps t,
                                               int64_t nlines, int nb) {
  __shared__ float line[PSF_ZLINES][128 + 2 * PSF_R];
  const int sub = threadIdx.x >> 7, z = threadIdx.x & 127;
  const int64_t l = (int64_t)blockIdx.x * PSF_ZLINES + sub;
  const bool live = l < nlines;
  const int64_t a = live ? l / nb : 0, b = live ? l - a * nb : 0;
  const int64_t base = a * NZ * nb + b;
  if (z < PSF_R) {
    line[sub][z] = 0.f;
    line[sub][128 + PSF_R + z] = 0.f;
  }
  line[sub][PSF_R + z] = (live && z < NZ) ? in[base + (int64_t)z * nb] : 0.f;
  __syncthreads();
  if (!live || z >= NZ) return;
  float s = 0.f;
#pragma unroll
  for (int j = -PSF_R; j <= PSF_R; j++) s += t.k[PSF_R - j] * line[sub][PSF_R + z + j];
  out[base + (int64_t)z * nb] = s;
}

// Strided axis of length n: element (a, i, c) at a * n * inner + i * inner + c.
// blockIdx.y selects the piece [y seg, (y + 1) seg) of every line (seg a
// multiple of 8): more threads in flight for the same traffic plus 16 halo
// reads per piece.
__global__ __launch_bounds__(256) void k_psf_line(float *__restrict__ out,
                                                  const float *__restrict__ in, PsfTaps t,
                                                  int n, int64_t inner, int64_t nlines,
                                                  int seg) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlines) return;
  const int64_t a = l / inner, c = l - a * inner;
  const float *src = in + a * n * inner + c;
  float *dst = out + a * n * inner + c;
  const int ib = blockIdx.y * seg, ie = min(n, ib + seg);
  // w[q] = in[i0 - 8 + q] for the current block of 8 outputs starting at i0
  float w[24];
#pragma unroll
  for (int q = 0; q < 24; q++) {
    const int i = ib + q - 8;
    w[q] = (i >= 0 && i < n) ? src[(int64_t)i * inner] : 0.f;
  }
  for (int i0 = ib; i0 < ie; i0 += 8) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      float s = 0.f;
#pragma unroll
      for (int j = -PSF_R; j <= PSF_R; j++) s += t.k[PSF_R - j] * w[8 + k + j];
      if (i0 + k < ie) dst[(int64_t)(i0 + k) * inner] = s;
    }
#pragma unroll
    for (int q = 0; q < 16; q++) w[q] = w[q + 8];
#pragma unroll
    for (int q = 16; q < 24; q++) {
      const int i = i0 + q;  // = (i0 + 8) - 8 + q
      w[q] = i < n ? src[(int64_t)i * inner] : 0.f;
    }
  }
}

int psf_conv(float *out, const float *in, float *tmp, const float *krnl, int nb,
             hipStream_t st) {
  PsfTaps tz, tx, ty;
  for (int i = 0; i < PSF_TAPS; i++) {
    tz.k[i] = krnl[i];
    tx.k[i] = krnl[PSF_TAPS + i];
    ty.k[i] = krnl[2 * PSF_TAPS + i];
  }
  // z: in -> out; x (axis u, inner 127 nb): out -> tmp; y (axis v): tmp -> out
  const int64_t zl = (int64_t)NX * NY * nb;
  k_psf_z<<<(unsigned)((zl + PSF_ZLINES - 1) / PSF_ZLINES), 256, 0, st>>>(out, in, tz, zl, nb);
  const int64_t in_x = (int64_t)NZ * nb, lx = (int64_t)NY * in_x;
  const dim3 gx((unsigned)((lx + 255) / 256), (NX + PSF_SEG - 1) / PSF_SEG);
  k_psf_line<<<gx, 256, 0, st>>>(tmp, out, tx, NX, in_x, lx, PSF_SEG);
  const int64_t in_y = (int64_t)NX * NZ * nb;
  const dim3 gy((unsigned)((in_y + 255) / 256), (NY + PSF_SEG - 1) / PSF_SEG);
  k_psf_line<<<gy, 256, 0, st>>>(out, tmp, ty, NY, in_y, in_y, PSF_SEG);
  return NIPET_OK;
}

}  // namespace nipet

using namespace nipet;

extern "C" {

int nipet_psf_conv(float *out, const float *in, float *tmp, const float *krnl, int nb,
                   void *stream) {
  if (!out || !in || !tmp || !krnl || nb < 1 || nb > NIPET_MAXBATCH || out == in ||
      tmp == out) {
    set_error("nipet_psf_conv: bad arguments (out distinct from in and tmp, host krnl[3][17], "
              "1 <= nb <= %d)", NIPET_MAXBATCH);
    return NIPET_EINVAL;
  }
  int rc = psf_conv(out, in, tmp, krnl, nb, (hipStream_t)stream);
  if (rc) return rc;
  NIPET_HIP(hipGetLastError());
  return NIPET_OK;
}

}  // extern "C"
