// Shared definitions of the gfx950 NIPET library (device image / sinogram
// geometry, the geometry object, error reporting).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nipet_hip.h"

namespace nipet {

constexpr int NX = 320;   // device image x (u)
constexpr int NY = 320;   // device image y (v)
constexpr int NZ = 127;   // device image z (w), fastest
// segments of one transaxial ray: one per voxel column it crosses, at most
// one more than the grid lines it crosses (checked when the geometry is built)
constexpr int MAX_SEG = NX + NY + 1;
constexpr int64_t NVOX = (int64_t)NX * NY * NZ;
constexpr int MAXNSN = 4096;  // largest sinogram count handled (span-1: 4084)
constexpr int LI_MAX = 2112;  // largest Michelogram (ring pair) count
constexpr float EPS = 1e-6f;  // OSEM division guard (prj/src/recon.cu:12)
constexpr int PSF_R = 8;      // PSF kernel half-width (17 taps)

// One transaxial segment: start t (cm from the first crystal, double so that
// z-boundary positions along the ray keep sub-nanometre precision), length
// (cm) and the z-column base offset 127*(u + 320 v) of its voxel column.
struct __align__(16) SegRec {
  double s;
  float len;
  int32_t colz;
};

// Geometry resident on one device. Transaxial intersection lists in CSR form:
// segments of ray aw are seg[off[aw] .. off[aw+1]).
struct Geom {
  int device = 0;
  int naw = 0, nli = 0, nsn = 0, ncrs = 0, ndirect = 0;
  int64_t nseg = 0;
  float vox_z = 0.203125f;
  int32_t *off = nullptr;    // [naw+1]
  SegRec *seg = nullptr;     // [nseg]
  double *atn = nullptr;     // [naw] crystal-to-crystal transaxial distance
  float *lmax = nullptr;     // [naw] longest segment of the ray
  float2 *li_z = nullptr;    // [nli] (zeta_a at t=0, zeta_b - zeta_a) in z-voxel units
  short2 *li_sn = nullptr;   // [nli] output sinogram of LOR a and of its mirror b
  float *li_nosi = nullptr;  // [nli] 1 / (span-1 LORs per output sinogram)
  // ring-difference view of the Michelogram: element li joins ring r0 (LOR a
  // starts there) and ring r0 + d
  int dmax = 0;                // largest ring difference d
  int16_t *dr_li = nullptr;    // [64 * 64] li of (d, r0) at d * 64 + r0, -1 if absent
  int32_t *sn_off = nullptr;   // [nsn + 1] CSR of the span-1 LORs of each output sinogram:
  int16_t *sn_mem = nullptr;   // [2 nli] 2 li (LOR a) or 2 li + 1 (mirror b)
  // forward-projector lane slots: slot s walks the LORs (d, r0 + k), k < FP_G,
  // of one ring difference; fp_slot[s] = d * 64 + r0
  int16_t *fp_slot = nullptr;
  int fp_nslot = 0;
  // back-projector LOR tables [3][2][BP_TABN] for 1, 2 and 4 images: byte
  // offset into the back projector's LDS sinogram row (sinogram * nb * 4) of
  // LOR (r0, d) and of its mirror at bp_row(d) + d + 1 + r0; the zero row
  // (nsn * nb * 4) where there is none
  uint16_t *bp_tab = nullptr;
};

// Rows of the back projector's LOR tables: row d holds r0 = -d-1 .. 64
// (66 + d entries), so that the gather's r0 - h and r0 - e stay in the row.
constexpr int BP_MAXD = 63;  // largest ring difference of the BP tables
__host__ __device__ constexpr int bp_row(int d) { return 66 * d + d * (d - 1) / 2; }
constexpr int BP_TABN = bp_row(BP_MAXD + 1);

constexpr int FP_G = 3;          // LORs per forward-projector lane
constexpr int FP_G_SLOTS = 768;  // lanes of that kernel

void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);

// Separable 17-tap convolution of nb replicate-innermost images, kernels
// krnl[3][17] (host) for z, x, y: in -> out, tmp a scratch image (out must
// differ from in and tmp; tmp may be in, which is then overwritten). Launches
// on st only (no synchronisation, no allocation).
int psf_conv(float *out, const float *in, float *tmp, const float *krnl, int nb,
             hipStream_t st);

// Opt-in kernel timing (nipet_prof_*): HIP events recorded around launches on
// the launch stream, resolved when read.
enum ProfKernel {
  PK_FPRJ = 0, PK_BPRJ, PK_OSEM_UPDATE, PK_FILL, PK_HIST, PK_OTHER, PK_BPRJ_PAIR, PK_COUNT
};
void prof_begin(int kernel, hipStream_t st);
void prof_end(int kernel, hipStream_t st);

}  // namespace nipet

struct NipetGeom : nipet::Geom {};

#define NIPET_HIP(call)                                              \
  do {                                                               \
    hipError_t e_ = (call);                                          \
    if (e_ != hipSuccess) return nipet::hip_fail(e_, #call);         \
  } while (0)
