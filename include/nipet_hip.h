/*
 * nipet_hip.h -- C ABI of the MI355X (gfx950) implementation of NIPET's
 * projection / reconstruction / list-mode histogramming path.
 *
 * Conventions
 *  - Plain pointers and sizes only. Every array argument documented "device"
 *    is a device pointer on the library's current device; "host" arrays are
 *    read (or written) before the call returns.
 *  - Every entry point returns 0 on success and a negative NIPET_E* code on
 *    failure; nipet_last_error() returns a message for the calling thread.
 *    The library never exits the process (the reference's HandleError calls
 *    exit(): niftypet/nipet/src/scanner_0.cu:13-18).
 *  - `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *    Launches are asynchronous on that stream unless stated otherwise.
 *  - Layouts are the reference's device layouts:
 *      image   float32 [320 (y)][320 (x)][127 (z)], z fastest
 *              (prj/src/prjf.cu:53-55);
 *      sinogram float32 [nprj][nsn], sinogram index fastest, rows in the
 *              order of the transaxial-bin list `subs` (prj/src/prjf.cu:90-94).
 */
#ifndef NIPET_HIP_H
#define NIPET_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NIPET_OK 0
#define NIPET_EINVAL (-1)   /* bad argument / shape */
#define NIPET_EHIP (-2)     /* HIP runtime error */
#define NIPET_ENOMEM (-3)   /* device allocation failed */
#define NIPET_ESTATE (-4)   /* object in wrong state */

typedef struct NipetGeom NipetGeom;

/* Scanner geometry handed over once per session. All pointers are host.
 * Replaces the per-call LUT upload + transaxial tracing of the reference
 * (prj/src/prj_module.cu:380-398, prj/src/tprj.cu:13-200). */
typedef struct {
  const float *crs;      /* [ncrs][4] crystal face end points x0,y0,x1,y1 [cm] (txLUT['crs']) */
  int ncrs;              /* 504 */
  const int16_t *s2c;    /* [naw][2] crystal pair of each transaxial bin (txLUT['s2c']) */
  int naw;               /* 68516 */
  const float *li2rng;   /* [nli][2] axial z start [cm] of ring r0, r1 (axLUT['li2rng']) */
  const int16_t *li2sn;  /* [nli][2] output sinogram of LOR r0->r1 and of its mirror
                            (axLUT['li2sn'] for span-11, ['li2sn1'] for span-1) */
  const int8_t *li2nos;  /* [nli] span-1 LORs per output sinogram (axLUT['li2nos']) */
  int nli;               /* 2074; the first ndirect entries are direct planes */
  int nsn;               /* 837 (span-11) or 4084 (span-1) */
  float ring_pitch;      /* 0.40625 cm (SZ_RING) */
  float vox_xy;          /* 0.208626 cm (SZ_VOXY) */
  float vox_z;           /* 0.203125 cm (SZ_VOXZ) */
  float tfov2;           /* 890 cm^2 (TFOV2) */
} NipetGeomDesc;

/* Build the geometry on the current device: the FP64 transaxial intersection
 * lists of all naw bins (kept device-resident in CSR form) and the axial tables.
 * Synchronous. Reference: petprj setup of frwd_prj/back_prj/osem_rec
 * (prj/src/prj_module.cu:232-425, 430-608, 613-839) + sddn_tx
 * (prj/src/tprj.cu:13-172). */
int nipet_geom_create(const NipetGeomDesc *desc, NipetGeom **out);
void nipet_geom_destroy(NipetGeom *g);

/* Sizes of the built geometry (host outputs, any may be NULL). */
int nipet_geom_info(const NipetGeom *g, int *naw, int64_t *nseg_total, int *nli, int *nsn);

/* The bins subs[0..n) (device int32) written to out[0..n) (device, distinct
 * from subs) in order of decreasing segment count, bins outside 0 .. naw-1
 * last; bins of equal count in no particular order. A projector dispatches
 * workgroups in bin order, so the long rays start first. No counterpart in
 * the reference (its OSEM walks subsets in table order); a device pass, so
 * that preparing a reconstruction from device subsets needs no host copy. */
int nipet_bins_longest_first(const NipetGeom *g, const int32_t *subs, int n, int32_t *out,
                             void *stream);

/* Copy the transaxial intersection lists to host arrays: offsets[naw+1],
 * col[nseg_total] (voxel column u + 320 v), t[nseg_total] (segment start, cm),
 * len[nseg_total] (segment length, cm). The equivalent of the reference's
 * tv/tt tables (petprj.tprj, prj/src/prj_module.cu:70-224). col, t and len
 * may all be NULL: then only the offsets are copied. Synchronous. */
int nipet_tx_lists(const NipetGeom *g, int64_t *offsets, int32_t *col, float *t, float *len);

/* Forward projection (petprj.fprj, prj/src/prj_module.cu:232-425;
 * gpu_fprj prj/src/prjf.cu:209-320).
 *   sino [nprj][nsn] float (device, fully overwritten), img (device),
 *   subs [nprj] int32 transaxial bins (device), or NULL for all naw bins in
 *   order (the reference's isub = [-1]).
 *   att = 0: emission (bins sum their span-1 LORs);
 *   att = 1: attenuation factors exp(-line integral), bins average their LORs. */
int nipet_fprj(const NipetGeom *g, float *sino, const float *img, const int32_t *subs, int nprj,
               int att, void *stream);

/* Back projection (petprj.bprj, prj/src/prj_module.cu:430-608;
 * gpu_bprj prj/src/prjb.cu:190-308): img (device) is overwritten with the
 * adjoint of the emission projector applied to sino [nprj][nsn] (device). */
int nipet_bprj(const NipetGeom *g, float *img, const float *sino, const int32_t *subs, int nprj,
               void *stream);

/* Gap removal: Siemens sinogram in[nsn][nbinang] (nbinang = 252*344; float32
 * when in_dtype = 0, uint16 when 1) -> device sinogram out[naw][nsn] float32,
 * out[aw][s] = in[s][aw2ali[aw]]. All device pointers. Replaces
 * mmr_auxe.rgaps / remove_gaps (src/scanner_0.cu:94-155, mmraux.py:1146-1160). */
int nipet_remgaps(float *out, const void *in, int in_dtype, const int32_t *aw2ali, int naw,
                  int nsn, int nbinang, void *stream);

/* Gap removal of a replicate-innermost histogram: in[(s * nbinang + ali) * nrep
 * + r] (uint32 counts, the layout nipet_lm_chunk writes for nrep replicates)
 * -> out[(aw * nsn + s) * out_nb + out_b0 + b] = min(in[.. r0 + b], 65535) as
 * uint16, b < nb (out_b0 + nb <= out_nb): columns out_b0 .. out_b0+nb-1 of the
 * batched prompts of nipet_osem_subset_b. The reference removes gaps per frame
 * (mmraux.py:1146-1160, src/scanner_0.cu:94-155) and hands petprj.osem uint16
 * prompts (prj/src/prj_module.cu:649-651). */
int nipet_remgaps_u16b(uint16_t *out, const uint32_t *in, const int32_t *aw2ali, int naw,
                       int nsn, int nbinang, int nrep, int r0, int nb, int out_nb, int out_b0,
                       void *stream);

/* nipet_remgaps without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight per-frame loops, harnesses
 * that forbid hipMalloc). No counterpart in the reference. The inverse bin
 * map lives in a static device array of NIPET_NOALLOC_MAXBINANG int32
 * (347 KB, held per device until the process ends); nbinang beyond that is
 * rejected with NIPET_EINVAL. Every call on one device shares that array and
 * rewrites it whole: calls that overlap on one device (different streams or
 * threads) race. Same arguments and result as nipet_remgaps. */
#define NIPET_NOALLOC_MAXBINANG (252 * 344)
#define NIPET_MAXDEVICES 64
int nipet_remgaps_noalloc(float *out, const void *in, int in_dtype, const int32_t *aw2ali,
                          int naw, int nsn, int nbinang, void *stream);

/* Gap insertion: device sinogram in[naw][nsn] -> Siemens out[nsn][nbinang]
 * (gaps zero), out[s][aw2ali[aw]] = in[aw][s]. Replaces mmr_auxe.pgaps /
 * put_gaps (src/scanner_0.cu:158-241) and the transpose of mmraux.py:1141. */
int nipet_putgaps(float *out, const float *in, const int32_t *aw2ali, int naw, int nsn,
                  int nbinang, void *stream);

/* ---- batches: nb replicates (bootstrap replicates, frames) projected together,
 * sharing the geometry traversal. Layouts are replicate-innermost:
 * images [320][320][127][nb], sinograms [nprj][nsn][nb]. nb <= NIPET_MAXBATCH;
 * span 11 only (nsn <= 837). Same semantics as the nb = 1 entry points. */
#define NIPET_MAXBATCH 64
int nipet_fprj_b(const NipetGeom *g, float *sino, const float *img, const int32_t *subs,
                 int nprj, int nb, int att, void *stream);
int nipet_bprj_b(const NipetGeom *g, float *img, const float *sino, const int32_t *subs,
                 int nprj, int nb, void *stream);

/* Sensitivity image of one subset (the images osemone back-projects before the
 * loop, prj/mmrrec.py:291-298, from the a n sinogram of recon.cu:381-398):
 * sens = BP_sub(an[subs]) with an the full device sinogram [naw][nsn] read at
 * the subset's bins (no gathered copy). sens (device image) is overwritten. */
int nipet_osem_sens(const NipetGeom *g, float *sens, const float *an, const int32_t *subs,
                    int nprj, void *stream);

/* OSEM corrections, computed once per reconstruction (prj/src/recon.cu:381-398):
 *   rho[i] = (r[i] + s[i]) / (a[i] n[i])  (0 where |a n| < 1e-6),  an[i] = a[i] n[i].
 * r or s may be NULL (zero). All device arrays of n elements. */
int nipet_osem_corr(float *rho, float *an, const float *a, const float *n, const float *r,
                    const float *s, int64_t nelem, void *stream);

/* One OSEM subset update (the body of the subset loop of petprj.osem,
 * prj/src/recon.cu:459-497), all device pointers:
 *   e = FP_sub(img) + rho[subs]; e = y[subs] / e (0 where |e| < 1e-6);
 *   b = BP_sub(e); img = mask ? img * (b / sens) : 0 (b/sens = 0 where |sens| < 1e-6).
 * prompts: uint16 [naw][nsn]; rho: float [naw][nsn]; sens: float image of this
 * subset; mask: uint8 image; win: NULL (every segment of every ray) or the
 * mask's segment windows from nipet_osem_window, valid when img is zero
 * outside the mask (after any earlier update with that mask) - the projectors
 * then skip the segments whose columns miss the mask (a window reaching
 * outside its bin's list is clamped into it); work_sino: float
 * [nprj][nsn]; work_img: float image distinct from img and sens (aliased work
 * buffers are rejected). */
int nipet_osem_subset(const NipetGeom *g, float *img, const uint16_t *prompts, const float *rho,
                      const float *sens, const uint8_t *mask, const int32_t *win,
                      const int32_t *subs, int nprj, float *work_sino, float *work_img,
                      void *stream);

/* Segment windows of a reconstruction mask for the OSEM updates (no
 * counterpart in the reference, which walks every segment): win[2 aw] and
 * win[2 aw + 1] = first segment and segment count of bin aw's transaxial list
 * whose columns hold a mask voxel (columns dilated by a box of half-width
 * `dilate` voxels: 8, the 17-tap reach, with resolution modelling); 0, 0 when
 * none. mask: device uint8 image; win: device int32 [naw][2]. Two 100 KB
 * temporaries are allocated and freed stream-ordered on `stream`. */
int nipet_osem_window(const NipetGeom *g, const uint8_t *mask, int dilate, int32_t *win,
                      void *stream);

/* nipet_osem_subset for a subset whose sensitivity image is not yet known
 * (the first OSEM iteration): sens = BP_sub(an[subs]) (the subset's image of
 * prj/mmrrec.py:291-298, which the reference computes before the loop) and
 * b = BP_sub(e) are back-projected together, as one two-image batch through
 * one traversal of the subset's rays; sens is written out and the update is
 * that of nipet_osem_subset. an: float [naw][nsn]; sens: float image (output);
 * work_sino: float [nprj][nsn]; work_sino2: float [nprj][nsn][2]; work_img2:
 * float [voxel][2] (work buffers distinct from img and sens). Span 11 only
 * (nsn <= 837). */
int nipet_osem_subset_sens(const NipetGeom *g, float *img, const uint16_t *prompts,
                           const float *rho, const float *an, float *sens, const uint8_t *mask,
                           const int32_t *win, const int32_t *subs, int nprj, float *work_sino,
                           float *work_sino2, float *work_img2, void *stream);

/* Batched OSEM subset update of nb images [voxel][nb]: prompts [naw][nsn][nb];
 * rho shared [naw][nsn] (rho_per_rep = 0) or [naw][nsn][nb]; sens shared
 * [voxel] (sens_per_rep = 0) or [voxel][nb]; mask shared; work_sino
 * [nprj][nsn][nb]; work_img [voxel][nb]. */
int nipet_osem_subset_b(const NipetGeom *g, float *img, const uint16_t *prompts,
                        const float *rho, int rho_per_rep, const float *sens, int sens_per_rep,
                        const uint8_t *mask, const int32_t *win, const int32_t *subs, int nprj,
                        int nb, float *work_sino, float *work_img, void *stream);

/* One OSEM iteration over all subsets: the extension call petprj.osem
 * (osem_rec, prj/src/prj_module.cu:613-839; osem, prj/src/recon.cu:318-521)
 * with the reference's arguments:
 *   rho = (r + s) / (a n) (0 where |a n| < 1e-6), formed once per call
 *   (recon.cu:381-398); then for subsets i = 0 .. nsub-1 in order the update of
 *   nipet_osem_subset with the sensitivity image sens + i * 320*320*127 (with
 *   resolution modelling, krnl[0] >= 0, that of nipet_osem_subset_rm_b with
 *   the convolved sensitivity image, recon.cu:441-445, 468-495).
 * img: device float image (in/out); prompts: device uint16 [naw][nsn];
 * r, s (either may be NULL = zero), n, a: device float [naw][nsn];
 * subs: HOST int32 subset table [nsub][stride], row i = (count_i, count_i bins),
 *   1 <= count_i <= stride - 1 (the (Nsub, Nprj) table osemone builds,
 *   prj/mmrrec.py:286-294);
 * sens: device float [nsub][320][320][127] (back projections of a n, not
 *   convolved); mask: device uint8 image; krnl: HOST float [3][17] or NULL.
 * Subsets after the first walk only the mask's segment windows
 * (nipet_osem_window): img is then zero outside the mask.
 * Temporaries (rho, a n, one work sinogram and image, the windows, two more
 * images with resolution modelling) are allocated and freed stream-ordered on
 * `stream`. */
int nipet_osem_iter(const NipetGeom *g, float *img, const uint16_t *prompts, const float *r,
                    const float *s, const float *n, const float *a, const int32_t *subs,
                    int nsub, int stride, const float *sens, const uint8_t *mask,
                    const float *krnl, void *stream);

/* Resolution modelling: separable 17-tap convolution of nb replicate-innermost
 * device images [voxel][nb] with the host kernels krnl[3][17] (rows: z, x, y
 * taps as psf_config builds them, prj/mmrrec.py:96-155), zero outside the
 * image, out[i] = sum_j k[8-j] in[i+j]. Replaces d_conv (prj/src/recon.cu:
 * 49-215). out must differ from in and tmp; tmp may be in (then overwritten). */
int nipet_psf_conv(float *out, const float *in, float *tmp, const float *krnl, int nb,
                   void *stream);

/* nipet_osem_subset_b with resolution modelling (petprj.osem with krnl[0][0]
 * >= 0, prj/src/recon.cu:468-500): the subset update of H x, back projection
 * convolved by H before the division by the (convolved) sensitivity. Two work
 * images [voxel][nb]. */
int nipet_osem_subset_rm_b(const NipetGeom *g, float *img, const uint16_t *prompts,
                           const float *rho, int rho_per_rep, const float *sens,
                           int sens_per_rep, const uint8_t *mask, const int32_t *win,
                           const int32_t *subs, int nprj, int nb, const float *krnl,
                           float *work_sino, float *work_img, float *work_img2, void *stream);

/* ---- list-mode histogramming (mmr_lmproc.hist: lm/src/lm_module.cu:174-391,
 * lmproc lm/src/lmproc.cu:12-230, gpu_hst / hst lm/src/hst.cu:37-528) ----
 * A session histograms one frame [tstart, tstop) (seconds from the first time
 * tag toff) of a word stream handed over in consecutive chunks; the time-tag
 * state and the word index (bootstrap counter) carry over between chunks.
 * Each event belongs to the most recent preceding time tag. */
typedef struct NipetLm NipetLm;

typedef struct {
  int toff;              /* first time tag of the acquisition [ms] (lminfo) */
  int nitag;             /* seconds covered: ((last tag - toff) + 999) / 1000 */
  int tstart, tstop;     /* frame [tstart, tstop) in seconds, 0 <= tstart <= tstop <= nitag */
  int span;              /* 11 (837 sinograms), 1 (4084) or 0 (SSRB, 127) */
  int btp;               /* 0: no bootstrap; 1: non-parametric (each event redrawn
                            uniformly from the events of its patch, the 8192-word
                            block of the session's stream that holds it; every chunk
                            but the last must then hold a multiple of 8192 words);
                            2: parametric (Poisson weights), lm/src/hst.cu:110-139 */
  double btprt;          /* Poisson mean of the bootstrap weights (Cnt['BTPRT']) */
  uint64_t seed;         /* Philox4x32-10 key; btp 2 draws word g of replicate r from
                            output word g % 4 of counter (g / 4, rep0 + r, 0) */
  int nrep;              /* replicates 1..64; nrep > 1 stores [bin][nrep] */
  int rep0;              /* replicate index of the first stored replicate (Philox counter
                            word 2 = rep0 + r), so sessions can split a replicate set */
  const int16_t *sn1_sn11;  /* host [4084] (axLUT['sn1_sn11']) */
  const int16_t *sn1_ssrb;  /* host [4084] (axLUT['sn1_ssrb']) */
  const int16_t *sn1_rno;   /* host [4084][2] (axLUT['sn1_rno']) */
  const int16_t *s2cF;      /* host [86688][2] (txLUT['s2cF']) */
} NipetLmConfig;

/* Device outputs, zeroed by the caller, accumulated by every chunk. Only
 * psino, dsino, phc and dhc hold every replicate; ssr, zR, zM, pvs and fan
 * are those of the first stored replicate (rep0): with nrep > 1, randoms from
 * `fan` are replicate rep0's estimate for every replicate. */
typedef struct {
  uint32_t *psino, *dsino;  /* prompts / delayeds [nbins][nrep] (nbins = nsinos*86688) */
  uint32_t *ssr;            /* SSRB prompts [127][252][344] (replicate rep0), by nipet_lm_finish */
  uint32_t *phc, *dhc;      /* head curves [nitag][nrep] */
  uint64_t *zR, *zM;        /* centre-of-mass sums [nitag] (replicate rep0) */
  uint32_t *pvs;            /* projection views [(min(nitag,5400)+3)/4][127][344] (replicate rep0) */
  uint32_t *fan;            /* delayed fan-sums [64][504] (replicate rep0) */
  uint32_t *bck;            /* singles buckets [2][nitag][224] (not resampled) */
} NipetLmOut;

int nipet_lm_create(const NipetLmConfig *cfg, NipetLm **out);
/* Histogram words[0..n) (device int32) on `stream`. A chunk that fails after
 * its first kernels were queued (a device allocation for the staging slots)
 * leaves the session unusable: later chunk / finish calls return an error. */
int nipet_lm_chunk(NipetLm *h, const int32_t *words, int64_t n, const NipetLmOut *out,
                   void *stream);
/* After the last chunk: the SSRB prompt sinogram out->ssr (overwritten) from
 * the replicate-0 prompt sinogram (each output sinogram lies on one SSRB plane). */
int nipet_lm_finish(NipetLm *h, const NipetLmOut *out, void *stream);
/* Words processed, words with an address beyond the span-1 range (skipped),
 * singles reports lost to list overflow. Synchronous. */
int nipet_lm_status(NipetLm *h, int64_t *words_done, unsigned int *bad_words,
                    unsigned int *singles_overflow);
void nipet_lm_destroy(NipetLm *h);

/* Time-tag index of a device word stream, what getLMinfo finds by scanning the
 * list-mode file (lm/src/lmaux.cu:17-200: first / last time tag, word positions
 * that split the stream by time), for words already in device memory.
 * span (device int64[2]): index of the first time-tag word (n if none) and
 * one past the index of the last one (0 if none). */
int nipet_lm_tag_span(const int32_t *words, int64_t n, int64_t *span, void *stream);
/* first (device int64[nsec + 1]): first[s] = index of the first time-tag word
 * whose time is at least toff + 1000 s ms (n if none), s = 0 .. nsec; the
 * words of seconds [t0, t1) are words[first[t0] .. first[t1]). Tags past
 * toff + 1000 nsec count for nsec, tags before toff for no second. */
int nipet_lm_second_index(const int32_t *words, int64_t n, int toff, int nsec, int64_t *first,
                          void *stream);

/* ---- randoms from delayed fan-sums (mmr_lmproc.rand: lm/src/lm_module.cu:404-545,
 * gpu_randoms lm/src/rnd.cu:180-420; Python rand/randoms lm/mmrhist.py:188-221) ----
 * Crystal singles s[c][r] (448 active crystals x 64 rings) solve
 * fansum = s * (F s), F summing over the crystals in coincidence (cij) and the
 * rings within MRD, by 10 (niter) damped fixed-point steps from sqrt(f / F 1);
 * the randoms sinogram of LOR (c0, r0; c1, r1) is s[c0][r0] s[c1][r1], span-11
 * sinograms sum their span-1 members. */
typedef struct NipetRnd NipetRnd;

typedef struct {
  int naw;                  /* active bins (txLUT['Naw'], 68516) */
  int nsn;                  /* output sinograms: 4084 (span 1) or 837 (span 11) */
  int nsn1;                 /* 4084 */
  int mrd;                  /* maximum ring difference (Cnt['MRD'], 60) */
  const int8_t *cij;        /* host [448][448] (txLUT['cij']) */
  const int16_t *crsri;     /* host [504] (txLUT['crsri']: reduced crystal, -1 in gaps) */
  const int16_t *s2cr;      /* host [naw][2] (txLUT['s2cr']) */
  const int16_t *sn1_rno;   /* host [4084][2] (axLUT['sn1_rno']) */
  const int16_t *sn1_sn11;  /* host [4084] (axLUT['sn1_sn11']), required for span 11 */
} NipetRndDesc;

int nipet_rnd_create(const NipetRndDesc *desc, NipetRnd **out);
/* fansums: device [64][504] delayed fan-sums (NipetLmOut.fan). Outputs
 * (device, either may be NULL): rsino [naw][nsn] randoms sinogram in the
 * device layout (Siemens layout through nipet_putgaps), cmap [504][64]
 * singles map with zero gap crystals. Asynchronous on `stream`. */
int nipet_rnd_estimate(NipetRnd *h, const uint32_t *fansums, int niter, float *rsino,
                       float *cmap, void *stream);
void nipet_rnd_destroy(NipetRnd *h);

/* ---- normalisation sinogram from components (mmr_auxe.norm: src/aux_module.cu:81-220,
 * norm_from_components / dev_norm src/norm.cu:13-220; Python get_norm_sino
 * mmrnorm.py:155-195) ----
 * Factor of span-1 LOR (bin w + 344 a, crystals c1, c2 with gaps, rings r0, r1):
 * geo[w] cinf[w][a % 9] E(c1, r0) E(c2, r1), E(c, r) = ceff[r][c] exp(q dtp[r] /
 * (1 + q dtnp[r])) / (1 + q dtnp[r]), q = buckets[c / 18 + 28 (r / 8)] / 2;
 * span 1: times axf1[si] / axe1[sn1_sn11[si]]; span 11: mean over its span-1
 * LORs divided by axe1[s11]. */
typedef struct NipetNorm NipetNorm;

typedef struct {
  int naw;                    /* active bins (68516) */
  int span;                   /* 1 (4084 sinograms) or 11 (837) */
  const int32_t *aw2ali;      /* host [naw] (txLUT['aw2ali'] = w + 344 a) */
  const int16_t *sn1_rno;     /* host [4084][2] (axLUT['sn1_rno']) */
  const int16_t *sn1_sn11;    /* host [4084] (axLUT['sn1_sn11']) */
  const int8_t *sn1_sn11no;   /* host [4084] (axLUT['sn1_sn11no']) */
} NipetNormDesc;

/* Host normalisation components (mmrnorm.get_components, mmrnorm.py:19-112). */
typedef struct {
  const float *geo;   /* [344] geometric effects (first row of the (127, 344) array) */
  const float *cinf;  /* [344][9] crystal interference */
  const float *ceff;  /* [64][504] crystal efficiencies */
  const float *axe1;  /* [837] axial effects */
  const float *axf1;  /* [4084] span-1 axial factors (span 1 only; may be NULL for span 11) */
  const float *dtp;   /* [64] paralysing ring dead time */
  const float *dtnp;  /* [64] non-paralysing ring dead time */
} NipetNormComp;

int nipet_norm_create(const NipetNormDesc *desc, NipetNorm **out);
/* out: device [naw][nsn] (the device layout). buckets: host [224] singles
 * rates (mmrhist()['buckets']). The host arrays are consumed before return. */
int nipet_norm_sino(NipetNorm *h, const NipetNormComp *comp, const int32_t *buckets, float *out,
                    void *stream);
void nipet_norm_destroy(NipetNorm *h);

/* Online mean / variance update with realisation b (0-based) over n device
 * floats (mmr_auxe.varon: src/aux_module.cu:491-556, var src/auxmath.cu:15-22):
 * m1 += (x - m1) / (b + 1); m2 += (x - m1_old) (x - m1_new). */
int nipet_varon(float *m1, float *m2, const float *x, int b, int64_t n, void *stream);

/* Philox4x32-10 on the device for one counter/key (host arrays), for
 * known-answer checks of the bootstrap generator. Synchronous. */
int nipet_philox4x32(const uint32_t *ctr, const uint32_t *key, uint32_t *out);

/* Kernel timing, the counterpart of the reference's cudaEvent timers printed
 * at LOG <= LOGDEBUG (prj/src/prjf.cu:274-317, prj/src/prjb.cu:246-307,
 * lm/src/hst.cu:452-501). When enabled, every launch is bracketed by HIP
 * events on its stream; nipet_prof_read resolves them (synchronising) and
 * returns the accumulated milliseconds and launch count of one kernel class. */
#define NIPET_PROF_FPRJ 0
#define NIPET_PROF_BPRJ 1
#define NIPET_PROF_OSEM_UPDATE 2
#define NIPET_PROF_FILL 3
#define NIPET_PROF_HIST 4
#define NIPET_PROF_OTHER 5      /* small helper kernels */
#define NIPET_PROF_BPRJ_PAIR 6  /* two-image BP of nipet_osem_subset_sens */
int nipet_prof_enable(int on);
/* Largest replicate sub-batch (4, 2 or 1) that one FP (kernel =
 * NIPET_PROF_FPRJ) or BP (NIPET_PROF_BPRJ) launch covers in the batched entry
 * points; batches are split greedily into sub-batches of at most that width.
 * Process-wide; set by the autotuner (no counterpart in the reference, whose
 * projectors take one image). */
int nipet_tune_set(int kernel, int width);
int nipet_tune_get(int kernel, int *width);
/* Segments per wave run of the back projector at sub-batch width `width` (1, 2
 * or 4): one of the two run lengths compiled for that width (1: 8 or 6;
 * 2: 4 or 6; 4: 6 or 4). Process-wide; set by the autotuner. nipet_tune_get_run
 * also returns the other compiled length in *alt (may be NULL). No counterpart
 * in the reference (its ray blocks are fixed at compile time). */
int nipet_tune_set_run(int width, int spw);
int nipet_tune_get_run(int width, int *spw, int *alt);
int nipet_prof_reset(void);
int nipet_prof_read(int kernel, double *total_ms, int64_t *count);

const char *nipet_last_error(void);

/* Library version string. */
const char *nipet_version(void);

#ifdef __cplusplus
}
#endif
#endif /* NIPET_HIP_H */
