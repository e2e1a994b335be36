/*
 * ebsd_radon.h -- C ABI of the MI355X (gfx950) band-detection library.
 *
 * The library replaces the compute inside PyEBSDIndex's BandDetect.find_bands
 * (pyebsdindex/band_detect.py:411-526): the fast Radon transform
 * (pyebsdindex/radon_fast.py:242-310), the Radon-space padding and convolution
 * (band_detect.py:562-637), the local-max search (band_detect.py:639-664) and the
 * peak labelling with sub-pixel fit (band_detect.py:667-809), plus the normmax and
 * theta/rho conversion of band_detect.py:457-460,521-524.
 *
 * Conventions: every pointer named d_* is device memory; every other pointer is host
 * memory.  Functions return 0 on success and a negative code on failure; the message
 * is available from ebsd_last_error() (per thread).  A plan is immutable after set-up
 * except through the ebsd_plan_set_* calls, which fail while a batch (ebsd_batch_create)
 * holds the plan; calls that use one plan from several host threads at once must be
 * serialised by the caller.
 */
#ifndef EBSD_RADON_H
#define EBSD_RADON_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ebsd_plan ebsd_plan;

/* Pattern element types accepted by ebsd_find_bands (numpy uint8 / uint16 / float32).
 * OR-ing EBSD_FLAG_NOBG into the dtype argument ignores the plan's background for that
 * call (used when the caller has already subtracted it, e.g. for float64 input). */
enum { EBSD_U8 = 0, EBSD_U16 = 1, EBSD_F32 = 2, EBSD_FLAG_NOBG = 0x100 };

/* Output fields of one find_bands call, all device pointers, row-major (N, nBands[,2]).
 * Mirrors the numeric fields of BandDetect.dataType (band_detect.py:110-115). */
typedef struct ebsd_bands {
    float  *d_max;      /* (N, nB)     peak value of the convolved Radon            */
    float  *d_normmax;  /* (N, nB)     (max - min) / clip(mean - min, 1e-7)          */
    float  *d_avemax;   /* (N, nB)     mean of the 3x3 neighbourhood                 */
    float  *d_maxloc;   /* (N, nB, 2)  integer (rho, theta) bin, padding removed     */
    float  *d_aveloc;   /* (N, nB, 2)  sub-pixel (rho, theta) bin                    */
    float  *d_width;    /* (N, nB)     FWHM-style width in rho bins                  */
    float  *d_theta;    /* (N, nB)     radians, detector convention                  */
    float  *d_rho;      /* (N, nB)     pixels, detector convention                   */
    int8_t *d_valid;    /* (N, nB)                                                   */
    float  *d_stats;    /* (N, 3)      min, mean, max of the convolved Radon region  */
} ebsd_bands;

/* Build a device plan from the reference's own set-up products.
 *   index_plan: Radon.indexPlan, int64 (nRho, nTheta, plan_len), each line sorted,
 *               entries >= H*W ending the line      (radon_fast.py:104-165)
 *   rdnmask:    BandDetect.rdnmask as uint8 (nRho, nTheta)        (band_detect.py:203-251)
 *   kernel:     BandDetect.kernel[0], float32 (kr, kt), odd sizes  (band_detect.py:277-285)
 *   pad_r/t:    BandDetect.padding; peak_r/t: BandDetect.peakPad   (band_detect.py:287-290)
 *   theta_deg:  Radon.theta (nTheta,) float32; rho: Radon.rho (nRho,) float32
 * Replaces: the plan upload of the OpenCL backend (opencl/band_detect_cl.py:284-291),
 * which re-sends indexPlan every chunk; here it is compiled once per pattern size. */
int ebsd_plan_create(int H, int W, int nTheta, int nRho, const int64_t *index_plan, int plan_len,
                     const uint8_t *rdnmask, const float *kernel, int kr, int kt, int pad_r,
                     int pad_t, int peak_r, int peak_t, int nBands, const float *theta_deg,
                     const float *rho, ebsd_plan **out);

int ebsd_plan_destroy(ebsd_plan *plan);

/* Background subtracted before the Radon (BandDetect.backgroundsub, radon_fast.py:255-259).
 * bg: host float64 (H*W) or NULL to clear.  exact=1 stages float32(x - bg) per pixel and
 * accumulates in float64 exactly as the reference; exact=0 subtracts the per-line float64
 * sum of bg from the exact integer line sum (a float64-exact form of the same quantity). */
int ebsd_plan_set_background(ebsd_plan *plan, const double *bg, int exact);

/* Kernel launch shape.  Any argument <= 0 keeps the current value.
 *   gather_pack:    patterns interleaved per Radon work-group (u8: 2 x gather_pack up to 8,
 *                   u16: 1/2/4/8, f32: gather_pack / 2 up to 4)
 *   gather_threads: threads per Radon work-group (multiple of 64, <= 1024)
 *   peak_tile:      theta columns per tile of the convolution/peak kernel
 *   peak_threads:   threads per convolution/peak work-group (multiple of 64, <= 512) */
int ebsd_plan_set_tuning(ebsd_plan *plan, int gather_pack, int gather_threads, int peak_tile,
                         int peak_threads);

/* LDS words per staged pattern row in the Radon gather (>= W; <= 0 restores W).  Padding the
 * rows changes which LDS banks the pixels of a line fall in; results are identical. */
int ebsd_plan_set_stage_pitch(ebsd_plan *plan, int pitch);

/* Row segments of the integer (uint8/uint16, fast background) Radon gather: the pattern's
 * rows are staged in `segments` consecutive ranges, one launch each, whose exact integer
 * partial line sums add up in the output -- so that more patterns share an LDS word when a
 * whole pattern would leave room for one or two.  0 = automatic (the fewest of 1, 2, 4, 8 at
 * which the configured pack size fits), 1..8 = that many.  Results are identical. */
int ebsd_plan_set_row_segments(ebsd_plan *plan, int segments);

/* Line order of the Radon gather: lines in units of `unit` neighbouring Radon cells (a power
 * of two <= 64; 0 = the default 8, one 32-byte write sector of float32 values), the units
 * sorted by their longest line.  Larger units make a batch's output stores cover whole
 * sectors (less HBM write traffic) at the price of padding shorter lines to the unit's
 * longest; 1 sorts single lines.  Results are identical.  No counterpart in the reference
 * (a launch-shape knob of this implementation). */
int ebsd_plan_set_line_unit(ebsd_plan *plan, int unit);

/* Which convolution/peak kernel runs the stages of band_detect.py:562-809 (same results):
 *   0 automatic (the streamed kernel where it applies), 1 theta-tiled work-group kernel,
 *   2 streamed kernel (one wavefront per block of theta columns; needs a 9x5 or 5x5 kernel
 *     with peakPad = 2*ksz+1 and at most 16 wavefronts a pattern). */
int ebsd_plan_set_peak_kernel(ebsd_plan *plan, int kind);

/* Theta columns per wavefront of the streamed peak kernel (its theta block; the pattern's
 * wavefronts = ceil(nTheta / theta_cols), at most 16), a launch-shape knob for the
 * autotuner with the same results for every value.  0 = the default (the fewest
 * wavefronts whose blocks fit 64 lanes less the halo).  Errors where the streamed kernel
 * does not apply or the value is out of range.  No counterpart in the reference (whose
 * OpenCL kernels take a fixed work-group shape). */
int ebsd_plan_set_peak_block(ebsd_plan *plan, int theta_cols);

/* Workspace (device bytes) that ebsd_find_bands needs to process `chunk` patterns per pass. */
size_t ebsd_workspace_bytes(const ebsd_plan *plan, int64_t chunk);

/* BandDetect.find_bands on N patterns resident on the device (band_detect.py:411-526).
 * d_pats: (N, H, W) of `dtype`; d_work: workspace of ws_bytes (>= ebsd_workspace_bytes(plan,1));
 * the patterns run in passes of equal size whose Radons fit it (a larger workspace, fewer
 * and fuller passes).  stream: a hipStream_t (NULL = default).
 * Replaces: band_detect_cl.BandDetect.find_bands (opencl/band_detect_cl.py:51-199). */
int ebsd_find_bands(ebsd_plan *plan, const void *d_pats, int dtype, int64_t N,
                    const ebsd_bands *out, void *d_work, size_t ws_bytes, void *stream);

/* ebsd_find_bands without a workspace and without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Each chunk's Radon (and,
 * where the gather runs in row segments, their partial sums) goes to a file-scope device
 * array of 64 MiB that every GPU the library runs on holds until the process ends; patterns
 * are processed in chunks that fit it (1,035 per chunk at 120x120 with a 180x90 Radon).  A
 * pattern whose Radon alone exceeds 64 MiB is refused with an error.  The call uses only
 * structures built by the plan's set-up calls: row segments already compiled for the
 * pattern type (the uint16 ones are compiled by ebsd_plan_create), otherwise one
 * whole-pattern pass with a smaller pack.  Every call on one GPU shares that GPU's array
 * and leaves its contents behind (each call overwrites what it reads, nothing needs
 * clearing): calls that overlap on one GPU -- on different streams or host threads -- race.
 * The plan's device must be current.  Results equal ebsd_find_bands's bit for bit. */
int ebsd_find_bands_noalloc(ebsd_plan *plan, const void *d_pats, int dtype, int64_t N,
                            const ebsd_bands *out, void *stream);

/* Stage-level entry points (the reference's per-stage methods, for tests and adjoints). */
/* Radon of N patterns into d_radon (N, nRho, nTheta) float32, -1 where a line is empty.
 * Replaces: Radon.radon_faster (radon_fast.py:242-279) without padding. */
int ebsd_radon(ebsd_plan *plan, const void *d_pats, int dtype, int64_t N, float *d_radon,
               void *stream);
/* Padding + convolution + local max + labelling of N unpadded Radons (N, nRho, nTheta).
 * Replaces: rdn_conv + rdn_local_max + band_label (band_detect.py:562-809). */
int ebsd_peaks(ebsd_plan *plan, const float *d_radon, int64_t N, const ebsd_bands *out,
               void *stream);
/* Padded, convolved Radon (N, nRho+2*pad_r, nTheta+2*pad_t) float32 for parity checks.
 * Replaces: the rdnConv array of rdn_conv (band_detect.py:562-637). */
int ebsd_conv(ebsd_plan *plan, const float *d_radon, int64_t N, float *d_conv, void *stream);

/* Batches of mixed pattern sizes (each set with its own plan): persistent Radon launches
 * (one per LDS-occupancy class) and peak launches (one per Radon geometry) over
 * heterogeneous work lists; a set of 1,024 patterns or more runs its own per-size gather
 * and peak launches, one after the other, in the same sequence (run once by
 * ebsd_batch_create, so that later launches allocate nothing).  The batch binds the given
 * device pointers; ebsd_batch_graph_launch captures the launch sequence into a hipGraph on
 * its first call (on a non-default stream) and replays it afterwards.  The batch's launches
 * and graph hold device pointers into its plans, so while the batch exists every
 * ebsd_plan_set_* call on those plans fails; ebsd_find_bands on them stays allowed (buffers
 * it outgrows are kept until the plan is destroyed).
 * Replaces: calling find_bands once per pattern size (one BandDetect/plan per size, as
 * EBSDIndexer re-plans on a size change, pyebsdindex/_ebsd_index_single.py:696-700). */
typedef struct ebsd_batch ebsd_batch;
size_t ebsd_batch_workspace_bytes(int nsets, ebsd_plan *const *plans, const int64_t *N);
int ebsd_batch_create(int nsets, ebsd_plan *const *plans, const void *const *d_pats, const int *dtypes,
                      const int64_t *N, const ebsd_bands *outs, void *d_work, size_t ws_bytes,
                      ebsd_batch **out);
int ebsd_batch_run(ebsd_batch *batch, void *stream);
int ebsd_batch_graph_launch(ebsd_batch *batch, void *stream);
int ebsd_batch_destroy(ebsd_batch *batch);

/* Plan geometry queries: out10 = cells, line-pixel pairs per pattern, gather batches, tiled
 * peak kernel's theta tile, its tiles, its LDS bytes, background mode, compute units, entry
 * slots per pattern of the whole-pattern gather (its lines padded to their batch's group
 * count), line unit. */
int ebsd_plan_info(const ebsd_plan *plan, int64_t *out10);

/* ---- adjoints: gradients of the band positions --------------------------------------
 * The reference has no derivatives; these are the adjoints of its forward stages, so that
 * a loss on the band (rho, theta) can be differentiated w.r.t. the patterns and the
 * pattern centre.  All outputs are written in full (no zeroing by the caller) and are
 * deterministic (float64 gathers in a fixed order). */

/* dL/dC of the convolved padded Radon (N, nRp, nTp) from dL/drho, dL/dtheta (N, nB) of
 * the bands ebsd_find_bands returned (maxloc, aveloc, valid): the Jacobian of the 3x3
 * sub-pixel fit of band_detect.py:764-798 (with the branch the forward took), the slope of
 * the rho/theta interpolation of band_detect.py:521-524.  d_radon: the (recomputed)
 * unpadded Radon (N, nR, nT). */
int ebsd_band_grad(ebsd_plan *plan, const float *d_radon, int64_t N, const float *d_maxloc,
                   const float *d_aveloc, const int8_t *d_valid, const float *d_grad_rho,
                   const float *d_grad_theta, float *d_grad_conv, void *stream);

/* Adjoint of padding + 'same' convolution (band_detect.py:573-580, 605-606):
 * dR (N, nR, nT) from dC (N, nRp, nTp). */
int ebsd_conv_transpose(ebsd_plan *plan, const float *d_grad_conv, int64_t N, float *d_grad_radon,
                        void *stream);

/* Adjoint of the normalised Radon (radon_fast.py:288-310), the backprojection:
 * d(pixels) (N, H, W) float32 from dR (N, nR, nT). */
int ebsd_backproject(ebsd_plan *plan, const float *d_grad_radon, int64_t N, float *d_grad_pats,
                     void *stream);

/* Predicted bands: the inverse of radon2pole (radon_fast.py:314-397).  For detector-frame
 * plane normals (N, M, 3) and pattern centres (3, or N x 3 when pc_per_pattern), the band
 * (rho, theta) (N, M) and d rho / d PC (N, M, 3); theta does not depend on PC.
 * vendor: 0 EDAX, 1 OXFORD, 2 BRUKER/KIKUCHIPY (radon_fast.py:361-370). */
int ebsd_gnomonic_bands(int64_t N, int M, const double *d_normals, const double *d_pc,
                        int pc_per_pattern, int H, int W, int vendor, double *d_rho,
                        double *d_theta, double *d_drho_dpc, void *stream);

/* Radon.radon2pole (radon_fast.py:314-397) on the device: band (theta, rho) (N, nB) ->
 * unit plane normals (N, nB, 3) float32 (float64 arithmetic); d_dn_dpc (N, nB, 3, 3) or
 * NULL: the Jacobian of each normal w.r.t. the pattern centre. */
int ebsd_radon2pole(int64_t N, int nB, const float *d_theta, const float *d_rho,
                    const double *d_pc, int pc_per_pattern, int H, int W, int vendor,
                    float *d_normals, double *d_dn_dpc, void *stream);

/* The band records of a find_bands result in the byte layout of BandDetect.dataType
 * (band_detect.py:110-115, numpy's packed layout): N*nB records of 49 + 16*nPhases bytes
 * (id, max, normmax, maxloc[2], avemax, aveloc[2], pqmax, width, theta, rho, valid,
 * band_match_index[nPhases], bandfit[nPhases]) into the device buffer d_records.  The
 * fields find_bands does not compute are zero, as in the reference's result (its records
 * come from a zeroed chunk array, band_detect.py:425, :466).  Replaces the host-side
 * assembly of the structured array (band_detect.py:466), so that a host caller receives the
 * array with one copy. */
int ebsd_pack_band_records(const ebsd_bands *bands, int64_t N, int nB, int nPhases, void *d_records,
                           void *stream);

const char *ebsd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
