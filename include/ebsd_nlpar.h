/*
 * ebsd_nlpar.h -- C ABI of the MI355X (gfx950) NLPAR library (non-local pattern averaging).
 *
 * The library replaces the two Numba kernels of PyEBSDIndex's NLPAR CPU path:
 * NLPAR.sigma_numba (pyebsdindex/nlpar_cpu.py:753-819), the per-pattern noise estimate
 * from the nearest scan neighbours, and NLPAR.nlpar_nb (nlpar_cpu.py:821-935), the
 * weighted average of each pattern's search window.  The host driver (file chunks,
 * lambda optimisation, output file) stays in Python: pyebsdindex_amd/nlpar.py.
 *
 * Layout: a chunk of the scan is nr x nc scan points in row-major scan order; point
 * (j, i) is pattern j*nc + i; a pattern is P = H*W pixels, row-major.  Pixel values are
 * uint8, uint16 or float32 (EBSD_NL_U8/U16/F32) and are read as float32, as the reference
 * reads its chunks with convertToFloat=True.  `idx` lists the M mask pixels (ascending
 * flat indices, the non-zero entries of the mask) that the distances run over.
 *
 * Conventions: every pointer named d_* is device memory; functions return 0 on success
 * and a negative code on failure, with the message from ebsd_nlpar_last_error() (per
 * thread).  Work is enqueued on `stream` (a hipStream_t; NULL = the default stream).
 */
#ifndef EBSD_NLPAR_H
#define EBSD_NLPAR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { EBSD_NL_U8 = 0, EBSD_NL_U16 = 1, EBSD_NL_F32 = 2 };

/* Largest pixel value of n patterns of P pixels, as float32, into *d_max (device).
 * Replaces the np.max(data) of each chunk (nlpar_cpu.py:764, :869). */
int ebsd_nlpar_max(const void *d_pats, int dtype, int64_t n, int64_t P, float *d_max, void *stream);

/* Bytes of the optional workspace of ebsd_nlpar_sigma / ebsd_nlpar for a chunk of n
 * patterns of P pixels.  With uint8 or uint16 patterns (P a multiple of 4) and this
 * workspace the kernels compute the distances as exact integer sums over packed words
 * (uint8: byte dot products), taking out the pixels saturated in either pattern from a
 * per-pattern list; patterns with more than 64 saturated pixels, float32 patterns, or no
 * workspace (NULL): pixel by pixel.  No counterpart in the reference (its Numba kernels
 * allocate nothing on a device). */
size_t ebsd_nlpar_workspace_bytes(int64_t n, int64_t P);

/* Noise estimate (nlpar_cpu.py:753-819) for every point of an nr x nc chunk, neighbours
 * within +-nn rows/columns clipped to the chunk.  A pixel pair counts when both values
 * are below `sat` (the reference's mxval, compared in float32: pass the smallest float32
 * not below it).  Writes
 *   d_sigma (nr*nc):          sqrt(min over neighbours with d2 >= 1e-3 of d2 / (2 n2)),
 *                             1e12 where there is none;
 *   d_dout, d_nout (nr*nc*(2nn+1)^2), optional (both NULL to skip): per window entry in
 *                             row-major window order, the squared distance normalised as
 *                             nlpar_cpu.py:810-818 with the sigmas of both points, and
 *                             the pixel count; the centre entry has d2 = 0 and count P;
 *                             entries beyond a clipped window are 0. */
int ebsd_nlpar_sigma(const void *d_pats, int dtype, int nr, int nc, int64_t P, const int32_t *d_idx,
                     int M, int nn, float sat, float *d_sigma, float *d_dout, float *d_nout,
                     void *d_work, size_t work_bytes, void *stream);

/* The averaged patterns (nlpar_cpu.py:821-935) of the points in rows [r0, r0+nrc) and
 * columns [c0, c0+ncc) of an nr x nc chunk (the reference's calclim), over a
 * (2sr+1)^2 window that is shifted, not clipped, at the chunk edges (needs nr, nc >=
 * 2sr+1).  Distances as for the sigma (pixel pairs below `sat`), normalised with d_sigma;
 * weights exp(-max(d - dthresh, 0) / lam^2) with the centre's d = -1e6, plus diff_offset
 * on the centre, divided by their sum.  d_out receives (nrc*ncc, P) values of out_dtype:
 * float32, or uint8/uint16 clipped to the type's range and truncated as the reference's
 * pat_flt2int(method='clip') (ebsd_pattern.py:159-162).  d_weights, optional
 * (nrc*ncc*(2sr+1)^2 floats or NULL), receives the normalised weights.  d_work /
 * work_bytes: the optional workspace of ebsd_nlpar_workspace_bytes(nr*nc, P). */
int ebsd_nlpar(const void *d_pats, int dtype, int nr, int nc, int64_t P, const int32_t *d_idx, int M,
               const float *d_sigma, int sr, float lam, float dthresh, float diff_offset, float sat,
               int r0, int c0, int nrc, int ncc, void *d_out, int out_dtype, float *d_weights,
               void *d_work, size_t work_bytes, void *stream);

/* Which kernel ebsd_nlpar runs (process-wide; the results are identical): 1 (the default)
 * a 2 x 2 tile of scan points per work-group, whose weighted average loads each pattern
 * of the tile's windows once for all four points (where the patterns start on 16-byte
 * boundaries and P is a multiple of 16), 0 one point per work-group.  No counterpart in the
 * reference (a launch-shape switch of this implementation). */
int ebsd_nlpar_set_tiled(int on);

/* Bytes of the pair table the tiled kernel uses when ebsd_nlpar's workspace holds it after
 * the integer part (work_bytes >= ebsd_nlpar_workspace_bytes(nr*nc, P) rounded up to 256,
 * plus this): the squared distance and pixel count of every (point, window entry) of the
 * npoints = nrc*ncc calculation points, each unordered pair computed once -- half the
 * distance work, the same results.  No counterpart in the reference. */
size_t ebsd_nlpar_pair_bytes(int64_t npoints, int sr);

/* ---- the options of NLPAR.calcnlpar (nlpar_cpu.py:283-540) ---------------------------- */

/* Smallest pixel value of n patterns, as float32, into *d_min (device): the file minimum
 * that stem_scale subtracts (NLPAR._getdatamaxmin, nlpar_cpu.py:948-969). */
int ebsd_nlpar_min(const void *d_pats, int dtype, int64_t n, int64_t P, float *d_min, void *stream);

/* The float32 patterns calcnlpar hands its kernels (nlpar_cpu.py:470-484): the values as
 * float32; stem != 0: sqrt(v - dmin) in float32 (:472-476); d_back (P float64, or NULL):
 * minus the fitted background, float32(v - back) (:481-482, NLPAR.backsub :679-735).
 * d_out: n*P float32 (may not alias d_in unless dtype is float32). */
int ebsd_nlpar_prepare(const void *d_in, int dtype, int64_t n, int64_t P, int stem, float dmin,
                       const double *d_back, float *d_out, void *stream);

/* d_acc[p] += d_x[i*P + p] for i = 0..n-1 in order, float32: the pixel sums of numpy's mean
 * over the patterns (np.mean(data, axis=0), nlpar_cpu.py:697), accumulated across calls. */
int ebsd_nlpar_colsum(const float *d_x, int64_t n, int64_t P, float *d_acc, void *stream);

/* Averaged float32 patterns to the output file's values (nlpar_cpu.py:511-526): stem != 0:
 * v*v + dmin (:511-513); rescale_top > 0: per pattern v -= min, v *= top / max (float32,
 * :516-521); then out_dtype float32, or uint8/uint16 clipped to the type's range and
 * truncated (pat_flt2int 'clip'; NaN from a flat rescaled pattern becomes 0). */
int ebsd_nlpar_finish(const float *d_x, int64_t n, int64_t P, int stem, float dmin, float rescale_top,
                      void *d_out, int out_dtype, void *stream);

const char *ebsd_nlpar_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
