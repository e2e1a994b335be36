/*
 * ebsd_tripvote.h -- C ABI of the MI355X (gfx950) triplet-vote library.
 *
 * The library replaces BandIndexer._tripvote_numba (pyebsdindex/tripletvote.py:1019-1207),
 * the first step of BandIndexer.bandindex (tripletvote.py:523-524): every triplet of a
 * pattern's bands is matched against a phase's library of pole-family angle triplets,
 * and each match votes for a family on each of the three bands.  The library itself
 * (BandIndexer.build_trip_lib, tripletvote.py:320-491) is built once per phase on the
 * host and passed in.
 *
 * Conventions: every pointer named d_* is device memory; functions return 0 on success
 * and a negative code on failure, with the message from ebsd_tripvote_last_error() (per
 * thread).  Work is enqueued on `stream` (a hipStream_t; NULL = the default stream).
 */
#ifndef EBSD_TRIPVOTE_H
#define EBSD_TRIPVOTE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Votes of npats patterns of nbands bands each (nbands <= 64).
 *   d_bandnorms  (npats, nbands, 3) float32 unit band normals (radon2pole's output);
 *   d_intensity  (npats, nbands) float32; a band below 1e-6 takes no part;
 *   d_tripang    (ntrip, 3) float64 library angles in degrees, ascending in each row;
 *   d_tripid     (ntrip, 3) int32 family of each vertex, in [0, nfam) (nfam >= 2, and
 *                nbands^2 + 2 nfam nbands floats within 64 KiB; a row with an id outside
 *                [0, nfam) votes for nothing);
 *   angtol       degrees;
 *   d_bin_start, d_bin_rows, nbins, bin_width: an optional row index (NULL = test every
 *                row).  Bin b holds the rows d_bin_rows[d_bin_start[b] .. d_bin_start[b+1])
 *                in ascending order; it must contain every row whose first angle lies
 *                within angtol + 1e-3 of some x in [b bin_width, (b+1) bin_width) -- the
 *                first bin extended to -inf and the last to +inf (the kernel finds the bin
 *                with a float32 product).  A triplet whose smallest angle x0
 *                falls in bin b tests only that list; the outputs are the same as without
 *                the index.  No counterpart in the reference (its loop tests every row).
 * Outputs, the five arrays of the reference in its order:
 *   d_accw  (npats, nfam, nbands) float32 weighted votes (weight angtol minus the mean
 *           angle error of the two triangle edges at the band),
 *   d_fam   (npats, nbands) int32 family with the most weighted votes (first on ties),
 *   d_rank  (npats, nbands) float32 ((nbands - b) / nbands) * cm * trunc(max vote),
 *   d_cm    (npats, nbands) float32 (largest - second largest) / trunc(vote sum), 0
 *           below one vote,
 *   d_acc   (npats, nfam, nbands) float32 vote counts.
 * The triplet table the reference passes as LUT (tripletvote.py:207-213) is fixed and
 * built into the kernel. */
int ebsd_tripvote(const float *d_bandnorms, const float *d_intensity, int64_t npats, int nbands,
                  const double *d_tripang, const int32_t *d_tripid, int ntrip, int nfam, double angtol,
                  const int32_t *d_bin_start, const int32_t *d_bin_rows, int nbins, double bin_width,
                  float *d_accw, int32_t *d_fam, float *d_rank, float *d_cm, float *d_acc, void *stream);

const char *ebsd_tripvote_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
