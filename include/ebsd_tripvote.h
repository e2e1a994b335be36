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

/* The rest of BandIndexer.bandindex after the vote (pyebsdindex/tripletvote.py:526-633):
 * band assignment (_assign_bands_nb2, :1483-1688) -- pairs of bands in rank order, each
 * library pole of the second band's family at the pair's angle giving a TRIAD rotation that
 * is scored on every band, the best pair kept, an early exit at n_band_early matched bands
 * --, the QUEST weights (_calc_quest_weights, :855-897: the nfit best-scoring matched bands
 * keep their intensity), the QUEST orientation (_orientation_quest_nb, :920-1018) and the
 * confidence metric (_calc_cm, :2056-2075).
 *   d_bandnorms, d_intensity: as for ebsd_tripvote;
 *   d_accw, d_fam, d_rank, d_acc: ebsd_tripvote's outputs for the same bands;
 *   the phase's complete pole list (npoles poles, nfam families): d_polescart (npoles, 3)
 *   float32 and d_polescart64 float64 unit Cartesian poles, d_famid (npoles) int32 family of
 *   each pole, d_angrows (nfam, npoles) float64 degrees between each family's first pole and
 *   every pole (rows famIndex of the reference's angTable), d_famindex / d_famcount (nfam)
 *   int32 first pole and pole count of each family;
 *   angtol degrees; n_band_early, nfit (6 in the reference); simple_weights 0 or 1.
 * Outputs, as the reference returns them: d_quat (npats, 4) float64, d_fit (npats) float64
 * degrees, d_cm (npats) float32, d_polematch (npats, nbands) int64 (-1000 where no pair
 * matched), d_nmatch (npats) int64, d_ij (npats, 4) int64, d_acc_correct (npats) int32,
 * d_bandfit (npats, nbands) float64 degrees; and, may be NULL: d_polevalid (npats, nbands)
 * uint8, d_assign_fit (npats) float32 and d_fitb (npats, nbands) float32, the assignment's
 * own fit.  The library must fit 64 KiB of LDS (npoles (40 B) + nfam npoles (8 B) + 8 nfam).
 * Replaces: the numba kernels of BandIndexer.bandindex after _tripvote_numba. */
int ebsd_bandindex(const float *d_bandnorms, const float *d_intensity, int64_t npats, int nbands,
                   const float *d_accw, const int32_t *d_fam, const float *d_rank, const float *d_acc, int nfam,
                   const float *d_polescart, const double *d_polescart64, const int32_t *d_famid, int npoles,
                   const double *d_angrows, const int32_t *d_famindex, const int32_t *d_famcount, double angtol,
                   int n_band_early, int nfit, int simple_weights, double *d_quat, double *d_fit, float *d_cm,
                   int64_t *d_polematch, int64_t *d_nmatch, int64_t *d_ij, int32_t *d_acc_correct,
                   double *d_bandfit, uint8_t *d_polevalid, float *d_assign_fit, float *d_fitb, void *stream);

const char *ebsd_tripvote_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
