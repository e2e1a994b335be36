/*
 * Oracle (test infrastructure only): a plain-C restatement of the triplet vote of
 * PyEBSDIndex's band indexer.  Only tests/ and bench.py's CPU-baseline leg load it.
 *
 *   ora_tripvote  <- pyebsdindex/tripletvote.py:1019-1207  (BandIndexer._tripvote_numba)
 *
 * Arithmetic types follow Numba's typing of the reference source:
 *   - the inter-band cosines are float32 dot products of the float32 band normals,
 *     evaluated here as ((x0 y0) + (x1 y1)) + (x2 y2) without contraction.  PARITY
 *     UNPINNED: the reference's 2-D .dot (tripletvote.py:1192) goes to a BLAS sgemm whose
 *     summation order and use of fused multiply-adds are not restated; an angle at the
 *     tolerance edge can vote differently (tests/test_oracle_tripvote.py); made
 *     absolute and clipped to [-1, 1]; the angle is the float32 arccos times RADEG in
 *     float64, rounded to float32 where the triplet is formed.  The float32 arccos is
 *     taken as the float64 arccos rounded to float32 (the correctly rounded value; the
 *     reference's float32 libm call agrees with it to an ulp);
 *   - a band with intensity below 1e-6 has angle 10000 to every band;
 *   - the test of a library triplet is float64 (a float64 library angle minus a float32
 *     triplet angle); the kept differences are float32; a weight is angTol - 0.5 (t_a +
 *     t_b) with a float32 sum and a float64 result; each vote adds to a float32
 *     accumulator through float64; the symmetry tests compare float32 differences;
 *   - the per-band metrics: a float32 running sum and maximum over the families, both
 *     truncated to int32 (votes, mxvote); cm = (largest - second largest) / votes,
 *     float32 difference, float64 quotient, 0 below one vote; the family is the first
 *     index of the largest; rank = f32(((nb - b) / nb) * cm * mxvote) in float64.
 * The loops run in the reference's order; OpenMP takes the place of numba.prange over
 * patterns.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const double RADEG = 180.0 / 3.14159265358979323846;

static void sort3_idx(const float *a, int64_t s[3])
{
    int64_t t;
    s[0] = 0; s[1] = 1; s[2] = 2;
    if (a[s[0]] > a[s[2]]) { t = s[2]; s[2] = s[0]; s[0] = t; }
    if (a[s[0]] > a[s[1]]) { t = s[1]; s[1] = s[0]; s[0] = t; }
    if (a[s[1]] > a[s[2]]) { t = s[2]; s[2] = s[1]; s[1] = t; }
}

static void sort3_idx_i(const int64_t *a, int64_t s[3])
{
    int64_t t;
    s[0] = 0; s[1] = 1; s[2] = 2;
    if (a[s[0]] > a[s[2]]) { t = s[2]; s[2] = s[0]; s[0] = t; }
    if (a[s[0]] > a[s[1]]) { t = s[1]; s[1] = s[0]; s[0] = t; }
    if (a[s[1]] > a[s[2]]) { t = s[2]; s[2] = s[1]; s[1] = t; }
}

static int cmp_float(const void *x, const void *y)
{
    float a = *(const float *)x, b = *(const float *)y;
    return (a > b) - (a < b);
}

/* bandnorms (npats, nb, 3) float32; intensity (npats, nb) float32; lut (3,3,3,3) int64;
 * tripAng (ntrip, 3) float64; tripID (ntrip, 3) int64.  Outputs: accW and acc (npats,
 * nfam, nb) float32, bandFam (npats, nb) int32, bandRank and bandCm (npats, nb) float32. */
void ora_tripvote(const float *bandnorms, const float *intensity, int64_t npats, int64_t nb,
                  const int64_t *lut, double angTol, const double *tripAng, const int64_t *tripID,
                  int64_t ntrip, int64_t nfam, float *accW, int32_t *bandFam, float *bandRank,
                  float *bandCm, float *acc)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t p = 0; p < npats; p++) {
        float *angs = (float *)malloc(sizeof(float) * nb * nb);
        float *colv = (float *)malloc(sizeof(float) * nfam);
        float *aW = accW + p * nfam * nb, *aC = acc + p * nfam * nb;
        const float *bn = bandnorms + p * nb * 3;
        memset(aW, 0, sizeof(float) * nfam * nb);
        memset(aC, 0, sizeof(float) * nfam * nb);
        for (int64_t a = 0; a < nb; a++)
            for (int64_t b = 0; b < nb; b++) {
                float d0 = bn[a * 3 + 0] * bn[b * 3 + 0];
                float d1 = bn[a * 3 + 1] * bn[b * 3 + 1];
                float d2 = bn[a * 3 + 2] * bn[b * 3 + 2];
                float c = fabsf((d0 + d1) + d2);
                if (c > 1.0f) c = 1.0f;
                float ac = (float)acos((double)c);
                angs[a * nb + b] = (float)((double)ac * RADEG);
            }
        for (int64_t q = 0; q < nb; q++)
            if ((double)intensity[p * nb + q] < 1e-6)
                for (int64_t t = 0; t < nb; t++) angs[q * nb + t] = angs[t * nb + q] = 10000.0f;

        float angTest0[3] = {0.0f, 0.0f, 0.0f};
        for (int64_t i = 0; i < nb; i++)
            for (int64_t j = i + 1; j < nb; j++)
                for (int64_t k = j + 1; k < nb; k++) {
                    float angtri[3] = {angs[i * nb + j], angs[i * nb + k], angs[j * nb + k]};
                    int64_t srt[3], unsrt[3], srt2[3];
                    sort3_idx(angtri, srt);
                    for (int m = 0; m < 3; m++) srt2[m] = lut[m * 27 + srt[0] * 9 + srt[1] * 3 + srt[2]];
                    sort3_idx_i(srt2, unsrt);
                    float s0 = angtri[srt[0]], s1 = angtri[srt[1]], s2 = angtri[srt[2]];
                    for (int64_t qq = 0; qq < ntrip; qq++) {
                        double t1 = fabs(tripAng[qq * 3 + 0] - (double)s0);
                        if (t1 > angTol) continue;
                        angTest0[0] = (float)t1;
                        double t2 = fabs(tripAng[qq * 3 + 1] - (double)s1);
                        if (t2 > angTol) continue;
                        angTest0[1] = (float)t2;
                        double t3 = fabs(tripAng[qq * 3 + 2] - (double)s2);
                        if (t3 > angTol) continue;
                        angTest0[2] = (float)t3;
                        int64_t f0 = tripID[qq * 3 + unsrt[0]], f1 = tripID[qq * 3 + unsrt[1]],
                                f2 = tripID[qq * 3 + unsrt[2]];
                        double w1 = angTol - 0.5 * (double)(angTest0[0] + angTest0[1]);
                        double w2 = angTol - 0.5 * (double)(angTest0[0] + angTest0[2]);
                        double w3 = angTol - 0.5 * (double)(angTest0[1] + angTest0[2]);
#define VOTE(f, b, w)                                                        \
    do {                                                                     \
        aW[(f) * nb + (b)] = (float)((double)aW[(f) * nb + (b)] + (w));      \
    } while (0)
#define COUNT(f, b)                                                          \
    do {                                                                     \
        aC[(f) * nb + (b)] = (float)((double)aC[(f) * nb + (b)] + 1.0);      \
    } while (0)
                        VOTE(f0, i, w1); VOTE(f1, j, w2); VOTE(f2, k, w3);
                        COUNT(f0, i); COUNT(f1, j); COUNT(f2, k);
                        int e1 = 0, e2 = 0, e3 = 0;
                        if (fabsf(s0 - s1) < angTol) {
                            VOTE(f0, j, w1); VOTE(f1, i, w2); VOTE(f2, k, w3);
                            COUNT(f0, j); COUNT(f1, i); COUNT(f2, k);
                            e1 = 1;
                        }
                        if (fabsf(s1 - s2) < angTol) {
                            VOTE(f0, i, w1); VOTE(f1, k, w2); VOTE(f2, j, w3);
                            COUNT(f0, i); COUNT(f1, k); COUNT(f2, j);
                            e2 = 1;
                        }
                        if (fabsf(s2 - s0) < angTol) {
                            VOTE(f0, k, w1); VOTE(f1, j, w2); VOTE(f2, i, w3);
                            COUNT(f0, k); COUNT(f1, j); COUNT(f2, i);
                            e3 = 1;
                        }
                        if (e1 && e2 && e3) {
                            VOTE(f0, k, w1); VOTE(f1, i, w2); VOTE(f2, j, w3);
                            VOTE(f0, j, w1); VOTE(f1, k, w2); VOTE(f2, i, w3);
                            COUNT(f0, k); COUNT(f1, i); COUNT(f2, j);
                            COUNT(f0, j); COUNT(f1, k); COUNT(f2, i);
                        }
#undef VOTE
#undef COUNT
                    }
                }

        for (int64_t b = 0; b < nb; b++) {
            float mx = -1.0e12f, sum = 0.0f;
            for (int64_t f = 0; f < nfam; f++) {
                float v = aW[f * nb + b];
                sum += v;
                if (v > mx) mx = v;
            }
            int32_t mxvote = (int32_t)mx, tvotes = (int32_t)sum;
            float cm = 0.0f;
            if (tvotes >= 1) {
                for (int64_t f = 0; f < nfam; f++) colv[f] = aW[f * nb + b];
                qsort(colv, (size_t)nfam, sizeof(float), cmp_float);
                float d = colv[nfam - 1] - colv[nfam - 2];
                cm = (float)((double)d / (double)tvotes);
            }
            int32_t fam = 0;
            float m2 = -1.0e12f;
            for (int64_t f = 0; f < nfam; f++)
                if (aW[f * nb + b] > m2) { m2 = aW[f * nb + b]; fam = (int32_t)f; }
            bandFam[p * nb + b] = fam;
            bandCm[p * nb + b] = cm;
            bandRank[p * nb + b] = (float)((double)(nb - b) / (double)nb * (double)cm * (double)mxvote);
        }
        free(angs);
        free(colv);
    }
}
