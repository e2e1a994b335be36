/*
 * Oracle (test infrastructure only): a plain-C restatement of the band indexer's steps after
 * the vote.  Only tests/ and bench.py's CPU-baseline leg load it.
 *
 *   ora_assign_bands   <- pyebsdindex/tripletvote.py:1483-1688  (_assign_bands_nb2)
 *   ora_quest_weights  <- pyebsdindex/tripletvote.py:855-897    (_calc_quest_weights)
 *   ora_quest          <- pyebsdindex/tripletvote.py:920-1018   (_orientation_quest_nb)
 *   ora_calc_cm        <- pyebsdindex/tripletvote.py:2056-2075  (_calc_cm)
 *
 * Arithmetic types follow Numba's typing of the reference source: the band normals, library
 * poles, TRIAD matrices and the pole-fit products are float32; a float32 array divided by a
 * float64 scalar (p2 / (|p2| + 1e-35)) is float64, as are the cross products taken with it;
 * arccos of a float32 is float32 (taken here as the float64 arccos rounded to float32, the
 * correctly rounded value) and times RADEG float64; the quest weights and QUEST itself are
 * float64.  Sums run in index order; products are not contracted.
 * PARITY UNPINNED: the reference's np.dot / @ on 2-D arrays and np.linalg.norm go to BLAS,
 * whose summation order and fused multiply-adds are not restated (here ((a0 b0) + (a1 b1)) +
 * (a2 b2)); np.linalg.det is LAPACK's LU (here the cofactor expansion); argsort ties are
 * ordered stably (numpy's own argsort of band ranks is not stable on every CPU).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const double RADEG_B = 180.0 / 3.14159265358979323846;

static float dot3f(const float *a, const float *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static void cross3f(const float *a, const float *b, float *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static void cross3d(const double *a, const double *b, double *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static float norm3f(const float *a) { return sqrtf((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]); }
static double norm3d(const double *a) { return sqrt((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]); }
static float acosf_cr(float x) { return (float)acos((double)x); }

/* One (ii, jj) band pair of one pattern: __assign_triadfit (tripletvote.py:1486-1590). */
static void triadfit(int64_t ii, int64_t jj, int64_t nb, const float *bn, const int64_t *rank_arg,
                     const int32_t *bandFam, const float *pflt, int64_t npl, const double *angTable,
                     const int64_t *famIndx, const int64_t *nFam, double angTol, int64_t *whGood,
                     int64_t *nGood, float *angFit, int64_t *polematch, uint8_t *polevalid, float *R,
                     int64_t *bnd1o, int64_t *bnd2o, float *test /* npl * nb scratch */,
                     double *angTry /* nb scratch */)
{
    const float eps = 1.0e-12f;
    for (int m = 0; m < 9; m++) R[m] = 0.0f;
    for (int64_t j = 0; j < nb; j++) {
        angFit[j] = 0.0f;
        whGood[j] = 0;
        polematch[j] = -1;
        polevalid[j] = 0;
    }
    *nGood = 0;
    const int64_t bnd1 = rank_arg[nb - 1 - ii], bnd2 = rank_arg[nb - 1 - jj];
    *bnd1o = bnd1;
    *bnd2o = bnd2;
    const float *v1 = bn + bnd1 * 3, *v2 = bn + bnd2 * 3;
    const int64_t f1 = bandFam[bnd1], f2 = bandFam[bnd2];
    float ang01 = dot3f(v1, v2);
    if (ang01 > 1.0f) ang01 = (float)(1.0 - (double)eps);
    if (ang01 < -1.0f) ang01 = (float)(-1.0 + (double)eps);
    const double paralleltest = (double)acosf_cr(fabsf(ang01)) * RADEG_B;
    if (!(paralleltest > angTol)) return;
    const double a01 = (double)acosf_cr(ang01) * RADEG_B;
    const int64_t r1 = famIndx[f1];
    int64_t n12 = 0;
    int64_t *cand = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nFam[f2] > 0 ? nFam[f2] : 1));
    for (int64_t k = famIndx[f2]; k < famIndx[f2] + nFam[f2]; k++)
        if (fabs(angTable[r1 * npl + k] - a01) < angTol) cand[n12++] = k;
    if (n12 > 0) {
        const float *p1 = pflt + r1 * 3;
        float v1v2c[3], b2[3];
        cross3f(v1, v2, v1v2c);
        const float nn = norm3f(v1v2c);
        for (int m = 0; m < 3; m++) v1v2c[m] /= nn;
        cross3f(v1, v1v2c, b2);
        float B[9], A[9];
        for (int m = 0; m < 3; m++) {
            B[0 * 3 + m] = v1[m];
            B[1 * 3 + m] = v1v2c[m];
            B[2 * 3 + m] = b2[m];
            A[m * 3 + 0] = p1[m];
        }
        double score = -1.0;
        const double p1d[3] = {p1[0], p1[1], p1[2]};
        for (int64_t c = 0; c < n12; c++) {
            const float *p2f = pflt + cand[c] * 3;
            double nt = (double)norm3f(p2f) + 1.0e-35;
            double p2[3], pc[3], pc2[3];
            for (int m = 0; m < 3; m++) p2[m] = (double)p2f[m] / nt;
            cross3d(p1d, p2, pc);
            nt = norm3d(pc) + 1.0e-35;
            for (int m = 0; m < 3; m++) pc[m] /= nt;
            cross3d(p1d, pc, pc2);
            for (int m = 0; m < 3; m++) {
                A[m * 3 + 1] = (float)pc[m];
                A[m * 3 + 2] = (float)pc2[m];
            }
            float AB[9];
            for (int r = 0; r < 3; r++)
                for (int q = 0; q < 3; q++)
                    AB[r * 3 + q] = (A[r * 3 + 0] * B[0 * 3 + q] + A[r * 3 + 1] * B[1 * 3 + q]) + A[r * 3 + 2] * B[2 * 3 + q];
            /* testp = AB . bndnorm (3 x nb); test = pflt . testp (npl x nb) */
            for (int64_t j = 0; j < nb; j++) {
                const float *nj = bn + j * 3;
                float tp[3];
                for (int r = 0; r < 3; r++) tp[r] = (AB[r * 3 + 0] * nj[0] + AB[r * 3 + 1] * nj[1]) + AB[r * 3 + 2] * nj[2];
                for (int64_t k = 0; k < npl; k++) {
                    const float *pk = pflt + k * 3;
                    test[k * nb + j] = (pk[0] * tp[0] + pk[1] * tp[1]) + pk[2] * tp[2];
                }
            }
            int64_t nmatch = 0;
            double sum = 0.0;
            for (int64_t j = 0; j < nb; j++) {
                float mx = test[j];
                for (int64_t k = 1; k < npl; k++)
                    if (test[k * nb + j] > mx) mx = test[k * nb + j];
                if (mx < -1.0f) mx = -1.0f;
                if (mx > 1.0f) mx = 1.0f;
                angTry[j] = (double)acosf_cr(mx) * RADEG_B;
                if (angTry[j] < angTol) {
                    nmatch++;
                    sum += angTry[j];
                }
            }
            const double scoreTry = (double)(float)nmatch / (sum / (double)nmatch + 1e-6);
            if (scoreTry > score) {
                score = scoreTry;
                for (int64_t j = 0; j < nb; j++) {
                    angFit[j] = (float)angTry[j];
                    int64_t am = 0;
                    float mx = test[j];
                    for (int64_t k = 1; k < npl; k++)
                        if (test[k * nb + j] > mx) { mx = test[k * nb + j]; am = k; }
                    polematch[j] = am;
                    polevalid[j] = (uint8_t)(angTry[j] < angTol);
                }
                for (int m = 0; m < 9; m++) R[m] = AB[m];
            }
        }
        int64_t n = 0;
        for (int64_t j = 0; j < nb; j++)
            if ((double)angFit[j] < angTol) whGood[n++] = j;
        *nGood = n;
    }
    free(cand);
}

/* _assign_bands_nb2's driver loop (tripletvote.py:1592-1688). */
void ora_assign_bands(const float *pflt, int64_t npl, const double *angTable, const int64_t *famIndx,
                      const int64_t *nFam, double angTol, int64_t n_band_early, const float *bandnorms,
                      const int64_t *rank_arg, const int32_t *bandFam, int64_t npats, int64_t nb, float *fitout,
                      int64_t *polematch_out, uint8_t *polevalid_out, int64_t *nMatch, int64_t *whGood_out,
                      int64_t *ij, float *Rout, float *fitbout)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t p = 0; p < npats; p++) {
        const float *bn = bandnorms + p * nb * 3;
        const int64_t *ra = rank_arg + p * nb;
        const int32_t *bf = bandFam + p * nb;
        float *test = (float *)malloc(sizeof(float) * (size_t)(npl * nb));
        double *angTry = (double *)malloc(sizeof(double) * (size_t)nb);
        int64_t *whGood = (int64_t *)malloc(sizeof(int64_t) * (size_t)nb);
        float *angFit = (float *)malloc(sizeof(float) * (size_t)nb);
        int64_t *pm = (int64_t *)malloc(sizeof(int64_t) * (size_t)nb);
        uint8_t *pv = (uint8_t *)malloc((size_t)nb);
        float R[9];
        float fit_p = 360.0f;
        int64_t nMatch_p = 0;
        float *fitb_p = fitbout + p * nb;
        int64_t *wg_p = whGood_out + p * nb, *pm_p = polematch_out + p * nb, *ij_p = ij + p * 4;
        uint8_t *pv_p = polevalid_out + p * nb;
        float *R_p = Rout + p * 9;
        for (int64_t j = 0; j < nb; j++) {
            fitb_p[j] = 360.0f;
            wg_p[j] = -1;
            pm_p[j] = -1000;
            pv_p[j] = 0;
        }
        for (int m = 0; m < 9; m++) R_p[m] = 0.0f;
        for (int m = 0; m < 4; m++) ij_p[m] = -1;
        for (int64_t ii = 0; ii < nb - 1; ii++) {
            for (int64_t jj = ii + 1; jj < nb; jj++) {
                int64_t nGood = 0, bnd1 = 0, bnd2 = 0;
                triadfit(ii, jj, nb, bn, ra, bf, pflt, npl, angTable, famIndx, nFam, angTol, whGood, &nGood, angFit,
                         pm, pv, R, &bnd1, &bnd2, test, angTry);
                if (nGood < 3) continue;
                float fit = 0.0f;
                for (int64_t q = 0; q < nGood; q++) fit += angFit[whGood[q]];
                fit /= (float)nGood;
                int take = 0;
                if (nMatch_p < nGood) take = 1;
                else if (nMatch_p == nGood && fit_p > fit) take = 1;
                if (take) {
                    fit_p = fit;
                    for (int64_t j = 0; j < nb; j++) {
                        fitb_p[j] = angFit[j];
                        pm_p[j] = pm[j];
                        pv_p[j] = pv[j];
                    }
                    nMatch_p = nGood;
                    for (int64_t q = 0; q < nGood; q++) wg_p[q] = whGood[q];
                    for (int m = 0; m < 9; m++) R_p[m] = R[m];
                    ij_p[0] = ii;
                    ij_p[1] = jj;
                    ij_p[2] = bnd1;
                    ij_p[3] = bnd2;
                }
                if (nMatch_p >= n_band_early) break;
            }
            if (nMatch_p >= n_band_early) break;
        }
        fitout[p] = fit_p;
        nMatch[p] = nMatch_p;
        free(test);
        free(angTry);
        free(whGood);
        free(angFit);
        free(pm);
        free(pv);
    }
}

/* Stable ascending argsort of n floats (insertion sort). */
static void argsort_f(const float *v, int64_t n, int64_t *idx)
{
    for (int64_t i = 0; i < n; i++) idx[i] = i;
    for (int64_t i = 1; i < n; i++) {
        const int64_t k = idx[i];
        int64_t j = i;
        while (j > 0 && v[k] < v[idx[j - 1]]) {
            idx[j] = idx[j - 1];
            j--;
        }
        idx[j] = k;
    }
}

/* _calc_quest_weights (tripletvote.py:855-897). */
void ora_quest_weights(const int32_t *libFamID, const float *acc, const float *acc_nw, const int64_t *polematch,
                       const uint8_t *polevalid, const float *band_intensity, int64_t npats, int64_t nfam,
                       int64_t nb, int64_t nfit, int simple, float *weights)
{
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < npats; p++) {
        double *w = (double *)calloc((size_t)nb, sizeof(double));
        float *score = (float *)malloc(sizeof(float) * (size_t)nb);
        int64_t *srt = (int64_t *)malloc(sizeof(int64_t) * (size_t)nb);
        int64_t *whg = (int64_t *)malloc(sizeof(int64_t) * (size_t)nb);
        for (int64_t j = 0; j < nb; j++) {
            score[j] = -1.0f;
            weights[p * nb + j] = 0.0f;
        }
        int64_t n = 0;
        for (int64_t j = 0; j < nb; j++)
            if (polevalid[p * nb + j] > 0) whg[n++] = j;
        if (n >= 3) {
            const float *a = acc + p * nfam * nb, *anw = acc_nw + p * nfam * nb;
            for (int64_t q = 0; q < n; q++) {
                const int64_t j = whg[q];
                if (!simple) {
                    const int64_t f = libFamID[polematch[p * nb + j]];
                    const double d = fmax((double)anw[f * nb + j], 1.0e-12);
                    score[j] = (float)((double)a[f * nb + j] / d);
                } else {
                    score[j] = band_intensity[p * nb + q]; /* the reference indexes with q here */
                }
            }
            argsort_f(score, nb, srt);
            const int64_t k = n < nfit ? n : nfit;
            for (int64_t t = 0; t < k; t++) {
                const int64_t s = srt[nb - 1 - t]; /* np.flip of the ascending order */
                w[s] = (double)band_intensity[p * nb + s];
            }
            double sum = 0.0;
            for (int64_t j = 0; j < nb; j++) sum += w[j];
            for (int64_t j = 0; j < nb; j++) weights[p * nb + j] = (float)(w[j] / sum);
        }
        free(w);
        free(score);
        free(srt);
        free(whg);
    }
}

/* _orientation_quest_nb (tripletvote.py:920-1018): polescart (npats, nb, 3) the matched
 * library poles, bandnorms (npats, nb, 3), weights (npats, nb) normalised per pattern. */
void ora_quest(const double *polescart, const double *bandnorms, const double *weights, int64_t npats, int64_t nb,
               double *qout, double *fitout, double *madout, double *bdot)
{
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < npats; p++) {
        double *q = qout + p * 4;
        q[0] = 1.0;
        q[1] = q[2] = q[3] = 0.0;
        fitout[p] = 3.14159265358979323846;
        madout[p] = -1.0;
        for (int64_t j = 0; j < nb; j++) bdot[p * nb + j] = -1.0;
        const double *wp = weights + p * nb;
        int64_t n = 0;
        for (int64_t j = 0; j < nb; j++)
            if (wp[j] > 1.0e-7) n++;
        if (n < 3) continue;
        int64_t *whg = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
        double *wn = (double *)malloc(sizeof(double) * (size_t)n);
        n = 0;
        for (int64_t j = 0; j < nb; j++)
            if (wp[j] > 1.0e-7) { whg[n] = j; wn[n] = wp[j]; n++; }
        double ws = 0.0;
        for (int64_t j = 0; j < n; j++) ws += wn[j];
        for (int64_t j = 0; j < n; j++) wn[j] /= ws;
        double B[9] = {0}, z[3] = {0}, S[9], S2[9];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                double s = 0.0;
                for (int64_t j = 0; j < n; j++)
                    s += (wn[j] * bandnorms[(p * nb + whg[j]) * 3 + r]) * polescart[(p * nb + whg[j]) * 3 + c];
                B[r * 3 + c] = s;
            }
        for (int64_t j = 0; j < n; j++) {
            double cr[3];
            cross3d(bandnorms + (p * nb + whg[j]) * 3, polescart + (p * nb + whg[j]) * 3, cr);
            for (int m = 0; m < 3; m++) z[m] += wn[j] * cr[m];
        }
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) S[r * 3 + c] = B[r * 3 + c] + B[c * 3 + r];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++)
                S2[r * 3 + c] = (S[r * 3 + 0] * S[0 * 3 + c] + S[r * 3 + 1] * S[1 * 3 + c]) + S[r * 3 + 2] * S[2 * 3 + c];
        const double det = S[0] * (S[4] * S[8] - S[5] * S[7]) - S[1] * (S[3] * S[8] - S[5] * S[6]) +
                           S[2] * (S[3] * S[7] - S[4] * S[6]);
        const double k = (S[4] * S[8] - S[5] * S[7]) + (S[0] * S[8] - S[2] * S[6]) + (S[0] * S[4] - S[3] * S[1]);
        const double sig = 0.5 * ((S[0] + S[4]) + S[8]);
        const double sig2 = sig * sig;
        double t[3];
        for (int c = 0; c < 3; c++) t[c] = (z[0] * S2[0 * 3 + c] + z[1] * S2[1 * 3 + c]) + z[2] * S2[2 * 3 + c];
        const double d = (t[0] * z[0] + t[1] * z[1]) + t[2] * z[2];
        for (int c = 0; c < 3; c++) t[c] = (z[0] * S[0 * 3 + c] + z[1] * S[1 * 3 + c]) + z[2] * S[2 * 3 + c];
        const double cc = det + ((t[0] * z[0] + t[1] * z[1]) + t[2] * z[2]);
        const double b = sig2 + ((z[0] * z[0] + z[1] * z[1]) + z[2] * z[2]);
        const double a = sig2 - k;
        double lam = 1.0, dlam = 1e6;
        int iter = 0;
        while (dlam > 1.0e-12 && iter < 10) {
            const double lam0 = lam, l2 = lam * lam;
            lam = lam - (l2 * l2 - (a + b) * l2 - cc * lam + (a * b + cc * sig - d)) /
                            (4.0 * (l2 * lam) - 2.0 * (a + b) * lam - cc);
            dlam = fabs(lam0 - lam);
            iter++;
        }
        const double beta = lam - sig, alpha = lam * lam - sig2 + k, gamma = (lam + sig) * alpha - det;
        double M[9], X[3];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) M[r * 3 + c] = ((r == c ? alpha : 0.0) + beta * S[r * 3 + c]) + S2[r * 3 + c];
        for (int r = 0; r < 3; r++) X[r] = (M[r * 3 + 0] * z[0] + M[r * 3 + 1] * z[1]) + M[r * 3 + 2] * z[2];
        double qn = 0.0;
        qn += ((gamma * gamma + X[0] * X[0]) + X[1] * X[1]) + X[2] * X[2];
        qn = sqrt(qn);
        q[0] = gamma / qn;
        q[1] = X[0] / qn;
        q[2] = X[1] / qn;
        q[3] = X[2] / qn;
        if (gamma < 0.0)
            for (int m = 0; m < 4; m++) q[m] = -q[m];
        fitout[p] = lam;
        /* bdot: the band normals rotated by q (convention p = +1), dotted with their poles */
        double msum = 0.0;
        for (int64_t j = 0; j < n; j++) {
            const double *v = bandnorms + (p * nb + whg[j]) * 3, *pc = polescart + (p * nb + whg[j]) * 3;
            const double r0 = q[1], r1 = q[2], r2 = q[3], w = q[0];
            double vt[3], vt2[3], o[3];
            vt[0] = r1 * v[2] - r2 * v[1];
            vt[1] = r2 * v[0] - r0 * v[2];
            vt[2] = r0 * v[1] - r1 * v[0];
            for (int m = 0; m < 3; m++) vt[m] += w * v[m];
            vt2[0] = r1 * vt[2] - r2 * vt[1];
            vt2[1] = r2 * vt[0] - r0 * vt[2];
            vt2[2] = r0 * vt[1] - r1 * vt[0];
            for (int m = 0; m < 3; m++) o[m] = 2.0 * vt2[m] + v[m];
            const double bd = (o[0] * pc[0] + o[1] * pc[1]) + o[2] * pc[2];
            bdot[p * nb + whg[j]] = bd;
            msum += bd;
        }
        madout[p] = msum / (double)n;
        free(whg);
        free(wn);
    }
}

/* _calc_cm (tripletvote.py:2056-2075): accumulator = the weighted votes. */
void ora_calc_cm(const float *acc, const int64_t *polematch, const uint8_t *polevalid, const int32_t *libFamID,
                 int64_t npats, int64_t nfam, int64_t nb, float *cm2)
{
    for (int64_t p = 0; p < npats; p++) {
        int64_t n = 0;
        for (int64_t j = 0; j < nb; j++) n += polevalid[p * nb + j] > 0;
        cm2[p] = -1.0f;
        if (n < 3) continue;
        const float *a = acc + p * nfam * nb;
        float c = 0.0f;
        for (int64_t j = 0; j < nb; j++)
            if (polevalid[p * nb + j] > 0) c += a[(int64_t)libFamID[polematch[p * nb + j]] * nb + j];
        float den = 0.0f;
        for (int64_t t = 0; t < nfam * nb; t++) den += a[t] < 1.0f ? 1.0f : a[t];
        cm2[p] = c / den;
    }
}
