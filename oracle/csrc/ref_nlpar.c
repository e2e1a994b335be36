/*
 * Oracle (test infrastructure only): a plain-C restatement of the two Numba kernels of
 * PyEBSDIndex's NLPAR CPU path.  Only tests/ and bench.py's CPU-baseline leg load it.
 *
 *   ora_nlpar_sigma  <- pyebsdindex/nlpar_cpu.py:753-819  (NLPAR.sigma_numba)
 *   ora_nlpar        <- pyebsdindex/nlpar_cpu.py:821-935  (NLPAR.nlpar_nb)
 *
 * data: (nrows*ncols, P) float32, scan point (j, i) at row j*ncols + i.  Arithmetic types
 * follow Numba's typing of the reference source: a variable that is assigned both a
 * float32 and a float64 value is float64 (mxval and n2 of sigma_numba), float32 ** int is
 * float32, float32 mixed with a Python float literal is float64.  The loops run in the
 * reference's order; OpenMP takes the place of numba.prange over rows (sigma) or columns
 * (nlpar).  The reference's per-column cache of pair distances (pairdict) only avoids
 * recomputing a symmetric value and is not restated.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

static float maxf_all(const float *data, int64_t n)
{
    float m = data[0];
    for (int64_t t = 1; t < n; t++)
        if (data[t] > m) m = data[t];
    return m;
}

/* nlpar_cpu.py:753-819.  sigma (nrows*ncols); dout, nout (nrows*ncols*(2nn+1)^2), zero-filled
 * by the caller. */
void ora_nlpar_sigma(const float *data, int64_t nrows, int64_t ncols, int64_t P, const int64_t *indices,
                     int64_t M, int64_t nn, int sat_protect, float *sigma, float *dout, float *nout)
{
    const int64_t win = (2 * nn + 1) * (2 * nn + 1);
    const float n0 = (float)P;
    double mxval = (double)maxf_all(data, nrows * ncols * P);   /* float64 by unification */
    if (!sat_protect) mxval += 1.0;
    else mxval *= 0.9961;
    int64_t *dij = (int64_t *)malloc(sizeof(int64_t) * nrows * ncols * win * 2);
#pragma omp parallel for schedule(dynamic)
    for (int64_t j = 0; j < nrows; j++) {
        int64_t r0 = j - nn >= 0 ? j - nn : 0;
        int64_t r1 = (j + nn < nrows ? j + nn : nrows - 1) + 1;
        for (int64_t i = 0; i < ncols; i++) {
            int64_t c0 = i - nn >= 0 ? i - nn : 0;
            int64_t c1 = (i + nn < ncols ? i + nn : ncols - 1) + 1;
            float mind = 1e24f;
            int64_t p = i + ncols * j;
            int64_t count = 0;
            for (int64_t jn = r0; jn < r1; jn++) {
                for (int64_t in = c0; in < c1; in++) {
                    int64_t e = p * win + count;
                    dij[2 * e] = jn;
                    dij[2 * e + 1] = in;
                    int64_t q = in + ncols * jn;
                    float d2 = 0.0f;
                    double n2 = 1.0e-12;
                    nout[e] = n0;
                    if (!(i == in && j == jn)) {
                        for (int64_t k = 0; k < M; k++) {
                            float d0 = data[p * P + indices[k]];
                            float d1 = data[q * P + indices[k]];
                            if ((double)d1 < mxval && (double)d0 < mxval) {
                                n2 += 1.0;
                                float d = d0 - d1;
                                d2 += d * d;
                            }
                        }
                        nout[e] = (float)n2;
                        if (d2 >= 1.e-3f) {
                            float s0 = d2 / (float)(n2 * 2.0);
                            if (s0 < mind) mind = s0;
                        }
                    }
                    dout[e] = d2;
                    count++;
                }
            }
            sigma[p] = sqrtf(mind);
        }
    }
    /* normalisation (nlpar_cpu.py:810-818) */
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < nrows * ncols; p++) {
        for (int64_t e = 0; e < win; e++) {
            int64_t t = p * win + e;
            if (nout[t] > 0) {
                int64_t jj = dij[2 * t], ii = dij[2 * t + 1];
                float s2a = sigma[p] * sigma[p];
                float s2b = sigma[jj * ncols + ii] * sigma[jj * ncols + ii];
                float s12 = s2a + s2b;
                dout[t] -= nout[t] * s12;
                dout[t] = (float)((double)dout[t] / ((double)s12 * sqrt(2.0 * (double)nout[t])));
            }
        }
    }
    free(dij);
}

/* nlpar_cpu.py:821-935 over calclim = (cstart, rstart, ncol, nrow) of the chunk.
 * dataout (nrows*ncols*P float32) zero-filled by the caller; weights_out, if not NULL,
 * (nrows*ncols*(2sr+1)^2) receives each computed point's normalised weights. */
void ora_nlpar(const float *data, float lam, int64_t sr, float dthresh, const float *sigma, int64_t nrows,
               int64_t ncols, int64_t P, const int64_t *indices, int64_t M, int sat_protect, float diff_offset,
               int64_t cstart, int64_t rstart, int64_t ncalc, int64_t nrcalc, float *dataout, float *weights_out)
{
    const float lamsq = lam * lam;                 /* lam**2: float32 ** int -> float32 */
    const double lam2 = 1.0 / (double)lamsq;
    const int64_t winsz = (2 * sr + 1) * (2 * sr + 1);
    float mxval = maxf_all(data, nrows * ncols * P);
    if (!sat_protect) mxval += 1.0f;
    else mxval *= 0.999f;
#pragma omp parallel for schedule(dynamic)
    for (int64_t ii = 0; ii < ncalc; ii++) {
        int64_t i = cstart + ii;
        int64_t wx0 = (i - sr > 0 ? i - sr : 0) - (i + sr - (ncols - 1) > 0 ? i + sr - (ncols - 1) : 0);
        int64_t wx1 = (i + sr < ncols - 1 ? i + sr : ncols - 1) + (sr - i > 0 ? sr - i : 0) + 1;
        float *weights = (float *)malloc(sizeof(float) * winsz);
        int64_t *pindx = (int64_t *)malloc(sizeof(int64_t) * winsz);
        float *diff_step = (float *)malloc(sizeof(float) * winsz);
        for (int64_t jj = 0; jj < nrcalc; jj++) {
            int64_t j = rstart + jj;
            int64_t wy0 = (j - sr > 0 ? j - sr : 0) - (j + sr - (nrows - 1) > 0 ? j + sr - (nrows - 1) : 0);
            int64_t wy1 = (j + sr < nrows - 1 ? j + sr : nrows - 1) + (sr - j > 0 ? sr - j : 0) + 1;
            int64_t p = i + ncols * j;
            float sigma0 = sigma[p] * sigma[p];
            int64_t counter = 0;
            for (int64_t jn = wy0; jn < wy1; jn++) {
                for (int64_t in = wx0; in < wx1; in++) {
                    int64_t q = in + ncols * jn;
                    pindx[counter] = q;
                    /* the centre's offset, added once per pattern (the reference keeps
                       diff_step in an array shared by the prange workers) */
                    diff_step[counter] = 0.0f;
                    if (q == p) {
                        weights[counter] = -1.0e6f;
                        diff_step[counter] = diff_offset;
                    } else {
                        float sigma1 = sigma[q] * sigma[q];
                        float d2 = 0.0f, n2 = 0.0f;
                        for (int64_t k = 0; k < M; k++) {
                            float d0 = data[p * P + indices[k]];
                            float d1 = data[q * P + indices[k]];
                            if (d1 < mxval && d0 < mxval) {
                                n2 += 1.0f;
                                float d = d0 - d1;
                                d2 += d * d;
                            }
                        }
                        d2 -= n2 * (sigma0 + sigma1);
                        float dnorm = (sigma1 + sigma0) * sqrtf(2.0f * n2);
                        if (dnorm > 1.e-8f) d2 /= dnorm;
                        else d2 = 1e6f * n2;
                        weights[counter] = d2;
                    }
                    counter++;
                }
            }
            float sum = 0.0f;
            for (int64_t k = 0; k < winsz; k++) {
                float w = weights[k] - dthresh;
                w = w > 0.0f ? w : 0.0f;
                w = (float)exp(-1.0 * (double)w * lam2);
                w += diff_step[k];
                weights[k] = w;
                sum += w;
            }
            for (int64_t k = 0; k < winsz; k++) {
                weights[k] /= sum;
                if (weights_out) weights_out[p * winsz + k] = weights[k];
                const float *src = data + pindx[k] * P;
                float *dst = dataout + p * P;
                for (int64_t x = 0; x < P; x++) dst[x] += src[x] * weights[k];
            }
        }
        free(weights);
        free(pindx);
        free(diff_step);
    }
}
