/*
 * Oracle (test infrastructure only): a plain-C restatement of the three Numba
 * loops on PyEBSDIndex's CPU band-detection path.  Only tests/, the smoke check
 * and bench.py's CPU-baseline leg may load this library; the product never does.
 *
 *   ora_rdn_loops      <- pyebsdindex/radon_fast.py:281-310   (Radon.rdn_loops)
 *   ora_grey_dilation  <- pyebsdindex/band_detect.py:811-843  (BandDetect._grey_dilation)
 *   ora_band_label     <- pyebsdindex/band_detect.py:691-809  (BandDetect.band_label_numba)
 *
 * Array layouts follow the reference: Radon-space arrays are (nRho, nTheta, nPats)
 * float32 with the pattern index innermost.  Arithmetic types follow Numba's
 * typing of the reference source (float64 accumulators and float64 promotion of
 * float32 values mixed with Python float literals); fastmath re-association
 * is not reproduced.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* radon_fast.py:281-310.  images: nIm*nPx float32, index: nRho*nTheta*nIndex int64
 * (sorted per line, entries >= nPx terminate the line), radon: nRp*nTp*nIm float32,
 * pre-filled by the caller with the missing value. */
void ora_rdn_loops(const float *images, const int64_t *index, int64_t nIm, int64_t nPx,
                   int64_t nRho, int64_t nTheta, int64_t nIndex, float *radon,
                   int64_t nRp, int64_t nTp, int64_t pad0, int64_t pad1, float norm)
{
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < nIm; q++) {
        const float *im = images + q * nPx;
        for (int64_t i = 0; i < nRho; i++) {
            int64_t ip = i + pad0;
            for (int64_t j = 0; j < nTheta; j++) {
                int64_t jp = j + pad1;
                double count = 0.0, sum = 0.0;
                const int64_t *line = index + (i * nTheta + j) * nIndex;
                for (int64_t k = 0; k < nIndex; k++) {
                    int64_t idx = line[k];
                    if (idx >= nPx) break;
                    sum += (double)im[idx];
                    count += 1.0;
                }
                if (count >= 1.0) {
                    if (norm < 0.001f) count = 1.0;
                    radon[(ip * nTp + jp) * nIm + q] = (float)(sum / count);
                }
            }
        }
    }
}

/* band_detect.py:811-843.  Brute-force max over the window clamped to the array. */
void ora_grey_dilation(const float *image, float *out, int64_t n0, int64_t n1, int64_t nq,
                       int64_t s0, int64_t s1)
{
    int64_t k0 = s0 / 2, k1 = s1 / 2;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < nq; q++) {
        for (int64_t j = 0; j < n1; j++) {
            int64_t j0 = j - k1 < 0 ? 0 : j - k1;
            int64_t j1 = (j + k1 < n1 - 1 ? j + k1 : n1 - 1) + 1;
            for (int64_t i = 0; i < n0; i++) {
                int64_t i0 = i - k0 < 0 ? 0 : i - k0;
                int64_t i1 = (i + k0 < n0 - 1 ? i + k0 : n0 - 1) + 1;
                float mx = -INFINITY;
                for (int64_t ii = i0; ii < i1; ii++)
                    for (int64_t jj = j0; jj < j1; jj++) {
                        float v = image[(ii * n1 + jj) * nq + q];
                        if (v > mx) mx = v;
                    }
                out[(i * n1 + j) * nq + q] = mx;
            }
        }
    }
}

typedef struct { float v; int64_t pos; } cand_t;

/* Ascending by value; equal values keep row-major order (a stable argsort).  Numba's
 * argsort is an unstable quicksort whose tie order is implementation-defined; the
 * stable order is the documented choice for exact ties. */
static int cand_cmp(const void *a, const void *b)
{
    const cand_t *x = (const cand_t *)a, *y = (const cand_t *)b;
    if (x->v < y->v) return -1;
    if (x->v > y->v) return 1;
    return (x->pos > y->pos) - (x->pos < y->pos);
}

#define AT(r, c) conv[((int64_t)(r) * n1 + (c)) * nq + q]

/* band_detect.py:691-809.  conv/lmax: (n0, n1, nq).  Outputs are (nq, nB[,2]) and must
 * be pre-initialised by the caller exactly as band_detect.py:700-707 does
 * (max = avemax = -2e6, everything else 0). */
void ora_band_label(int64_t nB, int64_t nq, const float *conv, const uint8_t *lmax,
                    int64_t n0, int64_t n1, float *bmax, float *bavemax, float *bmaxloc,
                    float *baveloc, int8_t *bvalid, float *bwidth)
{
#pragma omp parallel
    {
        cand_t *cands = (cand_t *)malloc(sizeof(cand_t) * (size_t)(n0 * n1));
#pragma omp for schedule(static)
        for (int64_t q = 0; q < nq; q++) {
            int64_t nc = 0;
            for (int64_t r = 0; r < n0; r++)
                for (int64_t c = 0; c < n1; c++)
                    if (lmax[(r * n1 + c) * nq + q]) {
                        cands[nc].v = AT(r, c);
                        cands[nc].pos = r * n1 + c;
                        nc++;
                    }
            qsort(cands, (size_t)nc, sizeof(cand_t), cand_cmp);
            int64_t nBq = nc > nB ? nB : nc;
            for (int64_t i = 0; i < nBq; i++) {
                int64_t pos = cands[nc - 1 - i].pos;
                int32_t r = (int32_t)(pos / n1), c = (int32_t)(pos % n1);
                if (!(AT(r, c) > 0.0f)) continue;
                int64_t o = q * nB + i;
                bmaxloc[2 * o + 0] = (float)r;
                bmaxloc[2 * o + 1] = (float)c;
                float mx = AT(r, c);
                bmax[o] = mx;
                /* band_detect.py:748-753: FWHM-style width from the rho neighbours. */
                bwidth[o] = 0.0f;
                float mn = AT(r - 1, c);
                if (AT(r, c) < mn) mn = AT(r, c);
                if (AT(r + 1, c) < mn) mn = AT(r + 1, c);
                if (mn > 0.0f) {
                    float lp = logf(AT(r + 1, c)), lm = logf(AT(r - 1, c)), l0 = logf(mx);
                    double a = (double)(lp - lm) * 0.5 - (double)l0;
                    if (a < -1.e-8) bwidth[o] = (float)(2.0 * sqrt(log(2.0) / (-1.0 * a)));
                }
                /* band_detect.py:764-798: 3x3 quadratic (Taylor) sub-pixel refinement. */
                float nn[3][3];
                float s = 0.0f;
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) { nn[a][b] = AT(r - 1 + a, c - 1 + b); s += nn[a][b]; }
                double sumnn = s > 1e-12 ? (double)s : 1e-12;
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) nn[a][b] = (float)((double)nn[a][b] / sumnn);
                bavemax[o] = (float)(sumnn / 9.0);
                double dx = 0.5 * (double)(nn[1][2] - nn[1][0]);
                double dy = 0.5 * (double)(nn[2][1] - nn[0][1]);
                double dxx = (double)(nn[1][2] + nn[1][0]) - 2.0 * (double)nn[1][1];
                double dyy = (double)(nn[2][1] + nn[0][1]) - 2.0 * (double)nn[1][1];
                double dxy = 0.25 * (double)(((nn[2][2] - nn[0][2]) - nn[2][0]) + nn[0][0]);
                double det = dxx * dyy - dxy * dxy;
                det = fabs(det) > 1e-12 ? det : 1.0e-12;
                det = 1.0 / det;
                double dc = (dyy * dx - dxy * dy) * det;
                double rc = (dxx * dy - dxy * dx) * det;
                if (fabs(dc) > 0.875 || fabs(rc) > 0.875) {
                    det = dxx * dyy;
                    det = fabs(det) > 1e-12 ? det : 1.0e-12;
                    det = 1.0 / det;
                    dc = (dyy * dx) * det;
                    rc = (dxx * dy) * det;
                    if (fabs(dc) > 0.875 || fabs(rc) > 0.875) { dc = 0.0; rc = 0.0; }
                }
                baveloc[2 * o + 0] = (float)((double)r - rc);
                baveloc[2 * o + 1] = (float)((double)c - dc);
                bvalid[o] = 1;
            }
        }
        free(cands);
    }
}
