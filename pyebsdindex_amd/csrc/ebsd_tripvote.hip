// Triplet voting of band normals against a phase's angle-triplet library, for gfx950.
//
// Replaces BandIndexer._tripvote_numba (pyebsdindex/tripletvote.py:1019-1207).
//
// k_tripvote_group<G> (up to 32 bands): 64 / G patterns per wavefront, one per group of G
// lanes (G = 16 up to 16 bands, else 32); k_tripvote (more bands): one pattern per
// wavefront.  The work-group copies the library into LDS as columns when it fits in 64 KiB
// with the per-pattern state.  A group computes its pattern's inter-band angles into LDS;
// its lanes then build the records of G band triplets at a time (ascending angles, vertex
// order, near-equality flags) and the group walks them in the reference's (i < j < k)
// order.  For each triplet the lanes test G library rows at once (the optional row index
// narrows the rows to those whose first angle can pass); the rows that pass all three
// tests are taken in ascending row order from the group's bits of the ballot mask; three
// lanes of the group add a match's votes to the pattern's LDS cells, one lane per band of
// the triplet, in the reference's statement order, so every accumulator receives the same
// float additions in the same order as the reference's sequential loop.
// The per-band metrics are computed one band per lane.
//
// Arithmetic (the types Numba gives the reference):
//   cosine   float32 ((x0 y0) + (x1 y1)) + (x2 y2), no contraction; |.|, clipped at 1
//   angle    f32(f64(f32 arccos) * RADEG); the float32 arccos is the float64 arccos
//            rounded to float32 (the correctly rounded value);  10000 for a band with
//            intensity below 1e-6
//   tests    float64 |library angle - triplet angle| <= angtol; kept as float32
//   weights  angtol - 0.5 * f32(t_a + t_b) in float64, added to float32 accumulators
//            through float64
//   metrics  float32 running sum and maximum over families, truncated to int32;
//            cm = f32(f64(largest - second largest) / votes); rank in float64
//
// Work per pattern: C(nb, 3) triplets x the listed rows (three LDS reads and three float64
// differences per lane and row), plus one lane's LDS updates per match (about 100 matches
// per pattern for 9 bands of an FCC scan).

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "ebsd_tripvote.h"

namespace {

constexpr int kWave = 64;
constexpr int kMaxBands = 64;
constexpr size_t kLdsBudget = 64 * 1024;   // library + per-wave state of up to 4 waves
constexpr double kRadeg = 180.0 / 3.14159265358979323846;
constexpr float kInvalidAngle = 10000.0f;

static thread_local std::string g_err;

static int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

// index network of the reference (tripletvote.py:1035-1041): srt such that a[srt] is
// ascending, the first place winning ties as in the reference
template <typename T>
__device__ __forceinline__ void sort3(const T a[3], int s[3])
{
    s[0] = 0; s[1] = 1; s[2] = 2;
    int t;
    if (a[s[0]] > a[s[2]]) { t = s[2]; s[2] = s[0]; s[0] = t; }
    if (a[s[0]] > a[s[1]]) { t = s[1]; s[1] = s[0]; s[0] = t; }
    if (a[s[1]] > a[s[2]]) { t = s[2]; s[2] = s[1]; s[1] = t; }
}

// LDS layout per work-group: the library as columns (a0, a1, a2 float64, then id0, id1, id2
// int32) when it fits, then per wave the nb x nb band angles and the two nfam x nb
// accumulators.
//
// Triplet records: lane t of a wave computes the record of triplet c*64 + t of the
// lexicographic order (i < j < k): the ascending angles x0 <= x1 <= x2, and packed into
// one word i, j, k (6 bits each), the vertex order u (2 bits each), the three
// near-equality flags and a valid bit.  The wave then walks the 64 records in order,
// reading each with v_readlane.
struct TripletRecord {
    float x0, x1, x2;
    uint32_t meta;
};

__device__ __forceinline__ TripletRecord make_record(const float *angs, int nb, int i, int j, int k, double tol,
                                                     bool skip_invalid)
{
    const float a3[3] = {angs[i * nb + j], angs[i * nb + k], angs[j * nb + k]};
    TripletRecord r;
    int s[3];
    sort3(a3, s);
    // the reference's LUT entry for s (the vertex order of the sorted triangle), then
    // its argsort
    const int lv[3] = {2 - s[2], 2 - s[1], 2 - s[0]};
    int u[3];
    sort3(lv, u);
    r.x0 = a3[s[0]];
    r.x1 = a3[s[1]];
    r.x2 = a3[s[2]];
    const bool e1 = fabsf(r.x0 - r.x1) < tol;
    const bool e2 = fabsf(r.x1 - r.x2) < tol;
    const bool e3 = fabsf(r.x2 - r.x0) < tol;
    // a triplet with an invalid band (angle 10000) can match no library row while the
    // tolerance is below 9000 degrees
    const bool valid = !(skip_invalid && (a3[0] == kInvalidAngle || a3[1] == kInvalidAngle || a3[2] == kInvalidAngle));
    r.meta = (uint32_t)i | ((uint32_t)j << 6) | ((uint32_t)k << 12) | ((uint32_t)u[0] << 18) |
             ((uint32_t)u[1] << 20) | ((uint32_t)u[2] << 22) | ((uint32_t)e1 << 24) | ((uint32_t)e2 << 25) |
             ((uint32_t)e3 << 26) | ((uint32_t)valid << 27);
    return r;
}

// a cell of the accumulators: the weighted vote and the count, side by side (one 8-byte
// LDS read and write per vote)
__device__ __forceinline__ void add_vote(float2 *cell, int idx, double w)
{
    float2 v = cell[idx];
    v.x = (float)((double)v.x + w);
    v.y = v.y + 1.0f;
    cell[idx] = v;
}

template <bool kLdsLib>
__global__ __launch_bounds__(256) void k_tripvote(const float *__restrict__ bn, const float *__restrict__ inten,
                                                  int64_t npats, int nb, const double *__restrict__ ta,
                                                  const int32_t *__restrict__ ti, int ntrip, int nfam, double tol,
                                                  const int32_t *__restrict__ bin_start,
                                                  const int32_t *__restrict__ bin_rows, int nbins, double bin_width,
                                                  float *__restrict__ accw, int32_t *__restrict__ fam,
                                                  float *__restrict__ rank, float *__restrict__ cmo,
                                                  float *__restrict__ acc)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int waves = blockDim.x / kWave;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    double *la0 = reinterpret_cast<double *>(lds_raw);
    double *la1 = la0 + (kLdsLib ? ntrip : 0);
    double *la2 = la1 + (kLdsLib ? ntrip : 0);
    int32_t *li = reinterpret_cast<int32_t *>(la2 + (kLdsLib ? ntrip : 0));
    float *wave_lds = reinterpret_cast<float *>(li + (kLdsLib ? 3 * ((ntrip + 3) & ~3) : 0));
    const int per = ((nb * nb + 1) & ~1) + 2 * nfam * nb;
    float *angs = wave_lds + wave * per;
    float2 *cell = reinterpret_cast<float2 *>(angs + ((nb * nb + 1) & ~1));
    const int64_t p = (int64_t)blockIdx.x * waves + wave;
    const bool live = p < npats;             // uniform over the wave

    if (kLdsLib) {
        for (int t = threadIdx.x; t < ntrip; t += blockDim.x) {
            la0[t] = ta[(int64_t)t * 3 + 0];
            la1[t] = ta[(int64_t)t * 3 + 1];
            la2[t] = ta[(int64_t)t * 3 + 2];
            li[t] = ti[(int64_t)t * 3 + 0];
            li[ntrip + t] = ti[(int64_t)t * 3 + 1];
            li[2 * ntrip + t] = ti[(int64_t)t * 3 + 2];
        }
    }
    if (live) {
        for (int t = lane; t < nfam * nb; t += kWave) cell[t] = make_float2(0.0f, 0.0f);
        const float *b = bn + p * nb * 3;
        const float *in = inten + p * nb;
        for (int t = lane; t < nb * nb; t += kWave) {
            const int r = t / nb, c = t - r * nb;
            float d = __fadd_rn(__fadd_rn(__fmul_rn(b[r * 3], b[c * 3]), __fmul_rn(b[r * 3 + 1], b[c * 3 + 1])),
                                __fmul_rn(b[r * 3 + 2], b[c * 3 + 2]));
            d = fabsf(d);
            if (d > 1.0f) d = 1.0f;
            const float a = (float)acos((double)d);
            float v = (float)((double)a * kRadeg);
            if ((double)in[r] < 1e-6 || (double)in[c] < 1e-6) v = kInvalidAngle;
            angs[t] = v;
        }
    }
    __syncthreads();
    if (!live) return;

    const bool skip_invalid = tol < 9000.0;
    const float inv_bin_width = bin_start ? (float)(1.0 / bin_width) : 0.0f;
    const int ntri = nb * (nb - 1) * (nb - 2) / 6;
    // lane's first triplet: the lexicographic combination of rank `lane`
    int ci = 0, cj = 1, ck = 2;
    {
        int rem = lane;
        while (ci < nb - 2) {
            const int cnt = (nb - 1 - ci) * (nb - 2 - ci) / 2;   // triplets starting at ci
            if (rem < cnt) break;
            rem -= cnt;
            ci++;
        }
        cj = ci + 1;
        while (cj < nb - 1) {
            const int cnt = nb - 1 - cj;                           // triplets starting at (ci, cj)
            if (rem < cnt) break;
            rem -= cnt;
            cj++;
        }
        ck = cj + 1 + rem;
    }
    for (int c0 = 0; c0 < ntri; c0 += kWave) {
        TripletRecord rec = {0.0f, 0.0f, 0.0f, 0u};
        if (c0 + lane < ntri) rec = make_record(angs, nb, ci, cj, ck, tol, skip_invalid);
        // advance the lane's combination by 64 ranks for the next chunk
        for (int step = 0; step < kWave && c0 + kWave + lane < ntri; step++) {
            if (++ck < nb) continue;
            if (++cj < nb - 1) {
                ck = cj + 1;
                continue;
            }
            ci++;
            cj = ci + 1;
            ck = cj + 1;
        }
        const int nchunk = min(kWave, ntri - c0);
        for (int t = 0; t < nchunk; t++) {
            const uint32_t meta = (uint32_t)__builtin_amdgcn_readlane((int)rec.meta, t);
            if (!(meta >> 27 & 1u)) continue;
            const float x0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rec.x0), t));
            const float x1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rec.x1), t));
            const float x2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rec.x2), t));
            const int i = meta & 63, j = meta >> 6 & 63, k = meta >> 12 & 63;
            const int u0 = meta >> 18 & 3, u1 = meta >> 20 & 3, u2 = meta >> 22 & 3;
            const bool e1 = meta >> 24 & 1u, e2 = meta >> 25 & 1u, e3 = meta >> 26 & 1u;
            const double dx0 = x0, dx1 = x1, dx2 = x2;
            // the rows to test: the whole library, or the index's list for the bin of x0
            // (ascending rows, a superset of those within tol of x0); NaN angles pass every
            // test of the reference, so they take the whole library
            int first = 0, count = ntrip;
            const bool indexed = bin_start != nullptr && x0 == x0 && x1 == x1 && x2 == x2;
            if (indexed) {
                // float32 product: within 2e-5 bins of the exact quotient, covered by the
                // index's 1e-3 degree slack at the bin edges
                int bin = (int)floorf(x0 * inv_bin_width);
                bin = bin < 0 ? 0 : (bin >= nbins ? nbins - 1 : bin);
                first = bin_start[bin];
                count = bin_start[bin + 1] - first;
            }
            for (int base = 0; base < count; base += kWave) {
                const int slot = base + lane;
                int q = slot;
                if (indexed && slot < count) q = bin_rows[first + slot];
                bool pass = false;
                float t1f = 0.0f, t2f = 0.0f, t3f = 0.0f;
                int fid0 = 0, fid1 = 0, fid2 = 0;
                if (slot < count) {
                    double r0, r1, r2;
                    if (kLdsLib) {
                        r0 = la0[q];
                        r1 = la1[q];
                        r2 = la2[q];
                    } else {
                        r0 = ta[(int64_t)q * 3 + 0];
                        r1 = ta[(int64_t)q * 3 + 1];
                        r2 = ta[(int64_t)q * 3 + 2];
                    }
                    const double t1 = fabs(r0 - dx0), t2 = fabs(r1 - dx1), t3 = fabs(r2 - dx2);
                    // the reference stops at the first failing test; outcome and kept
                    // values are the same
                    pass = !(t1 > tol) && !(t2 > tol) && !(t3 > tol);
                    t1f = (float)t1;
                    t2f = (float)t2;
                    t3f = (float)t3;
                    if (pass) {
                        // the family at each vertex of the sorted triangle
                        int id[3];
                        if (kLdsLib) {
                            id[0] = li[q];
                            id[1] = li[ntrip + q];
                            id[2] = li[2 * ntrip + q];
                        } else {
                            id[0] = ti[(int64_t)q * 3 + 0];
                            id[1] = ti[(int64_t)q * 3 + 1];
                            id[2] = ti[(int64_t)q * 3 + 2];
                        }
                        fid0 = id[u0];
                        fid1 = id[u1];
                        fid2 = id[u2];
                    }
                }
                uint64_t m = __ballot(pass);
                while (m) {
                    const int l = __builtin_ctzll(m);
                    m &= m - 1;
                    const float g0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t1f), l));
                    const float g1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t2f), l));
                    const float g2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t3f), l));
                    const int f0 = __builtin_amdgcn_readlane(fid0, l);
                    const int f1 = __builtin_amdgcn_readlane(fid1, l);
                    const int f2 = __builtin_amdgcn_readlane(fid2, l);
                    if (lane == 0) {
                        if ((unsigned)f0 < (unsigned)nfam && (unsigned)f1 < (unsigned)nfam &&
                            (unsigned)f2 < (unsigned)nfam) {
                            const double w1 = tol - 0.5 * (double)(g0 + g1);
                            const double w2 = tol - 0.5 * (double)(g0 + g2);
                            const double w3 = tol - 0.5 * (double)(g1 + g2);
                            add_vote(cell, f0 * nb + i, w1);
                            add_vote(cell, f1 * nb + j, w2);
                            add_vote(cell, f2 * nb + k, w3);
                            if (e1) {
                                add_vote(cell, f0 * nb + j, w1);
                                add_vote(cell, f1 * nb + i, w2);
                                add_vote(cell, f2 * nb + k, w3);
                            }
                            if (e2) {
                                add_vote(cell, f0 * nb + i, w1);
                                add_vote(cell, f1 * nb + k, w2);
                                add_vote(cell, f2 * nb + j, w3);
                            }
                            if (e3) {
                                add_vote(cell, f0 * nb + k, w1);
                                add_vote(cell, f1 * nb + j, w2);
                                add_vote(cell, f2 * nb + i, w3);
                            }
                            if (e1 && e2 && e3) {
                                // the reference adds these six weights before the six
                                // counts; the two arrays are separate, so per array the
                                // order is the same
                                add_vote(cell, f0 * nb + k, w1);
                                add_vote(cell, f1 * nb + i, w2);
                                add_vote(cell, f2 * nb + j, w3);
                                add_vote(cell, f0 * nb + j, w1);
                                add_vote(cell, f1 * nb + k, w2);
                                add_vote(cell, f2 * nb + i, w3);
                            }
                        }
                    }
                }
            }
        }
    }
    // lane 0's LDS writes become visible to the other lanes of the wave
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    if (lane < nb) {
        const int b = lane;
        float mx = -1.0e12f, sum = 0.0f;
        float top = -INFINITY, second = -INFINITY;
        int best = 0;
        for (int f = 0; f < nfam; f++) {
            const float v = cell[f * nb + b].x;
            sum += v;
            if (v > mx) {
                mx = v;
                best = f;
            }
            if (v > top) {
                second = top;
                top = v;
            } else if (v > second) {
                second = v;
            }
        }
        const int mxvote = (int)mx, votes = (int)sum;
        float c = 0.0f;
        if (votes >= 1) c = (float)((double)(top - second) / (double)votes);
        fam[p * nb + b] = best;
        cmo[p * nb + b] = c;
        rank[p * nb + b] = (float)((double)(nb - b) / (double)nb * (double)c * (double)mxvote);
    }
    float *ow = accw + p * nfam * nb, *oc = acc + p * nfam * nb;
    for (int t = lane; t < nfam * nb; t += kWave) {
        const float2 v = cell[t];
        ow[t] = v.x;
        oc[t] = v.y;
    }
}

// Several patterns per wavefront, one per group of G lanes (G = 32 or 16, nbands <= G): the
// same walk as k_tripvote, but every instruction of the triplet walk, the row tests and the
// votes serves 64 / G patterns.  Values that k_tripvote reads with v_readlane (uniform over
// the wave) are per group here and move with ds_bpermute (__shfl).  Loops run to the
// largest of the groups' bounds, the other groups masked; a group with no pattern has no
// valid triplet.
template <int G, bool kLdsLib>
__global__ __launch_bounds__(256) void k_tripvote_group(const float *__restrict__ bn, const float *__restrict__ inten,
                                                       int64_t npats, int nb, const double *__restrict__ ta,
                                                       const int32_t *__restrict__ ti, int ntrip, int nfam,
                                                       double tol, const int32_t *__restrict__ bin_start,
                                                       const int32_t *__restrict__ bin_rows, int nbins,
                                                       double bin_width, float *__restrict__ accw,
                                                       int32_t *__restrict__ fam, float *__restrict__ rank,
                                                       float *__restrict__ cmo, float *__restrict__ acc)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    constexpr int kGroups = kWave / G;           // patterns per wave
    constexpr uint64_t kGroupMask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    const int halves = blockDim.x / G;
    const int half = threadIdx.x / G, hl = threadIdx.x % G;
    const int lane = threadIdx.x % kWave;
    const int hbase = lane - hl;                 // first lane of this group in the wave
    double *la0 = reinterpret_cast<double *>(lds_raw);
    double *la1 = la0 + (kLdsLib ? ntrip : 0);
    double *la2 = la1 + (kLdsLib ? ntrip : 0);
    int32_t *li = reinterpret_cast<int32_t *>(la2 + (kLdsLib ? ntrip : 0));
    float *half_lds = reinterpret_cast<float *>(li + (kLdsLib ? 3 * ((ntrip + 3) & ~3) : 0));
    const int per = ((nb * nb + 1) & ~1) + 2 * nfam * nb;
    float *angs = half_lds + half * per;
    float2 *cell = reinterpret_cast<float2 *>(angs + ((nb * nb + 1) & ~1));
    const int64_t p = (int64_t)blockIdx.x * halves + half;
    const bool live = p < npats;                 // uniform over the half

    if (kLdsLib) {
        for (int t = threadIdx.x; t < ntrip; t += blockDim.x) {
            la0[t] = ta[(int64_t)t * 3 + 0];
            la1[t] = ta[(int64_t)t * 3 + 1];
            la2[t] = ta[(int64_t)t * 3 + 2];
            li[t] = ti[(int64_t)t * 3 + 0];
            li[ntrip + t] = ti[(int64_t)t * 3 + 1];
            li[2 * ntrip + t] = ti[(int64_t)t * 3 + 2];
        }
    }
    if (live) {
        for (int t = hl; t < nfam * nb; t += G) cell[t] = make_float2(0.0f, 0.0f);
        const float *b = bn + p * nb * 3;
        const float *in = inten + p * nb;
        for (int t = hl; t < nb * nb; t += G) {
            const int r = t / nb, c = t - r * nb;
            float d = __fadd_rn(__fadd_rn(__fmul_rn(b[r * 3], b[c * 3]), __fmul_rn(b[r * 3 + 1], b[c * 3 + 1])),
                                __fmul_rn(b[r * 3 + 2], b[c * 3 + 2]));
            d = fabsf(d);
            if (d > 1.0f) d = 1.0f;
            const float a = (float)acos((double)d);
            float v = (float)((double)a * kRadeg);
            if ((double)in[r] < 1e-6 || (double)in[c] < 1e-6) v = kInvalidAngle;
            angs[t] = v;
        }
    }
    __syncthreads();
    // the groups of a wave share every loop below (shuffles and ballots need all of them):
    // only a wave with no pattern at all leaves
    if (__builtin_amdgcn_readfirstlane((int)(blockIdx.x * (int64_t)halves + (threadIdx.x / kWave) * kGroups >= npats)))
        return;

    const bool skip_invalid = tol < 9000.0;
    const float inv_bin_width = bin_start ? (float)(1.0 / bin_width) : 0.0f;
    const int ntri = nb * (nb - 1) * (nb - 2) / 6;
    // the half-lane's first triplet: the lexicographic combination of rank hl
    int ci = 0, cj = 1, ck = 2;
    {
        int rem = hl;
        while (ci < nb - 2) {
            const int cnt = (nb - 1 - ci) * (nb - 2 - ci) / 2;
            if (rem < cnt) break;
            rem -= cnt;
            ci++;
        }
        cj = ci + 1;
        while (cj < nb - 1) {
            const int cnt = nb - 1 - cj;
            if (rem < cnt) break;
            rem -= cnt;
            cj++;
        }
        ck = cj + 1 + rem;
    }
    for (int c0 = 0; c0 < ntri; c0 += G) {
        TripletRecord rec = {0.0f, 0.0f, 0.0f, 0u};
        if (live && c0 + hl < ntri) rec = make_record(angs, nb, ci, cj, ck, tol, skip_invalid);
        for (int step = 0; step < G && c0 + G + hl < ntri; step++) {
            if (++ck < nb) continue;
            if (++cj < nb - 1) {
                ck = cj + 1;
                continue;
            }
            ci++;
            cj = ci + 1;
            ck = cj + 1;
        }
        const int nchunk = min(G, ntri - c0);
        for (int t = 0; t < nchunk; t++) {
            const int src = hbase + t;
            const uint32_t meta = (uint32_t)__shfl((int)rec.meta, src);
            const float x0 = __shfl(rec.x0, src), x1 = __shfl(rec.x1, src), x2 = __shfl(rec.x2, src);
            const bool valid = meta >> 27 & 1u;
            const int i = meta & 63, j = meta >> 6 & 63, k = meta >> 12 & 63;
            const int u0 = meta >> 18 & 3, u1 = meta >> 20 & 3, u2 = meta >> 22 & 3;
            const bool e1 = meta >> 24 & 1u, e2 = meta >> 25 & 1u, e3 = meta >> 26 & 1u;
            const double dx0 = x0, dx1 = x1, dx2 = x2;
            int first = 0, count = valid ? ntrip : 0;
            const bool indexed = valid && bin_start != nullptr && x0 == x0 && x1 == x1 && x2 == x2;
            if (indexed) {
                // float32 product: within 2e-5 bins of the exact quotient, covered by the
                // index's 1e-3 degree slack at the bin edges
                int bin = (int)floorf(x0 * inv_bin_width);
                bin = bin < 0 ? 0 : (bin >= nbins ? nbins - 1 : bin);
                first = bin_start[bin];
                count = bin_start[bin + 1] - first;
            }
            int cmax = 0;
#pragma unroll
            for (int g = 0; g < kGroups; g++) cmax = max(cmax, __builtin_amdgcn_readlane(count, g * G));
            for (int base = 0; base < cmax; base += G) {
                const int slot = base + hl;
                int q = slot;
                if (indexed && slot < count) q = bin_rows[first + slot];
                bool pass = false;
                float t1f = 0.0f, t2f = 0.0f, t3f = 0.0f;
                int fid0 = 0, fid1 = 0, fid2 = 0;
                if (slot < count) {
                    double r0, r1, r2;
                    if (kLdsLib) {
                        r0 = la0[q];
                        r1 = la1[q];
                        r2 = la2[q];
                    } else {
                        r0 = ta[(int64_t)q * 3 + 0];
                        r1 = ta[(int64_t)q * 3 + 1];
                        r2 = ta[(int64_t)q * 3 + 2];
                    }
                    const double t1 = fabs(r0 - dx0), t2 = fabs(r1 - dx1), t3 = fabs(r2 - dx2);
                    pass = !(t1 > tol) && !(t2 > tol) && !(t3 > tol);
                    t1f = (float)t1;
                    t2f = (float)t2;
                    t3f = (float)t3;
                    if (pass) {
                        int id[3];
                        if (kLdsLib) {
                            id[0] = li[q];
                            id[1] = li[ntrip + q];
                            id[2] = li[2 * ntrip + q];
                        } else {
                            id[0] = ti[(int64_t)q * 3 + 0];
                            id[1] = ti[(int64_t)q * 3 + 1];
                            id[2] = ti[(int64_t)q * 3 + 2];
                        }
                        fid0 = id[u0];
                        fid1 = id[u1];
                        fid2 = id[u2];
                    }
                }
                // each group takes its own passing rows in ascending order, all groups at once
                uint64_t rem = __ballot(pass);
                while (rem) {
                    const uint64_t mine = (rem >> hbase) & kGroupMask;
                    const int from = hbase + (mine ? __builtin_ctzll(mine) : 0);
                    const float g0 = __shfl(t1f, from), g1 = __shfl(t2f, from), g2 = __shfl(t3f, from);
                    const int f0 = __shfl(fid0, from), f1 = __shfl(fid1, from), f2 = __shfl(fid2, from);
                    // each block of the reference's vote statements gives every band of the
                    // triplet exactly one vote: lanes 0, 1, 2 of the group take bands i, j, k
                    // and apply their band's votes in the block order, so each cell still
                    // receives the reference's additions in the reference's order
                    if (hl < 3 && mine != 0ull && (unsigned)f0 < (unsigned)nfam && (unsigned)f1 < (unsigned)nfam &&
                        (unsigned)f2 < (unsigned)nfam) {
                        const double w1 = tol - 0.5 * (double)(g0 + g1);
                        const double w2 = tol - 0.5 * (double)(g0 + g2);
                        const double w3 = tol - 0.5 * (double)(g1 + g2);
                        const int band = hl == 0 ? i : (hl == 1 ? j : k);
                        // block: votes (family index 0..2 into f, weight index 0..2 into w)
                        // for bands (i, j, k); main (0,1,2), e1 (1,0,2), e2 (0,2,1),
                        // e3 (2,1,0), all three (1,2,0) then (2,0,1)
                        const int fv[3] = {f0, f1, f2};
                        const double wv[3] = {w1, w2, w3};
                        add_vote(cell, fv[hl] * nb + band, wv[hl]);
                        if (e1) {
                            const int x = hl == 0 ? 1 : (hl == 1 ? 0 : 2);
                            add_vote(cell, fv[x] * nb + band, wv[x]);
                        }
                        if (e2) {
                            const int x = hl == 0 ? 0 : (hl == 1 ? 2 : 1);
                            add_vote(cell, fv[x] * nb + band, wv[x]);
                        }
                        if (e3) {
                            const int x = hl == 0 ? 2 : (hl == 1 ? 1 : 0);
                            add_vote(cell, fv[x] * nb + band, wv[x]);
                        }
                        if (e1 && e2 && e3) {
                            const int x = hl == 0 ? 1 : (hl == 1 ? 2 : 0);
                            add_vote(cell, fv[x] * nb + band, wv[x]);
                            const int y = hl == 0 ? 2 : (hl == 1 ? 0 : 1);
                            add_vote(cell, fv[y] * nb + band, wv[y]);
                        }
                    }
                    // clear the lowest set bit of every group
                    uint64_t next = 0;
#pragma unroll
                    for (int g = 0; g < kGroups; g++) {
                        const uint64_t x = (rem >> (g * G)) & kGroupMask;
                        next |= (x & (x - 1ull)) << (g * G);
                    }
                    rem = next;
                }
            }
        }
    }
    // the group's lane 0 wrote the cells; the other lanes of the wave read them
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (!live) return;

    if (hl < nb) {
        const int b = hl;
        float mx = -1.0e12f, sum = 0.0f;
        float top = -INFINITY, second = -INFINITY;
        int best = 0;
        for (int f = 0; f < nfam; f++) {
            const float v = cell[f * nb + b].x;
            sum += v;
            if (v > mx) {
                mx = v;
                best = f;
            }
            if (v > top) {
                second = top;
                top = v;
            } else if (v > second) {
                second = v;
            }
        }
        const int mxvote = (int)mx, votes = (int)sum;
        float c = 0.0f;
        if (votes >= 1) c = (float)((double)(top - second) / (double)votes);
        fam[p * nb + b] = best;
        cmo[p * nb + b] = c;
        rank[p * nb + b] = (float)((double)(nb - b) / (double)nb * (double)c * (double)mxvote);
    }
    float *ow = accw + p * nfam * nb, *oc = acc + p * nfam * nb;
    for (int t = hl; t < nfam * nb; t += G) {
        const float2 v = cell[t];
        ow[t] = v.x;
        oc[t] = v.y;
    }
}

// ---------------------------------------------------------------------------------------
// k_bandindex: the steps of BandIndexer.bandindex after the vote (tripletvote.py:526-633) --
// band assignment (_assign_bands_nb2, :1483-1688), the QUEST weights (:855-897), the QUEST
// fit (:920-1018) and the confidence metric (:2056-2075).  One wavefront per pattern, lane
// j holding band j (nbands <= 64); the (ii, jj) pair loop and each pair's candidate loop are
// uniform over the wavefront, and every sum the reference takes over bands runs in band
// order through v_readlane, so the float results follow the reference's operation order.
// Arithmetic types follow Numba's typing: band normals, library poles, TRIAD matrices and
// the pole-fit products float32 (products not contracted); p2 / (|p2| + 1e-35) and the cross
// products with it float64; arccos of a float32 as the float64 arccos rounded to float32;
// the weights and QUEST float64.
// ---------------------------------------------------------------------------------------
struct BandIndexArgs {
    const float *bn, *inten, *accw, *rank, *acc;
    const int32_t *fam;
    int64_t npats;
    int nb, nfam, npl;
    const float *pflt;     // (npl, 3) unit Cartesian library poles, float32
    const double *pc64;    // the same, float64 (QUEST)
    const int32_t *famid;  // (npl) family of each pole
    const double *angrows; // (nfam, npl) degrees between family f's first pole and every pole
    const int32_t *famidx, *famcnt; // (nfam) first pole and number of poles of each family
    double angtol;
    int n_band_early, nfit, simple;
    double *quat, *fit, *bandfit;
    float *cm, *afit, *fitb;
    int64_t *pm, *nmatch, *ij;
    int32_t *acc_correct;
    uint8_t *pv;
};

__device__ __forceinline__ float rl_f(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ __forceinline__ int rl_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ double rl_d(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float acosf_cr(float x) { return (float)acos((double)x); }
__device__ __forceinline__ void cross3(const double *a, const double *b, double *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ int first_lane(uint64_t m) { return __ffsll((unsigned long long)m) - 1; }

__global__ __launch_bounds__(256) void k_bandindex(BandIndexArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int npl = a.npl, nfam = a.nfam, nb = a.nb;
    double *l_pc64 = reinterpret_cast<double *>(lds_raw);
    double *l_ang = l_pc64 + 3 * npl;
    float *l_pf = reinterpret_cast<float *>(l_ang + (size_t)nfam * npl);
    int32_t *l_famid = reinterpret_cast<int32_t *>(l_pf + 3 * npl);
    int32_t *l_fidx = l_famid + npl;
    int32_t *l_fcnt = l_fidx + nfam;
    for (int t = threadIdx.x; t < 3 * npl; t += blockDim.x) {
        l_pc64[t] = a.pc64[t];
        l_pf[t] = a.pflt[t];
    }
    for (int t = threadIdx.x; t < nfam * npl; t += blockDim.x) l_ang[t] = a.angrows[t];
    for (int t = threadIdx.x; t < npl; t += blockDim.x) l_famid[t] = a.famid[t];
    for (int t = threadIdx.x; t < nfam; t += blockDim.x) {
        l_fidx[t] = a.famidx[t];
        l_fcnt[t] = a.famcnt[t];
    }
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t p = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
    if (p >= a.npats) return;
    const bool act = lane < nb;
    const int j = act ? lane : 0;
    const int G = nb <= 16 ? 4 : (nb <= 32 ? 2 : 1); // lanes per band in the nearest-pole search
    const float *bnp = a.bn + (p * nb + j) * 3;
    const float n0 = bnp[0], n1 = bnp[1], n2 = bnp[2];
    const int fam_j = a.fam[p * nb + j];
    const float rank_j = act ? a.rank[p * nb + j] : 0.0f;
    const float inten_j = a.inten[p * nb + j];
    // stable ascending argsort of the ranks: band j sits at place pos
    int pos = 0;
    for (int i = 0; i < nb; ++i) {
        const float ri = rl_f(rank_j, i);
        pos += (ri < rank_j) || (ri == rank_j && i < lane);
    }
    const float eps = 1.0e-12f;
    // the pattern's best assignment so far
    float best_fit = 360.0f, best_fitb = 360.0f;
    int best_n = 0, best_pm = -1000, best_ij0 = -1, best_ij1 = -1, best_ij2 = -1, best_ij3 = -1;
    bool best_pv = false;
    for (int ii = 0; ii < nb - 1; ++ii) {
        for (int jj = ii + 1; jj < nb; ++jj) {
            const int bnd1 = first_lane(__ballot(act && pos == nb - 1 - ii));
            const int bnd2 = first_lane(__ballot(act && pos == nb - 1 - jj));
            const float v1[3] = {rl_f(n0, bnd1), rl_f(n1, bnd1), rl_f(n2, bnd1)};
            const float v2[3] = {rl_f(n0, bnd2), rl_f(n1, bnd2), rl_f(n2, bnd2)};
            const int f1 = rl_i(fam_j, bnd1), f2 = rl_i(fam_j, bnd2);
            float ang01 = (v1[0] * v2[0] + v1[1] * v2[1]) + v1[2] * v2[2];
            if (ang01 > 1.0f) ang01 = (float)(1.0 - (double)eps);
            if (ang01 < -1.0f) ang01 = (float)(-1.0 + (double)eps);
            // this pair's best candidate (the reference's zero-initialised arrays)
            float pf_fit = 0.0f;
            int pf_pm = -1;
            bool pf_pv = false;
            int nGood = 0;
            const double ptest = (double)acosf_cr(fabsf(ang01)) * kRadeg;
            if (ptest > a.angtol) {
                const double a01 = (double)acosf_cr(ang01) * kRadeg;
                const int r1 = l_fidx[f1], base = l_fidx[f2], cnt = l_fcnt[f2];
                const float *p1 = l_pf + 3 * r1;
                float c12[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2],
                                v1[0] * v2[1] - v1[1] * v2[0]};
                const float nn = sqrtf((c12[0] * c12[0] + c12[1] * c12[1]) + c12[2] * c12[2]);
                for (int m = 0; m < 3; ++m) c12[m] /= nn;
                const float b2[3] = {v1[1] * c12[2] - v1[2] * c12[1], v1[2] * c12[0] - v1[0] * c12[2],
                                     v1[0] * c12[1] - v1[1] * c12[0]};
                const float B[9] = {v1[0], v1[1], v1[2], c12[0], c12[1], c12[2], b2[0], b2[1], b2[2]};
                const double p1d[3] = {p1[0], p1[1], p1[2]};
                double score = -1.0;
                bool any = false;
                for (int c0 = 0; c0 < cnt; c0 += kWave) {
                    const int kk = base + c0 + lane;
                    uint64_t hits = __ballot(c0 + lane < cnt && fabs(l_ang[(size_t)f1 * npl + kk] - a01) < a.angtol);
                    while (hits) {
                        const int t = first_lane(hits);
                        hits &= hits - 1;
                        any = true;
                        const float *p2f = l_pf + 3 * (base + c0 + t);
                        double nt = (double)sqrtf((p2f[0] * p2f[0] + p2f[1] * p2f[1]) + p2f[2] * p2f[2]) + 1.0e-35;
                        double p2[3], pc[3], pc2[3];
                        for (int m = 0; m < 3; ++m) p2[m] = (double)p2f[m] / nt;
                        cross3(p1d, p2, pc);
                        nt = sqrt((pc[0] * pc[0] + pc[1] * pc[1]) + pc[2] * pc[2]) + 1.0e-35;
                        for (int m = 0; m < 3; ++m) pc[m] /= nt;
                        cross3(p1d, pc, pc2);
                        float A[9];
                        for (int m = 0; m < 3; ++m) {
                            A[m * 3 + 0] = p1[m];
                            A[m * 3 + 1] = (float)pc[m];
                            A[m * 3 + 2] = (float)pc2[m];
                        }
                        float AB[9];
                        for (int r = 0; r < 3; ++r)
                            for (int q = 0; q < 3; ++q)
                                AB[r * 3 + q] = (A[r * 3 + 0] * B[0 * 3 + q] + A[r * 3 + 1] * B[1 * 3 + q]) +
                                                A[r * 3 + 2] * B[2 * 3 + q];
                        float tp[3];
                        for (int r = 0; r < 3; ++r) tp[r] = (AB[r * 3 + 0] * n0 + AB[r * 3 + 1] * n1) + AB[r * 3 + 2] * n2;
                        // the nearest library pole of every band, G lanes per band: lane L scans
                        // poles L % G, L % G + G, ... for band L / G, and the G lanes keep the
                        // larger value, or on equal (or unordered) values the smaller pole --
                        // the first index of the maximum, as the reference's scan with strict >
                        const int bsrc = lane / G;
                        const float s0 = __shfl(tp[0], bsrc), s1 = __shfl(tp[1], bsrc), s2 = __shfl(tp[2], bsrc);
                        float gmx = 0.0f;
                        int gam = -1;
                        for (int k = lane % G; k < npl; k += G) {
                            const float *pk = l_pf + 3 * k;
                            const float tv = (pk[0] * s0 + pk[1] * s1) + pk[2] * s2;
                            if (gam < 0 || tv > gmx) { gmx = tv; gam = k; }
                        }
                        for (int off = 1; off < G; off <<= 1) {
                            const float omx = __shfl_xor(gmx, off);
                            const int oam = __shfl_xor(gam, off);
                            if (oam >= 0 && (gam < 0 || omx > gmx || (!(gmx > omx) && oam < gam))) {
                                gmx = omx;
                                gam = oam;
                            }
                        }
                        const int gsrc = (lane < kWave / G ? lane : 0) * G;
                        float mx = __shfl(gmx, gsrc);
                        const int am = __shfl(gam, gsrc);
                        if (mx < -1.0f) mx = -1.0f;
                        if (mx > 1.0f) mx = 1.0f;
                        const double angTry = (double)acosf_cr(mx) * kRadeg;
                        const bool match = act && angTry < a.angtol;
                        const uint64_t mm = __ballot(match);
                        double sum = 0.0;
                        for (int i = 0; i < nb; ++i) {
                            const double ai = rl_d(angTry, i);
                            if ((mm >> i) & 1ull) sum += ai;
                        }
                        const int nmatch = __popcll((unsigned long long)mm);
                        const double scoreTry = (double)(float)nmatch / (sum / (double)nmatch + 1e-6);
                        if (scoreTry > score) {
                            score = scoreTry;
                            pf_fit = (float)angTry;
                            pf_pm = am;
                            pf_pv = angTry < a.angtol;
                        }
                    }
                }
                if (any) nGood = __popcll((unsigned long long)__ballot(act && (double)pf_fit < a.angtol));
            }
            if (nGood < 3) continue;
            const uint64_t gm = __ballot(act && (double)pf_fit < a.angtol);
            float fit = 0.0f;
            for (int i = 0; i < nb; ++i) {
                const float fi = rl_f(pf_fit, i);
                if ((gm >> i) & 1ull) fit += fi;
            }
            fit /= (float)nGood;
            if (best_n < nGood || (best_n == nGood && best_fit > fit)) {
                best_fit = fit;
                best_fitb = pf_fit;
                best_n = nGood;
                best_pm = pf_pm;
                best_pv = pf_pv;
                best_ij0 = ii; best_ij1 = jj; best_ij2 = bnd1; best_ij3 = bnd2;
            }
            if (best_n >= a.n_band_early) break;
        }
        if (best_n >= a.n_band_early) break;
    }
    const int pmc = best_pm < 0 ? 0 : best_pm;
    const bool valid = act && best_pv;
    const uint64_t vmask = __ballot(valid);
    const int nvalid = __popcll((unsigned long long)vmask);
    const int corr = __popcll((unsigned long long)__ballot(valid && l_famid[pmc] == fam_j));
    // QUEST weights: the nfit best-scoring matched bands keep their intensity, normalised
    float w32 = 0.0f;
    const float *accw = a.accw + p * (int64_t)nfam * nb;
    const float *accn = a.acc + p * (int64_t)nfam * nb;
    if (nvalid >= 3) {
        float score = -1.0f;
        if (valid) {
            if (!a.simple) {
                const int f = l_famid[pmc];
                score = (float)((double)accw[f * nb + j] / fmax((double)accn[f * nb + j], 1.0e-12));
            } else {
                // the reference scores matched band whg with the intensity of band q, its
                // place among the matched bands
                const int q = __popcll((unsigned long long)(vmask & ((1ull << lane) - 1ull)));
                score = a.inten[p * nb + q];
            }
        }
        int spos = 0;
        for (int i = 0; i < nb; ++i) {
            const float si = rl_f(score, i);
            spos += (si < score) || (si == score && i < lane);
        }
        const int kk = nvalid < a.nfit ? nvalid : a.nfit;
        const double w = (act && spos >= nb - kk) ? (double)inten_j : 0.0;
        double ws = 0.0;
        for (int i = 0; i < nb; ++i) ws += rl_d(w, i);
        w32 = act ? (float)(w / ws) : 0.0f;
    }
    // QUEST on the normalised weights (float64).  The reference normalises them with numpy's
    // np.sum over the bands, whose contiguous reduction sums 8 or more terms pairwise: eight
    // running sums over blocks of eight, combined as a tree, then the remainder in order
    double wsum = 0.0;
    if (nb < 8) {
        for (int i = 0; i < nb; ++i) wsum += (double)rl_f(w32, i);
    } else {
        double r8[8];
        for (int k = 0; k < 8; ++k) r8[k] = (double)rl_f(w32, k);
        int i = 8;
        for (; i < nb - (nb % 8); i += 8)
            for (int k = 0; k < 8; ++k) r8[k] += (double)rl_f(w32, i + k);
        wsum = ((r8[0] + r8[1]) + (r8[2] + r8[3])) + ((r8[4] + r8[5]) + (r8[6] + r8[7]));
        for (; i < nb; ++i) wsum += (double)rl_f(w32, i);
    }
    const double wnorm = (double)w32 / fmax(wsum, 1.0e-12);
    const uint64_t qm = __ballot(act && wnorm > 1.0e-7);
    const int nq = __popcll((unsigned long long)qm);
    double q[4] = {1.0, 0.0, 0.0, 0.0};
    double lam = 3.14159265358979323846, mad = -1.0, bdot = -1.0;
    if (nq >= 3) {
        double ws = 0.0;
        for (int i = 0; i < nb; ++i)
            if ((qm >> i) & 1ull) ws += rl_d(wnorm, i);
        const double wn = wnorm / ws;
        const double bd[3] = {(double)n0, (double)n1, (double)n2};
        const double pf[3] = {l_pc64[3 * pmc], l_pc64[3 * pmc + 1], l_pc64[3 * pmc + 2]};
        double B[9], z[3] = {0.0, 0.0, 0.0}, S[9], S2[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                const double term = (wn * bd[r]) * pf[c];
                double s = 0.0;
                for (int i = 0; i < nb; ++i)
                    if ((qm >> i) & 1ull) s += rl_d(term, i);
                B[r * 3 + c] = s;
            }
        double cr[3];
        cross3(bd, pf, cr);
        for (int m = 0; m < 3; ++m) {
            const double term = wn * cr[m];
            for (int i = 0; i < nb; ++i)
                if ((qm >> i) & 1ull) z[m] += rl_d(term, i);
        }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) S[r * 3 + c] = B[r * 3 + c] + B[c * 3 + r];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                S2[r * 3 + c] = (S[r * 3 + 0] * S[0 * 3 + c] + S[r * 3 + 1] * S[1 * 3 + c]) + S[r * 3 + 2] * S[2 * 3 + c];
        const double det = S[0] * (S[4] * S[8] - S[5] * S[7]) - S[1] * (S[3] * S[8] - S[5] * S[6]) +
                           S[2] * (S[3] * S[7] - S[4] * S[6]);
        const double kq = (S[4] * S[8] - S[5] * S[7]) + (S[0] * S[8] - S[2] * S[6]) + (S[0] * S[4] - S[3] * S[1]);
        const double sig = 0.5 * ((S[0] + S[4]) + S[8]);
        const double sig2 = sig * sig;
        double t[3];
        for (int c = 0; c < 3; ++c) t[c] = (z[0] * S2[0 * 3 + c] + z[1] * S2[1 * 3 + c]) + z[2] * S2[2 * 3 + c];
        const double d = (t[0] * z[0] + t[1] * z[1]) + t[2] * z[2];
        for (int c = 0; c < 3; ++c) t[c] = (z[0] * S[0 * 3 + c] + z[1] * S[1 * 3 + c]) + z[2] * S[2 * 3 + c];
        const double cc = det + ((t[0] * z[0] + t[1] * z[1]) + t[2] * z[2]);
        const double bq = sig2 + ((z[0] * z[0] + z[1] * z[1]) + z[2] * z[2]);
        const double aq = sig2 - kq;
        double dl = 1e6;
        int iter = 0;
        lam = 1.0;
        while (dl > 1.0e-12 && iter < 10) {
            const double l0 = lam, l2 = lam * lam;
            lam = lam - (l2 * l2 - (aq + bq) * l2 - cc * lam + (aq * bq + cc * sig - d)) /
                            (4.0 * (l2 * lam) - 2.0 * (aq + bq) * lam - cc);
            dl = fabs(l0 - lam);
            ++iter;
        }
        const double beta = lam - sig, alpha = lam * lam - sig2 + kq, gamma = (lam + sig) * alpha - det;
        double X[3];
        for (int r = 0; r < 3; ++r) {
            double M0 = ((r == 0 ? alpha : 0.0) + beta * S[r * 3 + 0]) + S2[r * 3 + 0];
            double M1 = ((r == 1 ? alpha : 0.0) + beta * S[r * 3 + 1]) + S2[r * 3 + 1];
            double M2 = ((r == 2 ? alpha : 0.0) + beta * S[r * 3 + 2]) + S2[r * 3 + 2];
            X[r] = (M0 * z[0] + M1 * z[1]) + M2 * z[2];
        }
        double qn = 0.0;
        qn += ((gamma * gamma + X[0] * X[0]) + X[1] * X[1]) + X[2] * X[2];
        qn = sqrt(qn);
        q[0] = gamma / qn; q[1] = X[0] / qn; q[2] = X[1] / qn; q[3] = X[2] / qn;
        if (gamma < 0.0)
            for (int m = 0; m < 4; ++m) q[m] = -q[m];
        // band normals rotated by q (convention p = +1), dotted with their poles
        double vt[3], vt2[3], o[3];
        vt[0] = q[2] * bd[2] - q[3] * bd[1];
        vt[1] = q[3] * bd[0] - q[1] * bd[2];
        vt[2] = q[1] * bd[1] - q[2] * bd[0];
        for (int m = 0; m < 3; ++m) vt[m] += q[0] * bd[m];
        vt2[0] = q[2] * vt[2] - q[3] * vt[1];
        vt2[1] = q[3] * vt[0] - q[1] * vt[2];
        vt2[2] = q[1] * vt[1] - q[2] * vt[0];
        for (int m = 0; m < 3; ++m) o[m] = 2.0 * vt2[m] + bd[m];
        const double bdj = (o[0] * pf[0] + o[1] * pf[1]) + o[2] * pf[2];
        const bool inq = act && ((qm >> lane) & 1ull);
        if (inq) bdot = bdj;
        double msum = 0.0;
        for (int i = 0; i < nb; ++i)
            if ((qm >> i) & 1ull) msum += rl_d(bdj, i);
        mad = msum / (double)nq;
    }
    // the confidence metric: matched votes over all votes, each cell at least 1 (float32)
    float cm2 = -1.0f;
    if (nvalid >= 3) {
        const float mv = valid ? accw[l_famid[pmc] * nb + j] : 0.0f;
        float cs = 0.0f;
        for (int i = 0; i < nb; ++i) {
            const float vi = rl_f(mv, i);
            if ((vmask >> i) & 1ull) cs += vi;
        }
        float den = 0.0f;
        for (int t = 0; t < nfam * nb; ++t) {
            const float v = accw[t];
            den += v < 1.0f ? 1.0f : v;
        }
        cm2 = cs / den;
    }
    if (act) {
        a.pm[p * nb + j] = best_pm;
        a.bandfit[p * nb + j] = acos(fmin(fmax(bdot, -1.0), 1.0)) * kRadeg;
        if (a.pv) a.pv[p * nb + j] = (uint8_t)best_pv;
        if (a.fitb) a.fitb[p * nb + j] = best_fitb;
    }
    if (lane < 4) a.quat[p * 4 + lane] = q[lane == 0 ? 0 : lane == 1 ? 1 : lane == 2 ? 2 : 3];
    if (lane == 0) {
        a.fit[p] = acos(fmin(fmax(mad, -1.0), 1.0)) * kRadeg;
        a.cm[p] = cm2;
        a.nmatch[p] = best_n;
        a.ij[p * 4 + 0] = best_ij0;
        a.ij[p * 4 + 1] = best_ij1;
        a.ij[p * 4 + 2] = best_ij2;
        a.ij[p * 4 + 3] = best_ij3;
        a.acc_correct[p] = corr;
        if (a.afit) a.afit[p] = best_fit;
    }
}

}  // namespace

extern "C" const char *ebsd_tripvote_last_error(void) { return g_err.c_str(); }

extern "C" int ebsd_tripvote(const float *d_bandnorms, const float *d_intensity, int64_t npats, int nbands,
                             const double *d_tripang, const int32_t *d_tripid, int ntrip, int nfam, double angtol,
                             const int32_t *d_bin_start, const int32_t *d_bin_rows, int nbins, double bin_width,
                             float *d_accw, int32_t *d_fam, float *d_rank, float *d_cm, float *d_acc, void *stream)
{
    if (npats < 0) return fail(-1, "npats must be >= 0");
    if (nbands < 1 || nbands > kMaxBands) return fail(-1, "nbands must be in [1, 64]");
    if (nfam < 2) return fail(-1, "nfam must be >= 2 (the confidence metric takes the second-best family)");
    if (ntrip < 0) return fail(-1, "ntrip must be >= 0");
    if (d_bin_start && (nbins < 1 || !(bin_width > 0.0) || !d_bin_rows))
        return fail(-1, "a row index needs nbins >= 1, bin_width > 0 and its row list");
    const size_t per_bytes = ((size_t)((nbands * nbands + 1) & ~1) + 2 * (size_t)nfam * nbands) * sizeof(float);
    const size_t lib_bytes = (size_t)ntrip * 3 * sizeof(double) + (size_t)((ntrip + 3) & ~3) * 3 * sizeof(int32_t);
    if (per_bytes > kLdsBudget) return fail(-1, "nbands^2 + 2 nfam nbands floats exceed the 64 KiB of LDS per pattern");
    if (npats == 0) return 0;
    if (!d_bandnorms || !d_intensity || !d_accw || !d_fam || !d_rank || !d_cm || !d_acc ||
        (ntrip > 0 && (!d_tripang || !d_tripid)))
        return fail(-1, "null device pointer");
    if (nbands <= 32) {
        // several patterns per wave: 16-lane groups up to 16 bands, else 32-lane halves;
        // up to 256 lanes per work-group as the LDS allows
        const int g = nbands <= 16 ? 16 : 32;
        int groups_per_wg = (int)(kLdsBudget / per_bytes);
        const int per_wave = kWave / g;
        groups_per_wg = groups_per_wg > 256 / g ? 256 / g : groups_per_wg - groups_per_wg % per_wave;
        if (groups_per_wg >= per_wave) {
            const bool lib_in = lib_bytes + groups_per_wg * per_bytes <= kLdsBudget;
            const size_t lds = (lib_in ? lib_bytes : 0) + groups_per_wg * per_bytes;
            const int64_t grid = (npats + groups_per_wg - 1) / groups_per_wg;
            if (grid > 0x7fffffff) return fail(-1, "too many patterns for one launch");
            const dim3 gd((unsigned)grid), bd(groups_per_wg * g);
            hipStream_t st = (hipStream_t)stream;
#define EBSD_TV_GROUP(G, L)                                                                                  \
    hipLaunchKernelGGL((k_tripvote_group<G, L>), gd, bd, lds, st, d_bandnorms, d_intensity, npats, nbands,   \
                       d_tripang, d_tripid, ntrip, nfam, angtol, d_bin_start, d_bin_rows, nbins, bin_width, \
                       d_accw, d_fam, d_rank, d_cm, d_acc)
            if (g == 16) {
                if (lib_in) EBSD_TV_GROUP(16, true);
                else EBSD_TV_GROUP(16, false);
            } else {
                if (lib_in) EBSD_TV_GROUP(32, true);
                else EBSD_TV_GROUP(32, false);
            }
#undef EBSD_TV_GROUP
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return fail(-2, std::string("k_tripvote_group launch: ") + hipGetErrorString(e));
            return 0;
        }
    }
    int waves = (int)(kLdsBudget / per_bytes);
    if (waves > 4) waves = 4;
    const bool lds_lib = lib_bytes + waves * per_bytes <= kLdsBudget;
    const size_t lds = (lds_lib ? lib_bytes : 0) + waves * per_bytes;
    const int64_t groups = (npats + waves - 1) / waves;
    if (groups > 0x7fffffff) return fail(-1, "too many patterns for one launch");
    if (lds_lib)
        hipLaunchKernelGGL(k_tripvote<true>, dim3((unsigned)groups), dim3(waves * kWave), lds, (hipStream_t)stream,
                           d_bandnorms, d_intensity, npats, nbands, d_tripang, d_tripid, ntrip, nfam, angtol,
                           d_bin_start, d_bin_rows, nbins, bin_width, d_accw, d_fam, d_rank, d_cm, d_acc);
    else
        hipLaunchKernelGGL(k_tripvote<false>, dim3((unsigned)groups), dim3(waves * kWave), lds, (hipStream_t)stream,
                           d_bandnorms, d_intensity, npats, nbands, d_tripang, d_tripid, ntrip, nfam, angtol,
                           d_bin_start, d_bin_rows, nbins, bin_width, d_accw, d_fam, d_rank, d_cm, d_acc);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(-2, std::string("k_tripvote launch: ") + hipGetErrorString(e));
    return 0;
}

extern "C" int ebsd_bandindex(const float *d_bandnorms, const float *d_intensity, int64_t npats, int nbands,
                              const float *d_accw, const int32_t *d_fam, const float *d_rank, const float *d_acc,
                              int nfam, const float *d_polescart, const double *d_polescart64,
                              const int32_t *d_famid, int npoles, const double *d_angrows,
                              const int32_t *d_famindex, const int32_t *d_famcount, double angtol,
                              int n_band_early, int nfit, int simple_weights, double *d_quat, double *d_fit,
                              float *d_cm, int64_t *d_polematch, int64_t *d_nmatch, int64_t *d_ij,
                              int32_t *d_acc_correct, double *d_bandfit, uint8_t *d_polevalid,
                              float *d_assign_fit, float *d_fitb, void *stream)
{
    if (npats < 0) return fail(-1, "npats must be >= 0");
    if (nbands < 1 || nbands > kMaxBands) return fail(-1, "nbands must be in [1, 64]");
    if (nfam < 1 || npoles < 1) return fail(-1, "the library needs at least one family and one pole");
    const size_t lds = (size_t)npoles * (3 * sizeof(double) + 3 * sizeof(float) + sizeof(int32_t)) +
                       (size_t)nfam * npoles * sizeof(double) + (size_t)2 * nfam * sizeof(int32_t);
    if (lds > kLdsBudget) return fail(-1, "library too large for the band assignment's LDS (64 KiB)");
    if (npats == 0) return 0;
    if (!d_bandnorms || !d_intensity || !d_accw || !d_fam || !d_rank || !d_acc || !d_polescart || !d_polescart64 ||
        !d_famid || !d_angrows || !d_famindex || !d_famcount || !d_quat || !d_fit || !d_cm || !d_polematch ||
        !d_nmatch || !d_ij || !d_acc_correct || !d_bandfit)
        return fail(-1, "null device pointer");
    BandIndexArgs a;
    a.bn = d_bandnorms; a.inten = d_intensity; a.accw = d_accw; a.rank = d_rank; a.acc = d_acc; a.fam = d_fam;
    a.npats = npats; a.nb = nbands; a.nfam = nfam; a.npl = npoles;
    a.pflt = d_polescart; a.pc64 = d_polescart64; a.famid = d_famid; a.angrows = d_angrows;
    a.famidx = d_famindex; a.famcnt = d_famcount;
    a.angtol = angtol; a.n_band_early = n_band_early; a.nfit = nfit; a.simple = simple_weights;
    a.quat = d_quat; a.fit = d_fit; a.bandfit = d_bandfit; a.cm = d_cm; a.afit = d_assign_fit; a.fitb = d_fitb;
    a.pm = d_polematch; a.nmatch = d_nmatch; a.ij = d_ij; a.acc_correct = d_acc_correct; a.pv = d_polevalid;
    const int per_block = 4;
    const int64_t grid = (npats + per_block - 1) / per_block;
    if (grid > 0x7fffffff) return fail(-1, "too many patterns for one launch");
    hipLaunchKernelGGL(k_bandindex, dim3((unsigned)grid), dim3(per_block * kWave), lds, (hipStream_t)stream, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(-2, std::string("k_bandindex launch: ") + hipGetErrorString(e));
    return 0;
}
