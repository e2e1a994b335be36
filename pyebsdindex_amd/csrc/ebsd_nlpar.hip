// ebsd_nlpar.hip -- NLPAR (non-local pattern averaging) on MI355X (gfx950).
//
// Two kernels replace the Numba loops of pyebsdindex/nlpar_cpu.py:
//   k_nlpar_sigma  NLPAR.sigma_numba (:753-819): per scan point, squared distances to the
//                  neighbours within +-nn (clipped window), the minimum of d2 / 2n2 -> sigma,
//                  and (optionally) the normalised distances used by the lambda optimiser.
//   k_nlpar        NLPAR.nlpar_nb (:821-935): per scan point, the distances to its
//                  (2sr+1)^2 search window (shifted inside the chunk), their normalisation by
//                  both sigmas, exponential weights, and the weighted average of the window's
//                  patterns, in one work-group: the weights never leave LDS.
//
// One work-group of 256 threads handles one scan point.  Distances run over the mask
// pixels `idx` in groups of NG = 8 window entries: each thread walks its share of the mask
// pixels, reads the centre value once per pixel and the 8 neighbours' values, and keeps 8
// float32 partial sums and 8 pair counts; the partials are then summed over the work-group
// in float64 (the sums are order-independent up to rounding; the reference sums
// sequentially in float32).  The average reads the window's patterns with one vector load
// per thread per pattern (4 pixels) and accumulates in float32 in window order, product
// then sum, as the reference does.
//
// Work-groups are numbered so that the 8 XCDs each take a contiguous run of scan points
// (work-group b runs on XCD b mod 8): neighbouring points, whose windows share patterns,
// then run on the same XCD and find each other's patterns in that XCD's L2.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "ebsd_nlpar.h"

static thread_local std::string g_err;

static int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return fail(-3, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" const char *ebsd_nlpar_last_error(void) { return g_err.c_str(); }

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int NG = 8;          // window entries per distance pass
constexpr int kMaxWin = 441;   // (2*10+1)^2: search radius up to 10

template <typename T> __device__ __forceinline__ float ldf(const T *p, int64_t i) { return (float)p[i]; }

// Scan point of work-group b: XCD x = b mod 8 takes points [x*per, (x+1)*per).
__device__ __forceinline__ int64_t xcd_point(int64_t b, int64_t npts)
{
    int64_t per = (npts + 7) / 8;
    return (b % 8) * per + b / 8;
}

// Sum 8 float partials and 8 int counts over the work-group; thread 0 of the work-group
// gets the totals in d2[], n2[] (float64 / int).
__device__ __forceinline__ void block_sum8(const float (&part)[NG], const int (&cnt)[NG], double *red_d,
                                           int *red_n, double (&d2)[NG], int (&n2)[NG])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double dv[NG];
    int nv[NG];
#pragma unroll
    for (int k = 0; k < NG; ++k) {
        dv[k] = (double)part[k];
        nv[k] = cnt[k];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
        for (int k = 0; k < NG; ++k) {
            dv[k] += __shfl_xor(dv[k], off, 64);
            nv[k] += __shfl_xor(nv[k], off, 64);
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NG; ++k) {
            red_d[wave * NG + k] = dv[k];
            red_n[wave * NG + k] = nv[k];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NG; ++k) {
            double s = 0.0;
            int c = 0;
            for (int w = 0; w < kWaves; ++w) {
                s += red_d[w * NG + k];
                c += red_n[w * NG + k];
            }
            d2[k] = s;
            n2[k] = c;
        }
    }
    __syncthreads();
}

// Partial sums of one group of NG neighbours q[k] against the centre point p over this
// thread's mask pixels; a pair counts when both values are below sat.
template <typename T>
__device__ __forceinline__ void pair_partials(const T *__restrict__ pats, int64_t P, const int32_t *__restrict__ idx,
                                              int M, int64_t p, const int64_t (&q)[NG], float sat, float (&part)[NG],
                                              int (&cnt)[NG])
{
#pragma unroll
    for (int k = 0; k < NG; ++k) {
        part[k] = 0.f;
        cnt[k] = 0;
    }
    const T *c0 = pats + p * P;
    for (int m = threadIdx.x; m < M; m += kThreads) {
        const int64_t x = idx[m];
        const float v0 = ldf(c0, x);
        const bool ok0 = v0 < sat;
#pragma unroll
        for (int k = 0; k < NG; ++k) {
            const float v1 = ldf(pats + q[k] * P, x);
            const bool ok = ok0 && (v1 < sat);
            const float d = ok ? v0 - v1 : 0.f;
            part[k] = fmaf(d, d, part[k]);
            cnt[k] += ok ? 1 : 0;
        }
    }
}

// ---------------------------------------------------------------------------------------
// sigma (nlpar_cpu.py:753-808) and the normalised distances (:810-818)
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_nlpar_sigma(const T *__restrict__ pats, int nr, int nc, int64_t P,
                                                          const int32_t *__restrict__ idx, int M, int nn, float sat,
                                                          float *__restrict__ sigma, float *__restrict__ dout,
                                                          float *__restrict__ nout)
{
    __shared__ double red_d[kWaves * NG];
    __shared__ int red_n[kWaves * NG];
    __shared__ float s_mind;
    const int64_t npts = (int64_t)nr * nc;
    const int64_t p = xcd_point(blockIdx.x, npts);
    if (p >= npts) return;
    const int j = (int)(p / nc), i = (int)(p % nc);
    const int r0 = max(j - nn, 0), r1 = min(j + nn, nr - 1) + 1;
    const int cc0 = max(i - nn, 0), cc1 = min(i + nn, nc - 1) + 1;
    const int wn = cc1 - cc0;
    const int count = (r1 - r0) * wn;
    const int win = (2 * nn + 1) * (2 * nn + 1);
    if (threadIdx.x == 0) s_mind = 1e24f;
    for (int g0 = 0; g0 < count; g0 += NG) {
        int64_t q[NG];
#pragma unroll
        for (int k = 0; k < NG; ++k) {
            const int e = min(g0 + k, count - 1);
            q[k] = (int64_t)(r0 + e / wn) * nc + (cc0 + e % wn);
        }
        float part[NG];
        int cnt[NG];
        pair_partials(pats, P, idx, M, p, q, sat, part, cnt);
        double d2[NG];
        int n2[NG];
        block_sum8(part, cnt, red_d, red_n, d2, n2);
        if (threadIdx.x == 0) {
            for (int k = 0; k < NG && g0 + k < count; ++k) {
                const int e = g0 + k;
                float dd = 0.f, nnf = (float)P;
                if (q[k] != p) {
                    // n2 starts at 1e-12 in the reference; float32 d2 / (n2 * 2)
                    const double n2d = 1.0e-12 + (double)n2[k];
                    dd = (float)d2[k];
                    nnf = (float)n2d;
                    if (dd >= 1.e-3f) {
                        const float s0 = (float)((double)dd / (double)(float)(n2d * 2.0));
                        if (s0 < s_mind) s_mind = s0;
                    }
                }
                if (dout) {
                    dout[p * win + e] = dd;
                    nout[p * win + e] = nnf;
                }
            }
        }
    }
    if (threadIdx.x == 0) {
        sigma[p] = sqrtf(s_mind);
        if (dout) {
            for (int e = count; e < win; ++e) {
                dout[p * win + e] = 0.f;
                nout[p * win + e] = 0.f;
            }
        }
    }
}

// dout normalisation (nlpar_cpu.py:810-818), after every sigma is known
__global__ void k_nlpar_dnorm(int nr, int nc, int nn, const float *__restrict__ sigma, float *__restrict__ dout,
                              const float *__restrict__ nout)
{
    const int win = (2 * nn + 1) * (2 * nn + 1);
    const int64_t total = (int64_t)nr * nc * win;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = t / win;
        const int e = (int)(t % win);
        const float nv = nout[t];
        if (!(nv > 0.f)) continue;
        const int j = (int)(p / nc), i = (int)(p % nc);
        const int r0 = max(j - nn, 0);
        const int cc0 = max(i - nn, 0), cc1 = min(i + nn, nc - 1) + 1;
        const int wn = cc1 - cc0;
        const int64_t q = (int64_t)(r0 + e / wn) * nc + (cc0 + e % wn);
        const float s2a = sigma[p] * sigma[p], s2b = sigma[q] * sigma[q];
        const float s12 = s2a + s2b;
        float d = dout[t];
        d -= nv * s12;
        d = (float)((double)d / ((double)s12 * sqrt(2.0 * (double)nv)));
        dout[t] = d;
    }
}

// ---------------------------------------------------------------------------------------
// NLPAR (nlpar_cpu.py:821-935)
// ---------------------------------------------------------------------------------------
template <typename T> struct Vec4;
template <> struct Vec4<uint8_t> {
    typedef uint32_t raw;
    static __device__ __forceinline__ void get(raw r, float (&v)[4])
    {
        v[0] = (float)(r & 0xff);
        v[1] = (float)((r >> 8) & 0xff);
        v[2] = (float)((r >> 16) & 0xff);
        v[3] = (float)(r >> 24);
    }
};
template <> struct Vec4<uint16_t> {
    typedef uint2 raw;
    static __device__ __forceinline__ void get(raw r, float (&v)[4])
    {
        v[0] = (float)(r.x & 0xffff);
        v[1] = (float)(r.x >> 16);
        v[2] = (float)(r.y & 0xffff);
        v[3] = (float)(r.y >> 16);
    }
};
template <> struct Vec4<float> {
    typedef float4 raw;
    static __device__ __forceinline__ void get(raw r, float (&v)[4])
    {
        v[0] = r.x;
        v[1] = r.y;
        v[2] = r.z;
        v[3] = r.w;
    }
};

template <typename O> __device__ __forceinline__ O store_cvt(float x);
template <> __device__ __forceinline__ float store_cvt<float>(float x) { return x; }
// pat_flt2int 'clip': clip to the type's range, then truncate (numpy astype)
template <> __device__ __forceinline__ uint8_t store_cvt<uint8_t>(float x)
{
    return (uint8_t)(int)fminf(fmaxf(x, 0.f), 255.f);
}
template <> __device__ __forceinline__ uint16_t store_cvt<uint16_t>(float x)
{
    return (uint16_t)(int)fminf(fmaxf(x, 0.f), 65535.f);
}

template <typename T, typename O>
__global__ __launch_bounds__(kThreads) void k_nlpar(const T *__restrict__ pats, int nr, int nc, int64_t P,
                                                    const int32_t *__restrict__ idx, int M,
                                                    const float *__restrict__ sigma, int sr, double lam2, float dthresh,
                                                    float diff_offset, float sat, int r0, int c0, int nrc, int ncc,
                                                    int vec, O *__restrict__ out, float *__restrict__ wout)
{
    __shared__ double red_d[kWaves * NG];
    __shared__ int red_n[kWaves * NG];
    __shared__ float s_w[kMaxWin];
    __shared__ int64_t s_q[kMaxWin];
    const int64_t npts = (int64_t)nrc * ncc;
    const int64_t o = xcd_point(blockIdx.x, npts);
    if (o >= npts) return;
    const int j = r0 + (int)(o / ncc), i = c0 + (int)(o % ncc);
    const int64_t p = (int64_t)j * nc + i;
    const int ws = 2 * sr + 1, win = ws * ws;
    // window shifted inside the chunk (nlpar_cpu.py:872-876)
    const int wx0 = max(i - sr, 0) - max(i + sr - (nc - 1), 0);
    const int wy0 = max(j - sr, 0) - max(j + sr - (nr - 1), 0);
    for (int e = threadIdx.x; e < win; e += kThreads) s_q[e] = (int64_t)(wy0 + e / ws) * nc + (wx0 + e % ws);
    __syncthreads();
    const float sig0 = sigma[p];
    const float s0 = sig0 * sig0;
    // distances of the non-centre entries, NG at a time (the centre is skipped: it gets -1e6)
    int ecur = 0;
    while (ecur < win) {
        int64_t q[NG];
        int es[NG];
        int got = 0, e = ecur;
        while (got < NG && e < win) {
            if (s_q[e] != p) {
                es[got] = e;
                q[got] = s_q[e];
                ++got;
            }
            ++e;
        }
        ecur = e;
        if (got == 0) break;
        for (int k = got; k < NG; ++k) {
            es[k] = -1;
            q[k] = q[0];
        }
        float part[NG];
        int cnt[NG];
        pair_partials(pats, P, idx, M, p, q, sat, part, cnt);
        double d2[NG];
        int n2[NG];
        block_sum8(part, cnt, red_d, red_n, d2, n2);
        if (threadIdx.x == 0) {
            for (int k = 0; k < got; ++k) {
                const float sg1 = sigma[q[k]];
                const float s1 = sg1 * sg1;
                float dd = (float)d2[k];
                const float nf = (float)n2[k];
                dd -= nf * (s0 + s1);
                const float dnorm = (s1 + s0) * sqrtf(2.0f * nf);
                dd = dnorm > 1.e-8f ? dd / dnorm : 1.0e6f * nf;
                s_w[es[k]] = dd;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float sum = 0.f;
        for (int e = 0; e < win; ++e) {
            const bool self = (s_q[e] == p);
            float w = self ? -1.0e6f : s_w[e];
            w = fmaxf(w - dthresh, 0.f);
            w = (float)exp(-1.0 * (double)w * lam2);
            if (self) w += diff_offset;
            s_w[e] = w;
            sum += w;
        }
        for (int e = 0; e < win; ++e) {
            const float w = s_w[e] / sum;
            s_w[e] = w;
            if (wout) wout[o * win + e] = w;
        }
    }
    __syncthreads();
    // weighted average over all P pixels, 4 per thread per step
    typedef typename Vec4<T>::raw R;
    O *dst = out + o * P;
    const int64_t P4 = vec ? P / 4 : 0;
    for (int64_t v = threadIdx.x; v < P4; v += kThreads) {
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
        for (int e = 0; e < win; ++e) {
            const R r = reinterpret_cast<const R *>(pats + s_q[e] * P)[v];
            float x[4];
            Vec4<T>::get(r, x);
            const float w = s_w[e];
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = __fadd_rn(acc[t], __fmul_rn(x[t], w));
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) dst[4 * v + t] = store_cvt<O>(acc[t]);
    }
    for (int64_t x = 4 * P4 + threadIdx.x; x < P; x += kThreads) {
        float acc = 0.f;
        for (int e = 0; e < win; ++e) acc = __fadd_rn(acc, __fmul_rn(ldf(pats + s_q[e] * P, x), s_w[e]));
        dst[x] = store_cvt<O>(acc);
    }
}

// max of the pixel values as order-preserving integer keys
__device__ __forceinline__ unsigned fkey(float x)
{
    const unsigned u = __float_as_uint(x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fkey_inv(unsigned k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

template <typename T>
__global__ void k_max(const T *__restrict__ pats, int64_t n, unsigned *__restrict__ key)
{
    unsigned best = 0;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        best = max(best, fkey(ldf(pats, t)));
    for (int off = 32; off >= 1; off >>= 1) best = max(best, (unsigned)__shfl_xor((int)best, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(key, best);
}

// in place: the key and the float share the 4 bytes
__global__ void k_key_to_float(unsigned *key, float *out)
{
    const unsigned k = *key;
    *out = fkey_inv(k);
}

__global__ void k_zero_key(unsigned *__restrict__ key) { *key = 0u; }

} // namespace

static int grid_for(int64_t total)
{
    int64_t g = (total + 255) / 256;
    return (int)(g > 8192 ? 8192 : (g < 1 ? 1 : g));
}

extern "C" int ebsd_nlpar_max(const void *d_pats, int dtype, int64_t n, int64_t P, float *d_max, void *stream)
{
    if (n < 0 || P <= 0) return fail(-1, "ebsd_nlpar_max: bad size");
    if (!d_pats || !d_max) return fail(-1, "ebsd_nlpar_max: null pointer");
    if (dtype < 0 || dtype > 2) return fail(-1, "ebsd_nlpar_max: dtype is 0 (u8), 1 (u16) or 2 (f32)");
    hipStream_t s = (hipStream_t)stream;
    // *d_max holds the running maximum as an order-preserving key, then the float itself
    unsigned *key = reinterpret_cast<unsigned *>(d_max);
    hipLaunchKernelGGL(k_zero_key, dim3(1), dim3(1), 0, s, key);
    const int64_t total = n * P;
    if (total > 0) {
        if (dtype == EBSD_NL_U8)
            hipLaunchKernelGGL(k_max<uint8_t>, dim3(grid_for(total)), dim3(256), 0, s, (const uint8_t *)d_pats, total,
                               key);
        else if (dtype == EBSD_NL_U16)
            hipLaunchKernelGGL(k_max<uint16_t>, dim3(grid_for(total)), dim3(256), 0, s, (const uint16_t *)d_pats,
                               total, key);
        else
            hipLaunchKernelGGL(k_max<float>, dim3(grid_for(total)), dim3(256), 0, s, (const float *)d_pats, total, key);
    }
    hipLaunchKernelGGL(k_key_to_float, dim3(1), dim3(1), 0, s, key, d_max);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int ebsd_nlpar_sigma(const void *d_pats, int dtype, int nr, int nc, int64_t P, const int32_t *d_idx, int M,
                                int nn, float sat, float *d_sigma, float *d_dout, float *d_nout, void *stream)
{
    if (nr <= 0 || nc <= 0 || P <= 0 || M < 0 || nn < 1) return fail(-1, "ebsd_nlpar_sigma: bad size");
    if (M > P) return fail(-1, "ebsd_nlpar_sigma: more mask pixels than pattern pixels");
    if (!d_pats || !d_idx || !d_sigma) return fail(-1, "ebsd_nlpar_sigma: null pointer");
    if ((d_dout == nullptr) != (d_nout == nullptr)) return fail(-1, "ebsd_nlpar_sigma: d_dout and d_nout go together");
    if (dtype < 0 || dtype > 2) return fail(-1, "ebsd_nlpar_sigma: dtype is 0 (u8), 1 (u16) or 2 (f32)");
    hipStream_t s = (hipStream_t)stream;
    const int64_t npts = (int64_t)nr * nc;
    const int64_t grid = ((npts + 7) / 8) * 8;
    if (grid > INT32_MAX) return fail(-1, "ebsd_nlpar_sigma: chunk too large");
#define EBSD_NL_SIG(T)                                                                                     \
    hipLaunchKernelGGL(k_nlpar_sigma<T>, dim3((unsigned)grid), dim3(kThreads), 0, s, (const T *)d_pats, nr, nc, P, \
                       d_idx, M, nn, sat, d_sigma, d_dout, d_nout)
    if (dtype == EBSD_NL_U8)
        EBSD_NL_SIG(uint8_t);
    else if (dtype == EBSD_NL_U16)
        EBSD_NL_SIG(uint16_t);
    else
        EBSD_NL_SIG(float);
#undef EBSD_NL_SIG
    HIPCHK(hipGetLastError());
    if (d_dout) {
        const int64_t total = npts * (2 * nn + 1) * (2 * nn + 1);
        hipLaunchKernelGGL(k_nlpar_dnorm, dim3(grid_for(total)), dim3(256), 0, s, nr, nc, nn, d_sigma, d_dout, d_nout);
        HIPCHK(hipGetLastError());
    }
    return 0;
}

extern "C" int ebsd_nlpar(const void *d_pats, int dtype, int nr, int nc, int64_t P, const int32_t *d_idx, int M,
                          const float *d_sigma, int sr, float lam, float dthresh, float diff_offset, float sat, int r0,
                          int c0, int nrc, int ncc, void *d_out, int out_dtype, float *d_weights, void *stream)
{
    if (nr <= 0 || nc <= 0 || P <= 0 || M < 0) return fail(-1, "ebsd_nlpar: bad size");
    if (M > P) return fail(-1, "ebsd_nlpar: more mask pixels than pattern pixels");
    if (sr < 1 || (2 * sr + 1) * (2 * sr + 1) > kMaxWin) return fail(-1, "ebsd_nlpar: search radius must be 1..10");
    if (nr < 2 * sr + 1 || nc < 2 * sr + 1)
        return fail(-1, "ebsd_nlpar: the chunk must be at least 2*searchradius+1 points in each direction");
    if (r0 < 0 || c0 < 0 || nrc < 0 || ncc < 0 || r0 + nrc > nr || c0 + ncc > nc)
        return fail(-1, "ebsd_nlpar: calculation region outside the chunk");
    if (!d_pats || !d_idx || !d_sigma || !d_out) return fail(-1, "ebsd_nlpar: null pointer");
    if (dtype < 0 || dtype > 2 || out_dtype < 0 || out_dtype > 2)
        return fail(-1, "ebsd_nlpar: dtype is 0 (u8), 1 (u16) or 2 (f32)");
    if (!(lam > 0.f)) return fail(-1, "ebsd_nlpar: lambda must be positive");
    const int64_t npts = (int64_t)nrc * ncc;
    if (npts == 0) return 0;
    const int64_t grid = ((npts + 7) / 8) * 8;
    if (grid > INT32_MAX) return fail(-1, "ebsd_nlpar: region too large");
    hipStream_t s = (hipStream_t)stream;
    // lam2 = 1.0 / lam**2 (nlpar_cpu.py:844): float32 square, float64 reciprocal
    const double lam2 = 1.0 / (double)(lam * lam);
    // 4-pixel vector loads when every pattern starts on a 16-byte boundary
    const int vec = (P % 4 == 0) && ((uintptr_t)d_pats % 16 == 0);
#define EBSD_NL_RUN(T, O)                                                                                   \
    hipLaunchKernelGGL((k_nlpar<T, O>), dim3((unsigned)grid), dim3(kThreads), 0, s, (const T *)d_pats, nr, nc, P, \
                       d_idx, M, d_sigma, sr, lam2, dthresh, diff_offset, sat, r0, c0, nrc, ncc, vec, (O *)d_out, \
                       d_weights)
#define EBSD_NL_OUT(T)                 \
    if (out_dtype == EBSD_NL_U8)       \
        EBSD_NL_RUN(T, uint8_t);       \
    else if (out_dtype == EBSD_NL_U16) \
        EBSD_NL_RUN(T, uint16_t);      \
    else                               \
        EBSD_NL_RUN(T, float);
    if (dtype == EBSD_NL_U8) {
        EBSD_NL_OUT(uint8_t)
    } else if (dtype == EBSD_NL_U16) {
        EBSD_NL_OUT(uint16_t)
    } else {
        EBSD_NL_OUT(float)
    }
#undef EBSD_NL_OUT
#undef EBSD_NL_RUN
    HIPCHK(hipGetLastError());
    return 0;
}
