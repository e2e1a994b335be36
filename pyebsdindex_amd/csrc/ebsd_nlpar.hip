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
// One work-group of 256 threads handles one scan point.  Its four wavefronts take turns over
// the window's neighbours, a few at a time (4 in k_nlpar, 2 in k_nlpar_sigma), and each
// sums its neighbours' squared distances over all mask pixels by itself, so only a
// wavefront reduction follows.  Per-pixel path: the lanes walk the mask pixels `idx`, read
// the centre value once per pixel and each neighbour's, and sum the squared float32
// differences in float64 (the reference sums sequentially in float32).  uint8 path (with a
// workspace): a prepass stores each pattern's masked sum of squares and the list of its
// masked pixels at or above the saturation threshold (up to 64, one per lane); then
// d2 = sum a^2 + sum b^2 - 2 sum ab exactly, with sum ab from v_dot4_u32_u8 over 4 pixels of
// the byte-masked centre and the neighbour per instruction, minus the exact terms of the
// pixels saturated in either pattern (which the reference leaves out).  Pairs that touch a
// pattern with more saturated pixels take the per-pixel path.  Weights are then computed one window entry per thread;
// one thread adds them in window order (float32, as the reference).  The average reads the
// window's patterns with 16-byte vector loads (16 pixels per thread) and accumulates
// in float32 in window order with fused multiply-adds (the reference rounds product and
// sum separately).
//
// Work-groups are numbered so that the 8 XCDs each take a contiguous run of scan points in
// an order of 8 x 32-point blocks (work-group b runs on XCD b mod 8): the points an XCD has
// in flight form a compact patch whose windows share patterns and fit that XCD's L2.

#include <hip/hip_runtime.h>

#include <algorithm>
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
constexpr int kMaxWin = 441;   // (2*10+1)^2: search radius up to 10

template <typename T> __device__ __forceinline__ float ldf(const T *p, int64_t i) { return (float)p[i]; }

// Scan point of work-group b: XCD x = b mod 8 takes points [x*per, (x+1)*per).
__device__ __forceinline__ int64_t xcd_point(int64_t b, int64_t npts)
{
    int64_t per = (npts + 7) / 8;
    return (b % 8) * per + b / 8;
}

// Scan points in blocks of kBR x kBC: point ordinals run block by block (blocks row-major,
// points row-major in a block), and each XCD takes a contiguous run of ordinals.  The
// points one XCD has in flight then form a compact patch whose search windows fit its L2
// (a 60x60 uint8 patch of 8 x 32 points with its 7 x 7 windows: 1.6 MB of the 4 MB).
constexpr int kBR = 8, kBC = 32;

__device__ __forceinline__ int64_t blocked_items(int nrows, int ncols)
{
    return (int64_t)((nrows + kBR - 1) / kBR) * ((ncols + kBC - 1) / kBC) * (kBR * kBC);
}

// (row, column) of work-group b over an nrows x ncols region; false for padding
__device__ __forceinline__ bool blocked_point(int64_t b, int nrows, int ncols, int &row, int &col)
{
    const int64_t l = xcd_point(b, blocked_items(nrows, ncols));
    const int nbc = (ncols + kBC - 1) / kBC;
    const int64_t blk = l / (kBR * kBC);
    const int w = (int)(l % (kBR * kBC));
    row = (int)(blk / nbc) * kBR + w / kBC;
    col = (int)(blk % nbc) * kBC + w % kBC;
    return row < nrows && col < ncols;
}

template <typename V> __device__ __forceinline__ V wave_sum(V x)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// uint8 workspace: byte mask (0xff on mask pixels), per-pattern masked sum of squares, the
// number of masked pixels at or above the saturation threshold (kNoList when more than
// kSatList) and the list of those pixels.  Null mask4 = no workspace (per-pixel path only).
// 64: one list entry per lane in the corrections.  (With 32, the patterns of the bench's
// clipped-Poisson scan that hold 33-48 saturated pixels sent every group of pairs they were
// in down the per-pixel path: 2.74 against 2.50 ms per 65,536 points with 64.)
constexpr int kSatList = 64;
constexpr uint8_t kNoList = 255;
struct U8Work {
    const uint32_t *mask4; // P/4 words of the byte mask
    const double *sq;
    const uint8_t *nsat;
    const int32_t *list; // kSatList pixel indices per pattern
};

// Squared distances of pattern p to the NGW patterns q[k], summed over one wavefront's 64
// lanes; every lane gets d2[k] (float64) and the pair count n2[k].  The dot-product path is
// taken when T is uint8, the workspace is there and no pattern of the group has more than
// kSatList saturated mask pixels (the same choice in every lane: they read the same counts).
template <int NGW, typename T>
__device__ __forceinline__ void wave_pair_sums(const T *__restrict__ pats, int64_t P, const int32_t *__restrict__ idx,
                                               int M, int64_t p, const int64_t (&q)[NGW], float sat, U8Work wk,
                                               double (&d2)[NGW], int (&n2)[NGW])
{
    const int lane = threadIdx.x & 63;
    if constexpr (sizeof(T) == 1) {
        bool fast = wk.mask4 != nullptr && wk.nsat[p] != kNoList;
#pragma unroll
        for (int k = 0; k < NGW; ++k) fast = fast && wk.nsat[q[k]] != kNoList;
        if (fast) {
            uint32_t acc[NGW];
#pragma unroll
            for (int k = 0; k < NGW; ++k) acc[k] = 0u;
            const uint32_t *a4 = reinterpret_cast<const uint32_t *>(pats + p * P);
            const int64_t P4 = P / 4;
            // unrolled so that several iterations' loads are in flight
#pragma unroll 4
            for (int64_t v = lane; v < P4; v += 64) {
                const uint32_t am = a4[v] & wk.mask4[v];
#pragma unroll
                for (int k = 0; k < NGW; ++k) {
                    const uint32_t b = reinterpret_cast<const uint32_t *>(pats + q[k] * P)[v];
                    acc[k] = __builtin_amdgcn_udot4(am, b, acc[k], false);
                }
            }
            // sum over all mask pixels: sum a^2 + sum b^2 - 2 sum ab; then take out the pixels
            // saturated in either pattern (one list entry per lane), which the reference does
            // not count.  Lane value (-2 ab - removed) * 256 + removed count: one exact float64
            // wave sum carries both (the count is below 2 kSatList < 256).
            const T *pp = pats + p * P;
            const int nsp = wk.nsat[p];
#pragma unroll
            for (int k = 0; k < NGW; ++k) {
                const T *pq = pats + q[k] * P;
                int64_t corr = 0;
                int cnt = 0;
                if (lane < nsp) {
                    const int32_t x = wk.list[p * kSatList + lane];
                    const int d = (int)pp[x] - (int)pq[x];
                    corr += d * d;
                    ++cnt;
                }
                if (lane < (int)wk.nsat[q[k]]) {
                    const int32_t x = wk.list[q[k] * kSatList + lane];
                    if ((float)pp[x] < sat) { // else already taken out with p's list
                        const int d = (int)pp[x] - (int)pq[x];
                        corr += d * d;
                        ++cnt;
                    }
                }
                const int64_t lv = (-2 * (int64_t)acc[k] - corr) * 256 + cnt;
                const int64_t X = (int64_t)wave_sum((double)lv);
                const int64_t c = ((X % 256) + 256) % 256;
                d2[k] = wk.sq[p] + wk.sq[q[k]] + (double)((X - c) / 256);
                n2[k] = M - (int)c;
            }
            return;
        }
    }
    if constexpr (sizeof(T) == 2) {
        bool fast = wk.mask4 != nullptr && wk.nsat[p] != kNoList;
#pragma unroll
        for (int k = 0; k < NGW; ++k) fast = fast && wk.nsat[q[k]] != kNoList;
        if (fast) {
            // two pixels per dword; masked-out pixels of both patterns set to 0 by a packed
            // 0xffff-per-pixel mask, then |a - b| squared and summed exactly in 64 bits
            uint64_t acc[NGW];
#pragma unroll
            for (int k = 0; k < NGW; ++k) acc[k] = 0;
            const uint32_t *a2 = reinterpret_cast<const uint32_t *>(pats + p * P);
            const uint16_t *mb2 = reinterpret_cast<const uint16_t *>(wk.mask4); // 2 mask bytes per dword
            const int64_t P2 = P / 2;
#pragma unroll 2
            for (int64_t v = lane; v < P2; v += 64) {
                const uint32_t mb = mb2[v];
                const uint32_t m = ((mb & 0xffu) ? 0xffffu : 0u) | ((mb & 0xff00u) ? 0xffff0000u : 0u);
                const uint32_t aw = a2[v] & m;
                const uint32_t alo = aw & 0xffffu, ahi = aw >> 16;
#pragma unroll
                for (int k = 0; k < NGW; ++k) {
                    const uint32_t bw = reinterpret_cast<const uint32_t *>(pats + q[k] * P)[v] & m;
                    const uint32_t blo = bw & 0xffffu, bhi = bw >> 16;
                    const uint32_t dlo = alo > blo ? alo - blo : blo - alo;
                    const uint32_t dhi = ahi > bhi ? ahi - bhi : bhi - ahi;
                    acc[k] += (uint64_t)dlo * dlo;
                    acc[k] += (uint64_t)dhi * dhi;
                }
            }
            // take out the pixels saturated in either pattern, as for uint8; the sums stay
            // below 2^53 (at most 2^22 pixels of 65535^2), so float64 carries them exactly
            const T *pp = pats + p * P;
            const int nsp = wk.nsat[p];
#pragma unroll
            for (int k = 0; k < NGW; ++k) {
                const T *pq = pats + q[k] * P;
                int64_t corr = 0;
                int cnt = 0;
                if (lane < nsp) {
                    const int32_t x = wk.list[p * kSatList + lane];
                    const int64_t d = (int64_t)pp[x] - (int64_t)pq[x];
                    corr += d * d;
                    ++cnt;
                }
                if (lane < (int)wk.nsat[q[k]]) {
                    const int32_t x = wk.list[q[k] * kSatList + lane];
                    if ((float)pp[x] < sat) {
                        const int64_t d = (int64_t)pp[x] - (int64_t)pq[x];
                        corr += d * d;
                        ++cnt;
                    }
                }
                d2[k] = wave_sum((double)((int64_t)acc[k] - corr));
                n2[k] = M - wave_sum(cnt);
            }
            return;
        }
    }
    double part[NGW];
    int cnt[NGW];
#pragma unroll
    for (int k = 0; k < NGW; ++k) {
        part[k] = 0.0;
        cnt[k] = 0;
    }
    const T *c0 = pats + p * P;
#pragma unroll 4
    for (int m = lane; m < M; m += 64) {
        const int64_t x = idx[m];
        const float v0 = ldf(c0, x);
        const bool ok0 = v0 < sat;
#pragma unroll
        for (int k = 0; k < NGW; ++k) {
            const float v1 = ldf(pats + q[k] * P, x);
            const bool ok = ok0 && (v1 < sat);
            // the float32 difference, as the reference forms it; its square and the sum in float64
            const double d = ok ? (double)(v0 - v1) : 0.0;
            part[k] = fma(d, d, part[k]);
            cnt[k] += ok ? 1 : 0;
        }
    }
#pragma unroll
    for (int k = 0; k < NGW; ++k) {
        d2[k] = wave_sum(part[k]);
        n2[k] = wave_sum(cnt[k]);
    }
}

// uint8 prepass: masked sum of squares (exact, float64), saturated mask pixels (count and
// list) of every pattern, one work-group per pattern; with the byte mask (mask4), word by
// word with packed byte dot products, else pixel by pixel through idx
template <typename T>
__global__ __launch_bounds__(kThreads) void k_int_prep(const T *__restrict__ pats, int64_t n, int64_t P,
                                                     const int32_t *__restrict__ idx, int M,
                                                     const uint32_t *__restrict__ mask4, float sat,
                                                     double *__restrict__ sq, uint8_t *__restrict__ nsat,
                                                     int32_t *__restrict__ list)
{
    __shared__ double s_sq[kWaves];
    __shared__ int s_ns;
    const int64_t p = blockIdx.x;
    if (p >= n) return;
    if (threadIdx.x == 0) s_ns = 0;
    __syncthreads();
    double acc = 0.0; // integer squares, exact in float64
    // the smallest saturated value: a pixel value b is saturated when !(b < sat)
    const float tf = ceilf(sat);
    if (sizeof(T) == 1 && mask4 != nullptr && tf >= 1.f) {
        // uint8: word by word over the byte-masked pattern
        // word by word over the byte-masked pattern: masked-out bytes are 0, below tf
        const unsigned tint = tf > 255.f ? 256u : (unsigned)tf;
        const uint32_t *a4 = reinterpret_cast<const uint32_t *>(pats + p * P);
        unsigned wacc = 0; // at most ceil(P / 1024) words of 4 x 255^2 per thread
        for (int64_t v = threadIdx.x; v < P / 4; v += kThreads) {
            const uint32_t am = a4[v] & mask4[v];
            wacc = __builtin_amdgcn_udot4(am, am, wacc, false);
            const unsigned mx = max(max(am & 0xff, (am >> 8) & 0xff), max((am >> 16) & 0xff, am >> 24));
            if (mx >= tint) {
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (((am >> (8 * k)) & 0xff) >= tint) {
                        const int slot = atomicAdd(&s_ns, 1);
                        if (slot < kSatList) list[p * kSatList + slot] = (int32_t)(4 * v + k);
                    }
            }
        }
        acc = (double)wacc;
    } else {
        for (int m = threadIdx.x; m < M; m += kThreads) {
            const int32_t x = idx[m];
            const unsigned v = pats[p * P + x];
            acc += (double)v * (double)v;
            if (!((float)v < sat)) {
                const int slot = atomicAdd(&s_ns, 1);
                if (slot < kSatList) list[p * kSatList + slot] = x;
            }
        }
    }
    double dacc = acc;
    for (int off = 32; off >= 1; off >>= 1) dacc += __shfl_xor(dacc, off, 64);
    if ((threadIdx.x & 63) == 0) s_sq[threadIdx.x >> 6] = dacc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kWaves; ++w) t += s_sq[w];
        sq[p] = t;
        nsat[p] = s_ns > kSatList ? kNoList : (uint8_t)s_ns;
    }
}

__global__ void k_mask_zero(int64_t P, uint8_t *__restrict__ mask)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < P; t += (int64_t)gridDim.x * blockDim.x)
        mask[t] = 0;
}

__global__ void k_mask_scatter(const int32_t *__restrict__ idx, int M, uint8_t *__restrict__ mask)
{
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < M; t += gridDim.x * blockDim.x) mask[idx[t]] = 0xff;
}

// ---------------------------------------------------------------------------------------
// sigma (nlpar_cpu.py:753-808) and the normalised distances (:810-818)
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_nlpar_sigma(const T *__restrict__ pats, int nr, int nc, int64_t P,
                                                          const int32_t *__restrict__ idx, int M, int nn, float sat,
                                                          U8Work wk, float *__restrict__ sigma,
                                                          float *__restrict__ dout, float *__restrict__ nout)
{
    constexpr int NGW = 2; // 8 neighbours at nn = 1: one pair per wavefront
    __shared__ float s_cand[kMaxWin];
    int j, i;
    if (!blocked_point(blockIdx.x, nr, nc, j, i)) return;
    const int64_t p = (int64_t)j * nc + i;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r0 = max(j - nn, 0), r1 = min(j + nn, nr - 1) + 1;
    const int cc0 = max(i - nn, 0), cc1 = min(i + nn, nc - 1) + 1;
    const int wn = cc1 - cc0;
    const int count = (r1 - r0) * wn;
    const int win = (2 * nn + 1) * (2 * nn + 1);
    const int eself = (j - r0) * wn + (i - cc0);
    const int ne = count - 1, ngroups = (ne + NGW - 1) / NGW;
    for (int g = wave; g < ngroups; g += kWaves) {
        int64_t q[NGW];
        int es[NGW];
#pragma unroll
        for (int k = 0; k < NGW; ++k) {
            const int t = g * NGW + k, tt = min(t, ne - 1);
            const int e = tt < eself ? tt : tt + 1;
            es[k] = t < ne ? e : -1;
            q[k] = (int64_t)(r0 + e / wn) * nc + (cc0 + e % wn);
        }
        double d2[NGW];
        int n2[NGW];
        wave_pair_sums<NGW>(pats, P, idx, M, p, q, sat, wk, d2, n2);
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < NGW; ++k) {
                if (es[k] < 0) continue;
                // n2 starts at 1e-12 in the reference (float64); d2 is float32
                const double n2d = 1.0e-12 + (double)n2[k];
                const float dd = (float)d2[k];
                float cand = 1e24f;
                if (dd >= 1.e-3f) cand = (float)((double)dd / (double)(float)(n2d * 2.0));
                s_cand[es[k]] = cand;
                if (dout) {
                    dout[p * win + es[k]] = dd;
                    nout[p * win + es[k]] = (float)n2d;
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float mind = 1e24f;
        for (int e = 0; e < count; ++e)
            if (e != eself && s_cand[e] < mind) mind = s_cand[e];
        sigma[p] = sqrtf(mind);
        if (dout) {
            dout[p * win + eself] = 0.f;
            nout[p * win + eself] = (float)P;
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
// 16 consecutive pixels from a 16-byte-aligned address as float32
template <typename T> __device__ __forceinline__ void load16(const T *__restrict__ src, float (&x)[16])
{
    const uint4 *p4 = reinterpret_cast<const uint4 *>(src);
    if constexpr (sizeof(T) == 1) {
        const uint4 r = p4[0];
        const uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            x[4 * k + 0] = (float)(w[k] & 0xff);
            x[4 * k + 1] = (float)((w[k] >> 8) & 0xff);
            x[4 * k + 2] = (float)((w[k] >> 16) & 0xff);
            x[4 * k + 3] = (float)(w[k] >> 24);
        }
    } else if constexpr (sizeof(T) == 2) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint4 r = p4[h];
            const uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                x[8 * h + 2 * k] = (float)(w[k] & 0xffff);
                x[8 * h + 2 * k + 1] = (float)(w[k] >> 16);
            }
        }
    } else {
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const uint4 r = p4[h];
            x[4 * h + 0] = __uint_as_float(r.x);
            x[4 * h + 1] = __uint_as_float(r.y);
            x[4 * h + 2] = __uint_as_float(r.z);
            x[4 * h + 3] = __uint_as_float(r.w);
        }
    }
}

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

// 16 output values to a 16-byte-aligned address
template <typename O> __device__ __forceinline__ void store16(O *__restrict__ dst, const float (&acc)[16])
{
    uint4 *d4 = reinterpret_cast<uint4 *>(dst);
    if constexpr (sizeof(O) == 1) {
        uint32_t w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            w[k] = (uint32_t)store_cvt<O>(acc[4 * k]) | ((uint32_t)store_cvt<O>(acc[4 * k + 1]) << 8) |
                   ((uint32_t)store_cvt<O>(acc[4 * k + 2]) << 16) | ((uint32_t)store_cvt<O>(acc[4 * k + 3]) << 24);
        d4[0] = make_uint4(w[0], w[1], w[2], w[3]);
    } else if constexpr (sizeof(O) == 2) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            uint32_t w[4];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                w[k] = (uint32_t)store_cvt<O>(acc[8 * h + 2 * k]) |
                       ((uint32_t)store_cvt<O>(acc[8 * h + 2 * k + 1]) << 16);
            d4[h] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    } else {
#pragma unroll
        for (int h = 0; h < 4; ++h)
            d4[h] = make_uint4(__float_as_uint(acc[4 * h]), __float_as_uint(acc[4 * h + 1]),
                               __float_as_uint(acc[4 * h + 2]), __float_as_uint(acc[4 * h + 3]));
    }
}

template <typename T, typename O>
__global__ __launch_bounds__(kThreads) void k_nlpar(const T *__restrict__ pats, int nr, int nc, int64_t P,
                                                    const int32_t *__restrict__ idx, int M,
                                                    const float *__restrict__ sigma, int sr, double lam2, float dthresh,
                                                    float diff_offset, float sat, int r0, int c0, int nrc, int ncc,
                                                    int vec, U8Work wk, O *__restrict__ out, float *__restrict__ wout)
{
    __shared__ float s_w[kMaxWin];
    __shared__ int64_t s_q[kMaxWin];
    __shared__ int64_t s_off[kMaxWin]; // element offset of each window pattern: q * P
    __shared__ float s_sum;
    int jr, ic;
    if (!blocked_point(blockIdx.x, nrc, ncc, jr, ic)) return;
    const int64_t o = (int64_t)jr * ncc + ic;
    const int j = r0 + jr, i = c0 + ic;
    const int64_t p = (int64_t)j * nc + i;
    const int ws = 2 * sr + 1, win = ws * ws;
    // window shifted inside the chunk (nlpar_cpu.py:872-876)
    const int wx0 = max(i - sr, 0) - max(i + sr - (nc - 1), 0);
    const int wy0 = max(j - sr, 0) - max(j + sr - (nr - 1), 0);
    for (int e = threadIdx.x; e < win; e += kThreads) {
        const int64_t q = (int64_t)(wy0 + e / ws) * nc + (wx0 + e % ws);
        s_q[e] = q;
        s_off[e] = q * P;
    }
    __syncthreads();
    const float sig0 = sigma[p];
    const float s0 = sig0 * sig0;
    // distances of the non-centre entries, NGW per wavefront pass, waves taking turns (the
    // centre gets d = -1e6 without a distance)
    constexpr int NGW = 4;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int eself = (j - wy0) * ws + (i - wx0);
    const int ne = win - 1, ngroups = (ne + NGW - 1) / NGW;
    for (int g = wave; g < ngroups; g += kWaves) {
        int64_t q[NGW];
        int es[NGW];
#pragma unroll
        for (int k = 0; k < NGW; ++k) {
            const int t = g * NGW + k, tt = min(t, ne - 1);
            const int e = tt < eself ? tt : tt + 1;
            es[k] = t < ne ? e : -1;
            q[k] = s_q[e];
        }
        double d2[NGW];
        int n2[NGW];
        wave_pair_sums<NGW>(pats, P, idx, M, p, q, sat, wk, d2, n2);
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < NGW; ++k) {
                if (es[k] < 0) continue;
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
    // weights, one entry per thread; their float32 sum in window order by one thread
    for (int e = threadIdx.x; e < win; e += kThreads) {
        const bool self = (e == eself);
        float w = self ? -1.0e6f : s_w[e];
        w = fmaxf(w - dthresh, 0.f);
        w = (float)exp(-1.0 * (double)w * lam2);
        if (self) w += diff_offset;
        s_w[e] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float sum = 0.f;
        for (int e = 0; e < win; ++e) sum += s_w[e];
        s_sum = sum;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < win; e += kThreads) {
        const float w = s_w[e] / s_sum;
        s_w[e] = w;
        if (wout) wout[o * win + e] = w;
    }
    __syncthreads();
    // weighted average over all P pixels, 16 per thread per step (16-byte loads), window
    // entries unrolled so that several loads are in flight
    O *dst = out + o * P;
    const int64_t P16 = vec ? P / 16 : 0;
    for (int64_t v = threadIdx.x; v < P16; v += kThreads) {
        float acc[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) acc[t] = 0.f;
#pragma unroll 4
        for (int e = 0; e < win; ++e) {
            float x[16];
            load16(pats + s_off[e] + 16 * v, x);
            const float w = s_w[e];
#pragma unroll
            for (int t = 0; t < 16; ++t) acc[t] = fmaf(x[t], w, acc[t]);
        }
        store16(dst + 16 * v, acc);
    }
    for (int64_t x = 16 * P16 + threadIdx.x; x < P; x += kThreads) {
        float acc = 0.f;
        for (int e = 0; e < win; ++e) acc = fmaf(ldf(pats + s_off[e], x), s_w[e], acc);
        dst[x] = store_cvt<O>(acc);
    }
}

// ---------------------------------------------------------------------------------------
// k_nlpar_pairs: the squared distances and pixel counts of every pair (point, window entry)
// the averaging needs, each unordered pair once.  A point of the calculation region computes
// the entries of its (shifted) window except those whose partner q is in the region, holds
// the point in its own window and has the smaller scan index: q computes that pair and
// writes it to both points' slots.  The distance of a pair does not depend on which of the
// two computes it (exact integer sums, or float64 sums of exact squares of float32
// differences in the same pixel order), so the table holds what k_nlpar computes itself.
// tab[o * win + e] = (float(d2), n2) for region point o and window entry e.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void shifted_window(int j, int i, int nr, int nc, int sr, int &wy0, int &wx0)
{
    wx0 = max(i - sr, 0) - max(i + sr - (nc - 1), 0);
    wy0 = max(j - sr, 0) - max(j + sr - (nr - 1), 0);
}

template <typename T, int NGW>
__global__ __launch_bounds__(kThreads) void k_nlpar_pairs(const T *__restrict__ pats, int nr, int nc, int64_t P,
                                                          const int32_t *__restrict__ idx, int M, int sr, float sat,
                                                          int r0, int c0, int nrc, int ncc, U8Work wk,
                                                          float2 *__restrict__ tab)
{
    __shared__ int s_e[kMaxWin];
    __shared__ int s_n;
    int jr, ic;
    if (!blocked_point(blockIdx.x, nrc, ncc, jr, ic)) return;
    const int j = r0 + jr, i = c0 + ic;
    const int64_t p = (int64_t)j * nc + i;
    const int64_t o = (int64_t)jr * ncc + ic;
    const int ws = 2 * sr + 1, win = ws * ws;
    int wy0, wx0;
    shifted_window(j, i, nr, nc, sr, wy0, wx0);
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    for (int e = threadIdx.x; e < win; e += kThreads) {
        const int qy = wy0 + e / ws, qx = wx0 + e % ws;
        if (qy == j && qx == i) continue;
        const int qr = qy - r0, qc = qx - c0;
        bool skip = false;
        if (qr >= 0 && qr < nrc && qc >= 0 && qc < ncc) {
            int qy0, qx0;
            shifted_window(qy, qx, nr, nc, sr, qy0, qx0);
            const bool holds = j >= qy0 && j < qy0 + ws && i >= qx0 && i < qx0 + ws;
            skip = holds && (int64_t)qy * nc + qx < p;
        }
        if (!skip) s_e[atomicAdd(&s_n, 1)] = e;
    }
    __syncthreads();
    const int n = s_n; // sr (sr + 1) / 2 * 4 for an interior point: one group of NGW per wave
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int g = wave; g * NGW < n; g += kWaves) {
        int64_t q[NGW];
        int es[NGW];
#pragma unroll
        for (int k = 0; k < NGW; ++k) {
            const int t = g * NGW + k;
            const int e = s_e[min(t, n - 1)];
            es[k] = t < n ? e : -1;
            q[k] = (int64_t)(wy0 + e / ws) * nc + (wx0 + e % ws);
        }
        double d2[NGW];
        int n2[NGW];
        wave_pair_sums<NGW>(pats, P, idx, M, p, q, sat, wk, d2, n2);
        // lane k writes pair k: this point's slot, and the partner's when it needs the pair
        double dk = d2[0];
        int nk = n2[0], e = es[0];
#pragma unroll
        for (int k = 1; k < NGW; ++k)
            if (lane == k) {
                dk = d2[k];
                nk = n2[k];
                e = es[k];
            }
        if (lane < NGW && e >= 0) {
            const float2 v = make_float2((float)dk, (float)nk);
            tab[o * win + e] = v;
            const int qy = wy0 + e / ws, qx = wx0 + e % ws, qr = qy - r0, qc = qx - c0;
            if (qr >= 0 && qr < nrc && qc >= 0 && qc < ncc) {
                int qy0, qx0;
                shifted_window(qy, qx, nr, nc, sr, qy0, qx0);
                if (j >= qy0 && j < qy0 + ws && i >= qx0 && i < qx0 + ws)
                    tab[((int64_t)qr * ncc + qc) * win + (j - qy0) * ws + (i - qx0)] = v;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// k_nlpar_tile: k_nlpar for a tile of TR x TC scan points per work-group.  The distances and
// weights of each point are those of k_nlpar (the same code per point); the weighted
// average walks the union of the tile's windows once in row-major order, loading and
// converting each pattern's 16-pixel vector once for all the points whose window holds it.
// A point's window entries then still come in row-major (window) order, so every point sees
// the same fused multiply-adds in the same order as in k_nlpar -- bit-identical results --
// while the patterns read per point drop from (2sr+1)^2 to about (2sr+TR)(2sr+TC) / (TR TC).
// ---------------------------------------------------------------------------------------
constexpr int kTBR = 4, kTBC = 16; // tiles per XCD block (8 x 32 points for 2 x 2 tiles)

__device__ __forceinline__ bool blocked_tile(int64_t b, int ntr, int ntc, int &ty, int &tx)
{
    const int64_t items = (int64_t)((ntr + kTBR - 1) / kTBR) * ((ntc + kTBC - 1) / kTBC) * (kTBR * kTBC);
    const int64_t l = xcd_point(b, items);
    const int nbc = (ntc + kTBC - 1) / kTBC;
    const int64_t blk = l / (kTBR * kTBC);
    const int w = (int)(l % (kTBR * kTBC));
    ty = (int)(blk / nbc) * kTBR + w / kTBC;
    tx = (int)(blk % nbc) * kTBC + w % kTBC;
    return ty < ntr && tx < ntc;
}

template <typename T, typename O, int TR, int TC>
__global__ __launch_bounds__(kThreads) void k_nlpar_tile(const T *__restrict__ pats, int nr, int nc, int64_t P,
                                                         const int32_t *__restrict__ idx, int M,
                                                         const float *__restrict__ sigma, int sr, double lam2,
                                                         float dthresh, float diff_offset, float sat, int r0, int c0,
                                                         int nrc, int ncc, U8Work wk, const float2 *__restrict__ tab,
                                                         O *__restrict__ out, float *__restrict__ wout)
{
    constexpr int NT = TR * TC;
    __shared__ float s_w[NT][kMaxWin];
    __shared__ float s_sum[NT];
    int ty, tx;
    if (!blocked_tile(blockIdx.x, (nrc + TR - 1) / TR, (ncc + TC - 1) / TC, ty, tx)) return;
    const int ws = 2 * sr + 1, win = ws * ws;
    // point t of the tile: scan position, window origin (shifted inside the chunk,
    // nlpar_cpu.py:872-876) and whether it lies in the calculation region
    auto geom = [&](int t, int &j, int &i, int &wy0, int &wx0) -> bool {
        const int jr = ty * TR + t / TC, ic = tx * TC + t % TC;
        j = r0 + min(jr, nrc - 1);
        i = c0 + min(ic, ncc - 1);
        wx0 = max(i - sr, 0) - max(i + sr - (nc - 1), 0);
        wy0 = max(j - sr, 0) - max(j + sr - (nr - 1), 0);
        return jr < nrc && ic < ncc;
    };
    // ---- distances of the non-centre entries: from the pair table, or point by point, NGW
    //      per wavefront pass (the same normalisation either way) ----
    constexpr int NGW = 4;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ne = win - 1, ngroups = (ne + NGW - 1) / NGW;
    if (tab) {
        for (int f = threadIdx.x; f < NT * win; f += kThreads) {
            const int t = f / win, e = f - t * win;
            int j, i, wy0, wx0;
            if (!geom(t, j, i, wy0, wx0)) continue;
            if (e == (j - wy0) * ws + (i - wx0)) continue;
            const float2 v = tab[((int64_t)(j - r0) * ncc + (i - c0)) * win + e];
            const float sig0 = sigma[(int64_t)j * nc + i];
            const float s0 = sig0 * sig0;
            const float sg1 = sigma[(int64_t)(wy0 + e / ws) * nc + (wx0 + e % ws)];
            const float s1 = sg1 * sg1;
            float dd = v.x;
            const float nf = v.y;
            dd -= nf * (s0 + s1);
            const float dnorm = (s1 + s0) * sqrtf(2.0f * nf);
            dd = dnorm > 1.e-8f ? dd / dnorm : 1.0e6f * nf;
            s_w[t][e] = dd;
        }
    }
    for (int it = wave; it < (tab ? 0 : NT * ngroups); it += kWaves) {
        const int t = it / ngroups, g = it - t * ngroups;
        int j, i, wy0, wx0;
        if (!geom(t, j, i, wy0, wx0)) continue;
        const int64_t p = (int64_t)j * nc + i;
        const int eself = (j - wy0) * ws + (i - wx0);
        int64_t q[NGW];
        int es[NGW];
#pragma unroll
        for (int k = 0; k < NGW; ++k) {
            const int e0 = g * NGW + k, tt = min(e0, ne - 1);
            const int e = tt < eself ? tt : tt + 1;
            es[k] = e0 < ne ? e : -1;
            q[k] = (int64_t)(wy0 + e / ws) * nc + (wx0 + e % ws);
        }
        double d2[NGW];
        int n2[NGW];
        wave_pair_sums<NGW>(pats, P, idx, M, p, q, sat, wk, d2, n2);
        if (lane == 0) {
            const float sig0 = sigma[p];
            const float s0 = sig0 * sig0;
#pragma unroll
            for (int k = 0; k < NGW; ++k) {
                if (es[k] < 0) continue;
                const float sg1 = sigma[q[k]];
                const float s1 = sg1 * sg1;
                float dd = (float)d2[k];
                const float nf = (float)n2[k];
                dd -= nf * (s0 + s1);
                const float dnorm = (s1 + s0) * sqrtf(2.0f * nf);
                dd = dnorm > 1.e-8f ? dd / dnorm : 1.0e6f * nf;
                s_w[t][es[k]] = dd;
            }
        }
    }
    __syncthreads();
    // ---- weights, one (point, entry) per thread; each point's sum in window order ----
    for (int f = threadIdx.x; f < NT * win; f += kThreads) {
        const int t = f / win, e = f - t * win;
        int j, i, wy0, wx0;
        if (!geom(t, j, i, wy0, wx0)) continue;
        const bool self = e == (j - wy0) * ws + (i - wx0);
        float w = self ? -1.0e6f : s_w[t][e];
        w = fmaxf(w - dthresh, 0.f);
        w = (float)exp(-1.0 * (double)w * lam2);
        if (self) w += diff_offset;
        s_w[t][e] = w;
    }
    __syncthreads();
    if (threadIdx.x < NT) {
        float sum = 0.f;
        for (int e = 0; e < win; ++e) sum += s_w[threadIdx.x][e];
        s_sum[threadIdx.x] = sum;
    }
    __syncthreads();
    for (int f = threadIdx.x; f < NT * win; f += kThreads) {
        const int t = f / win, e = f - t * win;
        int j, i, wy0, wx0;
        if (!geom(t, j, i, wy0, wx0)) continue;
        const float w = s_w[t][e] / s_sum[t];
        s_w[t][e] = w;
        if (wout) wout[((int64_t)(j - r0) * ncc + (i - c0)) * win + e] = w;
    }
    __syncthreads();
    // ---- weighted average over the union of the tile's windows ----
    int pj[NT], pi[NT], py[NT], px[NT];
    bool pv[NT];
    int uy0 = nr, uy1 = 0, ux0 = nc, ux1 = 0;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        pv[t] = geom(t, pj[t], pi[t], py[t], px[t]);
        if (pv[t]) {
            uy0 = min(uy0, py[t]);
            uy1 = max(uy1, py[t] + ws);
            ux0 = min(ux0, px[t]);
            ux1 = max(ux1, px[t] + ws);
        }
    }
    const int64_t P16 = P / 16;
    for (int64_t v = threadIdx.x; v < P16; v += kThreads) {
        float acc[NT][16];
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int k = 0; k < 16; ++k) acc[t][k] = 0.f;
        for (int uy = uy0; uy < uy1; ++uy) {
            for (int ux = ux0; ux < ux1; ++ux) {
                float x[16];
                load16(pats + ((int64_t)uy * nc + ux) * P + 16 * v, x);
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    if (pv[t] && uy >= py[t] && uy < py[t] + ws && ux >= px[t] && ux < px[t] + ws) {
                        const float w = s_w[t][(uy - py[t]) * ws + (ux - px[t])];
#pragma unroll
                        for (int k = 0; k < 16; ++k) acc[t][k] = fmaf(x[k], w, acc[t][k]);
                    }
                }
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t)
            if (pv[t]) store16(out + ((int64_t)(pj[t] - r0) * ncc + (pi[t] - c0)) * P + 16 * v, acc[t]);
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
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t head = 0;
    if (((uintptr_t)pats % 16) == 0) {
        // 16 values per thread per step from 16-byte loads, the tail element by element
        const int64_t n16 = n / 16;
        for (int64_t t = t0; t < n16; t += stride) {
            float x[16];
            load16(pats + 16 * t, x);
#pragma unroll
            for (int k = 0; k < 16; ++k) best = max(best, fkey(x[k]));
        }
        head = 16 * n16;
    }
    for (int64_t t = head + t0; t < n; t += stride) best = max(best, fkey(ldf(pats, t)));
    for (int off = 32; off >= 1; off >>= 1) best = max(best, (unsigned)__shfl_xor((int)best, off, 64));
    // one atomic per work-group (atomics on one address serialise)
    __shared__ unsigned s_best[kWaves];
    if ((threadIdx.x & 63) == 0) s_best[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned b = s_best[0];
        for (int w = 1; w < kWaves; ++w) b = max(b, s_best[w]);
        atomicMax(key, b);
    }
}

// in place: the key and the float share the 4 bytes
__global__ void k_key_to_float(unsigned *key, float *out)
{
    const unsigned k = *key;
    *out = fkey_inv(k);
}

__global__ void k_zero_key(unsigned *__restrict__ key) { *key = 0u; }

// ---- the options of calcnlpar (nlpar_cpu.py:283-540): stem_scale, backsub, rescale ----

// the smallest value as an order-preserving key (atomicMin from all ones)
template <typename T>
__global__ void k_min(const T *__restrict__ pats, int64_t n, unsigned *__restrict__ key)
{
    unsigned best = 0xffffffffu;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += stride)
        best = min(best, fkey(ldf(pats, t)));
    for (int off = 32; off >= 1; off >>= 1) best = min(best, (unsigned)__shfl_xor((int)best, off, 64));
    __shared__ unsigned s_best[kWaves];
    if ((threadIdx.x & 63) == 0) s_best[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned b = s_best[0];
        for (int w = 1; w < kWaves; ++w) b = min(b, s_best[w]);
        atomicMin(key, b);
    }
}

__global__ void k_ones_key(unsigned *__restrict__ key) { *key = 0xffffffffu; }

// float32 patterns as calcnlpar feeds them to its kernels: the file's values as float32;
// stem_scale: sqrt(v - dmin) (a float32 difference, the correctly rounded float32 root --
// the float64 root rounded to float32 is that, a 53-bit result rounded to 24 bits being
// innocuous for sqrt); backsub: float32(float64(v) - back[pixel]), as the in-place
// subtraction of a float64 background from a float32 array rounds.
template <typename T>
__global__ void k_prepare(const T *__restrict__ in, int64_t total, int64_t P, int stem, float dmin,
                          const double *__restrict__ back, float *__restrict__ out)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += stride) {
        float v = ldf(in, t);
        if (stem) {
            const float d = v - dmin;
            v = (float)sqrt((double)d);
        }
        if (back) v = (float)((double)v - back[t % P]);
        out[t] = v;
    }
}

// column sums over patterns in order, float32, one thread per pixel: acc[p] += x[i][p] for
// i = 0..n-1 -- numpy's mean over axis 0 of a float32 (n, H, W) array adds row after row
__global__ void k_colsum(const float *__restrict__ x, int64_t n, int64_t P, float *__restrict__ acc)
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= P) return;
    float a = acc[p];
    for (int64_t i = 0; i < n; ++i) a = a + x[i * P + p];
    acc[p] = a;
}

// One work-group per pattern: stem_scale's inverse v*v + dmin (float32, not fused), then
// rescale's v -= min(v); v *= f32(top / max(v)) over the pattern (float32, as
// nlpar_cpu.py:511-521), then the output type: float32, or clipped to [0, type max] and
// truncated (pat_flt2int 'clip', ebsd_pattern.py:159-162); NaN (a flat pattern rescaled)
// becomes 0.
template <typename O>
__global__ __launch_bounds__(kThreads) void k_finish(const float *__restrict__ x, int64_t P, int stem, float dmin,
                                                     float top, O *__restrict__ out, float omax)
{
    const float *src = x + blockIdx.x * P;
    O *dst = out + blockIdx.x * P;
    __shared__ unsigned s_key[kWaves];
    float mn = 0.0f, scale = 1.0f;
    if (top > 0.0f) {
        unsigned kmin = 0xffffffffu;
        for (int64_t i = threadIdx.x; i < P; i += kThreads) {
            float v = src[i];
            if (stem) v = __fadd_rn(__fmul_rn(v, v), dmin);
            kmin = min(kmin, fkey(v));
        }
        for (int off = 32; off >= 1; off >>= 1) kmin = min(kmin, (unsigned)__shfl_xor((int)kmin, off, 64));
        if ((threadIdx.x & 63) == 0) s_key[threadIdx.x >> 6] = kmin;
        __syncthreads();
        kmin = s_key[0];
        for (int w = 1; w < kWaves; ++w) kmin = min(kmin, s_key[w]);
        mn = fkey_inv(kmin);
        __syncthreads();
        unsigned kmax = 0u;
        for (int64_t i = threadIdx.x; i < P; i += kThreads) {
            float v = src[i];
            if (stem) v = __fadd_rn(__fmul_rn(v, v), dmin);
            kmax = max(kmax, fkey(__fsub_rn(v, mn)));
        }
        for (int off = 32; off >= 1; off >>= 1) kmax = max(kmax, (unsigned)__shfl_xor((int)kmax, off, 64));
        if ((threadIdx.x & 63) == 0) s_key[threadIdx.x >> 6] = kmax;
        __syncthreads();
        kmax = s_key[0];
        for (int w = 1; w < kWaves; ++w) kmax = max(kmax, s_key[w]);
        scale = (float)((double)top / (double)fkey_inv(kmax));
    }
    for (int64_t i = threadIdx.x; i < P; i += kThreads) {
        float v = src[i];
        if (stem) v = __fadd_rn(__fmul_rn(v, v), dmin);
        if (top > 0.0f) v = __fmul_rn(__fsub_rn(v, mn), scale);
        if (omax > 0.0f) {
            v = v != v ? 0.0f : fminf(fmaxf(v, 0.0f), omax);
            dst[i] = (O)v;
        } else {
            dst[i] = (O)v;
        }
    }
}

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
        const unsigned g = (unsigned)std::min<int64_t>(1024, (total + 16 * kThreads - 1) / (16 * kThreads));
        if (dtype == EBSD_NL_U8)
            hipLaunchKernelGGL(k_max<uint8_t>, dim3(g), dim3(kThreads), 0, s, (const uint8_t *)d_pats, total, key);
        else if (dtype == EBSD_NL_U16)
            hipLaunchKernelGGL(k_max<uint16_t>, dim3(g), dim3(kThreads), 0, s, (const uint16_t *)d_pats, total, key);
        else
            hipLaunchKernelGGL(k_max<float>, dim3(g), dim3(kThreads), 0, s, (const float *)d_pats, total, key);
    }
    hipLaunchKernelGGL(k_key_to_float, dim3(1), dim3(1), 0, s, key, d_max);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int ebsd_nlpar_min(const void *d_pats, int dtype, int64_t n, int64_t P, float *d_min, void *stream)
{
    if (n < 0 || P <= 0) return fail(-1, "ebsd_nlpar_min: bad size");
    if (!d_pats || !d_min) return fail(-1, "ebsd_nlpar_min: null pointer");
    if (dtype < 0 || dtype > 2) return fail(-1, "ebsd_nlpar_min: dtype is 0 (u8), 1 (u16) or 2 (f32)");
    hipStream_t s = (hipStream_t)stream;
    unsigned *key = reinterpret_cast<unsigned *>(d_min);
    hipLaunchKernelGGL(k_ones_key, dim3(1), dim3(1), 0, s, key);
    const int64_t total = n * P;
    if (total > 0) {
        const unsigned g = (unsigned)std::min<int64_t>(1024, (total + 16 * kThreads - 1) / (16 * kThreads));
        if (dtype == EBSD_NL_U8)
            hipLaunchKernelGGL(k_min<uint8_t>, dim3(g), dim3(kThreads), 0, s, (const uint8_t *)d_pats, total, key);
        else if (dtype == EBSD_NL_U16)
            hipLaunchKernelGGL(k_min<uint16_t>, dim3(g), dim3(kThreads), 0, s, (const uint16_t *)d_pats, total, key);
        else
            hipLaunchKernelGGL(k_min<float>, dim3(g), dim3(kThreads), 0, s, (const float *)d_pats, total, key);
    }
    hipLaunchKernelGGL(k_key_to_float, dim3(1), dim3(1), 0, s, key, d_min);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int ebsd_nlpar_prepare(const void *d_in, int dtype, int64_t n, int64_t P, int stem, float dmin,
                                  const double *d_back, float *d_out, void *stream)
{
    if (n < 0 || P <= 0) return fail(-1, "ebsd_nlpar_prepare: bad size");
    if (!d_in || !d_out) return fail(-1, "ebsd_nlpar_prepare: null pointer");
    if (dtype < 0 || dtype > 2) return fail(-1, "ebsd_nlpar_prepare: dtype is 0 (u8), 1 (u16) or 2 (f32)");
    const int64_t total = n * P;
    if (total == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    const unsigned g = (unsigned)std::min<int64_t>(8192, (total + kThreads - 1) / kThreads);
    if (dtype == EBSD_NL_U8)
        hipLaunchKernelGGL(k_prepare<uint8_t>, dim3(g), dim3(kThreads), 0, s, (const uint8_t *)d_in, total, P, stem,
                           dmin, d_back, d_out);
    else if (dtype == EBSD_NL_U16)
        hipLaunchKernelGGL(k_prepare<uint16_t>, dim3(g), dim3(kThreads), 0, s, (const uint16_t *)d_in, total, P,
                           stem, dmin, d_back, d_out);
    else
        hipLaunchKernelGGL(k_prepare<float>, dim3(g), dim3(kThreads), 0, s, (const float *)d_in, total, P, stem,
                           dmin, d_back, d_out);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int ebsd_nlpar_colsum(const float *d_x, int64_t n, int64_t P, float *d_acc, void *stream)
{
    if (n < 0 || P <= 0) return fail(-1, "ebsd_nlpar_colsum: bad size");
    if (!d_x || !d_acc) return fail(-1, "ebsd_nlpar_colsum: null pointer");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_colsum, dim3((unsigned)((P + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                       (hipStream_t)stream, d_x, n, P, d_acc);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int ebsd_nlpar_finish(const float *d_x, int64_t n, int64_t P, int stem, float dmin, float rescale_top,
                                 void *d_out, int out_dtype, void *stream)
{
    if (n < 0 || P <= 0) return fail(-1, "ebsd_nlpar_finish: bad size");
    if (!d_x || !d_out) return fail(-1, "ebsd_nlpar_finish: null pointer");
    if (out_dtype < 0 || out_dtype > 2) return fail(-1, "ebsd_nlpar_finish: out_dtype is 0 (u8), 1 (u16) or 2 (f32)");
    if (n == 0) return 0;
    if (n > 0x7fffffff) return fail(-1, "ebsd_nlpar_finish: too many patterns for one launch");
    hipStream_t s = (hipStream_t)stream;
    const float top = rescale_top > 0.0f ? rescale_top : 0.0f;
    if (out_dtype == EBSD_NL_U8)
        hipLaunchKernelGGL(k_finish<uint8_t>, dim3((unsigned)n), dim3(kThreads), 0, s, d_x, P, stem, dmin, top,
                           (uint8_t *)d_out, 255.0f);
    else if (out_dtype == EBSD_NL_U16)
        hipLaunchKernelGGL(k_finish<uint16_t>, dim3((unsigned)n), dim3(kThreads), 0, s, d_x, P, stem, dmin, top,
                           (uint16_t *)d_out, 65535.0f);
    else
        hipLaunchKernelGGL(k_finish<float>, dim3((unsigned)n), dim3(kThreads), 0, s, d_x, P, stem, dmin, top,
                           (float *)d_out, 0.0f);
    HIPCHK(hipGetLastError());
    return 0;
}

static size_t al256(size_t x) { return (x + 255) / 256 * 256; }

extern "C" size_t ebsd_nlpar_pair_bytes(int64_t npoints, int sr)
{
    if (npoints <= 0 || sr < 1 || sr > 10) return 0;
    return al256((size_t)npoints * (size_t)((2 * sr + 1) * (2 * sr + 1)) * sizeof(float2));
}

// the tiled NLPAR kernel (k_nlpar_tile) where it applies; ebsd_nlpar_set_tiled switches it
static bool g_nlpar_tiled = true;

extern "C" int ebsd_nlpar_set_tiled(int on)
{
    g_nlpar_tiled = on != 0;
    return 0;
}

extern "C" size_t ebsd_nlpar_workspace_bytes(int64_t n, int64_t P)
{
    if (n <= 0 || P <= 0) return 0;
    return al256((size_t)P) + al256((size_t)n * sizeof(double)) + al256((size_t)n) +
           al256((size_t)n * kSatList * sizeof(int32_t));
}

// Set up the uint8 workspace for n patterns (mask bytes, sums of squares, saturated pixels)
// when it applies: uint8 patterns, P a multiple of 4, 4-byte aligned patterns, a large
// enough workspace.  Otherwise wk stays empty and the kernels take the per-pixel path.
static int u8_setup(const void *d_pats, int dtype, int64_t n, int64_t P, const int32_t *d_idx, int M, float sat,
                    void *d_work, size_t work_bytes, hipStream_t s, U8Work *wk)
{
    *wk = U8Work{nullptr, nullptr, nullptr, nullptr};
    if (dtype == EBSD_NL_F32 || !d_work || P % 4 != 0 || ((uintptr_t)d_pats % 4) != 0) return 0;
    if (work_bytes < ebsd_nlpar_workspace_bytes(n, P) || ((uintptr_t)d_work % 256) != 0) return 0;
    if (n > INT32_MAX) return 0;
    uint8_t *base = (uint8_t *)d_work;
    uint8_t *mask = base;
    size_t off = al256((size_t)P);
    double *sq = (double *)(base + off);
    off += al256((size_t)n * sizeof(double));
    uint8_t *nsat = base + off;
    off += al256((size_t)n);
    int32_t *list = (int32_t *)(base + off);
    hipLaunchKernelGGL(k_mask_zero, dim3(grid_for(P)), dim3(256), 0, s, P, mask);
    if (M > 0) hipLaunchKernelGGL(k_mask_scatter, dim3(grid_for(M)), dim3(256), 0, s, d_idx, M, mask);
    // P is a multiple of 4 and the patterns 4-byte aligned here: the prepass reads words
    if (dtype == EBSD_NL_U8)
        hipLaunchKernelGGL(k_int_prep<uint8_t>, dim3((unsigned)n), dim3(kThreads), 0, s, (const uint8_t *)d_pats, n, P,
                           d_idx, M, (const uint32_t *)mask, sat, sq, nsat, list);
    else
        hipLaunchKernelGGL(k_int_prep<uint16_t>, dim3((unsigned)n), dim3(kThreads), 0, s, (const uint16_t *)d_pats, n,
                           P, d_idx, M, (const uint32_t *)mask, sat, sq, nsat, list);
    HIPCHK(hipGetLastError());
    *wk = U8Work{(const uint32_t *)mask, sq, nsat, list};
    return 0;
}

extern "C" int ebsd_nlpar_sigma(const void *d_pats, int dtype, int nr, int nc, int64_t P, const int32_t *d_idx, int M,
                                int nn, float sat, float *d_sigma, float *d_dout, float *d_nout, void *d_work,
                                size_t work_bytes, void *stream)
{
    if (nr <= 0 || nc <= 0 || P <= 0 || M < 0) return fail(-1, "ebsd_nlpar_sigma: bad size");
    if (nn < 1 || (2 * nn + 1) * (2 * nn + 1) > kMaxWin) return fail(-1, "ebsd_nlpar_sigma: nn must be 1..10");
    if (M > P) return fail(-1, "ebsd_nlpar_sigma: more mask pixels than pattern pixels");
    if (!d_pats || !d_idx || !d_sigma) return fail(-1, "ebsd_nlpar_sigma: null pointer");
    if ((d_dout == nullptr) != (d_nout == nullptr)) return fail(-1, "ebsd_nlpar_sigma: d_dout and d_nout go together");
    if (dtype < 0 || dtype > 2) return fail(-1, "ebsd_nlpar_sigma: dtype is 0 (u8), 1 (u16) or 2 (f32)");
    hipStream_t s = (hipStream_t)stream;
    const int64_t npts = (int64_t)nr * nc;
    const int64_t grid = (((int64_t)((nr + kBR - 1) / kBR) * ((nc + kBC - 1) / kBC) * (kBR * kBC) + 7) / 8) * 8;
    if (grid > INT32_MAX) return fail(-1, "ebsd_nlpar_sigma: chunk too large");
    U8Work wk;
    if (int rc = u8_setup(d_pats, dtype, npts, P, d_idx, M, sat, d_work, work_bytes, s, &wk)) return rc;
#define EBSD_NL_SIG(T)                                                                                     \
    hipLaunchKernelGGL(k_nlpar_sigma<T>, dim3((unsigned)grid), dim3(kThreads), 0, s, (const T *)d_pats, nr, nc, P, \
                       d_idx, M, nn, sat, wk, d_sigma, d_dout, d_nout)
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
                          int c0, int nrc, int ncc, void *d_out, int out_dtype, float *d_weights, void *d_work,
                          size_t work_bytes, void *stream)
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
    const int64_t items = (int64_t)((nrc + kBR - 1) / kBR) * ((ncc + kBC - 1) / kBC) * (kBR * kBC);
    const int64_t grid = ((items + 7) / 8) * 8;
    if (grid > INT32_MAX) return fail(-1, "ebsd_nlpar: region too large");
    hipStream_t s = (hipStream_t)stream;
    // lam2 = 1.0 / lam**2 (nlpar_cpu.py:844): float32 square, float64 reciprocal
    const double lam2 = 1.0 / (double)(lam * lam);
    // 16-pixel vector loads and stores when every pattern starts on a 16-byte boundary
    const int vec = (P % 16 == 0) && ((uintptr_t)d_pats % 16 == 0) && ((uintptr_t)d_out % 16 == 0);
    const bool tiled = g_nlpar_tiled;
    U8Work wk;
    if (int rc = u8_setup(d_pats, dtype, (int64_t)nr * nc, P, d_idx, M, sat, d_work, work_bytes, s, &wk)) return rc;
#define EBSD_NL_RUN(T, O)                                                                                   \
    hipLaunchKernelGGL((k_nlpar<T, O>), dim3((unsigned)grid), dim3(kThreads), 0, s, (const T *)d_pats, nr, nc, P, \
                       d_idx, M, d_sigma, sr, lam2, dthresh, diff_offset, sat, r0, c0, nrc, ncc, vec, wk, (O *)d_out, \
                       d_weights)
#define EBSD_NL_OUT(T)                 \
    if (out_dtype == EBSD_NL_U8)       \
        EBSD_NL_RUN(T, uint8_t);       \
    else if (out_dtype == EBSD_NL_U16) \
        EBSD_NL_RUN(T, uint16_t);      \
    else                               \
        EBSD_NL_RUN(T, float);
    if (vec && tiled) {
        // 2 x 2 points per work-group (k_nlpar_tile); the same results.  With room in the
        // workspace after the integer part, each pair's distance is computed once first.
        float2 *tab = nullptr;
        const size_t base_bytes = al256(ebsd_nlpar_workspace_bytes((int64_t)nr * nc, P));
        if (d_work && ((uintptr_t)d_work % 256) == 0 && work_bytes >= base_bytes + ebsd_nlpar_pair_bytes(npts, sr))
            tab = (float2 *)((uint8_t *)d_work + base_bytes);
        if (tab) {
            // an interior point owns sr (sr + 1) / 2 * 4 pairs: groups of NGW pairs so that
            // each of the four wavefronts takes about one
#define EBSD_NLP_RUN(T, G)                                                                                      \
    hipLaunchKernelGGL((k_nlpar_pairs<T, G>), dim3((unsigned)grid), dim3(kThreads), 0, s, (const T *)d_pats, nr, \
                       nc, P, d_idx, M, sr, sat, r0, c0, nrc, ncc, wk, tab)
#define EBSD_NLP_G(T)            \
    if (sr == 1)                 \
        EBSD_NLP_RUN(T, 1);      \
    else if (sr == 2)            \
        EBSD_NLP_RUN(T, 3);      \
    else if (sr == 3)            \
        EBSD_NLP_RUN(T, 6);      \
    else                         \
        EBSD_NLP_RUN(T, 8);
            if (dtype == EBSD_NL_U8) {
                EBSD_NLP_G(uint8_t)
            } else if (dtype == EBSD_NL_U16) {
                EBSD_NLP_G(uint16_t)
            } else {
                EBSD_NLP_G(float)
            }
#undef EBSD_NLP_G
#undef EBSD_NLP_RUN
        }
        const int ntr = (nrc + 1) / 2, ntc = (ncc + 1) / 2;
        const int64_t titems = (int64_t)((ntr + kTBR - 1) / kTBR) * ((ntc + kTBC - 1) / kTBC) * (kTBR * kTBC);
        const int64_t tgrid = ((titems + 7) / 8) * 8;
        if (tgrid > INT32_MAX) return fail(-1, "ebsd_nlpar: region too large");
#define EBSD_NLT_RUN(T, O)                                                                                           \
    hipLaunchKernelGGL((k_nlpar_tile<T, O, 2, 2>), dim3((unsigned)tgrid), dim3(kThreads), 0, s, (const T *)d_pats, nr, \
                       nc, P, d_idx, M, d_sigma, sr, lam2, dthresh, diff_offset, sat, r0, c0, nrc, ncc, wk, tab,     \
                       (O *)d_out, d_weights)
#define EBSD_NLT_OUT(T)                \
    if (out_dtype == EBSD_NL_U8)       \
        EBSD_NLT_RUN(T, uint8_t);      \
    else if (out_dtype == EBSD_NL_U16) \
        EBSD_NLT_RUN(T, uint16_t);     \
    else                               \
        EBSD_NLT_RUN(T, float);
        if (dtype == EBSD_NL_U8) {
            EBSD_NLT_OUT(uint8_t)
        } else if (dtype == EBSD_NL_U16) {
            EBSD_NLT_OUT(uint16_t)
        } else {
            EBSD_NLT_OUT(float)
        }
#undef EBSD_NLT_OUT
#undef EBSD_NLT_RUN
    } else if (dtype == EBSD_NL_U8) {
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
