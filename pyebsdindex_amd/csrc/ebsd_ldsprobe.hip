// ebsd_ldsprobe.hip -- measurement only: the LDS read ceiling and the vector-instruction
// issue ceiling of the GPU the library runs on.
//
// The Radon gather (k_radon_gather) is bound by LDS reads and vector issue, so its roofline
// needs the LDS bandwidth this GPU actually delivers.  k_lds_read has every lane of 16-wave
// work-groups read words of W = 4, 8 or 16 bytes at conflict-free addresses (the 64 lanes of
// a wave read 64 consecutive words of a 64 KiB LDS array), 16 reads per loop step with
// immediate offsets, summed so that no read is dead.  The
// entry point times `iters` steps of `blocks` work-groups with HIP events and returns the
// bytes read per second.  No counterpart in the reference (a measurement of the hardware).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ebsd_probe.h"

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int W> struct ProbeWord;
template <> struct ProbeWord<4> { typedef uint32_t T; };
template <> struct ProbeWord<8> { typedef u32x2 T; };
template <> struct ProbeWord<16> { typedef u32x4 T; };

// one v_add3_u32 per 8 bytes read: the vector work stays below the LDS time
__device__ __forceinline__ uint32_t fold(uint32_t a, uint32_t w) { return a + w; }
__device__ __forceinline__ uint32_t fold(uint32_t a, u32x2 w) { return a + w.x + w.y; }
__device__ __forceinline__ uint32_t fold(uint32_t a, u32x4 w) { return (a + w.x + w.y) + w.z + w.w; }

constexpr int kProbeLdsBytes = 64 * 1024;

template <int W>
__global__ __launch_bounds__(1024) void k_lds_read(int iters, uint32_t *sink)
{
    typedef typename ProbeWord<W>::T Word;
    constexpr int NW = kProbeLdsBytes / W;     // words in the array
    // 16 reads per step from one base: lane l reads word base + k SP + l of read k.  SP = 257
    // words for 4- and 8-byte words keeps the compiler from pairing reads into ds_read2 forms
    // (their offsets span at most 255 words, or multiples of 64); 16-byte reads have no pair form
    constexpr int SP = W == 16 ? 64 : 257;
    constexpr int LIMIT = ((NW - 15 * SP - 64) / 64) * 64; // bases in [0, LIMIT) stay in bounds
    static_assert(LIMIT > 0, "the array holds a step");
    __shared__ Word buf[NW];
    for (int i = threadIdx.x; i < NW; i += blockDim.x) {
        Word v;
        __builtin_memset(&v, 0, sizeof(v));
        *reinterpret_cast<uint32_t *>(&v) = (uint32_t)i * 2654435761u;
        buf[i] = v;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the base moves by 64 words per step, wrapping below LIMIT, so no read can be hoisted
    int base = (wave * 64) % LIMIT;
    uint32_t acc = 0u;
    for (int it = 0; it < iters; ++it) {
        const Word *row = buf + base + lane;
#pragma unroll
        for (int k = 0; k < 16; ++k) acc = fold(acc, row[k * SP]);
        base += 64;
        if (base >= LIMIT) base -= LIMIT;
    }
    if (acc == 0x9e3779b9u) sink[blockIdx.x] = acc; // almost never: keeps the reads live
}

extern "C" int ebsd_probe_lds_read(int width_bytes, int blocks, int iters, double *bytes_per_s, void *stream)
{
    if (!bytes_per_s || blocks <= 0 || iters <= 0) return -1;
    hipStream_t s = (hipStream_t)stream;
    uint32_t *sink = nullptr;
    if (hipMalloc((void **)&sink, (size_t)blocks * sizeof(uint32_t)) != hipSuccess) return -3;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = -3;
    auto launch = [&](int n) {
        if (width_bytes == 4) hipLaunchKernelGGL(k_lds_read<4>, dim3(blocks), dim3(1024), 0, s, n, sink);
        else if (width_bytes == 8) hipLaunchKernelGGL(k_lds_read<8>, dim3(blocks), dim3(1024), 0, s, n, sink);
        else hipLaunchKernelGGL(k_lds_read<16>, dim3(blocks), dim3(1024), 0, s, n, sink);
    };
    if (!rc && width_bytes != 4 && width_bytes != 8 && width_bytes != 16) rc = -1;
    if (!rc) {
        launch(iters);  // warm-up
        float ms = 0.f;
        if (hipEventRecord(e0, s) != hipSuccess) rc = -3;
        launch(iters);
        if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess || hipGetLastError() != hipSuccess))
            rc = -3;
        if (!rc) *bytes_per_s = (double)blocks * 1024.0 * 16.0 * (double)iters * width_bytes / (ms * 1e-3);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(sink);
    return rc;
}

// VALU issue ceiling: every lane of 256-thread work-groups (8 per CU at 256 CUs fill every
// SIMD with 8 waves) runs 16 independent chains per step -- kind 0: x = x * a + b, one
// v_fma_f32 each (16 vector instructions a step); kind 1: y = (y ^ a) + b, a v_xor_b32 and a
// v_add_u32 each (32); kind 2: y = max(y, next lane's y) - b, a v_max_i32 with a DPP source
// and a v_sub_u32 each (32); kind 3: the float64 x * a + b, one v_fma_f64 each (16) -- which
// the compiler cannot fold across steps.  The loop control is scalar.
template <int KIND>
__global__ __launch_bounds__(256) void k_valu_issue(int iters, float fa, float fb, uint32_t *sink)
{
    float x[16];
    uint32_t y[16];
    double z[16];
    const uint32_t ua = __float_as_uint(fa), ub = __float_as_uint(fb);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        x[k] = (float)(threadIdx.x + k);
        y[k] = threadIdx.x * 2654435761u + (uint32_t)k;
        z[k] = (double)x[k];
    }
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if constexpr (KIND == 0) x[k] = __builtin_fmaf(x[k], fa, fb);
            else if constexpr (KIND == 1) y[k] = (y[k] ^ ua) + ub;
            else if constexpr (KIND == 2) // the maximum with the next lane's value: v_max_i32 with a DPP source
                y[k] = max((int)y[k], __builtin_amdgcn_update_dpp(0, (int)y[k], 0x130, 0xf, 0xf, true)) - (int)ub;
            else z[k] = __builtin_fma(z[k], (double)fa, (double)fb);
        }
    }
    uint32_t acc = 0u;
#pragma unroll
    for (int k = 0; k < 16; ++k)
        acc += KIND == 0 ? __float_as_uint(x[k]) : (KIND == 3 ? (uint32_t)__double2loint(z[k]) : y[k]);
    if (acc == 0x9e3779b9u) sink[blockIdx.x] = acc; // almost never: keeps the chains live
}

extern "C" int ebsd_probe_valu_issue(int kind, int blocks, int iters, double *instr_per_s, void *stream)
{
    if (!instr_per_s || blocks <= 0 || iters <= 0 || kind < 0 || kind > 3) return -1;
    hipStream_t s = (hipStream_t)stream;
    uint32_t *sink = nullptr;
    if (hipMalloc((void **)&sink, (size_t)blocks * sizeof(uint32_t)) != hipSuccess) return -3;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = -3;
    auto launch = [&](int n) {
        if (kind == 0) hipLaunchKernelGGL(k_valu_issue<0>, dim3(blocks), dim3(256), 0, s, n, 0.99999f, 0.5f, sink);
        else if (kind == 1) hipLaunchKernelGGL(k_valu_issue<1>, dim3(blocks), dim3(256), 0, s, n, 0.99999f, 0.5f, sink);
        else if (kind == 2) hipLaunchKernelGGL(k_valu_issue<2>, dim3(blocks), dim3(256), 0, s, n, 0.99999f, 0.5f, sink);
        else hipLaunchKernelGGL(k_valu_issue<3>, dim3(blocks), dim3(256), 0, s, n, 0.99999f, 0.5f, sink);
    };
    if (!rc) {
        launch(iters); // warm-up
        float ms = 0.f;
        if (hipEventRecord(e0, s) != hipSuccess) rc = -3;
        launch(iters);
        if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess || hipGetLastError() != hipSuccess))
            rc = -3;
        // wave-instructions: blocks x 4 waves x (16 or 32) per step
        if (!rc) *instr_per_s = (double)blocks * 4.0 * (kind == 0 || kind == 3 ? 16.0 : 32.0) * (double)iters / (ms * 1e-3);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(sink);
    return rc;
}
