// ebsd_radon.hip -- MI355X (gfx950) band detection: fast Radon gather-sum, Radon-space
// padding + convolution, masked local-max peak search with sub-pixel fit.
//
// Reference path: PyEBSDIndex BandDetect.find_bands (pyebsdindex/band_detect.py:411-526)
// on its CPU/Numba backend: radon_fast.py:242-310, band_detect.py:562-809.
//
// Two kernels per chunk of patterns:
//   k_radon_gather  one work-group holds P patterns interleaved pixel-major in LDS (one
//                   LDS word = the same pixel of P patterns) and walks the compiled index
//                   plan: every lane owns one Radon line, reads its 8-entry plan group with
//                   one 16-byte global load and gathers P patterns per LDS read.
//   k_peaks         one work-group per pattern, theta-tiled: stage the padded Radon tile,
//                   separable convolution, theta/rho dilation, equality local max under the
//                   rho mask, candidate merge into a running top-nBands list, then the
//                   reference's width / 3x3 sub-pixel fit / normmax / theta-rho conversion.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>
#include <mutex>
#include <vector>

#include "ebsd_radon.h"

// ---------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return fail(-3, std::string(#x) + ": " + hipGetErrorString(e_));              \
    } while (0)

extern "C" const char *ebsd_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------------------
static constexpr int KMAX = 16;      // max separable taps per axis
static constexpr int CAND_CAP = 256;  // candidates per tile pass held in LDS
static constexpr int NB_MAX = 64;     // max bands per pattern

struct PeakGeo {
    int nR, nT, p0, p1, nRp, nTp;
    int kr, kt, hr, ht;   // conv sizes, dilation half-windows
    int cr0, cr1;         // convolved rows kept (padded coordinates)
    int mr0, mr1;         // rows that can hold a candidate
    int sr0, sr1, sc0, sc1; // stats region
    int tile, ntiles;
    int rh, apitch, bpitch, maskw; // LDS tile rows, pitches of the two tile buffers, mask words per row
    int nB;
    int sbw;              // streamed peak kernel: theta columns per wavefront (0: not used)
    float u[KMAX], v[KMAX]; // C = sum_a u[a] sum_b v[b] Rpad[r + kr/2 - a][c + kt/2 - b]
};

struct ebsd_plan {
    int device = 0, ncu = 0, lds_max = 0;
    int H = 0, W = 0, nPx = 0, nT = 0, nR = 0, nCells = 0;
    // compiled gather plan
    int spitch = 0;              // LDS words per staged pattern row (>= W); plan entries use it
    int lunit = 8;               // lines ordered in units of this many neighbouring cells
    std::vector<int> h_cnt;      // per cell entry count
    std::vector<int64_t> h_off;  // per cell first entry in h_ent
    std::vector<uint16_t> h_ent; // per cell valid entries, plan order
    // device streams, one layout per count of (virtual) wave slots: a work-group of nw waves
    // that splits each pack's lines S ways walks them as nw * S slots
    struct Layout {
        int nv, kind; // kind 0: entries in plan order, 1..4: bank-scheduled for LDS words of
                      // 2, 4, 8 or 16 bytes (sched_kind)
        uint4 *gidx;
        int *wofs, *wcnt;
    };
    // The lines' entries in staged rows [r0, r1), cut into 64-line batches: the whole pattern,
    // or one of the row segments whose integer partial sums add up to the line sums.
    struct GSeg {
        int r0 = 0, r1 = 0;
        int nwords = 0;              // (r1 - r0) * spitch: the zero word's index
        int nbatch = 0;
        uint2 *d_bhdr = nullptr;     // per batch: first block, number of blocks
        int2 *d_lane = nullptr;      // per batch lane: cell id (-1 = none), entries of the whole line
        double *d_lrc = nullptr;     // per batch lane: 1 / entries (IEEE, 0 for an empty line)
        std::vector<uint16_t> h_groups[5]; // batch-major groups [block][lane][8]: plan order, then
                                           // bank-scheduled per word size (built on first use)
        std::vector<uint2> h_bhdr;   // per batch: first block in h_groups, number of blocks
        std::vector<Layout> layouts;
    };
    GSeg full;                   // all rows
    std::vector<GSeg> rowsets[9]; // row segments by their count S (2..8), compiled on first use and
                                  // kept: a batch's launches may hold pointers into any of them
    int rseg = 0;                // row segments for integer sums: 0 automatic, else 1..8
    uint32_t *d_part = nullptr;  // row segments' partial sums, part_cap elements (grown on demand;
    size_t part_cap = 0;         // calls on one plan are ordered on the caller's stream)
    std::vector<uint32_t *> part_retired; // outgrown partial-sum buffers, freed with the plan:
                                          // launches still queued on any stream may use them
    int pins = 0;                // batches (ebsd_batch_create) holding pointers into this plan
    void *scratch = nullptr;     // device address of this GPU's no-allocation scratch array
    // background
    int bg_mode = 0;             // 0 none, 1 per-line sums, 2 exact per-pixel
    double *d_linebg = nullptr;  // nCells
    double *d_bg = nullptr;      // nPx
    // peaks
    PeakGeo g{};
    uint8_t *d_mask = nullptr;   // nR*nT
    std::vector<uint8_t> h_rdnmask; // nR*nT, the peak mask (host copy for the lane-mask tables)
    // streamed peak kernel: per theta block width sbw, the rho-mask as 64-bit lane masks,
    // [wavefront][padded row d + 8] for d in [-8, nRp + 8); built when sbw is set, kept until
    // the plan is destroyed (launches queued on any stream may hold an older one)
    std::map<int, uint64_t *> mlane_tabs;
    uint64_t *d_mlane = nullptr; // the table of the current sbw
    uint32_t *d_mrow = nullptr;  // rho-mask bits per rho row (PeakGeo::maskw words each)
    // adjoint structures, built on first use
    int *d_px_off = nullptr, *d_px_cell = nullptr; // pixel -> lines through it (CSR)
    uint32_t *d_bp_tab = nullptr; // the same, padded to bp_m2 pairs of u16 cells per pixel,
    int bp_m2 = 0;                // pair-major [bp_m2][nPx] (0: not built, nCells >= 65535)
    double *d_icnt = nullptr;                      // 1 / pixels per line
    int *d_pre_off = nullptr, *d_pre = nullptr;    // Radon cell -> padded cells copying it (CSR)
    double *d_tscr = nullptr;    // transposed convolution: a padded array of float64 per work-group,
    size_t tscr_cap = 0;         //   tscr_cap doubles (grown on demand)
    double *d_grid = nullptr;    // theta_deg[nT] then rho[nR], as float64
    // tuning
    int gpack_u8 = 8, gpack_u16 = 4, gpack_f32 = 2, gthreads = 1024;
    int ptile = 45, pthreads = 512;
    int pkind = 0; // peak kernel: 0 automatic, 1 tiled (k_peaks), 2 streamed (k_peaks_stream)
};

// n / d for n, d < 2^16 by multiply-high (m = ceil(2^32 / d)).
struct FDiv {
    uint32_t d, m;
};
__device__ __forceinline__ FDiv fdiv_make(uint32_t d)
{
    FDiv f;
    f.d = d;
    f.m = d > 1 ? 0xFFFFFFFFu / d + 1u : 0u;
    return f;
}
__device__ __forceinline__ uint32_t fdiv_q(uint32_t n, FDiv f) { return f.d > 1 ? __umulhi(n, f.m) : n; }

// ---------------------------------------------------------------------------------------
// k_radon_gather
// ---------------------------------------------------------------------------------------
struct GatherArgs {
    const void *pats;
    int64_t N;
    int nPx;
    int W, pitch;         // pattern row length; LDS words per staged row (>= W)
    int px0, rows;        // staged rows [px0 / W, px0 / W + rows): all rows, or one row segment
    int segmode;          // 0 whole lines; row segments: 1 first (store the partial sums),
                          // 2 middle (add), 3 last (add and finish the line values).  Integer
                          // sums add the earlier segments' sum at the line end; float64 sums
                          // start from it, so the sum runs in plan order across segments
    int nbatch;
    const uint4 *gidx;    // per-wave streams of plan groups
    const uint2 *bhdr;    // per batch: (unused, groups)
    const int2 *lane;     // per batch lane: (cell, count)
    const double *lrc;    // per batch lane: 1 / count, correctly rounded
    const int *wofs;      // per wave slot: first group of its stream
    const int *wcnt;      // per wave slot: groups in its stream
    int nslice;           // work-groups sharing a pack's lines (slot = slice * waves + wave)
    const double *linebg; // fast background mode
    const double *bg;     // exact background mode
    float *out;           // (N, nCells)
    uint32_t *part;       // row segments: partial sums (N, npos) by batch position, uint32
                          // (integer sums) or float64 (float sums, two words each)
    int npos;             //   (batch * 64 + lane; the segments share the line order)
    int nCells;
    int64_t npacks;
};

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
template <int WB> struct WordT;
template <> struct WordT<2> { typedef uint16_t T; };
template <> struct WordT<4> { typedef uint32_t T; };
template <> struct WordT<8> { typedef uint64_t T; };
template <> struct WordT<16> { typedef u32x4 T; };

template <int IN> struct InT;
template <> struct InT<EBSD_U8> { typedef uint8_t T; };
template <> struct InT<EBSD_U16> { typedef uint16_t T; };
template <> struct InT<EBSD_F32> { typedef float T; };

// Word layout: FLT=0: P integers of 8 (u8) or 16 (u16) bits packed little-endian;
//              FLT=1: P float32 values of float32(x - bg) (reference staging).
template <int IN, bool FLT, int P> struct GatherCfg {
    static constexpr int EB = FLT ? 4 : (IN == EBSD_U8 ? 1 : 2); // bytes per element
    static constexpr int WB = EB * P;                             // bytes per LDS word
    static constexpr int ND = WB >= 4 ? WB / 4 : 1;               // dwords per word (a 2-byte word in 1)
    static constexpr int VEC = 8 / (int)sizeof(typename InT<IN>::T); // pixels per 8-byte load
};

// LDS byte address of the low / high 16-bit plan entry of a dword, times the word size:
// v_mad_u32_u16 reads one 16-bit half (op_sel picks the high half) -- one VALU op per entry.
// The LDS base of the staged words is folded into the same instruction.
template <int WB> __device__ __forceinline__ uint32_t entry_addr_lo(uint32_t e, uint32_t base)
{
    uint32_t r;
    asm("v_mad_u32_u16 %0, %1, %2, %3" : "=v"(r) : "v"(e), "n"(WB), "s"(base));
    return r;
}
template <int WB> __device__ __forceinline__ uint32_t entry_addr_hi(uint32_t e, uint32_t base)
{
    uint32_t r;
    asm("v_mad_u32_u16 %0, %1, %2, %3 op_sel:[1,0,0,0]" : "=v"(r) : "v"(e), "n"(WB), "s"(base));
    return r;
}

template <typename T> __device__ __forceinline__ T lds_load(uint32_t addr)
{
    return *reinterpret_cast<const __attribute__((address_space(3))) T *>((size_t)addr);
}

// acc += field SEL of d (SEL 0..3: byte, 4/5: low/high 16 bits) -- one SDWA add.
template <int SEL> __device__ __forceinline__ void add_field(uint32_t &acc, uint32_t d)
{
    if constexpr (SEL == 0)
        asm("v_add_u32_sdwa %0, %1, %0 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_0 src1_sel:DWORD" : "+v"(acc) : "v"(d));
    else if constexpr (SEL == 1)
        asm("v_add_u32_sdwa %0, %1, %0 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_1 src1_sel:DWORD" : "+v"(acc) : "v"(d));
    else if constexpr (SEL == 2)
        asm("v_add_u32_sdwa %0, %1, %0 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_2 src1_sel:DWORD" : "+v"(acc) : "v"(d));
    else if constexpr (SEL == 3)
        asm("v_add_u32_sdwa %0, %1, %0 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_3 src1_sel:DWORD" : "+v"(acc) : "v"(d));
    else if constexpr (SEL == 4)
        asm("v_add_u32_sdwa %0, %1, %0 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:WORD_0 src1_sel:DWORD" : "+v"(acc) : "v"(d));
    else
        asm("v_add_u32_sdwa %0, %1, %0 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:WORD_1 src1_sel:DWORD" : "+v"(acc) : "v"(d));
}

template <int IN, bool FLT, int P, int K = 0>
__device__ __forceinline__ void accumulate_word(const uint32_t *d, uint32_t *acc, double *accd)
{
    if constexpr (K < P) {
        if constexpr (FLT) {
            accd[K] += (double)__uint_as_float(d[K]);
        } else if constexpr (IN == EBSD_U8) {
            add_field<K & 3>(acc[K], d[K >> 2]);
        } else {
            add_field<4 + (K & 1)>(acc[K], d[K >> 1]);
        }
        accumulate_word<IN, FLT, P, K + 1>(d, acc, accd);
    }
}

// The LDS words of one plan group (8 entries of a line) are issued one group ahead of
// their accumulation, so the LDS and VALU work of consecutive groups overlap in a wave.
template <int IN, bool FLT, int P>
__device__ __forceinline__ void group_issue(uint32_t base, const uint4 &e,
                                            typename WordT<GatherCfg<IN, FLT, P>::WB>::T *w)
{
    typedef GatherCfg<IN, FLT, P> C;
    typedef typename WordT<C::WB>::T Word;
    const uint32_t ew[4] = {e.x, e.y, e.z, e.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        w[2 * k] = lds_load<Word>(entry_addr_lo<C::WB>(ew[k], base));
        w[2 * k + 1] = lds_load<Word>(entry_addr_hi<C::WB>(ew[k], base));
    }
}

// Integer words of P >= 2 patterns: the accumulator of the first field of each dword j sums
// whole dwords (two entries per v_add3_u32), the other fields' accumulators sum their fields
// (one SDWA add each per entry).  The first field's sum is the dword sum minus the other
// fields' sums shifted into place, mod 2^32 (gather_decode): for uint16 1.5 VALU ops per dword
// and entry instead of 2, for uint8 3.5 instead of 4.
template <int IN, bool FLT, int P> struct PairSum {
    static constexpr bool value = !FLT && IN != EBSD_F32 && P >= 2;
    static constexpr int F = IN == EBSD_U8 ? 4 : 2; // fields per dword
    static constexpr int B = IN == EBSD_U8 ? 8 : 16; // bits per field
};

template <int IN, int P, int J, int I>
__device__ __forceinline__ void add_fields(uint32_t *acc, uint32_t da, uint32_t db)
{
    typedef PairSum<IN, false, P> S;
    if constexpr (I < S::F && J * S::F + I < P) {
        constexpr int SEL = IN == EBSD_U8 ? I : 4 + I;
        add_field<SEL>(acc[J * S::F + I], da);
        add_field<SEL>(acc[J * S::F + I], db);
        add_fields<IN, P, J, I + 1>(acc, da, db);
    }
}

template <int IN, int P, int J = 0>
__device__ __forceinline__ void add_pair(uint32_t *acc, const uint32_t *da, const uint32_t *db)
{
    typedef PairSum<IN, false, P> S;
    if constexpr (J * S::F < P) {
        acc[J * S::F] += da[J] + db[J];
        add_fields<IN, P, J, 1>(acc, da[J], db[J]);
        add_pair<IN, P, J + 1>(acc, da, db);
    }
}

template <int IN, bool FLT, int P>
__device__ __forceinline__ void group_accumulate(const typename WordT<GatherCfg<IN, FLT, P>::WB>::T *w,
                                                 uint32_t *acc, double *accd)
{
    typedef GatherCfg<IN, FLT, P> C;
    typedef typename WordT<C::WB>::T Word;
    if constexpr (PairSum<IN, FLT, P>::value) {
#pragma unroll
        for (int k = 0; k < 8; k += 2) {
            uint32_t da[C::ND] = {}, db[C::ND] = {};
            __builtin_memcpy(da, &w[k], sizeof(Word)); // no type punning through pointers
            __builtin_memcpy(db, &w[k + 1], sizeof(Word));
            add_pair<IN, P>(acc, da, db);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            uint32_t d[C::ND] = {};
            __builtin_memcpy(d, &w[k], sizeof(Word)); // no type punning through pointers
            accumulate_word<IN, FLT, P>(d, acc, accd);
        }
    }
}

// The P pattern sums from the accumulators (identity except for PairSum words).
template <int IN, bool FLT, int P> __device__ __forceinline__ void gather_decode(uint32_t *acc)
{
    typedef PairSum<IN, FLT, P> S;
    if constexpr (S::value) {
#pragma unroll
        for (int j = 0; j * S::F < P; ++j)
#pragma unroll
            for (int i = 1; i < S::F && j * S::F + i < P; ++i)
                acc[j * S::F] -= acc[j * S::F + i] << (S::B * i);
    }
}

// Narrow sums of uint16 words (P >= 2): when a pack's largest value V is at most 8,191, the
// two 16-bit fields of a dword can be summed over F >= 1 whole plan groups (8 F V <= 65,535)
// without a carry from the low field into the high one.  A per-dword accumulator then adds
// two entries per v_add3_u32, and a flush every F groups (and at each line end) moves its two
// fields into the per-pattern sums with two SDWA adds: 2 + 2/F VALU ops per dword and entry
// pair instead of 6.
template <int IN, bool FLT, int P> struct Narrow {
    static constexpr bool ok = !FLT && IN == EBSD_U16 && P >= 2;
    static constexpr int ND = P / 2; // dwords per word
};

// FRESH: the group's sums start from zero (a flush after every group), else they add to dacc.
template <int P, bool FRESH>
__device__ __forceinline__ void group_accumulate_narrow(const typename WordT<2 * P>::T *w, uint32_t *dacc)
{
    typedef typename WordT<2 * P>::T Word;
    constexpr int ND = P / 2;
#pragma unroll
    for (int k = 0; k < 8; k += 2) {
        uint32_t da[ND], db[ND];
        __builtin_memcpy(da, &w[k], sizeof(Word));
        __builtin_memcpy(db, &w[k + 1], sizeof(Word));
#pragma unroll
        for (int j = 0; j < ND; ++j) dacc[j] = (FRESH && k == 0) ? da[j] + db[j] : dacc[j] + da[j] + db[j];
    }
}

// Moves the fields of the dword sums into the per-pattern sums; RESET: zero the dword sums
// for groups that add to them.
template <int P, bool RESET> __device__ __forceinline__ void narrow_flush(uint32_t *dacc, uint32_t *acc)
{
#pragma unroll
    for (int j = 0; j < P / 2; ++j) {
        add_field<4>(acc[2 * j], dacc[j]);
        add_field<5>(acc[2 * j + 1], dacc[j]);
        if constexpr (RESET) dacc[j] = 0u;
    }
}

// Groups per flush of the narrow sums for a pack whose largest value is vmax (0: the narrow
// sums do not apply).
__device__ __forceinline__ int narrow_groups(uint32_t vmax)
{
    if (vmax > 8191u) return 0;
    return vmax == 0u ? (1 << 20) : (int)(65535u / (8u * vmax));
}

// Extra LDS bytes of a gather work-group past its staged words: the per-wave maxima of the
// staged values (16 waves of 4 bytes, 16-byte aligned).
static constexpr int kGatherLdsExtra = 128;

// Build the LDS word of pixel px from element values x[p] (p < P), as packed integers or
// float32(x - bg).
template <int IN, bool FLT, int P>
__device__ __forceinline__ void pack_word(uint32_t *d, const typename InT<IN>::T *x, const double *bg, int px)
{
    typedef GatherCfg<IN, FLT, P> C;
#pragma unroll
    for (int k = 0; k < C::ND; ++k) d[k] = 0u;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        if constexpr (FLT) {
            const float f = bg ? (float)((double)x[p] - bg[px]) : (float)x[p];
            d[p] = __float_as_uint(f);
        } else if constexpr (IN == EBSD_U8) {
            d[p >> 2] |= (uint32_t)x[p] << (8 * (p & 3));
        } else {
            d[p >> 1] |= (uint32_t)x[p] << (16 * (p & 1));
        }
    }
}

// One pack of P patterns: stage them in LDS, then every wave walks its stream of plan
// groups and writes the P Radon values of each of its lines.
template <int IN, bool FLT, int P>
__device__ __forceinline__ void gather_pack(const GatherArgs &a, int64_t pack, int slice, unsigned char *lds_raw)
{
    typedef GatherCfg<IN, FLT, P> C;
    typedef typename WordT<C::WB>::T Word;
    typedef typename InT<IN>::T Tin;
    constexpr int VEC = C::VEC;
    Word *lds = reinterpret_cast<Word *>(lds_raw);

    const Tin *pats = reinterpret_cast<const Tin *>(a.pats);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    // the wave index is the same on every lane: said so, the batch walk's indices, bounds
    // and line-end tests stay scalar (otherwise each test is an exec-masked vector branch)
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nwaves = blockDim.x >> 6;
    const int nPx = a.nPx;
    const int nst = a.rows * a.W; // staged pixels: global pixel a.px0 + j at local j
    const int nb = a.nbatch;
    const bool vec_ok = (a.W % VEC) == 0 && ((uintptr_t)a.pats % 8) == 0;
    const int nvec = vec_ok ? nst / VEC : 0;
    // pixel px is staged at word (px / W) * pitch + px % W (the plan's entries use the same map)
    const FDiv fW = fdiv_make(a.W);
    auto word_of = [&](int px) { const int y = (int)fdiv_q(px, fW); return y * a.pitch + (px - y * a.W); };
    const uint4 *gbase = a.gidx + lane;
    // LDS address of the staged words (the entries are word indices)
    const uint32_t lbase = (uint32_t)(size_t)(const __attribute__((address_space(3))) unsigned char *)(lds_raw);

    {
        const int64_t q0 = pack * P;
        constexpr bool NOK = Narrow<IN, FLT, P>::ok;
        uint32_t vmax = 0u; // largest staged value (narrow sums)
        // ---- stage P patterns pixel-major (8-byte loads per pattern), zero word at nst ----
        const Tin *src = pats + a.px0;
        for (int v = tid; v < nvec; v += blockDim.x) {
            Tin x[P][VEC];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                uint2 raw = make_uint2(0u, 0u);
                if (q0 + p < a.N)
                    raw = *reinterpret_cast<const uint2 *>(src + (q0 + p) * (int64_t)nPx + (int64_t)v * VEC);
                __builtin_memcpy(&x[p][0], &raw, 8);
            }
            const int w0 = word_of(v * VEC); // VEC divides W: a vector stays in one row
#pragma unroll
            for (int j = 0; j < VEC; ++j) {
                Tin xp[P];
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    xp[p] = x[p][j];
                    if constexpr (NOK) vmax = max(vmax, (uint32_t)xp[p]);
                }
                uint32_t d[C::ND];
                pack_word<IN, FLT, P>(d, xp, a.bg, a.px0 + v * VEC + j);
                Word w;
                __builtin_memcpy(&w, d, sizeof(Word));
                lds[w0 + j] = w;
            }
        }
        for (int px = nvec * VEC + tid; px <= nst; px += blockDim.x) {
            Tin xp[P];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                xp[p] = (px < nst && q0 + p < a.N) ? src[(q0 + p) * (int64_t)nPx + px] : (Tin)0;
                if constexpr (NOK) vmax = max(vmax, (uint32_t)xp[p]);
            }
            uint32_t d[C::ND];
            pack_word<IN, FLT, P>(d, xp, px < nst ? a.bg : nullptr, a.px0 + px);
            if (px == nst) {
#pragma unroll
                for (int k = 0; k < C::ND; ++k) d[k] = 0u;
            }
            Word w;
            __builtin_memcpy(&w, d, sizeof(Word));
            lds[word_of(px)] = w; // px == nst: the zero word rows * pitch
        }
        // the pack's largest value: per-wave maxima past the staged words, then every wave
        // takes the maximum of those (a uniform choice of the accumulation per pack)
        uint32_t *wmax = reinterpret_cast<uint32_t *>(
            lds_raw + ((((size_t)nst / a.W * a.pitch + 1) * C::WB + 15) & ~(size_t)15));
        if constexpr (NOK) {
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) vmax = max(vmax, (uint32_t)__shfl_xor((int)vmax, o));
            if (lane == 0) wmax[wave] = vmax;
        }
        __syncthreads();
        int F = 0; // narrow sums: groups per flush (0: the wide sums)
        if constexpr (NOK) {
            uint32_t m = 0u;
            for (int w2 = 0; w2 < nwaves; ++w2) m = max(m, wmax[w2]);
            F = __builtin_amdgcn_readfirstlane(narrow_groups(m));
        }

        // ---- gather: each lane owns one line of the wave's current 64-line batch; the
        //      wave walks its stream of plan groups with three groups in flight ----
        const int nslot = nwaves * a.nslice, slot = slice * nwaves + wave;
        // one walk, instantiated for the wide and the narrow sums (chosen per pack)
        // mode 0: wide sums; 1: narrow, flushed every group; 2: narrow, flushed every F groups.
        // SEG: row segments after the first (segmode 2, 3), whose lines start from the earlier
        // segments' partial sums -- a compile-time choice, so that the other walks hold no
        // load of them: a register loaded on one path and read after the paths merge makes
        // the wave wait there for every load in flight, at each line end (5-6% of the gather)
        auto walk = [&](auto mode_tag, auto seg_tag) {
          constexpr int MODE = decltype(mode_tag)::value;
          constexpr bool SEG = decltype(seg_tag)::value;
          constexpr bool NARROW = MODE > 0;
          int b = slot;
          if (b < nb) {
            const uint4 *S = a.gidx + (size_t)a.wofs[slot] * 64 + lane;
            const int G = a.wcnt[slot];
            // four plan groups in flight; each register is refilled right after use
            uint4 eq[4] = {S[0], S[64], S[128], S[192]}; // filler past the end keeps these in bounds
            int bend = (int)a.bhdr[b].y;
            int2 li = a.lane[(size_t)b * 64 + lane];
            double rcl = a.lrc[(size_t)b * 64 + lane], rcn = 0.0;
            int bn = b + nslot;
            int2 lin = make_int2(-1, 0);
            int ngn = 0;
            if (bn < nb) {
                lin = a.lane[(size_t)bn * 64 + lane];
                rcn = a.lrc[(size_t)bn * 64 + lane];
                ngn = (int)a.bhdr[bn].y;
            }
            uint32_t acc[P];
            double accd[P];
#pragma unroll
            for (int p = 0; p < P; ++p) { acc[p] = 0u; accd[p] = 0.0; }
            uint32_t dacc[NARROW ? P / 2 : 1];  // narrow sums: two fields per dword
#pragma unroll
            for (int j = 0; j < (NARROW ? P / 2 : 1); ++j) dacc[j] = 0u;
            int fc = 0;                         // groups since the last flush
            // row segments after the first: the integer partial sums of the segments before,
            // loaded when a batch starts so that the loads are done by its end
            uint32_t prev[P];
            auto load_prev = [&](int bb, int cell) {
#pragma unroll
                for (int p = 0; p < P; ++p)
                    prev[p] = (SEG && !FLT && cell >= 0 && q0 + p < a.N)
                                  ? a.part[(q0 + p) * (int64_t)a.npos + bb * 64 + lane] : 0u;
            };
            load_prev(b, li.x);
            // float sums in row segments after the first start from the earlier segments'
            // float64 sum (entries run in plan order, which is row-major): the next batch's
            // starting sums are loaded one batch ahead
            double *partd = reinterpret_cast<double *>(a.part);
            double pdn[P];
            auto load_pd = [&](double *d, int bb, int cell) {
#pragma unroll
                for (int p = 0; p < P; ++p)
                    d[p] = (SEG && FLT && bb < nb && cell >= 0 && q0 + p < a.N)
                               ? partd[(q0 + p) * (int64_t)a.npos + bb * 64 + lane] : 0.0;
            };
            if constexpr (FLT) {
                load_pd(accd, b, li.x);
                load_pd(pdn, bn, lin.x);
            }
            // words up to 8 bytes: the LDS reads of group g+1 are issued before group g is
            // accumulated, so a wave's reads overlap its own adds (two register sets).  Two
            // groups ahead (three sets, the walk unrolled 12 times) measured slower: 0.159
            // against 0.145 ms per 2,048 120x120 uint16 patterns.
            constexpr bool PIPE = C::WB <= 8;
            Word wq[PIPE ? 2 : 1][8];
            if constexpr (PIPE) group_issue<IN, FLT, P>(lbase, eq[0], wq[0]);
            for (int g0 = 0; g0 < G; g0 += 4) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int gi = g0 + u; // G is a multiple of 4 (the layout pads it)
                    {
                        const Word *wg;
                        if constexpr (PIPE) {
                            // group gi+1 is in eq[(u+1)&3] (past the end: a filler group)
                            group_issue<IN, FLT, P>(lbase, eq[(u + 1) & 3], wq[(u + 1) & 1]);
                            wg = wq[u & 1];
                        } else {
                            group_issue<IN, FLT, P>(lbase, eq[u], wq[0]);
                            wg = wq[0];
                        }
                        eq[u] = S[(size_t)(gi + 4) * 64];
                        const bool line_end = gi + 1 == bend;
                        if constexpr (MODE == 1) {
                            group_accumulate_narrow<P, true>(wg, dacc);
                            narrow_flush<P, false>(dacc, acc);
                        } else if constexpr (MODE == 2) {
                            group_accumulate_narrow<P, false>(wg, dacc);
                            if (++fc == F || line_end) {
                                narrow_flush<P, true>(dacc, acc);
                                fc = 0;
                            }
                        } else {
                            group_accumulate<IN, FLT, P>(wg, acc, accd);
                        }
                        if (line_end) {
                            if constexpr (!NARROW) gather_decode<IN, FLT, P>(acc);
                            if (!FLT && (a.segmode == 1 || a.segmode == 2) && li.x >= 0) {
#pragma unroll
                                for (int p = 0; p < P; ++p)
                                    if (q0 + p < a.N) a.part[(q0 + p) * (int64_t)a.npos + b * 64 + lane] = prev[p] + acc[p];
                            } else if (FLT && (a.segmode == 1 || a.segmode == 2) && li.x >= 0) {
#pragma unroll
                                for (int p = 0; p < P; ++p)
                                    if (q0 + p < a.N) partd[(q0 + p) * (int64_t)a.npos + b * 64 + lane] = accd[p];
                            } else if (li.x >= 0) {
                                // All P values first, then the stores back to back: a store
                                // between two loads' uses makes the wave wait for the store.
                                const double cnt = (double)li.y;
                                const double lb = (!FLT && a.linebg) ? a.linebg[li.x] : 0.0;
                                // x / cnt correctly rounded from one reciprocal per line:
                                // q = x * rc is within an ulp of x / cnt, the fma residual e is
                                // exact, and q + e * rc rounds to x / cnt (Markstein's theorem,
                                // rc the correctly rounded 1 / cnt, computed on the host).
                                const double rc = rcl;
                                float r[P];
#pragma unroll
                                for (int p = 0; p < P; ++p) {
                                    double x;
                                    if constexpr (FLT)
                                        x = accd[p];
                                    else
                                        x = (double)(prev[p] + acc[p]) - lb;
                                    const double q = x * rc;
                                    const double e = __builtin_fma(-q, cnt, x);
                                    r[p] = (float)__builtin_fma(e, rc, q);
                                    if (li.y == 0) r[p] = -1.0f;
                                }
#pragma unroll
                                for (int p = 0; p < P; ++p)
                                    if (q0 + p < a.N) a.out[(q0 + p) * (int64_t)a.nCells + li.x] = r[p];
                            }
#pragma unroll
                            for (int p = 0; p < P; ++p) { acc[p] = 0u; accd[p] = FLT ? pdn[p] : 0.0; }
                            b = bn;
                            li = lin;
                            bend += ngn;
                            bn = b + nslot;
                            // past the slot's last batch: no further line end (the walk then
                            // only runs the filler groups that round the stream up)
                            rcl = rcn;
                            if (bn < nb) {
                                lin = a.lane[(size_t)bn * 64 + lane];
                                rcn = a.lrc[(size_t)bn * 64 + lane];
                                ngn = (int)a.bhdr[bn].y;
                            } else {
                                lin = make_int2(-1, 0);
                                ngn = 0;
                            }
                            load_prev(b, b < nb ? li.x : -1);
                            if constexpr (FLT) load_pd(pdn, bn, lin.x);
                        }
                    }
                }
            }
          }
        };
        auto walk_seg = [&](auto mode_tag) {
            if (a.segmode >= 2) walk(mode_tag, std::true_type{});
            else walk(mode_tag, std::false_type{});
        };
        if constexpr (Narrow<IN, FLT, P>::ok) {
            if (F == 1) walk_seg(std::integral_constant<int, 1>{});
            else if (F > 1) walk_seg(std::integral_constant<int, 2>{});
            else walk_seg(std::integral_constant<int, 0>{});
        } else {
            walk_seg(std::integral_constant<int, 0>{});
        }
        __syncthreads();
    }
}

template <int IN, bool FLT, int P>
__global__ __launch_bounds__(1024) void k_radon_gather(GatherArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int64_t nitems = a.npacks * a.nslice;
    for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x)
        gather_pack<IN, FLT, P>(a, it / a.nslice, (int)(it % a.nslice), lds_raw);
}

// Variants of the gather a batch may mix (element type, float staging, patterns per word);
// ID is the variant's index in kVariants.  (Sixteen uint8 patterns per 16-byte word needed
// more than 128 VGPRs and spilled, in its own kernel and in the mixed kernel; it is not built.)
#define EBSD_GATHER_VARIANTS(X)                                                                    \
    X(0, EBSD_U8, false, 8) X(1, EBSD_U8, false, 4) X(2, EBSD_U8, false, 2)                         \
    X(3, EBSD_U16, false, 8) X(4, EBSD_U16, false, 4) X(5, EBSD_U16, false, 2) X(6, EBSD_U16, false, 1) \
    X(7, EBSD_U8, true, 4) X(8, EBSD_U8, true, 2) X(9, EBSD_U8, true, 1)                           \
    X(10, EBSD_U16, true, 4) X(11, EBSD_U16, true, 2) X(12, EBSD_U16, true, 1)                     \
    X(13, EBSD_F32, true, 4) X(14, EBSD_F32, true, 2) X(15, EBSD_F32, true, 1)

template <int IN, bool FLT, int P> struct GatherVariant;
#define EBSD_GV_ID(ID, IN, FLT, P) \
    template <> struct GatherVariant<IN, FLT, P> { static constexpr int id = ID; };
EBSD_GATHER_VARIANTS(EBSD_GV_ID)
#undef EBSD_GV_ID

struct GatherMultiArgs {
    const GatherArgs *sets; // per set
    const int *variant;     // per set
    const int2 *items;      // (set, pack), largest patterns first
    int64_t nitems;
};

// Persistent gather over a heterogeneous work list: each item is one pack of one set.
__global__ __launch_bounds__(1024) void k_radon_gather_multi(GatherMultiArgs m)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    for (int64_t it = blockIdx.x; it < m.nitems; it += gridDim.x) {
        const int2 item = m.items[it];
        const GatherArgs &a = m.sets[item.x];
        switch (m.variant[item.x]) {
#define EBSD_GV_CASE(ID, IN, FLT, P) \
    case ID: gather_pack<IN, FLT, P>(a, item.y / a.nslice, item.y % a.nslice, lds_raw); break;
            EBSD_GATHER_VARIANTS(EBSD_GV_CASE)
#undef EBSD_GV_CASE
        default: break;
        }
    }
}

// ---------------------------------------------------------------------------------------
// k_peaks
// ---------------------------------------------------------------------------------------
// Pointers of one pattern set of the peak kernel.
struct PeakPtrs {
    const float *radon; // (N, nR, nT)
    int64_t N;
    const uint8_t *mask; // (nR, nT)
    const double *grid;  // theta_deg[nT], rho[nR]
    const uint64_t *mlane; // streamed kernel: rho-mask lane masks [wavefront][padded row + 8]
    const uint32_t *mrow; // rho-mask bits per rho row (maskw words a row)
    ebsd_bands out;
};

// Source cell of padded Radon cell (r, c) in the unpadded (nR, nT) Radon: theta padding
// mirrors the whole padded rho axis (band_detect.py:573-575), then rho padding repeats
// the first/last interior rows (band_detect.py:578-580).
__device__ __forceinline__ int pad_src(const PeakGeo &g, int r, int c)
{
    int i = r - g.p0;
    i = i < 0 ? 0 : (i > g.nR - 1 ? g.nR - 1 : i);
    int j;
    if (c < g.p1) {
        j = g.nT - g.p1 + c;
        i = g.nR - 1 - i;
    } else if (c >= g.p1 + g.nT) {
        j = c - g.p1 - g.nT;
        i = g.nR - 1 - i;
    } else {
        j = c - g.p1;
    }
    return i * g.nT + j;
}

struct Key {
    float v;
    int pos;
};

__device__ __forceinline__ bool key_gt(float av, int ap, float bv, int bp)
{
    return av > bv || (av == bv && ap > bp);
}

// LDS layout of k_peaks (offsets in 4-byte words).  A holds the staged padded-Radon tile
// and later the convolved tile; B holds the theta-convolved tile and later the theta
// dilation; the pattern-independent rho mask is kept as bits.
struct PeakLds {
    int a_off, b_off, mask_off, cand_v, cand_p, cand_nn, run_v, run_p, run_nn, sel_v, sel_p, sel_nn, red,
        total;
};

__host__ __device__ inline PeakLds peak_lds_layout(const PeakGeo &g)
{
    PeakLds L;
    L.a_off = 0;
    L.b_off = g.rh * g.apitch;
    L.mask_off = L.b_off + g.rh * g.bpitch;
    L.cand_v = L.mask_off + g.nR * g.maskw;
    L.cand_p = L.cand_v + CAND_CAP;
    L.cand_nn = L.cand_p + CAND_CAP;
    L.run_v = L.cand_nn + 9 * CAND_CAP;
    L.run_p = L.run_v + NB_MAX;
    L.run_nn = L.run_p + NB_MAX;
    L.sel_v = L.run_nn + 9 * NB_MAX;
    L.sel_p = L.sel_v + NB_MAX;
    L.sel_nn = L.sel_p + NB_MAX;
    L.red = L.sel_nn + 9 * NB_MAX; // min[32], max[32], sum (double)[32]
    L.red += L.red & 1;            // 8-byte alignment of the double sums
    L.total = L.red + 4 * 32 + 8;  // + counters
    return L;
}

// Kernel arguments of k_peaks.  The geometry and LDS layout are launch-wide and travel by
// value (kernel arguments stay invariant across barriers); a heterogeneous launch adds a
// (set, pattern) work list over sets that share that geometry.
struct PeakArgs {
    PeakPtrs s;             // the one set of a plain launch
    const PeakPtrs *sets;   // heterogeneous launch: per-set pointers
    const int2 *items;      //   (set, pattern) work list
    int64_t nitems;
    PeakLds L;
    PeakGeo g;
};

__device__ __forceinline__ double np_interp_unit(double x, const double *fp, int n)
{
    // numpy.interp with xp = 0..n-1 (compiled_base.c: clamp outside, exact at nodes,
    // slope*(x - xp[j]) + fp[j] inside)
    if (!(x == x)) return x;
    if (x < 0.0) return fp[0];
    if (x > (double)(n - 1)) return fp[n - 1];
    int j = (int)floor(x);
    if (j >= n - 1) return fp[n - 1];
    if ((double)j == x) return fp[j];
    const double slope = fp[j + 1] - fp[j];
    return slope * (x - (double)j) + fp[j];
}


// Top-nB of (running list U candidate list) by (value, position) descending -- a total
// order, so ranks are unique: each entry counts the entries above it and, if its rank is
// below nB, writes itself to that slot.  Result replaces the running list; the candidate
// list is emptied.  All threads of the work-group call it.
__device__ __noinline__ void select_top(int tid, int nthr, int nB, int *ctr, float *run_v, int *run_p,
                                        float *run_nn, const float *cand_v, const int *cand_p,
                                        const float *cand_nn, float *sel_v, int *sel_p, float *sel_nn)
{
    __syncthreads();
    const int nrun = ctr[0];
    const int ncand = min(ctr[1], CAND_CAP);
    const int ntot = nrun + ncand;
    for (int i = tid; i < ntot; i += nthr) {
        const bool ir = i < nrun;
        const float v = ir ? run_v[i] : cand_v[i - nrun];
        const int p = ir ? run_p[i] : cand_p[i - nrun];
        int rank = 0;
        for (int j = 0; j < nrun; ++j) rank += key_gt(run_v[j], run_p[j], v, p);
        for (int j = 0; j < ncand && rank < nB; ++j) rank += key_gt(cand_v[j], cand_p[j], v, p);
        if (rank < nB) {
            sel_v[rank] = v;
            sel_p[rank] = p;
            const float *nn = ir ? run_nn + 9 * i : cand_nn + 9 * (i - nrun);
            for (int k = 0; k < 9; ++k) sel_nn[9 * rank + k] = nn[k];
        }
    }
    __syncthreads();
    const int nsel = min(nB, ntot);
    for (int i = tid; i < nsel * 11; i += nthr) {
        const int e = i / 11, k = i - e * 11;
        if (k == 0) run_v[e] = sel_v[e];
        else if (k == 1) run_p[e] = sel_p[e];
        else run_nn[9 * e + k - 2] = sel_nn[9 * e + k - 2];
    }
    __syncthreads();
    if (tid == 0) {
        ctr[0] = nsel;
        ctr[1] = 0;
    }
    __syncthreads();
}

// Geometry of theta tile t: interior columns [hc0, hc1) of the padded Radon, convolved
// columns [cc0, cc1) (dilation halo), staged columns starting at cb0 (convolution halo).
struct TileGeo {
    int hc0, hc1, HW, cc0, cc1, CW, cb0, RW, hoff;
};
__device__ __forceinline__ TileGeo tile_geo(const PeakGeo &g, int t, int ht, int ckt)
{
    TileGeo T;
    const int j0 = t * g.tile, j1 = min(j0 + g.tile, g.nT);
    T.hc0 = g.p1 + j0;
    T.hc1 = g.p1 + j1;
    T.HW = T.hc1 - T.hc0;
    T.cc0 = max(T.hc0 - ht, 0);
    T.cc1 = min(T.hc1 + ht, g.nTp);
    T.CW = T.cc1 - T.cc0;
    T.cb0 = T.cc0 - ckt;
    T.RW = T.CW + 2 * ckt;
    T.hoff = T.hc0 - ht - T.cc0;
    return T;
}

constexpr int PF_REGS = 16; // staged elements per thread held in registers across a tile

// Padded Radon value of staged element `it` of a tile (zero outside the padded array).
__device__ __forceinline__ float tile_elem(const PeakGeo &g, const float *R, const TileGeo &T, int y, int x)
{
    const int r = g.cr0 - g.kr / 2 + y, c = T.cb0 + x;
    if (r >= 0 && r < g.nRp && c >= 0 && c < g.nTp) return __builtin_nontemporal_load(R + pad_src(g, r, c));
    return 0.0f;
}

// Raw load for staged element (y, x), clamped into the padded array so that the address is
// always valid; tile_commit zeroes what lies outside.  No branches: the loads of a whole
// prefetch stay in flight together.
__device__ __forceinline__ float tile_load(const PeakGeo &g, const float *R, const TileGeo &T, int y, int x)
{
    int r = g.cr0 - g.kr / 2 + y, c = T.cb0 + x;
    r = r < 0 ? 0 : (r > g.nRp - 1 ? g.nRp - 1 : r);
    c = c < 0 ? 0 : (c > g.nTp - 1 ? g.nTp - 1 : c);
    return __builtin_nontemporal_load(R + pad_src(g, r, c));
}

__device__ __forceinline__ bool tile_inside(const PeakGeo &g, const TileGeo &T, int y, int x)
{
    const int r = g.cr0 - g.kr / 2 + y, c = T.cb0 + x;
    return r >= 0 && r < g.nRp && c >= 0 && c < g.nTp;
}

__device__ __forceinline__ void tile_fetch(const PeakGeo &g, const float *R, const TileGeo &T, int RH, int tid,
                                           int nthr, float *pf)
{
    const FDiv f = fdiv_make(T.RW);
    const int last = RH * T.RW - 1;
#pragma unroll
    for (int u = 0; u < PF_REGS; ++u) {
        const int it = min(tid + u * nthr, last);
        const int y = (int)fdiv_q(it, f), x = it - y * T.RW;
        pf[u] = tile_load(g, R, T, y, x);
    }
}

__device__ __forceinline__ void tile_commit(const PeakGeo &g, const float *R, const TileGeo &T, int RH, int tid,
                                            int nthr, const float *pf, float *A, int AP)
{
    const FDiv f = fdiv_make(T.RW);
#pragma unroll
    for (int u = 0; u < PF_REGS; ++u) {
        const int it = tid + u * nthr;
        if (it < RH * T.RW) {
            const int y = (int)fdiv_q(it, f), x = it - y * T.RW;
            A[y * AP + x] = tile_inside(g, T, y, x) ? pf[u] : 0.0f;
        }
    }
    // elements beyond the register window (large Radons), 8 loads in flight
    constexpr int U = 8;
    for (int it0 = tid + PF_REGS * nthr; it0 < RH * T.RW; it0 += U * nthr) {
        float val[U];
        int dst[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int it = it0 + u * nthr;
            dst[u] = -1;
            val[u] = 0.0f;
            if (it < RH * T.RW) {
                const int y = (int)fdiv_q(it, f), x = it - y * T.RW;
                dst[u] = y * AP + x;
                val[u] = tile_elem(g, R, T, y, x);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (dst[u] >= 0) A[dst[u]] = val[u];
    }
}

// One band slot of one pattern: width, 3x3 quadratic sub-pixel fit, normmax, theta/rho
// (band_detect.py:736-800, then :457-460 and :521-524).  Slot i of the top-nB list
// (v, p, nn) sorted by (value, position) descending; nsel entries are filled.
__device__ __forceinline__ void finalize_band(const PeakGeo &g, const PeakPtrs &S, int64_t q, int i, int nsel,
                                           const float *run_v, const int *run_p, const float *run_nn, float mn,
                                           float ave)
{
    const int64_t o = q * g.nB + i;
    float bmax = -2.0e6f, bavemax = -2.0e6f, width = 0.0f;
    float ml_r = 0.0f, ml_c = 0.0f, al_r = 0.0f, al_c = 0.0f;
    int8_t valid = 0;
    if (i < nsel && run_v[i] > 0.0f) {
        const int pos = run_p[i];
        const int r = pos / g.nTp, c = pos - (pos / g.nTp) * g.nTp;
        float nn[9];
        for (int k = 0; k < 9; ++k) nn[k] = run_nn[i * 9 + k];
        const float mx = run_v[i];
        ml_r = (float)r;
        ml_c = (float)c;
        bmax = mx;
        // width (band_detect.py:748-753)
        const float mnt = fminf(fminf(nn[1], nn[4]), nn[7]);
        if (mnt > 0.0f) {
            const float lp = logf(nn[7]), lm = logf(nn[1]), l0 = logf(mx);
            const double aa = (double)(lp - lm) * 0.5 - (double)l0;
            if (aa < -1.e-8) width = (float)(2.0 * sqrt(log(2.0) / (-1.0 * aa)));
        }
        // 3x3 quadratic sub-pixel fit (band_detect.py:764-798)
        float s = 0.0f;
        for (int k = 0; k < 9; ++k) s += nn[k];
        const double sumnn = s > 1e-12f ? (double)s : 1e-12;
        for (int k = 0; k < 9; ++k) nn[k] = (float)((double)nn[k] / sumnn);
        bavemax = (float)(sumnn / 9.0);
        const double dx = 0.5 * (double)(nn[5] - nn[3]);
        const double dy = 0.5 * (double)(nn[7] - nn[1]);
        const double dxx = (double)(nn[5] + nn[3]) - 2.0 * (double)nn[4];
        const double dyy = (double)(nn[7] + nn[1]) - 2.0 * (double)nn[4];
        const double dxy = 0.25 * (double)(((nn[8] - nn[2]) - nn[6]) + nn[0]);
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
        al_r = (float)((double)r - rc);
        al_c = (float)((double)c - dc);
        valid = 1;
    }
    // maxloc/aveloc -= padding (float32 - int64 in float64, stored as float32)
    ml_r = (float)((double)ml_r - (double)g.p0);
    ml_c = (float)((double)ml_c - (double)g.p1);
    al_r = (float)((double)al_r - (double)g.p0);
    al_c = (float)((double)al_c - (double)g.p1);
    float d = ave - mn;
    d = d < 1e-7f ? 1e-7f : d;
    const float nm = (bmax - mn) / d;
    const double *thg = S.grid, *rhg = S.grid + g.nT;
    const double th = 3.141592653589793 - np_interp_unit((double)al_c, thg, g.nT) /
                                              (180.0 / 3.141592653589793);
    const double rh = -1.0 * np_interp_unit((double)al_r, rhg, g.nR);
    S.out.d_max[o] = bmax;
    S.out.d_normmax[o] = nm;
    S.out.d_avemax[o] = bavemax;
    S.out.d_maxloc[2 * o + 0] = ml_r;
    S.out.d_maxloc[2 * o + 1] = ml_c;
    S.out.d_aveloc[2 * o + 0] = al_r;
    S.out.d_aveloc[2 * o + 1] = al_c;
    S.out.d_width[o] = width;
    S.out.d_theta[o] = (float)th;
    S.out.d_rho[o] = (float)rh;
    S.out.d_valid[o] = valid;
}

constexpr int CSTRIP = 8; // rows per thread in the rho passes
constexpr int CGRP = 4;   // columns per thread in the theta passes

// KR/KT > 0: compile-time kernel sizes (dilation half-windows equal to them, as
// band_detect.py:287-288 derives peakPad = 2*ksz+1); 0: run-time sizes.  MULTI walks a
// heterogeneous (set, pattern) work list, otherwise the patterns of one set.
template <int KR, int KT, bool MULTI>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_peaks(PeakArgs a)
{
    extern __shared__ __align__(16) float sm[];
    const PeakLds &L = a.L;
    float *A = sm + L.a_off;
    float *B = sm + L.b_off;
    uint32_t *mbits = reinterpret_cast<uint32_t *>(sm + L.mask_off);
    float *cand_v = sm + L.cand_v;
    int *cand_p = reinterpret_cast<int *>(sm + L.cand_p);
    float *cand_nn = sm + L.cand_nn;
    float *run_v = sm + L.run_v;
    int *run_p = reinterpret_cast<int *>(sm + L.run_p);
    float *run_nn = sm + L.run_nn;
    float *sel_v = sm + L.sel_v;
    int *sel_p = reinterpret_cast<int *>(sm + L.sel_p);
    float *sel_nn = sm + L.sel_nn;
    float *red = sm + L.red;
    int *ctr = reinterpret_cast<int *>(sm + L.total - 8); // [0] running count, [1] candidates

    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int64_t nitems = MULTI ? a.nitems : a.s.N;
    int cur_set = -1;

    // the padded-Radon block of the next tile is loaded into registers while the current
    // tile is processed (the Radon of a pattern comes from HBM/MALL)
    float pf[PF_REGS];
    if ((int64_t)blockIdx.x < nitems) {
        const int64_t it0 = blockIdx.x;
        const PeakGeo &g0 = a.g;
        const float *R0 = MULTI ? a.sets[a.items[it0].x].radon : a.s.radon;
        const int64_t q0 = MULTI ? (int64_t)a.items[it0].y : it0;
        const int ht0 = KT > 0 ? KT : g0.ht, ckt0 = (KT > 0 ? KT : g0.kt) / 2;
        tile_fetch(g0, R0 + q0 * (int64_t)(g0.nR * g0.nT), tile_geo(g0, 0, ht0, ckt0), g0.rh, tid, nthr, pf);
    }

    for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int sidx = MULTI ? a.items[it].x : 0;
        const PeakPtrs &S = MULTI ? a.sets[sidx] : a.s;
        const int64_t q = MULTI ? (int64_t)a.items[it].y : it;
        const PeakGeo &g = a.g;
        const int kr = KR > 0 ? KR : g.kr, kt = KT > 0 ? KT : g.kt;
        const int hr = KR > 0 ? KR : g.hr, ht = KT > 0 ? KT : g.ht;
        const int ckr = kr / 2, ckt = kt / 2;
        const int CH = g.cr1 - g.cr0;
        const int RH = g.rh;
        const int AP = g.apitch, BP = g.bpitch;
        const int nCells = g.nR * g.nT;
        (void)ckr;

        // ---- rho mask as bits, reloaded when the set changes ----
        if (sidx != cur_set) {
            for (int i2 = tid; i2 < g.nR * g.maskw; i2 += nthr) mbits[i2] = S.mrow[i2];
            cur_set = sidx;
        }

        const float *R = S.radon + q * (int64_t)nCells;
        float smin = INFINITY, smax = -INFINITY;
        double ssum = 0.0;
        if (tid == 0) { ctr[0] = 0; ctr[1] = 0; }
        __syncthreads();

        for (int t = 0; t < g.ntiles; ++t) {
            const TileGeo TG = tile_geo(g, t, ht, ckt);
            const int hc0 = TG.hc0, hc1 = TG.hc1, HW = TG.HW;
            const int cc0 = TG.cc0, CW = TG.CW;
            const int hoff = TG.hoff; // 0 unless the theta window is clipped

            // ---- stage the prefetched padded Radon tile, then start fetching the next ----
            tile_commit(g, R, TG, RH, tid, nthr, pf, A, AP);
            __syncthreads();
            if (t + 1 < g.ntiles) {
                tile_fetch(g, R, tile_geo(g, t + 1, ht, ckt), RH, tid, nthr, pf);
            } else if (it + gridDim.x < nitems) {
                const int64_t itn = it + gridDim.x;
                const float *Rn = MULTI ? a.sets[a.items[itn].x].radon : a.s.radon;
                const int64_t qn = MULTI ? (int64_t)a.items[itn].y : itn;
                tile_fetch(g, Rn + qn * (int64_t)nCells, tile_geo(g, 0, ht, ckt), RH, tid, nthr, pf);
            }
            // ---- theta pass: T[y][x] = sum_b v[b] * A[y][x + kt - 1 - b], 4 columns/thread ----
            {
                const int G = (CW + CGRP - 1) / CGRP;
                const FDiv f = fdiv_make(G);
                for (int it = tid; it < RH * G; it += nthr) {
                    const int y = (int)fdiv_q(it, f), x0 = (it - y * G) * CGRP;
                    const float *row = A + y * AP + x0;
                    float o[CGRP];
                    if constexpr (KT > 0) {
                        // 16-byte aligned reads (AP and x0 are multiples of 4): conflict-free
                        constexpr int NV = (CGRP + KT - 1 + 3) / 4;
                        float w[4 * NV];
#pragma unroll
                        for (int k = 0; k < NV; ++k) {
                            const float4 v4 = reinterpret_cast<const float4 *>(row)[k];
                            w[4 * k] = v4.x; w[4 * k + 1] = v4.y; w[4 * k + 2] = v4.z; w[4 * k + 3] = v4.w;
                        }
#pragma unroll
                        for (int c = 0; c < CGRP; ++c) {
                            float s = 0.0f;
#pragma unroll
                            for (int b = 0; b < KT; ++b) s = fmaf(g.v[b], w[c + KT - 1 - b], s);
                            o[c] = s;
                        }
                    } else {
                        for (int c = 0; c < CGRP; ++c) {
                            float s = 0.0f;
                            for (int b = 0; b < kt; ++b) s = fmaf(g.v[b], row[c + kt - 1 - b], s);
                            o[c] = s;
                        }
                    }
#pragma unroll
                    for (int c = 0; c < CGRP; ++c)
                        if (x0 + c < CW) B[y * BP + x0 + c] = o[c];
                }
            }
            __syncthreads();
            // ---- rho pass: C[y][x] = sum_a u[a] * T[y + kr - 1 - a][x], 8 rows/thread (into A) ----
            {
                const int NS = (CH + CSTRIP - 1) / CSTRIP;
                const FDiv f = fdiv_make(CW);
                for (int it = tid; it < NS * CW; it += nthr) {
                    const int s = (int)fdiv_q(it, f), x = it - s * CW;
                    const int y0 = s * CSTRIP;
                    float o[CSTRIP];
                    if constexpr (KR > 0) {
                        float w[CSTRIP + KR - 1];
#pragma unroll
                        for (int k = 0; k < CSTRIP + KR - 1; ++k)
                            w[k] = (y0 + k < RH) ? B[(y0 + k) * BP + x] : 0.0f;
#pragma unroll
                        for (int c = 0; c < CSTRIP; ++c) {
                            float acc = 0.0f;
#pragma unroll
                            for (int aa = 0; aa < KR; ++aa) acc = fmaf(g.u[aa], w[c + KR - 1 - aa], acc);
                            o[c] = acc;
                        }
                    } else {
                        for (int c = 0; c < CSTRIP; ++c) {
                            float acc = 0.0f;
                            if (y0 + c < CH)
                                for (int aa = 0; aa < kr; ++aa)
                                    acc = fmaf(g.u[aa], B[(y0 + c + kr - 1 - aa) * BP + x], acc);
                            o[c] = acc;
                        }
                    }
                    const int cglob = cc0 + x;
                    const bool scol = cglob >= hc0 && cglob < hc1;
#pragma unroll
                    for (int c = 0; c < CSTRIP; ++c) {
                        if (y0 + c < CH) {
                            A[(y0 + c) * AP + x] = o[c];
                            const int r = g.cr0 + y0 + c;
                            if (scol && r >= g.sr0 && r < g.sr1) {
                                smin = fminf(smin, o[c]);
                                smax = fmaxf(smax, o[c]);
                                ssum += (double)o[c];
                            }
                        }
                    }
                }
            }
            __syncthreads();
            // ---- theta dilation: H[y][x] = max C[y][x + hoff .. x + hoff + 2ht], 4 columns/thread (into B) ----
            {
                const int G = (HW + CGRP - 1) / CGRP;
                const FDiv f = fdiv_make(G);
                for (int it = tid; it < CH * G; it += nthr) {
                    const int y = (int)fdiv_q(it, f), x0 = (it - y * G) * CGRP;
                    const float *row = A + y * AP + x0 + hoff;
                    float o[CGRP];
                    if constexpr (KT > 0) {
                        // (hoff is 0: the padding p1 > ht keeps theta windows inside the array)
                        // windows w[c .. c + 2KT]; all four share w[3 .. 2KT]; 16-byte reads
                        constexpr int NV = (CGRP + 2 * KT + 3) / 4;
                        float w[4 * NV];
#pragma unroll
                        for (int k = 0; k < NV; ++k) {
                            const float4 v4 = reinterpret_cast<const float4 *>(row)[k];
                            w[4 * k] = v4.x; w[4 * k + 1] = v4.y; w[4 * k + 2] = v4.z; w[4 * k + 3] = v4.w;
                        }
                        float mid = w[CGRP - 1];
#pragma unroll
                        for (int k = CGRP; k <= 2 * KT; ++k) mid = fmaxf(mid, w[k]);
                        const float l2 = w[2], l1 = fmaxf(w[1], l2), l0 = fmaxf(w[0], l1);
                        const float r1 = w[2 * KT + 1], r2 = fmaxf(r1, w[2 * KT + 2]), r3 = fmaxf(r2, w[2 * KT + 3]);
                        o[0] = fmaxf(mid, l0);
                        o[1] = fmaxf(mid, fmaxf(l1, r1));
                        o[2] = fmaxf(mid, fmaxf(l2, r2));
                        o[3] = fmaxf(mid, r3);
                    } else {
                        for (int c = 0; c < CGRP; ++c) {
                            const int cg = hc0 + x0 + c;
                            const int lo = max(cg - ht, 0) - cc0, hi = min(cg + ht, g.nTp - 1) - cc0;
                            float m = -INFINITY;
                            for (int k = lo; k <= hi && cg < hc1; ++k) m = fmaxf(m, A[y * AP + k]);
                            o[c] = m;
                        }
                    }
#pragma unroll
                    for (int c = 0; c < CGRP; ++c)
                        if (x0 + c < HW) B[y * BP + x0 + c] = o[c];
                }
            }
            __syncthreads();

            // ---- rho dilation + equality under the mask -> candidates (with their 3x3
            //      neighbourhoods) appended to the pattern's list.  If a tile overflows the
            //      list (plateaus), its appends are dropped, the list is merged into the
            //      running top-nB and the tile is redone in row chunks that cannot overflow ----
            const int pre = ctr[1];
            int rstart = g.mr0;
            int rstep = g.mr1 - g.mr0;
            bool chunked = false;
            while (rstart < g.mr1) {
                const int rend = min(rstart + rstep, g.mr1);
                {
                    const int NS = (rend - rstart + CSTRIP - 1) / CSTRIP;
                    const FDiv f = fdiv_make(HW);
                    for (int it = tid; it < NS * HW; it += nthr) {
                        const int s = (int)fdiv_q(it, f), x = it - s * HW;
                        const int r0 = rstart + s * CSTRIP;
                        const int jc = hc0 + x - g.p1;
                        uint32_t any = 0u;
#pragma unroll
                        for (int c = 0; c < CSTRIP; ++c) {
                            const int r = r0 + c;
                            if (r < rend)
                                any |= ((mbits[(r - g.p0) * g.maskw + (jc >> 5)] >> (jc & 31)) & 1u) << c;
                        }
                        if (!any) continue;
                        float d[CSTRIP];
                        if constexpr (KR > 0) {
                            float w[CSTRIP + 2 * KR];
#pragma unroll
                            for (int k = 0; k < CSTRIP + 2 * KR; ++k) {
                                const int row = r0 - KR + k;
                                w[k] = (row >= 0 && row < g.nRp && row < g.cr1) ? B[(row - g.cr0) * BP + x] : -INFINITY;
                            }
                            // windows w[c .. c + 2KR]; all eight share w[7 .. 2KR]
                            float mid = w[CSTRIP - 1];
#pragma unroll
                            for (int k = CSTRIP; k <= 2 * KR; ++k) mid = fmaxf(mid, w[k]);
                            float lft[CSTRIP], rgt[CSTRIP];
                            lft[CSTRIP - 1] = -INFINITY;
#pragma unroll
                            for (int c = CSTRIP - 2; c >= 0; --c) lft[c] = fmaxf(w[c], lft[c + 1]);
                            rgt[0] = -INFINITY;
#pragma unroll
                            for (int c = 1; c < CSTRIP; ++c) rgt[c] = fmaxf(rgt[c - 1], w[2 * KR + c]);
#pragma unroll
                            for (int c = 0; c < CSTRIP; ++c) d[c] = fmaxf(mid, fmaxf(lft[c], rgt[c]));
                        } else {
                            for (int c = 0; c < CSTRIP; ++c) {
                                const int r = r0 + c;
                                const int lo = max(r - hr, 0) - g.cr0, hi = min(r + hr, g.nRp - 1) - g.cr0;
                                float m = -INFINITY;
                                if (r < rend)
                                    for (int k = lo; k <= hi; ++k) m = fmaxf(m, B[k * BP + x]);
                                d[c] = m;
                            }
                        }
#pragma unroll
                        for (int c = 0; c < CSTRIP; ++c) {
                            if (!((any >> c) & 1u)) continue;
                            const int r = r0 + c;
                            const int ax = hc0 + x - cc0;
                            const float cv = A[(r - g.cr0) * AP + ax];
                            if (d[c] == cv) {
                                const int slot = atomicAdd(&ctr[1], 1);
                                if (slot < CAND_CAP) {
                                    cand_v[slot] = cv;
                                    cand_p[slot] = r * g.nTp + hc0 + x;
#pragma unroll
                                    for (int dr = 0; dr < 3; ++dr)
#pragma unroll
                                        for (int dc = 0; dc < 3; ++dc)
                                            cand_nn[slot * 9 + dr * 3 + dc] = A[(r - 1 + dr - g.cr0) * AP + (ax - 1 + dc)];
                                }
                            }
                        }
                    }
                }
                __syncthreads();
                const int found = ctr[1];
                __syncthreads();
                if (found > CAND_CAP) {
                    // drop this pass's appends, fold the earlier ones into the running list
                    if (tid == 0) ctr[1] = chunked ? 0 : pre;
                    __syncthreads();
                    select_top(tid, nthr, g.nB, ctr, run_v, run_p, run_nn, cand_v, cand_p, cand_nn, sel_v,
                               sel_p, sel_nn);
                    chunked = true;
                    rstep = max(1, CAND_CAP / HW);
                    continue;
                }
                if (chunked) select_top(tid, nthr, g.nB, ctr, run_v, run_p, run_nn, cand_v, cand_p, cand_nn,
                                        sel_v, sel_p, sel_nn);
                rstart = rend;
            }
        }
        // ---- top-nB of the running list and all remaining candidates ----
        select_top(tid, nthr, g.nB, ctr, run_v, run_p, run_nn, cand_v, cand_p, cand_nn, sel_v, sel_p, sel_nn);

        // ---- stats reduction: min, max, sum over the work-group ----
        for (int off = 32; off > 0; off >>= 1) {
            smin = fminf(smin, __shfl_xor(smin, off, 64));
            smax = fmaxf(smax, __shfl_xor(smax, off, 64));
            ssum += __shfl_xor(ssum, off, 64);
        }
        double *redd = reinterpret_cast<double *>(red + 64);
        if (lane == 0) {
            red[wave] = smin;
            red[32 + wave] = smax;
            redd[wave] = ssum;
        }
        __syncthreads();
        if (tid == 0) {
            const int nw = nthr >> 6;
            float mn = red[0], mx = red[32];
            double sum = redd[0];
            for (int w = 1; w < nw; ++w) {
                mn = fminf(mn, red[w]);
                mx = fmaxf(mx, red[32 + w]);
                sum += redd[w];
            }
            const double cnt = (double)(g.sr1 - g.sr0) * (double)(g.sc1 - g.sc0);
            red[0] = mn;
            red[1] = (float)(sum / cnt);
            red[2] = mx;
            S.out.d_stats[q * 3 + 0] = mn;
            S.out.d_stats[q * 3 + 1] = red[1];
            S.out.d_stats[q * 3 + 2] = mx;
        }
        __syncthreads();

        // ---- finalize: band_detect.py:736-800, then :457-460 and :521-524 ----
        if (tid < g.nB) finalize_band(g, S, q, tid, ctr[0], run_v, run_p, run_nn, red[0], red[1]);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// k_peaks_stream: the same stages as k_peaks, streamed down the rho axis in registers.
// One work-group per pattern; wavefront w owns a block of `sbw` interior theta columns
// (one padded column per lane, with a halo of KT/2 + HT lanes each side).  Per strip of
// SROWS rows a lane stages its padded-Radon column, convolves along theta with DPP lane
// shifts and along rho from a register ring, then -- on rows that can hold a candidate --
// dilates along rho from a second ring and along theta with lane reads (the separable max in the
// other order: the same exact result) and tests D == C against the rho mask.  Only the
// candidate lists, the per-wave statistics and the final top-nB pass through LDS; the
// work-group synchronises twice per pattern.
// ---------------------------------------------------------------------------------------
constexpr int SROWS = 8; // rows per strip
constexpr int WCAP = 64; // candidates per wavefront list (>= the useful lanes of one row)

__device__ __forceinline__ float lane_next(float x) // lane i receives lane i+1 (column c+1)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_prev(float x) // lane i receives lane i-1 (column c-1)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ int lane_next_i(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xf, 0xf, true); }

// Order-preserving integer key of a float (an involution): the dilation maxima and the
// D == C test run on keys, exact and without the canonicalising moves float max needs
// for values that arrive through lane shifts.  Keys of -0 and +0 differ; only cells with
// C > 0 can become bands, so that never changes a result.
__device__ __forceinline__ int fkey(float x)
{
    const int b = __float_as_int(x);
    return b ^ ((b >> 31) & 0x7fffffff);
}
__device__ __forceinline__ float fkey_inv(int k) { return __int_as_float(k ^ ((k >> 31) & 0x7fffffff)); }
constexpr int KEY_MASKED = INT_MIN; // below the key of every float

// LDS layout of k_peaks_stream (4-byte words), per wavefront w of nw:
//   list: v[WCAP], p[WCAP], nn[9*WCAP]; run: v[nB], p[nB], nn[9*nB]; sel: same as run;
//   red: min, max, sum (double); cnt: running-list count
// then the final list (v, p, nn) of nB and the pattern statistics.
struct StreamLds {
    int per_wave, list_v, list_p, list_nn, run_v, run_p, run_nn, sel_v, sel_p, sel_nn, red, cnt, fin_v,
        fin_p, fin_nn, stats, total;
};
__host__ __device__ inline StreamLds stream_lds_layout(int nw, int nB)
{
    StreamLds L;
    L.list_v = 0;
    L.list_p = WCAP;
    L.list_nn = 2 * WCAP;
    L.run_v = 11 * WCAP;
    L.run_p = L.run_v + nB;
    L.run_nn = L.run_p + nB;
    L.sel_v = L.run_nn + 9 * nB;
    L.sel_p = L.sel_v + nB;
    L.sel_nn = L.sel_p + nB;
    L.red = L.sel_nn + 9 * nB;
    L.red += L.red & 1;        // 8-byte alignment of the double sum
    L.cnt = L.red + 4;         // running-list count
    L.per_wave = L.cnt + 2;    // even: every wave's red stays 8-byte aligned
    L.fin_v = nw * L.per_wave;
    L.fin_p = L.fin_v + nB;
    L.fin_nn = L.fin_p + nB;
    L.stats = L.fin_nn + 9 * nB;
    L.total = L.stats + 4;
    return L;
}

// Wave-local: the top-nB of (running list U candidate list) by (value, position) replace
// the running list; returns the new running count.  The whole wavefront calls it.
__device__ __forceinline__ int wave_merge(int lane, int nB, int nrun, int ncand, float *run_v, int *run_p,
                                       float *run_nn, const float *lst_v, const int *lst_p, const float *lst_nn,
                                       float *sel_v, int *sel_p, float *sel_nn)
{
    const int ntot = nrun + ncand;
    for (int i = lane; i < ntot; i += 64) {
        const bool ir = i < nrun;
        const float v = ir ? run_v[i] : lst_v[i - nrun];
        const int p = ir ? run_p[i] : lst_p[i - nrun];
        int rank = 0;
        for (int j = 0; j < nrun; ++j) rank += key_gt(run_v[j], run_p[j], v, p);
        for (int j = 0; j < ncand && rank < nB; ++j) rank += key_gt(lst_v[j], lst_p[j], v, p);
        if (rank < nB) {
            sel_v[rank] = v;
            sel_p[rank] = p;
            const float *nn = ir ? run_nn + 9 * i : lst_nn + 9 * (i - nrun);
            for (int k = 0; k < 9; ++k) sel_nn[9 * rank + k] = nn[k];
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const int nsel = min(nB, ntot);
    for (int i = lane; i < nsel * 11; i += 64) {
        const int e = i / 11, k = i - e * 11;
        if (k == 0) run_v[e] = sel_v[e];
        else if (k == 1) run_p[e] = sel_p[e];
        else run_nn[9 * e + k - 2] = sel_nn[9 * e + k - 2];
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    return nsel;
}

// Lane offsets of the windows of three lanes that, with the window at c, cover the theta
// dilation window c-HT .. c+HT: left of c they start at c-HT and step by 3 (the last may
// overlap the window at c); right of c+2 they step by 3 from c+3, the last moved back so
// that it ends at c+HT.
template <int N> struct StaticOffsets { int v[N > 0 ? N : 1]; __host__ __device__ constexpr int operator[](int i) const { return v[i]; } };
template <int N> __host__ __device__ constexpr StaticOffsets<N> theta_windows(int ht)
{
    StaticOffsets<N> o{};
    int n = 0;
    for (int s = -ht; s < 0; s += 3) o.v[n++] = s;
    for (int s = 3; s + 2 < ht + 3 && n < N; s += 3) o.v[n++] = s + 2 > ht ? ht - 2 : s;
    return o;
}

template <int KR, int KT, bool MULTI>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(MULTI ? 1 : 5))) void k_peaks_stream(PeakArgs a)
{
    constexpr int CKR = KR / 2, CKT = KT / 2, HR = KR, HT = KT; // dilation half-windows = kernel sizes
    constexpr int HALO = CKT + HT;
    // theta dilation window 2HT+1 from windows of three lanes: the one at c and NWIN more
    // at lanes c + WOFS[j]
    static_assert(HT >= 2 && HT <= 8, "theta dilation half-window outside 2 .. 8");
    constexpr int NLEFT = (HT + 2) / 3, NRIGHT = HT / 3; // windows left of c, right of c+2
    constexpr int NWIN = NLEFT + NRIGHT;
    constexpr StaticOffsets<NWIN> WOFS = theta_windows<NWIN>(HT);
    constexpr int TRING = SROWS + 2 * CKR;            // theta-convolved rows r0-CKR .. r0+SROWS-1+CKR
    constexpr int MRING = SROWS + 2 * HR;             // convolved rows r0-2HR .. r0+SROWS-1
    static_assert(2 * CKR <= SROWS && SROWS - 1 <= 2 * HR, "strip shorter than the rho windows");
    extern __shared__ __align__(16) float sm[];
    const PeakGeo &g = a.g;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    const int64_t it = blockIdx.x;
    const int sidx = MULTI ? a.items[it].x : 0;
    const PeakPtrs &S = MULTI ? a.sets[sidx] : a.s;
    const int64_t q = MULTI ? (int64_t)a.items[it].y : it;
    const int nB = g.nB;
    const StreamLds L = stream_lds_layout(nw, nB);
    float *wl = sm + wave * L.per_wave;
    float *lst_v = wl + L.list_v, *lst_nn = wl + L.list_nn;
    int *lst_p = reinterpret_cast<int *>(wl + L.list_p);
    float *run_v = wl + L.run_v, *run_nn = wl + L.run_nn;
    int *run_p = reinterpret_cast<int *>(wl + L.run_p);
    float *sel_v = wl + L.sel_v, *sel_nn = wl + L.sel_nn;
    int *sel_p = reinterpret_cast<int *>(wl + L.sel_p);

    // the pattern's Radon as a buffer resource (word 3: raw 32-bit access on gfx9), formed
    // before any lane-dependent control flow so that it stays in scalar registers; each
    // load's row, column and flip form a 32-bit lane offset
    const __amdgpu_buffer_rsrc_t Rrs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(S.radon + q * (int64_t)(g.nR * g.nT)), (short)0, g.nR * g.nT * 4, 0x00020000);

    // ---- this lane's padded column and its source in the unpadded Radon (selects only) ----
    const int c = g.p1 + wave * g.sbw - HALO + lane;
    const bool cval = c >= 0 && c < g.nTp;
    const int jint = c - g.p1;                              // interior column index
    const bool useful = lane >= HALO && lane < HALO + g.sbw && jint < g.nT;
    const bool left = c < g.p1, right = c >= g.p1 + g.nT;
    const int jc = !cval ? 0 : (left ? g.nT - g.p1 + c : (right ? c - g.p1 - g.nT : jint));
    const bool flip = cval && (left || right);
    const int jc4 = jc * 4;
    // this wavefront's rho-mask rows as 64-bit lane masks (useful lanes only), read with
    // scalar loads: padded row d at mlane[d + 8]
    const uint64_t *mlane = S.mlane + (size_t)__builtin_amdgcn_readfirstlane(wave) * (size_t)(g.nRp + 16) + 8;

    // staged value of padded row r (zero outside the padded array and in columns outside
    // it), issued unconditionally so that the loads of a whole strip stay in flight
    // together.  The zeros come from the buffer's range check: a read at or beyond
    // nR*nT*4 bytes returns 0, so a masked lane or row gets an offset far out of range
    // instead of a select.  Lane offset = (S ^ fm) + fb with S the row's byte offset
    // i*nT*4 (uniform): fm = 0, fb = jc4 reads row i; fm = -1, fb = (nR-1)*nT*4 + jc4 + 1
    // reads the flipped row nR-1-i (two's complement: (S ^ -1) + 1 = -S).  Every lane
    // offset is a single 32-bit VALU add of the uniform S (the range check ignores the
    // scalar offset, so the whole offset is in the lane register).
    const uint32_t fm = flip ? 0xffffffffu : 0u;
    const uint32_t fb = !cval ? 0x80000000u                      // column outside: >= 2^31
                              : (flip ? (uint32_t)((g.nR - 1) * g.nT * 4 + jc4 + 1) : (uint32_t)jc4);
    // row range tests as one unsigned compare each (r in [lo, lo+n) <=> (unsigned)(r-lo) < n),
    // so that each is one scalar compare and select rather than two and a mask operation
    const int nT4 = g.nT * 4;
    const uint32_t sr_n = (uint32_t)max(g.sr1 - g.sr0, 0), cr_n = (uint32_t)max(g.cr1 - g.cr0, 0);
    auto stage_load = [&](int r) -> float {
        const int i = min(max(r - g.p0, 0), g.nR - 1);
        // a row outside the padded array: bit 30 of S set puts every lane's offset at or
        // above 2^30 - (nR*nT*4) (flipped lanes) or 2^30 (others), far beyond the buffer
        // (computed without a branch: i*nT*4 < nR*nT*4 < 2^30)
        const uint32_t S = (uint32_t)(i * nT4) | ((uint32_t)((uint32_t)r >= (uint32_t)g.nRp) << 30);
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(Rrs, (int)((S ^ fm) + fb), 0, 0));
    };

    // ds_bpermute byte addresses of the theta dilation's lane reads, lanes c + WOFS[j]
    // taken modulo the wave
    int badr[NWIN];
#pragma unroll
    for (int j = 0; j < NWIN; ++j) badr[j] = ((lane + WOFS[j]) & 63) << 2;

    float Tr[TRING], pf[SROWS];
    int Mr[MRING]; // keys of the convolved rows (KEY_MASKED outside the array and the kept rows)
#pragma unroll
    for (int k = 0; k < TRING; ++k) Tr[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < MRING; ++k) Mr[k] = KEY_MASKED;
    int kmin = INT_MAX, kmax = KEY_MASKED; // statistics min/max as keys
    double ssum = 0.0;
    int nrun = 0, ncand = 0;

    const int rbeg = g.cr0 - SROWS, rend = max(g.cr1, g.mr1 + HR);
#pragma unroll
    for (int k = 0; k < SROWS; ++k) pf[k] = stage_load(rbeg + CKR + k);

    for (int r0 = rbeg; r0 < rend; r0 += SROWS) {
        const int d0 = r0 - HR; // this strip's dilation rows d0 .. d0+7
        // ---- stage rows r0+CKR .. r0+CKR+SROWS-1 (loaded as zero outside the padded
        //      array); fetch the next strip's ----
        float x[SROWS];
#pragma unroll
        for (int k = 0; k < SROWS; ++k) x[k] = pf[k];
#pragma unroll
        for (int k = 0; k < SROWS; ++k) pf[k] = stage_load(r0 + SROWS + CKR + k);
        // ---- theta convolution: T[c] = sum_b v[b] S[c + CKT - b] (fma order of k_peaks) ----
#pragma unroll
        for (int k = 0; k < SROWS; ++k) {
            float sh[KT];
            sh[CKT] = x[k];
#pragma unroll
            for (int d = 1; d <= CKT; ++d) {
                sh[CKT + d] = lane_next(sh[CKT + d - 1]);
                sh[CKT - d] = lane_prev(sh[CKT - d + 1]);
            }
            float s = 0.0f;
#pragma unroll
            for (int b = 0; b < KT; ++b) s = fmaf(g.v[b], sh[KT - 1 - b], s);
            Tr[2 * CKR + k] = s;
        }
        // ---- rho convolution: C[r] = sum_a u[a] T[r + CKR - a]; statistics; the ring keeps C
        //      with -inf outside the array and the kept rows (the dilation windows' clamp) ----
#pragma unroll
        for (int k = 0; k < SROWS; ++k) {
            float acc = 0.0f;
#pragma unroll
            for (int aa = 0; aa < KR; ++aa) acc = fmaf(g.u[aa], Tr[k + 2 * CKR - aa], acc);
            const int r = r0 + k;
            const int key = fkey(acc);
            // statistics rows [sr0, sr1), every lane (lanes outside the block are dropped
            // below).  The row test is uniform but taken as scalar factors, not a branch
            // (which the compiler turns into four selects per row): a row outside adds
            // acc * 0.0 (a zero: the float64 sum, which starts at +0, is unchanged; the
            // convolved values are finite) and clamps its key to the identity of min/max.
            const bool srow = (uint32_t)(r - g.sr0) < sr_n;
            kmin = min(kmin, max(key, srow ? INT_MIN : INT_MAX));
            kmax = max(kmax, min(key, srow ? INT_MAX : KEY_MASKED));
            ssum = fma((double)acc, srow ? 1.0 : 0.0, ssum);
            Mr[2 * HR + k] = (cval && (uint32_t)(r - g.cr0) < cr_n) ? key : KEY_MASKED;
        }
        // ---- rows d = d0 + k: rho dilation (windows Mr[k .. k + 2HR]), theta dilation,
        //      D == C under the mask -> candidates ----
        if (d0 + SROWS > g.mr0 && d0 < g.mr1) {
            // (in this branch -8 < d0 < nRp: rows d0 .. d0+7 lie in the lane-mask table)
            // all eight windows share Mr[SROWS-1 .. 2HR]
            int mid = Mr[SROWS - 1];
#pragma unroll
            for (int k = SROWS; k <= 2 * HR; ++k) mid = max(mid, Mr[k]);
            int lf[SROWS], rg[SROWS];
            lf[SROWS - 1] = KEY_MASKED;
#pragma unroll
            for (int k = SROWS - 2; k >= 0; --k) lf[k] = max(Mr[k], lf[k + 1]);
            rg[0] = KEY_MASKED;
#pragma unroll
            for (int k = 1; k < SROWS; ++k) rg[k] = max(rg[k - 1], Mr[2 * HR + k]);
            // theta dilation: w = the maximum over lanes c .. c+2 (two DPP lane shifts fused
            // into maxima), then D = the maximum of w at c and at lanes c + WOFS[j], windows of
            // three that together cover c-HT .. c+HT.  Those other lanes' w are ds_bpermute
            // lane reads, which issue to the LDS unit rather than the VALU.  A useful lane's
            // result depends on lanes c-HT .. c+HT only.
            // DGRP rows at a time: their lane reads all in flight, then their maxima (four rows
            // at a time need more than the 96 VGPRs of five waves per SIMD)
            uint64_t cand[SROWS], any = 0; // per row: the lanes whose cell is a candidate
            constexpr int DGRP = 2;
#pragma unroll
            for (int h = 0; h < SROWS; h += DGRP) {
                int w[DGRP], wr[DGRP][NWIN];
#pragma unroll
                for (int kk = 0; kk < DGRP; ++kk) {
                    const int vk = max(mid, max(lf[h + kk], rg[h + kk]));
                    w[kk] = max(vk, lane_next_i(max(vk, lane_next_i(vk))));
#pragma unroll
                    for (int j = 0; j < NWIN; ++j) wr[kk][j] = __builtin_amdgcn_ds_bpermute(badr[j], w[kk]);
                }
#pragma unroll
                for (int kk = 0; kk < DGRP; ++kk) {
                    const int k = h + kk;
                    int dk = w[kk];
#pragma unroll
                    for (int j = 0; j < NWIN; ++j) dk = max(dk, wr[kk][j]);
                    cand[k] = mlane[d0 + k] & __ballot(dk == Mr[k + HR]);
                    any |= cand[k];
                }
            }
            if (any) {
#pragma unroll
                for (int k = 0; k < SROWS; ++k) {
                    const uint64_t bal = cand[k];
                    if (bal == 0) continue;
                    const bool ck = (bal >> lane) & 1u;
                    const int n = __popcll(bal);
                    if (ncand + n > WCAP) {
                        nrun = wave_merge(lane, nB, nrun, ncand, run_v, run_p, run_nn, lst_v, lst_p, lst_nn, sel_v,
                                          sel_p, sel_nn);
                        ncand = 0;
                    }
                    // 3x3 neighbourhood: rows d-1, d, d+1 = Mr[k+HR-1 .. k+HR+1] (inside the kept
                    // rows and the array for every candidate); columns c-1, c, c+1
                    float nn[9];
#pragma unroll
                    for (int dr = 0; dr < 3; ++dr) {
                        const float cv = fkey_inv(Mr[k + HR - 1 + dr]);
                        nn[dr * 3 + 0] = lane_prev(cv);
                        nn[dr * 3 + 1] = cv;
                        nn[dr * 3 + 2] = lane_next(cv);
                    }
                    if (ck) {
                        const int slot = ncand + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                        lst_v[slot] = nn[4];
                        lst_p[slot] = (d0 + k) * g.nTp + c;
#pragma unroll
                        for (int e = 0; e < 9; ++e) lst_nn[9 * slot + e] = nn[e];
                    }
                    ncand += n;
                    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                }
            }
        }
        // ---- advance the rings by one strip ----
#pragma unroll
        for (int k = 0; k < TRING - SROWS; ++k) Tr[k] = Tr[k + SROWS];
#pragma unroll
        for (int k = 0; k < MRING - SROWS; ++k) Mr[k] = Mr[k + SROWS];
    }
    // ---- per wavefront: top-nB of its candidates, statistics ----
    nrun = wave_merge(lane, nB, nrun, ncand, run_v, run_p, run_nn, lst_v, lst_p, lst_nn, sel_v, sel_p, sel_nn);
    if (!useful) {
        kmin = INT_MAX;
        kmax = KEY_MASKED;
        ssum = 0.0;
    }
    for (int off = 32; off > 0; off >>= 1) {
        kmin = min(kmin, __shfl_xor(kmin, off, 64));
        kmax = max(kmax, __shfl_xor(kmax, off, 64));
        ssum += __shfl_xor(ssum, off, 64);
    }
    if (lane == 0) {
        wl[L.red + 0] = fkey_inv(kmin); // every wavefront has a column in the block
        wl[L.red + 1] = fkey_inv(kmax);
        *reinterpret_cast<double *>(wl + L.red + 2) = ssum;
        reinterpret_cast<int *>(wl)[L.cnt] = nrun;
    }
    __syncthreads();
    // ---- top-nB over the wavefronts' lists (wave 0), statistics (thread 0) ----
    float *fin_v = sm + L.fin_v, *fin_nn = sm + L.fin_nn, *stats = sm + L.stats;
    int *fin_p = reinterpret_cast<int *>(sm + L.fin_p);
    if (wave == 0) {
        int ntot = 0;
        for (int w = 0; w < nw; ++w) ntot += reinterpret_cast<const int *>(sm + w * L.per_wave)[L.cnt];
        for (int i = lane; i < ntot; i += 64) {
            int w = 0, e = i;
            for (;;) {
                const int nr = reinterpret_cast<const int *>(sm + w * L.per_wave)[L.cnt];
                if (e < nr) break;
                e -= nr;
                ++w;
            }
            const float *wv = sm + w * L.per_wave + L.run_v;
            const int *wp = reinterpret_cast<const int *>(sm + w * L.per_wave + L.run_p);
            const float v = wv[e];
            const int p = wp[e];
            int rank = 0;
            for (int w2 = 0; w2 < nw && rank < nB; ++w2) {
                const float *v2 = sm + w2 * L.per_wave + L.run_v;
                const int *p2 = reinterpret_cast<const int *>(sm + w2 * L.per_wave + L.run_p);
                const int nr2 = reinterpret_cast<const int *>(sm + w2 * L.per_wave)[L.cnt];
                for (int j = 0; j < nr2; ++j) rank += key_gt(v2[j], p2[j], v, p);
            }
            if (rank < nB) {
                fin_v[rank] = v;
                fin_p[rank] = p;
                const float *nn = sm + w * L.per_wave + L.run_nn + 9 * e;
                for (int k = 0; k < 9; ++k) fin_nn[9 * rank + k] = nn[k];
            }
        }
        if (lane == 0) {
            reinterpret_cast<int *>(stats)[3] = min(ntot, nB);
            float mn = sm[L.red], mx = sm[L.red + 1];
            double sum = *reinterpret_cast<const double *>(sm + L.red + 2);
            for (int w = 1; w < nw; ++w) {
                const float *r = sm + w * L.per_wave + L.red;
                mn = fminf(mn, r[0]);
                mx = fmaxf(mx, r[1]);
                sum += *reinterpret_cast<const double *>(r + 2);
            }
            const double cnt = (double)(g.sr1 - g.sr0) * (double)(g.sc1 - g.sc0);
            stats[0] = mn;
            stats[1] = (float)(sum / cnt);
            stats[2] = mx;
            S.out.d_stats[q * 3 + 0] = mn;
            S.out.d_stats[q * 3 + 1] = stats[1];
            S.out.d_stats[q * 3 + 2] = mx;
        }
    }
    __syncthreads();
    if (tid < nB)
        finalize_band(g, S, q, tid, reinterpret_cast<const int *>(stats)[3], fin_v, fin_p, fin_nn, stats[0],
                      stats[1]);
}

// Convolved padded Radon, every padded cell, same arithmetic order as k_peaks.
__global__ void k_conv_full(const float *radon, int64_t N, PeakGeo g, float *conv)
{
    const int64_t total = N * (int64_t)g.nRp * g.nTp;
    const int ckr = g.kr / 2, ckt = g.kt / 2;
    for (int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; id < total;
         id += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = id / ((int64_t)g.nRp * g.nTp);
        const int rem = (int)(id - q * (int64_t)g.nRp * g.nTp);
        const int r = rem / g.nTp, c = rem - (rem / g.nTp) * g.nTp;
        const float *R = radon + q * (int64_t)g.nR * g.nT;
        float s = 0.0f;
        for (int aa = 0; aa < g.kr; ++aa) {
            const int rr = r + ckr - aa;
            float t = 0.0f;
            for (int b = 0; b < g.kt; ++b) {
                const int cc = c + ckt - b;
                float val = 0.0f;
                if (rr >= 0 && rr < g.nRp && cc >= 0 && cc < g.nTp) val = R[pad_src(g, rr, cc)];
                t = fmaf(g.v[b], val, t);
            }
            s = fmaf(g.u[aa], t, s);
        }
        conv[id] = s;
    }
}

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------
template <typename T> static int dev_upload(T **dptr, const T *src, size_t n)
{
    if (n == 0) { *dptr = nullptr; return 0; }
    HIPCHK(hipMalloc((void **)dptr, n * sizeof(T)));
    HIPCHK(hipMemcpy(*dptr, src, n * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

static void free_seg(ebsd_plan::GSeg &g)
{
    for (ebsd_plan::Layout &l : g.layouts) {
        (void)hipFree(l.gidx);
        (void)hipFree(l.wofs);
        (void)hipFree(l.wcnt);
    }
    g.layouts.clear();
    (void)hipFree(g.d_bhdr);
    (void)hipFree(g.d_lane);
    (void)hipFree(g.d_lrc);
    g.d_lrc = nullptr;
    g.d_bhdr = nullptr;
    g.d_lane = nullptr;
}

static void free_rows(ebsd_plan *p)
{
    for (std::vector<ebsd_plan::GSeg> &rs : p->rowsets) {
        for (ebsd_plan::GSeg &g : rs) free_seg(g);
        rs.clear();
    }
}

static void plan_free(ebsd_plan *p)
{
    if (!p) return;
    free_seg(p->full);
    free_rows(p);
    (void)hipFree(p->d_part);
    for (uint32_t *r : p->part_retired) (void)hipFree(r);
    (void)hipFree(p->d_linebg);
    (void)hipFree(p->d_bg);
    (void)hipFree(p->d_mask);
    for (auto &t : p->mlane_tabs) (void)hipFree(t.second);
    (void)hipFree(p->d_mrow);
    (void)hipFree(p->d_px_off);
    (void)hipFree(p->d_px_cell);
    (void)hipFree(p->d_bp_tab);
    (void)hipFree(p->d_tscr);
    (void)hipFree(p->d_icnt);
    (void)hipFree(p->d_pre_off);
    (void)hipFree(p->d_pre);
    (void)hipFree(p->d_grid);
    delete p;
}

// Rank-1 split of the (exactly symmetric, Gaussian-derivative) kernel:
// u = centre column, v = centre row / centre value.
static void split_kernel(const float *k, int kr, int kt, float *u, float *v)
{
    const int cr = kr / 2, ct = kt / 2;
    const float kc = k[cr * kt + ct];
    for (int a = 0; a < kr; ++a) u[a] = k[a * kt + ct];
    for (int b = 0; b < kt; ++b) v[b] = (float)((double)k[cr * kt + b] / (double)kc);
}

// Streamed peak kernel: one wavefront per block of sbw theta columns, each lane one padded
// column, with KT/2 + HT halo lanes on either side; sbw = 0 where it does not apply
// (kernel sizes without an instantiation, or more than 16 wavefronts a pattern).
// theta columns a wavefront of the streamed kernel can own besides its halo lanes
static int stream_usable(const PeakGeo &g) { return 64 - 2 * (g.kt / 2 + g.ht); }

static void geo_stream(PeakGeo &g)
{
    g.sbw = 0;
    const bool fixed = g.hr == g.kr && g.ht == g.kt && ((g.kr == 9 && g.kt == 5) || (g.kr == 5 && g.kt == 5));
    if (!fixed) return;
    const int usable = stream_usable(g);
    const int nblk = (g.nT + usable - 1) / usable;
    if (nblk > 16) return;
    g.sbw = (g.nT + nblk - 1) / nblk;
}

static int stream_waves(const PeakGeo &g) { return g.sbw > 0 ? (g.nT + g.sbw - 1) / g.sbw : 0; }

static bool use_stream(const ebsd_plan *p) { return p->g.sbw > 0 && p->pkind != 1; }

// The streamed kernel's rho-mask for the plan's current theta block width: per wavefront w
// and padded row d in [-8, nRp + 8), bit `lane` is set where the lane owns an interior
// column (lanes [halo, halo + sbw), column w*sbw + lane - halo < nT) whose cell
// (d - p0, column) is in the mask.  One table per width, kept until the plan is destroyed.
static int build_mlane(ebsd_plan *p)
{
    const PeakGeo &g = p->g;
    if (g.sbw <= 0) { p->d_mlane = nullptr; return 0; }
    const auto it = p->mlane_tabs.find(g.sbw);
    if (it != p->mlane_tabs.end()) { p->d_mlane = it->second; return 0; }
    if (g.mr0 < 0) return fail(-3, "lane-mask table: mask rows start above the padded array");
    const int nw = stream_waves(g), rows = g.nRp + 16, halo = g.kt / 2 + g.ht;
    std::vector<uint64_t> t((size_t)nw * rows, 0ull);
    for (int w = 0; w < nw; ++w)
        for (int d = -8; d < g.nRp + 8; ++d) {
            const int i = d - g.p0;
            if (i < 0 || i >= g.nR) continue;
            uint64_t m = 0ull;
            for (int lane = halo; lane < halo + g.sbw && lane < 64; ++lane) {
                const int j = w * g.sbw + lane - halo;
                if (j < g.nT && p->h_rdnmask[(size_t)i * g.nT + j]) m |= 1ull << lane;
            }
            t[(size_t)w * rows + (d + 8)] = m;
        }
    uint64_t *d = nullptr;
    int rc;
    if ((rc = dev_upload(&d, t.data(), t.size()))) return rc;
    p->mlane_tabs[g.sbw] = d;
    p->d_mlane = d;
    return 0;
}

static int geo_tiles(ebsd_plan *p)
{
    PeakGeo &g = p->g;
    // a tile row never holds more cells than the candidate list (the overflow fallback
    // processes one row at a time)
    int tile = std::max(1, std::min(std::min(p->ptile, g.nT), CAND_CAP));
    g.rh = g.cr1 - g.cr0 + 2 * (g.kr / 2);
    g.maskw = (g.nT + 31) / 32;
    for (;;) {
        g.tile = tile;
        g.ntiles = (g.nT + tile - 1) / tile;
        const int cw = tile + 2 * g.ht;
        // the theta passes read whole 16-byte groups past the row end: 2*4 floats of slack
        g.apitch = ((cw + 2 * (g.kt / 2) + 3) / 4) * 4 + 8;
        g.bpitch = ((cw + 3) / 4) * 4;
        const size_t bytes = (size_t)peak_lds_layout(g).total * 4;
        if ((int)bytes <= p->lds_max || tile == 1) break;
        tile = std::max(1, tile * 3 / 4);
    }
    if ((size_t)peak_lds_layout(g).total * 4 > (size_t)p->lds_max)
        return fail(-2, "Radon too large for the peak kernel's LDS tiling");
    return 0;
}

// Lay the plan groups out as one contiguous stream per wave slot: slot v of nv walks batches
// v, v + nv, v + 2nv, ... back to back, so its next groups can be prefetched several
// iterations ahead without looking at batch boundaries.  Layouts are cached per nv.
// The schedule kind of LDS words of WB bytes (layout kinds 1..4).
static int sched_kind(int WB) { return WB <= 2 ? 1 : WB == 4 ? 2 : WB == 8 ? 3 : 4; }

static std::vector<uint16_t> bank_schedule(const std::vector<uint16_t> &g, const std::vector<uint2> &bhdr,
                                           uint16_t filler, int WB);

static int get_layout(ebsd_plan::GSeg &g, int nv, int kind, ebsd_plan::Layout *out, bool may_alloc = true)
{
    for (const ebsd_plan::Layout &l : g.layouts)
        if (l.nv == nv && l.kind == kind) { *out = l; return 0; }
    if (!may_alloc) {
        // any bank schedule gives the same integer sums: take one that is built
        for (const ebsd_plan::Layout &l : g.layouts)
            if (l.nv == nv && kind > 0 && l.kind > 0) { *out = l; return 0; }
        return fail(-2, "no-allocation path: the plan streams for " + std::to_string(nv) +
                            " wave slots are not built (the gather's work-group size changed?)");
    }
    if (kind > 0 && g.h_groups[kind].empty()) {
        static const int kWB[5] = {0, 2, 4, 8, 16};
        g.h_groups[kind] = bank_schedule(g.h_groups[0], g.h_bhdr, (uint16_t)g.nwords, kWB[kind]);
    }
    const std::vector<uint16_t> &hg = g.h_groups[kind];
    std::vector<uint16_t> st;
    std::vector<int> wofs(nv), wcnt(nv);
    st.reserve(hg.size() + (size_t)(nv + 1) * 64 * 8);
    int ofs = 0;
    for (int w = 0; w < nv; ++w) {
        wofs[w] = ofs;
        int cnt = 0;
        for (int b = w; b < g.nbatch; b += nv) {
            const uint2 h = g.h_bhdr[b];
            const size_t from = (size_t)h.x * 64 * 8;
            st.insert(st.end(), hg.begin() + from, hg.begin() + from + (size_t)h.y * 64 * 8);
            cnt += (int)h.y;
        }
        // a whole number of 4-group steps: the walk loop then has no per-group bound test
        const int pad = (4 - cnt % 4) % 4;
        st.resize(st.size() + (size_t)pad * 64 * 8, (uint16_t)g.nwords);
        cnt += pad;
        wcnt[w] = cnt;
        ofs += cnt;
    }
    // filler groups past the end keep the prefetch queue's reads in bounds
    st.resize(st.size() + 8 * 64 * 8, (uint16_t)g.nwords);
    ebsd_plan::Layout l{nv, kind, nullptr, nullptr, nullptr};
    if (dev_upload(reinterpret_cast<uint16_t **>(&l.gidx), st.data(), st.size()) ||
        dev_upload(&l.wofs, wofs.data(), wofs.size()) || dev_upload(&l.wcnt, wcnt.data(), wcnt.size())) {
        (void)hipFree(l.gidx);
        (void)hipFree(l.wofs);
        (void)hipFree(l.wcnt);
        return fail(-3, "plan stream layout: device allocation failed");
    }
    g.layouts.push_back(l);
    *out = l;
    return 0;
}

// The streams of one slice per pack for the configured work-group size, in both entry orders:
// built with the segment, so that the no-allocation path finds them.
static int prebuild_layouts(const ebsd_plan *p, ebsd_plan::GSeg &g)
{
    ebsd_plan::Layout l;
    int rc;
    if ((rc = get_layout(g, p->gthreads / 64, 0, &l))) return rc;
    return get_layout(g, p->gthreads / 64, sched_kind(8), &l); // the default packs' 8-byte words
}

// Slices per pack so that a launch of npacks packs has about two resident rounds of
// work-groups, keeping at least two batches per slot.
static int choose_slices(const ebsd_plan *p, const ebsd_plan::GSeg &g, int64_t npacks, int nw, int occ)
{
    const int64_t target = 2 * (int64_t)occ * p->ncu;
    if (npacks >= target) return 1;
    const int smax = std::max(1, g.nbatch / (2 * nw));
    return (int)std::max<int64_t>(1, std::min<int64_t>(smax, (target + npacks - 1) / npacks));
}

// Default staged row pitch: rows of W words put the pixels of a near-vertical line in a few
// LDS banks when W is a multiple of 8 or 16 (measured: 160x160 u16 gathers 2x faster at
// pitch 163 than at 160, 120x120 5% faster at 122 than at 120), so pad by 3 (W a multiple of
// 16) or 2 words -- unless the padding would push a word size that fits at W out of LDS.
static int default_pitch(const ebsd_plan *p)
{
    const int W = p->W, H = p->H;
    const int want = (W % 16 == 0) ? 3 : (W % 8 == 0 ? 2 : 0);
    for (int pad = want; pad > 0; --pad) {
        bool ok = H * (W + pad) <= 65535;
        for (int wb = 2; wb <= 16 && ok; wb *= 2) {
            const bool fit0 = (size_t)(H * W + 1) * wb + kGatherLdsExtra <= (size_t)p->lds_max;
            const bool fit1 = (size_t)(H * (W + pad) + 1) * wb + kGatherLdsExtra <= (size_t)p->lds_max;
            ok = !fit0 || fit1;
        }
        if (ok) return W + pad;
    }
    return W;
}

// Integer line sums do not depend on the order of their entries, so each line's entries can
// be reordered to spread the LDS words that a wave reads at one step over the banks.  The LDS
// serves 256 bytes a clock over 64 four-byte banks: a read of W-byte words is served G = 256 /
// max(W, 4) lanes at a time (64 for 2- and 4-byte words, 32 for 8, 16 for 16), and lane
// group's words collide when they differ but share a bank set -- word w's residue w mod G
// (a 2-byte word shares its dword, index w / 2, with its neighbour: same dword, no collision).
// Measured on the MI355X: conflict-free ds_read_b64 135 TB/s, ds_read_b128 150 TB/s, and the
// gather's bank-conflict cycles were 54% of its LDS cycles under the 16-lane model this
// replaces.  Greedy per step and lane group: lanes with the fewest distinct residues left
// choose first, each taking a residue that is least loaded at this step (ties: the residue
// the lane has most entries left in).  Float sums keep plan order.

static std::vector<uint16_t> bank_schedule(const std::vector<uint16_t> &g, const std::vector<uint2> &bhdr,
                                           uint16_t filler, int WB)
{
    const int G = 256 / std::max(WB, 4); // lanes served together = bank sets
    const int NR = G;
    const int shift = WB == 2 ? 1 : 0;   // 2-byte words: the dword is the bank item
    std::vector<uint16_t> out(g.size(), filler);
    std::vector<std::vector<std::vector<uint16_t>>> pool(G, std::vector<std::vector<uint16_t>>(NR));
    std::vector<int> nres(G), lanes(G), nused(NR);
    std::vector<std::vector<int>> used(NR);
    for (const uint2 &h : bhdr) {
        const int steps = (int)h.y * 8;
        for (int grp = 0; grp < 64 / G; ++grp) {
            for (int l = 0; l < G; ++l)
                for (int r = 0; r < NR; ++r) pool[l][r].clear();
            for (int l = 0; l < G; ++l) {
                const int lane = grp * G + l;
                for (int k = 0; k < steps; ++k) {
                    const uint16_t w = g[((size_t)(h.x + k / 8) * 64 + lane) * 8 + (k % 8)];
                    if (w != filler) pool[l][(w >> shift) % NR].push_back(w);
                }
            }
            for (int l = 0; l < G; ++l) {
                int n = 0;
                for (int r = 0; r < NR; ++r) n += !pool[l][r].empty();
                nres[l] = n; // distinct residues the lane has left (kept up to date below)
            }
            for (int k = 0; k < steps; ++k) {
                for (int l = 0; l < G; ++l) lanes[l] = l;
                std::stable_sort(lanes.begin(), lanes.end(), [&](int x, int y) {
                    const int ax = nres[x] ? nres[x] : 1 << 20, ay = nres[y] ? nres[y] : 1 << 20;
                    return ax < ay;
                });
                for (int r = 0; r < NR; ++r) {
                    nused[r] = 0;
                    used[r].clear();
                }
                auto load = [&](uint16_t w) {
                    const int item = w >> shift, r = item % NR;
                    for (int u : used[r])
                        if (u == item) return;
                    used[r].push_back(item);
                    ++nused[r];
                };
                for (int li = 0; li < G; ++li) {
                    const int l = lanes[li];
                    uint16_t w = filler;
                    if (nres[l]) {
                        int best = -1;
                        for (int r = 0; r < NR; ++r) {
                            if (pool[l][r].empty()) continue;
                            if (best < 0 || nused[r] < nused[best] ||
                                (nused[r] == nused[best] && pool[l][r].size() > pool[l][best].size()))
                                best = r;
                        }
                        w = pool[l][best].back();
                        pool[l][best].pop_back();
                        if (pool[l][best].empty()) --nres[l];
                    }
                    load(w);
                    out[((size_t)(h.x + k / 8) * 64 + grp * G + l) * 8 + (k % 8)] = w;
                }
            }
        }
    }
    return out;
}

// Count of each line's entries in staged rows [r0, r1).
static std::vector<int> row_counts(const ebsd_plan *p, int r0, int r1)
{
    const int e0 = r0 * p->W, e1 = r1 * p->W;
    std::vector<int> cs(p->nCells, 0);
    for (int cell = 0; cell < p->nCells; ++cell)
        for (int64_t k = p->h_off[cell]; k < p->h_off[cell + 1]; ++k)
            cs[cell] += (p->h_ent[k] >= e0 && p->h_ent[k] < e1);
    return cs;
}

// The gather's line order: lines in units of U neighbouring cells (U x 4 bytes of a
// pattern's Radon; U = 8 is one 32-byte write sector), the units by their largest group count
// in each row range (descending, the first range first) and then in cell order.  A batch's
// store of one pattern then writes whole sectors: with single lines sorted by group count a
// batch's 64 values fall into about twice as many 32-byte sectors as they fill, and the
// sectors are written piecewise (1.7x the Radon's bytes reach HBM).  The price is padding:
// a unit's lines run to its longest line's group count.  A partial last unit is filled with
// empty lanes (-1) so that no unit straddles a 64-line batch.
static std::vector<int> unit_order(int nCells, int U, const std::vector<std::vector<int>> &grp)
{
    const int nu = (nCells + U - 1) / U;
    const size_t S = grp.size();
    std::vector<int> key((size_t)nu * S, 0);
    for (int c = 0; c < nCells; ++c)
        for (size_t s = 0; s < S; ++s) {
            int &k = key[(size_t)(c / U) * S + s];
            k = std::max(k, grp[s][c]);
        }
    std::vector<int> units(nu);
    for (int u = 0; u < nu; ++u) units[u] = u;
    std::stable_sort(units.begin(), units.end(), [&](int x, int y) {
        for (size_t s = 0; s < S; ++s)
            if (key[(size_t)x * S + s] != key[(size_t)y * S + s]) return key[(size_t)x * S + s] > key[(size_t)y * S + s];
        return false;
    });
    std::vector<int> order;
    order.reserve((size_t)nu * U);
    for (int u : units)
        for (int j = 0; j < U; ++j) order.push_back(u * U + j < nCells ? u * U + j : -1);
    return order;
}

// The lines' entries in staged rows [r0, r1), lines in the given order (or in units ordered by
// their group count there) and cut into 64-line batches of 8-entry groups (one lane per line; -1 an empty lane); entries
// are LDS word indices of the staged pixels, (y - r0) * spitch + x.  Every line has a lane in
// every row range (a line with no pixel there adds zero), and the lane keeps the count of the
// whole line.  The device streams of a work-group size are laid out on first use.
static int compile_seg(ebsd_plan *p, ebsd_plan::GSeg &g, int r0, int r1, const std::vector<int> *given = nullptr)
{
    const int nCells = p->nCells, W = p->W;
    free_seg(g);
    g.r0 = r0;
    g.r1 = r1;
    g.nwords = (r1 - r0) * p->spitch;
    if (g.nwords > 65535) return fail(-2, "ebsd_plan: staged pattern exceeds 65,535 LDS words");
    const int e0 = r0 * W, e1 = r1 * W;
    const std::vector<int> cs = row_counts(p, r0, r1);
    std::vector<int> order;
    if (given) {
        order = *given;
    } else {
        // by group count (descending), cell order within a count: lines with equal group
        // counts pad alike, and a batch of neighbouring cells stores its line values into
        // few cache lines (0.28 -> 0.25 ms per 2,048 120x120 u16 patterns against a sort by
        // pixel count)
        std::vector<std::vector<int>> grp(1, cs);
        for (int &c : grp[0]) c = (c + 7) / 8;
        order = unit_order(nCells, p->lunit, grp);
    }
    const int nl = (int)order.size();
    g.nbatch = (nl + 63) / 64;
    std::vector<uint2> bhdr(g.nbatch);
    std::vector<int2> lanes((size_t)g.nbatch * 64);
    std::vector<uint16_t> gidx;
    uint32_t blocks = 0;
    for (int b = 0; b < g.nbatch; ++b) {
        int mx = 0;
        for (int l = 0; l < 64; ++l) {
            const int i = b * 64 + l;
            if (i < nl && order[i] >= 0) mx = std::max(mx, cs[order[i]]);
        }
        const int ngrp = std::max(1, (mx + 7) / 8); // an empty batch still writes its cells
        bhdr[b] = make_uint2(blocks, (uint32_t)ngrp);
        gidx.resize(gidx.size() + (size_t)ngrp * 64 * 8, (uint16_t)g.nwords);
        for (int l = 0; l < 64; ++l) {
            const int i = b * 64 + l;
            if (i >= nl || order[i] < 0) { lanes[(size_t)b * 64 + l] = make_int2(-1, 0); continue; }
            const int cell = order[i];
            lanes[(size_t)b * 64 + l] = make_int2(cell, p->h_cnt[cell]);
            int k = 0;
            for (int64_t j = p->h_off[cell]; j < p->h_off[cell + 1]; ++j) {
                const int e = p->h_ent[j];
                if (e < e0 || e >= e1) continue;
                const size_t at = ((size_t)(blocks + k / 8) * 64 + l) * 8 + (k % 8);
                gidx[at] = (uint16_t)((e / W - r0) * p->spitch + e % W);
                ++k;
            }
        }
        blocks += (uint32_t)ngrp;
    }
    for (int k = 1; k < 5; ++k) g.h_groups[k].clear();
    g.h_groups[0].swap(gidx);
    g.h_bhdr = bhdr;
    int rc;
    if ((rc = dev_upload(&g.d_bhdr, bhdr.data(), bhdr.size()))) return rc;
    if ((rc = dev_upload(&g.d_lane, lanes.data(), lanes.size()))) return rc;
    std::vector<double> lrc(lanes.size(), 0.0);
    for (size_t i = 0; i < lanes.size(); ++i)
        if (lanes[i].y > 0) lrc[i] = 1.0 / (double)lanes[i].y;
    if ((rc = dev_upload(&g.d_lrc, lrc.data(), lrc.size()))) return rc;
    return prebuild_layouts(p, g);
}

// S row segments of (nearly) equal height, segment s holding rows [s H / S, (s + 1) H / S).
// All segments share one line order, so that a line keeps its batch and lane -- its partial
// sum's position -- from segment to segment: units of lines sorted by their group counts per
// segment, segment 0 first (descending), which keeps the lines of a batch of similar length in
// every segment (about 96% of the entry slots used at 240x240 in two segments, 87% in four,
// with single-line units).
static int compile_rows(ebsd_plan *p, int S)
{
    if (S < 2 || S > 8) return fail(-1, "row segments must be in 2..8");
    std::vector<ebsd_plan::GSeg> &rows = p->rowsets[S];
    if ((int)rows.size() == S) return 0;
    for (ebsd_plan::GSeg &g : rows) free_seg(g);
    rows.clear();
    const int nCells = p->nCells;
    std::vector<std::vector<int>> q(S);
    for (int s = 0; s < S; ++s) {
        q[s] = row_counts(p, s * p->H / S, (s + 1) * p->H / S);
        for (int &c : q[s]) c = (c + 7) / 8;
    }
    const std::vector<int> order = unit_order(nCells, p->lunit, q);
    rows.resize(S);
    for (int s = 0; s < S; ++s) {
        const int rc = compile_seg(p, rows[s], s * p->H / S, (s + 1) * p->H / S, &order);
        if (rc) {
            for (ebsd_plan::GSeg &g : rows) free_seg(g);
            rows.clear();
            return rc;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------------------
// no-allocation scratch (ebsd_find_bands_noalloc): a file-scope device array, one copy per
// GPU, held until the process ends; it takes each chunk's Radon and, where the gather runs
// in row segments, their partial sums
// ---------------------------------------------------------------------------------------
static constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices]; // per device ordinal: the array's device address

// The device address of the current device's scratch array, resolved once per device with
// that device current.  In host code the array's name is only a host-side stand-in: kernels
// and copies take this address.
static int scratch_address(int device, void **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) return fail(-3, "no-allocation scratch: device ordinal out of range");
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(-3, std::string("no-allocation scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return 0;
}

static int row_segments(const ebsd_plan *p, int dtype);

// The whole-pattern batches, and the row segments that uint16 patterns (and float staging of
// the same word size) use where a whole pattern does not fit; other segment counts are
// compiled on first use.
static int compile_groups(ebsd_plan *p)
{
    free_rows(p);
    int rc = compile_seg(p, p->full, 0, p->H);
    if (!rc && row_segments(p, EBSD_U16) > 1) rc = compile_rows(p, row_segments(p, EBSD_U16));
    return rc;
}

extern "C" int ebsd_plan_create(int H, int W, int nTheta, int nRho, const int64_t *index_plan,
                                int plan_len, const uint8_t *rdnmask, const float *kernel, int kr,
                                int kt, int pad_r, int pad_t, int peak_r, int peak_t, int nBands,
                                const float *theta_deg, const float *rho, ebsd_plan **out)
{
    if (!out || !index_plan || !rdnmask || !kernel || !theta_deg || !rho)
        return fail(-1, "ebsd_plan_create: null argument");
    *out = nullptr;
    if (H <= 0 || W <= 0 || nTheta <= 0 || nRho <= 0 || plan_len <= 0)
        return fail(-1, "ebsd_plan_create: non-positive size");
    if ((int64_t)H * W + 1 > 65536)
        return fail(-1, "ebsd_plan_create: patterns above 65535 pixels are not supported");
    if (kr % 2 == 0 || kt % 2 == 0 || kr > KMAX || kt > KMAX)
        return fail(-1, "ebsd_plan_create: kernel sizes must be odd and <= 16");
    if (nBands <= 0 || nBands > NB_MAX) return fail(-1, "ebsd_plan_create: nBands must be in 1..64");
    if (pad_t > nTheta) return fail(-1, "ebsd_plan_create: theta padding larger than nTheta");
    if (peak_r % 2 == 0 || peak_t % 2 == 0)
        return fail(-1, "ebsd_plan_create: peakPad must be odd");

    ebsd_plan *p = new ebsd_plan();
    int rc;
    if (hipGetDevice(&p->device) != hipSuccess) { delete p; return fail(-3, "hipGetDevice failed"); }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, p->device) != hipSuccess) {
        delete p;
        return fail(-3, "hipGetDeviceProperties failed");
    }
    p->ncu = prop.multiProcessorCount;
    if ((rc = scratch_address(p->device, &p->scratch))) { delete p; return rc; }
    p->lds_max = (int)prop.sharedMemPerBlock;
    if (prop.maxSharedMemoryPerMultiProcessor > 0)
        p->lds_max = std::max(p->lds_max, (int)std::min<size_t>(prop.maxSharedMemoryPerMultiProcessor, 160 * 1024));
    p->H = H; p->W = W; p->nPx = H * W; p->nT = nTheta; p->nR = nRho; p->nCells = nTheta * nRho;

    // ---- compile the gather plan ----
    const int nPx = p->nPx, nCells = p->nCells;
    p->h_cnt.assign(nCells, 0);
    p->h_off.assign(nCells + 1, 0);
    p->h_ent.clear();
    p->h_ent.reserve((size_t)nCells * std::max(H, W));
    for (int cell = 0; cell < nCells; ++cell) {
        const int64_t *line = index_plan + (int64_t)cell * plan_len;
        p->h_off[cell] = (int64_t)p->h_ent.size();
        int n = 0;
        for (int k = 0; k < plan_len; ++k) {
            const int64_t id = line[k];
            if (id >= nPx) break; // rdn_loops stops at the first out-of-range entry
            if (id < 0) { plan_free(p); return fail(-1, "ebsd_plan_create: negative plan entry"); }
            p->h_ent.push_back((uint16_t)id);
            ++n;
        }
        p->h_cnt[cell] = n;
    }
    p->h_off[nCells] = (int64_t)p->h_ent.size();
    p->spitch = default_pitch(p);
    if ((rc = compile_groups(p))) { plan_free(p); return rc; }

    // ---- peak geometry ----
    PeakGeo &g = p->g;
    g.nR = nRho; g.nT = nTheta; g.p0 = pad_r; g.p1 = pad_t;
    g.nRp = nRho + 2 * pad_r; g.nTp = nTheta + 2 * pad_t;
    g.kr = kr; g.kt = kt; g.hr = peak_r / 2; g.ht = peak_t / 2;
    g.nB = nBands;
    split_kernel(kernel, kr, kt, g.u, g.v);
    int mlo = nRho, mhi = -1;
    for (int i = 0; i < nRho; ++i)
        for (int j = 0; j < nTheta; ++j)
            if (rdnmask[i * nTheta + j]) { mlo = std::min(mlo, i); mhi = std::max(mhi, i); }
    if (mhi < 0) { mlo = 0; mhi = -1; }
    g.mr0 = pad_r + mlo;
    g.mr1 = pad_r + mhi + 1;
    g.sr0 = pad_r;
    g.sr1 = std::min(g.nRp, g.nTp - pad_r);
    g.sc0 = pad_t;
    g.sc1 = g.nTp - pad_t;
    if (g.sr1 <= g.sr0 || g.sc1 <= g.sc0) { plan_free(p); return fail(-1, "ebsd_plan_create: empty stats region"); }
    int lo = g.sr0, hi = g.sr1;
    if (g.mr1 > g.mr0) {
        lo = std::min(lo, std::max(g.mr0 - g.hr, 0));
        hi = std::max(hi, std::min(g.mr1 + g.hr, g.nRp));
        lo = std::min(lo, g.mr0 - 1);
        hi = std::max(hi, g.mr1 + 1);
    }
    g.cr0 = std::max(lo, 0);
    g.cr1 = std::min(hi, g.nRp);
    if ((rc = geo_tiles(p))) { plan_free(p); return rc; }
    geo_stream(g);
    if ((rc = dev_upload(&p->d_mask, rdnmask, (size_t)nCells))) { plan_free(p); return rc; }
    {
        p->h_rdnmask.assign(rdnmask, rdnmask + nCells);
        if ((rc = build_mlane(p))) { plan_free(p); return rc; }
        // row-major mask bits; bit b of word w of row i is rdnmask[i][32 w + b]
        std::vector<uint32_t> mrow((size_t)nRho * g.maskw, 0u);
        for (int i = 0; i < nRho; ++i)
            for (int j = 0; j < nTheta; ++j)
                if (rdnmask[i * nTheta + j]) mrow[(size_t)i * g.maskw + (j >> 5)] |= 1u << (j & 31);
        if ((rc = dev_upload(&p->d_mrow, mrow.data(), mrow.size()))) { plan_free(p); return rc; }
    }
    std::vector<double> grid(nTheta + nRho);
    for (int j = 0; j < nTheta; ++j) grid[j] = (double)theta_deg[j];
    for (int i = 0; i < nRho; ++i) grid[nTheta + i] = (double)rho[i];
    if ((rc = dev_upload(&p->d_grid, grid.data(), grid.size()))) { plan_free(p); return rc; }
    *out = p;
    return 0;
}

extern "C" int ebsd_plan_destroy(ebsd_plan *plan)
{
    plan_free(plan);
    return 0;
}

// A batch keeps device pointers into its plans (plan streams, background, masks) in its
// kernel arguments and captured graph: while one exists, the plan's set-up is frozen.
static int check_unpinned(const ebsd_plan *p, const char *fn)
{
    if (p->pins > 0)
        return fail(-1, std::string(fn) + ": the plan is bound to a batch (ebsd_batch_create); "
                                          "destroy the batch before changing the plan");
    return 0;
}

extern "C" int ebsd_plan_set_background(ebsd_plan *p, const double *bg, int exact)
{
    if (!p) return fail(-1, "ebsd_plan_set_background: null plan");
    if (check_unpinned(p, "ebsd_plan_set_background")) return -1;
    (void)hipFree(p->d_linebg);
    (void)hipFree(p->d_bg);
    p->d_linebg = nullptr;
    p->d_bg = nullptr;
    p->bg_mode = 0;
    if (!bg) return 0;
    int rc;
    // the per-pixel background is always kept: float32 patterns stage float32(x - bg)
    if ((rc = dev_upload(&p->d_bg, bg, (size_t)p->nPx))) return rc;
    if (exact) {
        p->bg_mode = 2;
    } else {
        std::vector<double> lb(p->nCells, 0.0);
        for (int cell = 0; cell < p->nCells; ++cell) {
            double s = 0.0;
            for (int64_t k = p->h_off[cell]; k < p->h_off[cell + 1]; ++k) s += bg[p->h_ent[k]];
            lb[cell] = s;
        }
        if ((rc = dev_upload(&p->d_linebg, lb.data(), lb.size()))) return rc;
        p->bg_mode = 1;
    }
    return 0;
}

extern "C" int ebsd_plan_set_stage_pitch(ebsd_plan *p, int pitch)
{
    if (!p) return fail(-1, "ebsd_plan_set_stage_pitch: null plan");
    if (check_unpinned(p, "ebsd_plan_set_stage_pitch")) return -1;
    if (pitch <= 0) pitch = p->W;
    if (pitch < p->W) return fail(-1, "ebsd_plan_set_stage_pitch: pitch below the pattern width");
    if (pitch == p->spitch) return 0;
    const int old = p->spitch;
    p->spitch = pitch;
    const int rc = compile_groups(p);
    if (rc) {
        p->spitch = old;
        (void)compile_groups(p);
    }
    return rc;
}

extern "C" int ebsd_plan_set_line_unit(ebsd_plan *p, int unit)
{
    if (!p) return fail(-1, "ebsd_plan_set_line_unit: null plan");
    if (check_unpinned(p, "ebsd_plan_set_line_unit")) return -1;
    if (unit <= 0) unit = 8;
    if (unit > 64 || (unit & (unit - 1))) return fail(-1, "ebsd_plan_set_line_unit: unit must be a power of two <= 64");
    if (unit == p->lunit) return 0;
    const int old = p->lunit;
    p->lunit = unit;
    const int rc = compile_groups(p);
    if (rc) {
        p->lunit = old;
        (void)compile_groups(p);
    }
    return rc;
}

extern "C" int ebsd_plan_set_row_segments(ebsd_plan *p, int segments)
{
    if (!p) return fail(-1, "ebsd_plan_set_row_segments: null plan");
    if (check_unpinned(p, "ebsd_plan_set_row_segments")) return -1;
    if (segments < 0 || segments > 8) return fail(-1, "ebsd_plan_set_row_segments: segments must be in 0..8");
    p->rseg = segments;
    return 0;
}

extern "C" int ebsd_plan_set_peak_kernel(ebsd_plan *p, int kind)
{
    if (!p) return fail(-1, "ebsd_plan_set_peak_kernel: null plan");
    if (check_unpinned(p, "ebsd_plan_set_peak_kernel")) return -1;
    if (kind < 0 || kind > 2) return fail(-1, "ebsd_plan_set_peak_kernel: kind is 0, 1 or 2");
    if (kind == 2 && p->g.sbw <= 0) return fail(-1, "ebsd_plan_set_peak_kernel: the streamed kernel does not apply to this plan");
    p->pkind = kind;
    return 0;
}

extern "C" int ebsd_plan_set_peak_block(ebsd_plan *p, int theta_cols)
{
    if (!p) return fail(-1, "ebsd_plan_set_peak_block: null plan");
    if (check_unpinned(p, "ebsd_plan_set_peak_block")) return -1;
    PeakGeo g = p->g;
    geo_stream(g);                           // the default block (and whether it applies)
    if (theta_cols == 0) {
        p->g.sbw = g.sbw;
        return build_mlane(p);
    }
    if (g.sbw <= 0) return fail(-1, "ebsd_plan_set_peak_block: the streamed kernel does not apply to this plan");
    if (theta_cols < 1 || theta_cols > stream_usable(g))
        return fail(-1, "ebsd_plan_set_peak_block: theta_cols must be in 1.." + std::to_string(stream_usable(g)));
    if ((g.nT + theta_cols - 1) / theta_cols > 16)
        return fail(-1, "ebsd_plan_set_peak_block: more than 16 wavefronts a pattern");
    const int prev = p->g.sbw;
    p->g.sbw = theta_cols;
    const int rc = build_mlane(p);
    if (rc) {
        p->g.sbw = prev;
        (void)build_mlane(p); // the previous width's table exists: no allocation
    }
    return rc;
}

extern "C" int ebsd_plan_set_tuning(ebsd_plan *p, int gather_pack, int gather_threads, int peak_tile,
                                    int peak_threads)
{
    if (!p) return fail(-1, "ebsd_plan_set_tuning: null plan");
    if (check_unpinned(p, "ebsd_plan_set_tuning")) return -1;
    if (gather_pack > 0) {
        // one value per element type, interpreted by the type-specific instantiation
        p->gpack_u8 = std::min(8, std::max(2, gather_pack * 2));
        p->gpack_u16 = std::min(8, std::max(1, gather_pack));
        p->gpack_f32 = std::min(4, std::max(1, gather_pack / 2));
    }
    if (gather_threads > 0) {
        if (gather_threads % 64 || gather_threads > 1024) return fail(-1, "gather_threads must be a multiple of 64 <= 1024");
        p->gthreads = gather_threads;
        int rc = prebuild_layouts(p, p->full);
        for (std::vector<ebsd_plan::GSeg> &rs : p->rowsets)
            for (ebsd_plan::GSeg &g : rs)
                if (!rc) rc = prebuild_layouts(p, g);
        if (rc) return rc;
    }
    if (peak_threads > 0) {
        if (peak_threads % 64 || peak_threads > 512) return fail(-1, "peak_threads must be a multiple of 64 <= 512");
        p->pthreads = peak_threads;
    }
    if (peak_tile > 0) {
        p->ptile = peak_tile;
        int rc = geo_tiles(p);
        if (rc) return rc;
    }
    return 0;
}

extern "C" int ebsd_plan_info(const ebsd_plan *p, int64_t *o)
{
    if (!p || !o) return fail(-1, "ebsd_plan_info: null argument");
    int64_t total = (int64_t)p->h_ent.size();
    o[0] = p->nCells;
    o[1] = total;                 // line-pixel pairs per pattern
    o[2] = p->full.nbatch;
    o[3] = p->g.tile;
    o[4] = p->g.ntiles;
    o[5] = (int64_t)peak_lds_layout(p->g).total * 4;
    o[6] = p->bg_mode;
    o[7] = p->ncu;
    int64_t slots = 0;
    for (const uint2 &h : p->full.h_bhdr) slots += (int64_t)h.y * 64 * 8;
    o[8] = slots;
    o[9] = p->lunit;
    return 0;
}

extern "C" size_t ebsd_workspace_bytes(const ebsd_plan *p, int64_t chunk)
{
    if (!p || chunk <= 0) return 0;
    return (size_t)chunk * (size_t)p->nCells * sizeof(float);
}

// One launch over all packs for the row range of g (maxwords: the largest range's words, so
// that every segment of a launch sequence fits the same pack size).
template <int IN, bool FLT, int P>
static int launch_gather_t(ebsd_plan *p, ebsd_plan::GSeg &g, int maxwords, int segmode, uint32_t *part,
                           const void *d_pats, int64_t N, float *d_out, bool use_bg, bool noalloc,
                           hipStream_t s)
{
    typedef GatherCfg<IN, FLT, P> C;
    const size_t lds = (size_t)(maxwords + 1) * C::WB + kGatherLdsExtra;
    if ((int)lds > p->lds_max) return 1; // caller falls back to a smaller pack
    auto kern = k_radon_gather<IN, FLT, P>;
    // the cap is the device's LDS, not this launch's size: a graph that captured launches of
    // one instantiation with different sizes then replays each within the cap
    HIPCHK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, p->lds_max));
    int occ = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, p->gthreads, lds));
    if (occ < 1) return fail(-3, "radon gather kernel cannot be resident");
    GatherArgs a{};
    a.pats = d_pats;
    a.N = N;
    a.nPx = p->nPx;
    a.W = p->W;
    a.pitch = p->spitch;
    a.px0 = g.r0 * p->W;
    a.rows = g.r1 - g.r0;
    a.segmode = segmode;
    a.part = part;
    a.npos = g.nbatch * 64;
    a.nbatch = g.nbatch;
    a.bhdr = g.d_bhdr;
    a.lane = g.d_lane;
    a.lrc = g.d_lrc;
    a.npacks = (N + P - 1) / P;
    const int nw = p->gthreads / 64;
    // the no-allocation path walks the plan streams prebuilt for one slice per pack
    a.nslice = noalloc ? 1 : choose_slices(p, g, a.npacks, nw, occ);
    ebsd_plan::Layout lay;
    int rc;
    if ((rc = get_layout(g, nw * a.nslice, FLT ? 0 : sched_kind(C::WB), &lay, !noalloc))) return rc;
    a.gidx = lay.gidx;
    a.wofs = lay.wofs;
    a.wcnt = lay.wcnt;
    a.linebg = (use_bg && !FLT && p->bg_mode == 1) ? p->d_linebg : nullptr;
    a.bg = (use_bg && FLT) ? p->d_bg : nullptr;
    a.out = d_out;
    a.nCells = p->nCells;
    const int64_t grid = std::min<int64_t>(a.npacks * a.nslice, (int64_t)occ * p->ncu);
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(p->gthreads), lds, s, a);
    HIPCHK(hipGetLastError());
    return 0;
}

// Integer sums of P patterns per word over the whole pattern (S = 1) or over S row segments
// in S launches whose partial sums add up in the output (exact: integers).
static int ensure_part(ebsd_plan *p, size_t elems)
{
    if (p->part_cap >= elems) return 0;
    // grow geometrically; the outgrown buffer is kept until the plan is destroyed, so a launch
    // still queued on another stream never writes into freed memory (and no device-wide
    // synchronisation is needed here)
    const size_t want = std::max(elems, p->part_cap + p->part_cap / 2);
    uint32_t *fresh = nullptr;
    if (hipMalloc((void **)&fresh, want * sizeof(uint32_t)) != hipSuccess)
        return fail(-3, "row-segment partial sums: device allocation failed");
    if (p->d_part) p->part_retired.push_back(p->d_part);
    p->d_part = fresh;
    p->part_cap = want;
    return 0;
}

// Partial sums per launch sequence: at most this many bytes (larger N runs in pieces).
static const int64_t kPartBytes = 2ll << 30;

// Where a gather's row-segment partial sums live, and whether it may allocate.
struct GatherOpts {
    uint32_t *part = nullptr; // caller's partial-sum buffer (nullptr: the plan's, grown on demand)
    size_t part_bytes = 0;
    bool noalloc = false;     // no device allocation: prebuilt plan streams, one slice per pack
};

// Patterns of one launch sequence over S row segments whose partial sums (pbytes per line
// and pattern) fit in cap_bytes: whole packs, in equal pieces.
static int64_t seg_piece(int64_t N, int P, int64_t npos, int pbytes, size_t cap_bytes)
{
    const int64_t cap = std::max<int64_t>(P, (int64_t)(cap_bytes / (size_t)pbytes / (size_t)npos) / P * P);
    const int64_t npieces = (N + cap - 1) / cap;
    return std::min(cap, ((N + npieces - 1) / npieces + P - 1) / P * P);
}

template <int IN, bool FLT, int P>
static int launch_gather_rows(ebsd_plan *p, int S, const void *d_pats, int64_t N, float *d_out, bool use_bg,
                              const GatherOpts &o, hipStream_t s)
{
    if (S <= 1)
        return launch_gather_t<IN, FLT, P>(p, p->full, p->full.nwords, 0, nullptr, d_pats, N, d_out, use_bg,
                                           o.noalloc, s);
    std::vector<ebsd_plan::GSeg> &rows = p->rowsets[S];
    int maxwords = 0;
    for (const ebsd_plan::GSeg &g : rows) maxwords = std::max(maxwords, g.nwords);
    if ((size_t)(maxwords + 1) * GatherCfg<IN, FLT, P>::WB + kGatherLdsExtra > (size_t)p->lds_max) return 1;
    const int64_t npos = (int64_t)rows[0].nbatch * 64;
    const int pbytes = FLT ? 8 : 4;
    uint32_t *part = o.part;
    int64_t piece;
    if (part) {
        if (o.part_bytes < (size_t)P * npos * pbytes)
            return fail(-2, "row-segment partial sums: buffer smaller than one pack's");
        piece = seg_piece(N, P, npos, pbytes, std::min<size_t>(o.part_bytes, (size_t)kPartBytes));
    } else {
        if (o.noalloc) return fail(-2, "row-segment partial sums: no buffer on the no-allocation path");
        piece = seg_piece(N, P, npos, pbytes, (size_t)kPartBytes);
        int rc;
        if ((rc = ensure_part(p, (size_t)std::min(piece, N) * npos * (pbytes / 4)))) return rc;
        part = p->d_part;
    }
    const size_t esz = sizeof(typename InT<IN>::T);
    int rc;
    for (int64_t q0 = 0; q0 < N; q0 += piece) {
        const int64_t n = std::min(piece, N - q0);
        const void *src = (const char *)d_pats + (size_t)q0 * p->nPx * esz;
        float *dst = d_out + (size_t)q0 * p->nCells;
        for (int k = 0; k < S; ++k) {
            const int mode = k == 0 ? 1 : (k == S - 1 ? 3 : 2);
            rc = launch_gather_t<IN, FLT, P>(p, rows[k], maxwords, mode, part, src, n, dst, use_bg, o.noalloc, s);
            if (rc) return rc < 0 ? rc : fail(-3, "row-segment gather launch failed");
        }
    }
    return 0;
}

// Row segments: the requested count, or the fewest (up to 8) at which words of `WB` bytes
// fit in LDS (measured, uint16 240x240 with a 360x180 Radon: 2.39 ms per 1024 patterns
// whole, 1.09 ms in three segments of four patterns per word).
static int segments_for(const ebsd_plan *p, int WB)
{
    if (p->rseg > 0) return std::min(p->rseg, p->H);
    for (int S = 1; S <= 8; ++S) {
        const size_t words = (size_t)((p->H + S - 1) / S) * p->spitch + 1;
        if (words * WB + kGatherLdsExtra <= (size_t)p->lds_max) return std::min(S, p->H);
    }
    return std::min(8, p->H);
}

// Integer sums: words of the configured pack size.
static int row_segments(const ebsd_plan *p, int dtype)
{
    const int P = dtype == EBSD_U8 ? p->gpack_u8 : p->gpack_u16;
    return segments_for(p, dtype == EBSD_U8 ? P : 2 * P);
}

// Float sums (float32 patterns, exact background): float32 words of the configured pack size.
static int row_segments_flt(const ebsd_plan *p) { return segments_for(p, 4 * p->gpack_f32); }

// The segments a gather of this dtype uses: S > 1 needs the plan's row segments compiled for
// S (compiled here, or on the no-allocation path already compiled -- else one whole-pattern
// pass with a smaller pack, when one fits).
static int gather_segments(ebsd_plan *p, int S, const GatherOpts &o)
{
    if (S <= 1) return 1;
    if ((int)p->rowsets[S].size() == S) return S;
    if (o.noalloc) return 1;
    const int rc = compile_rows(p, S);
    return rc ? rc : S;
}

static int launch_gather(ebsd_plan *p, const void *d_pats, int dtype_flags, int64_t N, float *d_out,
                         hipStream_t s, const GatherOpts &o = GatherOpts())
{
    int rc = 1;
    const int dtype = dtype_flags & 0xff;
    const bool use_bg = !(dtype_flags & EBSD_FLAG_NOBG) && p->bg_mode != 0;
    // float32 input (or the exact background mode) stages float32(x - bg) per pixel
    const bool flt = (dtype == EBSD_F32) || (use_bg && p->bg_mode == 2);
    if (flt) {
        const int S = gather_segments(p, row_segments_flt(p), o);
        if (S < 0) return S;
        for (int P = p->gpack_f32; rc == 1 && P >= 1; P /= 2) {
            if (dtype == EBSD_U8) {
                if (P >= 4) rc = launch_gather_rows<EBSD_U8, true, 4>(p, S, d_pats, N, d_out, use_bg, o, s);
                else if (P >= 2) rc = launch_gather_rows<EBSD_U8, true, 2>(p, S, d_pats, N, d_out, use_bg, o, s);
                else rc = launch_gather_rows<EBSD_U8, true, 1>(p, S, d_pats, N, d_out, use_bg, o, s);
            } else if (dtype == EBSD_U16) {
                if (P >= 4) rc = launch_gather_rows<EBSD_U16, true, 4>(p, S, d_pats, N, d_out, use_bg, o, s);
                else if (P >= 2) rc = launch_gather_rows<EBSD_U16, true, 2>(p, S, d_pats, N, d_out, use_bg, o, s);
                else rc = launch_gather_rows<EBSD_U16, true, 1>(p, S, d_pats, N, d_out, use_bg, o, s);
            } else {
                if (P >= 4) rc = launch_gather_rows<EBSD_F32, true, 4>(p, S, d_pats, N, d_out, use_bg, o, s);
                else if (P >= 2) rc = launch_gather_rows<EBSD_F32, true, 2>(p, S, d_pats, N, d_out, use_bg, o, s);
                else rc = launch_gather_rows<EBSD_F32, true, 1>(p, S, d_pats, N, d_out, use_bg, o, s);
            }
        }
        if (rc == 1) return fail(-2, "pattern too large for the Radon kernel's LDS staging");
        return rc;
    }
    const int S = gather_segments(p, row_segments(p, dtype), o);
    if (S < 0) return S;
    if (dtype == EBSD_U8) {
        for (int P = p->gpack_u8; rc == 1 && P >= 2; P /= 2) {
            if (P >= 8) rc = launch_gather_rows<EBSD_U8, false, 8>(p, S, d_pats, N, d_out, use_bg, o, s);
            else if (P >= 4) rc = launch_gather_rows<EBSD_U8, false, 4>(p, S, d_pats, N, d_out, use_bg, o, s);
            else rc = launch_gather_rows<EBSD_U8, false, 2>(p, S, d_pats, N, d_out, use_bg, o, s);
        }
    } else {
        for (int P = p->gpack_u16; rc == 1 && P >= 1; P /= 2) {
            if (P >= 8) rc = launch_gather_rows<EBSD_U16, false, 8>(p, S, d_pats, N, d_out, use_bg, o, s);
            else if (P >= 4) rc = launch_gather_rows<EBSD_U16, false, 4>(p, S, d_pats, N, d_out, use_bg, o, s);
            else if (P >= 2) rc = launch_gather_rows<EBSD_U16, false, 2>(p, S, d_pats, N, d_out, use_bg, o, s);
            else rc = launch_gather_rows<EBSD_U16, false, 1>(p, S, d_pats, N, d_out, use_bg, o, s);
        }
    }
    if (rc == 1) return fail(-2, "pattern too large for the Radon kernel's LDS staging");
    return rc;
}

template <int KR, int KT>
static int launch_peaks_t(ebsd_plan *p, const float *d_radon, int64_t N, const ebsd_bands *out, hipStream_t s)
{
    PeakArgs a;
    a.s.radon = d_radon;
    a.s.N = N;
    a.s.mask = p->d_mask;
    a.s.grid = p->d_grid;
    a.s.mlane = p->d_mlane;
    a.s.mrow = p->d_mrow;
    a.s.out = *out;
    a.g = p->g;
    a.L = peak_lds_layout(p->g);
    const size_t lds = (size_t)a.L.total * 4;
    auto kern = k_peaks<KR, KT, false>;
    HIPCHK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int occ = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, p->pthreads, lds));
    if (occ < 1) return fail(-3, "peak kernel cannot be resident");
    const int64_t grid = std::min<int64_t>(N, (int64_t)occ * p->ncu);
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(p->pthreads), lds, s, a);
    HIPCHK(hipGetLastError());
    return 0;
}

template <int KR, int KT>
static int launch_peaks_stream_t(ebsd_plan *p, const float *d_radon, int64_t N, const ebsd_bands *out,
                                 hipStream_t s)
{
    PeakArgs a{};
    a.s.radon = d_radon;
    a.s.N = N;
    a.s.mask = p->d_mask;
    a.s.grid = p->d_grid;
    a.s.mlane = p->d_mlane;
    a.s.mrow = p->d_mrow;
    a.s.out = *out;
    a.g = p->g;
    const int nw = stream_waves(p->g);
    const size_t lds = (size_t)stream_lds_layout(nw, p->g.nB).total * 4;
    auto kern = k_peaks_stream<KR, KT, false>;
    HIPCHK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3((unsigned)N), dim3(64 * nw), lds, s, a);
    HIPCHK(hipGetLastError());
    return 0;
}

static int launch_peaks(ebsd_plan *p, const float *d_radon, int64_t N, const ebsd_bands *out, hipStream_t s)
{
    const PeakGeo &g = p->g;
    if (N <= 0) return 0;
    if (use_stream(p)) {
        if (g.kr == 9) return launch_peaks_stream_t<9, 5>(p, d_radon, N, out, s);
        return launch_peaks_stream_t<5, 5>(p, d_radon, N, out, s);
    }
    // compile-time kernel sizes when the dilation half-windows equal them (peakPad = 2*ksz+1)
    if (g.kr == 9 && g.kt == 5 && g.hr == 9 && g.ht == 5) return launch_peaks_t<9, 5>(p, d_radon, N, out, s);
    if (g.kr == 5 && g.kt == 5 && g.hr == 5 && g.ht == 5) return launch_peaks_t<5, 5>(p, d_radon, N, out, s);
    return launch_peaks_t<0, 0>(p, d_radon, N, out, s);
}

static ebsd_bands offset_bands(const ebsd_bands &o, int64_t q0, int nB)
{
    ebsd_bands r = o;
    const int64_t e = q0 * nB;
    r.d_max += e; r.d_normmax += e; r.d_avemax += e; r.d_width += e;
    r.d_theta += e; r.d_rho += e; r.d_valid += e;
    r.d_maxloc += 2 * e; r.d_aveloc += 2 * e;
    r.d_stats += 3 * q0;
    return r;
}

static bool bands_ok(const ebsd_bands *o)
{
    return o && o->d_max && o->d_normmax && o->d_avemax && o->d_maxloc && o->d_aveloc &&
           o->d_width && o->d_theta && o->d_rho && o->d_valid && o->d_stats;
}

extern "C" int ebsd_find_bands(ebsd_plan *p, const void *d_pats, int dtype, int64_t N,
                               const ebsd_bands *out, void *d_work, size_t ws_bytes, void *stream)
{
    if (!p) return fail(-1, "ebsd_find_bands: null plan");
    if (N < 0) return fail(-1, "ebsd_find_bands: negative N");
    if (N == 0) return 0;
    if (!d_pats || !d_work) return fail(-1, "ebsd_find_bands: null device pointer");
    if (!bands_ok(out)) return fail(-1, "ebsd_find_bands: null output pointer");
    if ((dtype & 0xff) > EBSD_F32 || (dtype & ~(0xff | EBSD_FLAG_NOBG)))
        return fail(-1, "ebsd_find_bands: bad dtype");
    const size_t per = (size_t)p->nCells * sizeof(float);
    const int64_t chunk = (int64_t)(ws_bytes / per);
    if (chunk < 1) return fail(-1, "ebsd_find_bands: workspace smaller than one Radon");
    hipStream_t s = (hipStream_t)stream;
    const size_t esz = (dtype & 0xff) == EBSD_U8 ? 1 : ((dtype & 0xff) == EBSD_U16 ? 2 : 4);
    // passes of equal size (whole packs of 8 where the workspace allows), not full passes
    // and a short last one: a short pass leaves most of the GPU idle
    const int64_t npass = (N + chunk - 1) / chunk;
    int64_t piece = (N + npass - 1) / npass;
    piece = std::min(chunk, (piece + 7) / 8 * 8);
    for (int64_t q0 = 0; q0 < N; q0 += piece) {
        const int64_t n = std::min(piece, N - q0);
        const void *src = (const char *)d_pats + (size_t)q0 * p->nPx * esz;
        int rc = launch_gather(p, src, dtype, n, (float *)d_work, s);
        if (rc) return rc;
        ebsd_bands o = offset_bands(*out, q0, p->g.nB);
        rc = launch_peaks(p, (const float *)d_work, n, &o, s);
        if (rc) return rc;
    }
    return 0;
}

extern "C" int ebsd_find_bands_noalloc(ebsd_plan *p, const void *d_pats, int dtype, int64_t N,
                                       const ebsd_bands *out, void *stream)
{
    if (!p) return fail(-1, "ebsd_find_bands_noalloc: null plan");
    if (N < 0) return fail(-1, "ebsd_find_bands_noalloc: negative N");
    if (N == 0) return 0;
    if (!d_pats) return fail(-1, "ebsd_find_bands_noalloc: null device pointer");
    if (!bands_ok(out)) return fail(-1, "ebsd_find_bands_noalloc: null output pointer");
    if ((dtype & 0xff) > EBSD_F32 || (dtype & ~(0xff | EBSD_FLAG_NOBG)))
        return fail(-1, "ebsd_find_bands_noalloc: bad dtype");
    if (!p->scratch) return fail(-3, "ebsd_find_bands_noalloc: no scratch address for the plan's device");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != p->device)
        return fail(-1, "ebsd_find_bands_noalloc: the plan's device is not the current device");
    // the gather's row segments, if it runs in them with segmentations already compiled
    const int dt = dtype & 0xff;
    const bool use_bg = !(dtype & EBSD_FLAG_NOBG) && p->bg_mode != 0;
    const bool flt = dt == EBSD_F32 || (use_bg && p->bg_mode == 2);
    const int S = flt ? row_segments_flt(p) : row_segments(p, dt);
    const bool seg = S > 1 && (int)p->rowsets[S].size() == S;
    // patterns per chunk: their Radons, then (row segments) their partial sums, in the scratch
    const size_t per = (size_t)p->nCells * sizeof(float);
    const size_t npos = seg ? (size_t)p->rowsets[S][0].nbatch * 64 : 0;
    const size_t pbytes = seg ? npos * (flt ? 8 : 4) : 0;
    int64_t chunk = (int64_t)(kScratchBytes / (per + pbytes));
    if (seg) chunk = chunk / 8 * 8; // whole packs of up to 8 patterns share the partial sums
    if (chunk < 1)
        return fail(-2, "ebsd_find_bands_noalloc: one pattern's Radon (" + std::to_string(per + pbytes) +
                            " bytes) exceeds the no-allocation scratch of " + std::to_string(kScratchBytes) +
                            " bytes");
    const size_t rbytes = (((size_t)chunk * per + 255) / 256) * 256;
    GatherOpts o;
    o.noalloc = true;
    if (seg) {
        o.part = (uint32_t *)((char *)p->scratch + rbytes);
        o.part_bytes = kScratchBytes - rbytes;
    }
    float *radon = (float *)p->scratch;
    hipStream_t s = (hipStream_t)stream;
    const size_t esz = dt == EBSD_U8 ? 1 : (dt == EBSD_U16 ? 2 : 4);
    for (int64_t q0 = 0; q0 < N; q0 += chunk) {
        const int64_t n = std::min(chunk, N - q0);
        const void *src = (const char *)d_pats + (size_t)q0 * p->nPx * esz;
        int rc = launch_gather(p, src, dtype, n, radon, s, o);
        if (rc) return rc;
        ebsd_bands ob = offset_bands(*out, q0, p->g.nB);
        rc = launch_peaks(p, radon, n, &ob, s);
        if (rc) return rc;
    }
    return 0;
}

extern "C" int ebsd_radon(ebsd_plan *p, const void *d_pats, int dtype, int64_t N, float *d_radon, void *stream)
{
    if (!p) return fail(-1, "ebsd_radon: null plan");
    if (N <= 0) return N == 0 ? 0 : fail(-1, "ebsd_radon: negative N");
    if (!d_pats || !d_radon) return fail(-1, "ebsd_radon: null device pointer");
    if ((dtype & 0xff) > EBSD_F32 || (dtype & ~(0xff | EBSD_FLAG_NOBG)))
        return fail(-1, "ebsd_radon: bad dtype");
    return launch_gather(p, d_pats, dtype, N, d_radon, (hipStream_t)stream);
}

extern "C" int ebsd_peaks(ebsd_plan *p, const float *d_radon, int64_t N, const ebsd_bands *out, void *stream)
{
    if (!p) return fail(-1, "ebsd_peaks: null plan");
    if (N <= 0) return N == 0 ? 0 : fail(-1, "ebsd_peaks: negative N");
    if (!d_radon || !bands_ok(out)) return fail(-1, "ebsd_peaks: null device pointer");
    return launch_peaks(p, d_radon, N, out, (hipStream_t)stream);
}

extern "C" int ebsd_conv(ebsd_plan *p, const float *d_radon, int64_t N, float *d_conv, void *stream)
{
    if (!p) return fail(-1, "ebsd_conv: null plan");
    if (N <= 0) return N == 0 ? 0 : fail(-1, "ebsd_conv: negative N");
    if (!d_radon || !d_conv) return fail(-1, "ebsd_conv: null device pointer");
    const int64_t total = N * (int64_t)p->g.nRp * p->g.nTp;
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_conv_full, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       d_radon, N, p->g, d_conv);
    HIPCHK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------------------
// Adjoints (differentiation of the band positions).  The band's sub-pixel position is a
// function of its 3x3 convolved-Radon neighbourhood (band_detect.py:764-798); the
// convolution is linear in the padded Radon (band_detect.py:605-606), the padding an index
// map (:573-580) and the Radon linear in the pixels (radon_fast.py:288-310).  The backward
// pass therefore runs: band-fit Jacobian -> dC (3x3 per band) -> transposed convolution
// and padding adjoint -> dR -> backprojection -> d(pixels).  All three are gathers with
// float64 sums in a fixed order: the gradients are deterministic.
// ---------------------------------------------------------------------------------------
struct BandGradArgs {
    const float *radon;    // (N, nR, nT), recomputed by the caller
    int64_t N;
    const float *maxloc;   // (N, nB, 2) integer peak bins, padding removed (find_bands output)
    const float *aveloc;   // (N, nB, 2) sub-pixel positions (find_bands output)
    const int8_t *valid;   // (N, nB)
    const float *g_rho;    // (N, nB) dL/drho
    const float *g_theta;  // (N, nB) dL/dtheta
    const double *grid;    // theta_deg[nT], rho[nR]
    float *dconv;          // (N, nRp, nTp) out: dL/dC, zero except at the bands' neighbourhoods
    PeakGeo g;
};

// Convolved padded Radon at (y, x) with the arithmetic of the peak kernels (theta taps
// then rho taps, fmaf chains from zero), so that the neighbourhood equals the forward's.
__device__ __forceinline__ float conv_at(const PeakGeo &g, const float *R, int y, int x)
{
    const int ckr = g.kr / 2, ckt = g.kt / 2;
    float acc = 0.0f;
    for (int aa = 0; aa < g.kr; ++aa) {
        const int yy = y + ckr - aa;
        float t = 0.0f;
        for (int b = 0; b < g.kt; ++b) {
            const int xx = x + ckt - b;
            const float s = (yy >= 0 && yy < g.nRp && xx >= 0 && xx < g.nTp) ? R[pad_src(g, yy, xx)] : 0.0f;
            t = fmaf(g.v[b], s, t);
        }
        acc = fmaf(g.u[aa], t, acc);
    }
    return acc;
}

// Slope of numpy.interp(x, 0..n-1, fp) (0 where it clamps).
__device__ __forceinline__ double np_interp_slope(double x, const double *fp, int n)
{
    if (!(x >= 0.0) || x > (double)(n - 1) || n < 2) return 0.0;
    int j = (int)floor(x);
    j = j > n - 2 ? n - 2 : j;
    return fp[j + 1] - fp[j];
}

// dL/d(nn) of one band: the Jacobian of band_detect.py:764-798 (including which branch
// -- full Hessian, diagonal fallback, zero -- the forward took) applied to
// dL/d(aveloc).  nn: raw 3x3 neighbourhood (row-major, rows along rho).
__device__ void band_fit_grad(const float *nn, double gr_al, double gc_al, double *dnn)
{
    // forward-identical normalisation and branch decisions (float32 sum, as the peak kernels)
    float s = 0.0f;
    for (int k = 0; k < 9; ++k) s += nn[k];
    const bool sclamp = !(s > 1e-12f);
    const double S = sclamp ? 1e-12 : (double)s;
    float nf[9];
    for (int k = 0; k < 9; ++k) nf[k] = (float)((double)nn[k] / S);
    auto terms = [](const double *n, double &dx, double &dy, double &dxx, double &dyy, double &dxy) {
        dx = 0.5 * (n[5] - n[3]);
        dy = 0.5 * (n[7] - n[1]);
        dxx = (n[5] + n[3]) - 2.0 * n[4];
        dyy = (n[7] + n[1]) - 2.0 * n[4];
        dxy = 0.25 * (((n[8] - n[2]) - n[6]) + n[0]);
    };
    double nd[9];
    for (int k = 0; k < 9; ++k) nd[k] = (double)nf[k];
    double dx, dy, dxx, dyy, dxy;
    terms(nd, dx, dy, dxx, dyy, dxy);
    int branch = 0; // 0 full, 1 diagonal, 2 zero
    bool dclamp = false;
    {
        double det = dxx * dyy - dxy * dxy;
        const bool cl0 = !(fabs(det) > 1e-12);
        det = cl0 ? 1.0e-12 : det;
        const double dc = (dyy * dx - dxy * dy) / det, rc = (dxx * dy - dxy * dx) / det;
        dclamp = cl0;
        if (fabs(dc) > 0.875 || fabs(rc) > 0.875) {
            double d2 = dxx * dyy;
            const bool cl1 = !(fabs(d2) > 1e-12);
            d2 = cl1 ? 1.0e-12 : d2;
            dclamp = cl1;
            const double dc2 = (dyy * dx) / d2, rc2 = (dxx * dy) / d2;
            branch = (fabs(dc2) > 0.875 || fabs(rc2) > 0.875) ? 2 : 1;
        }
    }
    for (int k = 0; k < 9; ++k) dnn[k] = 0.0;
    if (branch == 2) return;
    // derivatives at the exactly normalised neighbourhood
    double n64[9];
    for (int k = 0; k < 9; ++k) n64[k] = (double)nn[k] / S;
    terms(n64, dx, dy, dxx, dyy, dxy);
    // q = (dx, dy, dxx, dyy, dxy); dL/dq from dL/drc = -gr_al, dL/ddc = -gc_al
    const double Lrc = -gr_al, Ldc = -gc_al;
    double Lq[5];
    if (branch == 0) {
        const double D = dclamp ? 1.0e-12 : dxx * dyy - dxy * dxy;
        const double Nc = dyy * dx - dxy * dy, Nr = dxx * dy - dxy * dx;
        const double iD = 1.0 / D, iD2 = iD * iD;
        const double dD[5] = {0.0, 0.0, dclamp ? 0.0 : dyy, dclamp ? 0.0 : dxx, dclamp ? 0.0 : -2.0 * dxy};
        const double dNc[5] = {dyy, -dxy, 0.0, dx, -dy};
        const double dNr[5] = {-dxy, dxx, dy, 0.0, -dx};
        for (int i = 0; i < 5; ++i)
            Lq[i] = Ldc * (dNc[i] * iD - Nc * dD[i] * iD2) + Lrc * (dNr[i] * iD - Nr * dD[i] * iD2);
    } else {
        const double D = dclamp ? 1.0e-12 : dxx * dyy;
        const double Nc = dyy * dx, Nr = dxx * dy;
        const double iD = 1.0 / D, iD2 = iD * iD;
        const double dD[5] = {0.0, 0.0, dclamp ? 0.0 : dyy, dclamp ? 0.0 : dxx, 0.0};
        const double dNc[5] = {dyy, 0.0, 0.0, dx, 0.0};
        const double dNr[5] = {0.0, dxx, dy, 0.0, 0.0};
        for (int i = 0; i < 5; ++i)
            Lq[i] = Ldc * (dNc[i] * iD - Nc * dD[i] * iD2) + Lrc * (dNr[i] * iD - Nr * dD[i] * iD2);
    }
    // q as linear forms of the normalised neighbourhood
    double gn[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    gn[5] += 0.5 * Lq[0]; gn[3] -= 0.5 * Lq[0];
    gn[7] += 0.5 * Lq[1]; gn[1] -= 0.5 * Lq[1];
    gn[5] += Lq[2]; gn[3] += Lq[2]; gn[4] -= 2.0 * Lq[2];
    gn[7] += Lq[3]; gn[1] += Lq[3]; gn[4] -= 2.0 * Lq[3];
    gn[8] += 0.25 * Lq[4]; gn[2] -= 0.25 * Lq[4]; gn[6] -= 0.25 * Lq[4]; gn[0] += 0.25 * Lq[4];
    // through the normalisation n / S(n)
    double dot = 0.0;
    for (int k = 0; k < 9; ++k) dot += gn[k] * n64[k];
    for (int k = 0; k < 9; ++k) dnn[k] = sclamp ? gn[k] / S : (gn[k] - dot) / S;
}

// One work-group per pattern: zero its dC, then one thread per band writes dL/dC at the
// band's 3x3 neighbourhood (atomically: plateau candidates may be adjacent).
__global__ __launch_bounds__(256) void k_band_grad(BandGradArgs a)
{
    const PeakGeo &g = a.g;
    const int64_t q = blockIdx.x;
    const int nC = g.nRp * g.nTp;
    float *dC = a.dconv + q * (int64_t)nC;
    for (int i = threadIdx.x; i < nC; i += blockDim.x) dC[i] = 0.0f;
    __syncthreads();
    const float *R = a.radon + q * (int64_t)(g.nR * g.nT);
    for (int b = threadIdx.x; b < g.nB; b += blockDim.x) {
        const int64_t o = q * g.nB + b;
        if (!a.valid[o]) continue;
        const int r = (int)a.maxloc[2 * o] + g.p0, c = (int)a.maxloc[2 * o + 1] + g.p1;
        float nn[9];
        for (int dr = 0; dr < 3; ++dr)
            for (int dc = 0; dc < 3; ++dc) nn[dr * 3 + dc] = conv_at(g, R, r - 1 + dr, c - 1 + dc);
        // the forward's sub-pixel location selects the interpolation segments
        const double *thg = a.grid, *rhg = a.grid + g.nT;
        const double al_r = (double)a.aveloc[2 * o], al_c = (double)a.aveloc[2 * o + 1];
        double dnn0[9];
        // theta = pi - interp(al_c)/RADEG, rho = -interp(al_r)  (band_detect.py:521-524)
        const double gc_al = (double)a.g_theta[o] * (-np_interp_slope(al_c, thg, g.nT) / (180.0 / 3.141592653589793));
        const double gr_al = (double)a.g_rho[o] * (-np_interp_slope(al_r, rhg, g.nR));
        band_fit_grad(nn, gr_al, gc_al, dnn0);
        for (int dr = 0; dr < 3; ++dr)
            for (int dc = 0; dc < 3; ++dc)
                atomicAdd(&dC[(r - 1 + dr) * g.nTp + (c - 1 + dc)], (float)dnn0[dr * 3 + dc]);
    }
}

// dR = padding adjoint of the transposed convolution of dC: for every Radon cell, the sum
// over the padded cells that copy it (interior, replicated rho rows, mirrored theta
// columns) of sum_{a,b} u[a] v[b] dC[y - kr/2 + a][x - kt/2 + b].  LDS: one work-group per
// pattern stages its dC; GLOBAL: dC read from memory (padded Radons too large for LDS).
template <bool LDS>
__global__ __launch_bounds__(1024) void k_conv_transpose(const float *dconv, int64_t N, PeakGeo g, const int *pre_off,
                                                         const int *pre, float *dradon)
{
    extern __shared__ __align__(16) float sdc[];
    const int nCells = g.nR * g.nT, nC = g.nRp * g.nTp;
    const int ckr = g.kr / 2, ckt = g.kt / 2;
    auto cell_value = [&](const float *dC, int cell) -> double {
        double acc = 0.0;
        for (int k = pre_off[cell]; k < pre_off[cell + 1]; ++k) {
            const int y = pre[k] / g.nTp, x = pre[k] - (pre[k] / g.nTp) * g.nTp;
            for (int aa = 0; aa < g.kr; ++aa) {
                const int yy = y - ckr + aa;
                if (yy < 0 || yy >= g.nRp) continue;
                double t = 0.0;
                for (int b = 0; b < g.kt; ++b) {
                    const int xx = x - ckt + b;
                    if (xx >= 0 && xx < g.nTp) t += (double)g.v[b] * (double)dC[yy * g.nTp + xx];
                }
                acc += (double)g.u[aa] * t;
            }
        }
        return acc;
    };
    if constexpr (LDS) {
        const int64_t q = blockIdx.x;
        const float *src = dconv + q * (int64_t)nC;
        for (int i = threadIdx.x; i < nC; i += blockDim.x) sdc[i] = src[i];
        __syncthreads();
        for (int cell = threadIdx.x; cell < nCells; cell += blockDim.x)
            dradon[q * (int64_t)nCells + cell] = (float)cell_value(sdc, cell);
    } else {
        const int64_t total = N * (int64_t)nCells;
        for (int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; id < total;
             id += (int64_t)gridDim.x * blockDim.x) {
            const int64_t q = id / nCells;
            const int cell = (int)(id - q * nCells);
            dradon[id] = (float)cell_value(dconv + q * (int64_t)nC, cell);
        }
    }
}

// Transposed convolution, separable and streamed: thread per (padded column x, band of
// CT_ROWS rows) walks down the band keeping the theta-tap sums E[yy][x] = sum_b v[b]
// dC[yy][x - ckt + b] of rows y - ckr .. y + ckr in registers, so each E is computed once
// instead of once per rho tap; dP[y][x] = sum_a u[a] E[y - ckr + a] goes to the work-group's
// float64 scratch.  The same products and the same order as cell_value of k_conv_transpose
// (rows and columns outside the padded array skipped), so dP is that kernel's value for a
// single preimage; then each Radon cell sums the dP of the padded cells that copy it, in
// `pre` order (cells with several copies round once per copy instead of once per tap).
constexpr int CT_ROWS = 16;
template <int KR, int KT>
__global__ __launch_bounds__(1024) void k_conv_transpose_sep(const float *dconv, int64_t N, PeakGeo g,
                                                             const int *pre_off, const int *pre, double *scratch,
                                                             float *dradon)
{
    constexpr int CKR = KR / 2, CKT = KT / 2;
    const int nCells = g.nR * g.nT, nTp = g.nTp, nRp = g.nRp, nC = nRp * nTp;
    double *dP = scratch + (size_t)blockIdx.x * nC;
    const int nbands = (nRp + CT_ROWS - 1) / CT_ROWS;
    for (int64_t q = blockIdx.x; q < N; q += gridDim.x) {
        const float *src = dconv + q * (int64_t)nC;
        for (int it = threadIdx.x; it < nTp * nbands; it += blockDim.x) {
            const int x = it % nTp, y0 = (it / nTp) * CT_ROWS, y1 = min(y0 + CT_ROWS, nRp);
            auto erow = [&](int yy) -> double {
                double t = 0.0;
#pragma unroll
                for (int b = 0; b < KT; ++b) {
                    const int xx = x - CKT + b;
                    if (xx >= 0 && xx < nTp) t += (double)g.v[b] * (double)src[yy * nTp + xx];
                }
                return t;
            };
            double E[KR]; // E[a] holds row y - CKR + a
#pragma unroll
            for (int a = 0; a < KR - 1; ++a) {
                const int yy = y0 - CKR + a;
                E[a] = (yy >= 0 && yy < nRp) ? erow(yy) : 0.0;
            }
            for (int y = y0; y < y1; ++y) {
                const int yn = y + CKR;
                E[KR - 1] = (yn < nRp) ? erow(yn) : 0.0;
                double acc = 0.0;
#pragma unroll
                for (int a = 0; a < KR; ++a) {
                    const int yy = y - CKR + a;
                    if (yy >= 0 && yy < nRp) acc += (double)g.u[a] * E[a];
                }
                dP[y * nTp + x] = acc;
#pragma unroll
                for (int a = 0; a < KR - 1; ++a) E[a] = E[a + 1];
            }
        }
        __syncthreads(); // the work-group's dP writes before its reads
        for (int cell = threadIdx.x; cell < nCells; cell += blockDim.x) {
            double acc = 0.0;
            for (int k = pre_off[cell]; k < pre_off[cell + 1]; ++k) acc += dP[pre[k]];
            dradon[q * (int64_t)nCells + cell] = (float)acc;
        }
        __syncthreads(); // dP reads before the next pattern's writes
    }
}

// Radon adjoint: d(pixel) = sum over the lines through the pixel of dR[line] / count[line]
// (radon_fast.py:288-310 divides each line sum by its pixel count).  LDS: one work-group
// per pattern stages dR[line] / count[line] in float64 once (and a zero at cell nCells);
// each thread takes a pixel and sums its lines from the pair-major table (coalesced u16
// cell pairs) into two float64 accumulators, even and odd list positions, added at the end:
// a fixed order, deterministic.  GLOBAL: one thread per pixel reads the lines from memory
// (Radons too large for LDS).
template <bool LDS>
__global__ __launch_bounds__(1024) void k_backproject(const float *dradon, int64_t N, int nPx, int nCells,
                                                      const int *px_off, const int *px_cell, const double *icnt,
                                                      const uint32_t *tab, int m2, float *dpats)
{
    if constexpr (LDS) {
        extern __shared__ __align__(16) double sdr[];
        const int64_t q = blockIdx.x;
        const float *dR = dradon + q * (int64_t)nCells;
        for (int c = threadIdx.x; c < nCells; c += blockDim.x) sdr[c] = (double)dR[c] * icnt[c];
        if (threadIdx.x == 0) sdr[nCells] = 0.0;
        __syncthreads();
        for (int px = threadIdx.x; px < nPx; px += blockDim.x) {
            const uint32_t *t = tab + px;
            double a0 = 0.0, a1 = 0.0;
#pragma unroll 4
            for (int k2 = 0; k2 < m2; ++k2) {
                const uint32_t v = t[(size_t)k2 * nPx];
                a0 += sdr[v & 0xFFFFu];
                a1 += sdr[v >> 16];
            }
            dpats[q * (int64_t)nPx + px] = (float)(a0 + a1);
        }
    } else {
        const int64_t total = N * (int64_t)nPx;
        for (int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; id < total;
             id += (int64_t)gridDim.x * blockDim.x) {
            const int64_t q = id / nPx;
            const int px = (int)(id - q * nPx);
            const float *dR = dradon + q * (int64_t)nCells;
            double acc = 0.0;
            for (int k = px_off[px]; k < px_off[px + 1]; ++k) {
                const int cell = px_cell[k];
                acc = fma((double)dR[cell], icnt[cell], acc);
            }
            dpats[id] = (float)acc;
        }
    }
}

// Detector-frame translation t = a * PC + b of the gnomonic geometry (radon_fast.py:361-370).
__device__ __forceinline__ bool vendor_t(int vendor, int H, int W, double *a, double *b)
{
    const double h = H, w = W, mn = H < W ? H : W;
    b[0] = b[1] = b[2] = 0.0;
    if (vendor == 0) { a[0] = w; a[1] = h; a[2] = -mn; return true; }          // EDAX
    if (vendor == 1) { a[0] = w; a[1] = w; a[2] = -w; return true; }           // OXFORD
    if (vendor == 2) { a[0] = w; a[1] = -h; a[2] = -h; b[1] = h; return true; } // BRUKER / KIKUCHIPY
    return false;
}

// Predicted band of a detector-frame plane normal n at pattern centre PC: the trace of the
// plane through t with normal n on the detector, in the (rho, theta) of radon2pole
// (radon_fast.py:379-396): cos(th) (x - W/2) + sin(th) (y - H/2) = rho, theta in (0, pi].
// d rho / d PC is returned; theta does not depend on PC.
__global__ void k_gnomonic_bands(int64_t N, int M, const double *normals, const double *pc, int pc_per_pattern,
                                 int H, int W, int vendor, double *rho, double *theta, double *drho_dpc)
{
    const int64_t total = N * (int64_t)M;
    for (int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; id < total;
         id += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = id / M;
        const double *n = normals + 3 * id;
        const double *P = pc + (pc_per_pattern ? 3 * q : 0);
        double va[3], vb[3];
        vendor_t(vendor, H, W, va, vb);
        const double t0 = va[0] * P[0] + vb[0], t1 = va[1] * P[1] + vb[1], t2 = va[2] * P[2] + vb[2];
        const double m = sqrt(n[0] * n[0] + n[1] * n[1]);
        double sg = n[1] > 0.0 ? 1.0 : (n[1] < 0.0 ? -1.0 : (n[0] < 0.0 ? 1.0 : -1.0));
        if (!(m > 0.0)) {
            rho[id] = NAN;
            theta[id] = NAN;
            drho_dpc[3 * id] = drho_dpc[3 * id + 1] = drho_dpc[3 * id + 2] = 0.0;
            continue;
        }
        const double d = n[0] * t0 + n[1] * t1 + n[2] * t2;
        rho[id] = sg * (d - n[0] * 0.5 * W - n[1] * 0.5 * H) / m;
        theta[id] = atan2(sg * n[1], sg * n[0]);
        drho_dpc[3 * id + 0] = sg * n[0] * va[0] / m;
        drho_dpc[3 * id + 1] = sg * n[1] * va[1] / m;
        drho_dpc[3 * id + 2] = sg * n[2] * va[2] / m;
    }
}

// radon2pole (radon_fast.py:314-397): band (theta, rho) -> unit plane normal in the detector
// frame, float64 arithmetic, float32 result; dn/dPC (3x3 per band) when d_dn_dpc is set.
__global__ void k_radon2pole(int64_t N, int nB, const float *theta, const float *rho, const double *pc,
                             int pc_per_pattern, int H, int W, int vendor, float *normals, double *dn_dpc)
{
    const int64_t total = N * (int64_t)nB;
    for (int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; id < total;
         id += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = id / nB;
        const double *P = pc + (pc_per_pattern ? 3 * q : 0);
        double va[3], vb[3];
        vendor_t(vendor, H, W, va, vb);
        const double t[3] = {va[0] * P[0] + vb[0], va[1] * P[1] + vb[1], va[2] * P[2] + vb[2]};
        const double th = theta[id], rh = rho[id];
        const double st = sin(th), ct = cos(th);
        const double r[3] = {-st, ct, 0.0};
        const double p[3] = {rh * ct + 0.5 * W, rh * st + 0.5 * H, 0.0};
        const double d[3] = {p[0] - t[0], p[1] - t[1], p[2] - t[2]};
        const double n[3] = {r[1] * d[2] - r[2] * d[1], r[2] * d[0] - r[0] * d[2], r[0] * d[1] - r[1] * d[0]};
        double nrm = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        const bool clip = !(nrm > 1e-12);
        nrm = clip ? 1e-12 : nrm;
        for (int k = 0; k < 3; ++k) normals[3 * id + k] = (float)(n[k] / nrm);
        if (dn_dpc) {
            // n = r x (p - t): dn/dt_j = -(r x e_j); normalised: (I - u u^T) dn / |n|
            const double u[3] = {n[0] / nrm, n[1] / nrm, n[2] / nrm};
            for (int j = 0; j < 3; ++j) {
                double e[3] = {0.0, 0.0, 0.0};
                e[j] = 1.0;
                const double dn[3] = {-(r[1] * e[2] - r[2] * e[1]), -(r[2] * e[0] - r[0] * e[2]),
                                      -(r[0] * e[1] - r[1] * e[0])};
                const double ud = clip ? 0.0 : u[0] * dn[0] + u[1] * dn[1] + u[2] * dn[2];
                for (int k = 0; k < 3; ++k) dn_dpc[9 * id + 3 * k + j] = (dn[k] - u[k] * ud) / nrm * va[j];
            }
        }
    }
}

static int ensure_adjoint(ebsd_plan *p)
{
    if (p->d_px_off) return 0;
    const int nPx = p->nPx, nCells = p->nCells;
    // pixel -> lines through it, by ascending line index
    std::vector<int> px_off(nPx + 1, 0), px_cell(p->h_ent.size());
    std::vector<double> icnt(nCells);
    for (size_t k = 0; k < p->h_ent.size(); ++k) ++px_off[p->h_ent[k] + 1];
    for (int i = 0; i < nPx; ++i) px_off[i + 1] += px_off[i];
    std::vector<int> fill(px_off.begin(), px_off.end() - 1);
    for (int cell = 0; cell < nCells; ++cell) {
        icnt[cell] = p->h_cnt[cell] > 0 ? 1.0 / (double)p->h_cnt[cell] : 0.0;
        for (int64_t k = p->h_off[cell]; k < p->h_off[cell + 1]; ++k) px_cell[fill[p->h_ent[k]]++] = cell;
    }
    // Radon cell -> padded cells that copy it (the padding index map inverted)
    const PeakGeo &g = p->g;
    const int nC = g.nRp * g.nTp;
    std::vector<int> pre_off(nCells + 1, 0), pre(nC), src(nC);
    for (int y = 0; y < g.nRp; ++y)
        for (int x = 0; x < g.nTp; ++x) {
            int i = y - g.p0;
            i = i < 0 ? 0 : (i > g.nR - 1 ? g.nR - 1 : i);
            int j;
            if (x < g.p1) { j = g.nT - g.p1 + x; i = g.nR - 1 - i; }
            else if (x >= g.p1 + g.nT) { j = x - g.p1 - g.nT; i = g.nR - 1 - i; }
            else j = x - g.p1;
            src[y * g.nTp + x] = i * g.nT + j;
            ++pre_off[i * g.nT + j + 1];
        }
    for (int i = 0; i < nCells; ++i) pre_off[i + 1] += pre_off[i];
    std::vector<int> pf(pre_off.begin(), pre_off.end() - 1);
    for (int k = 0; k < nC; ++k) pre[pf[src[k]]++] = k;
    // the lines of pixel px as u16 cell pairs, pair-major so that a wavefront of consecutive
    // pixels reads each pair step coalesced; short lists padded with cell nCells (a zero)
    std::vector<uint32_t> tab;
    int m2 = 0;
    if (nCells < 65535) {
        int m = 0;
        for (int i = 0; i < nPx; ++i) m = std::max(m, px_off[i + 1] - px_off[i]);
        m2 = (m + 1) / 2;
        tab.assign((size_t)m2 * nPx, (uint32_t)nCells | ((uint32_t)nCells << 16));
        for (int i = 0; i < nPx; ++i)
            for (int k = px_off[i]; k < px_off[i + 1]; ++k) {
                const int j = k - px_off[i];
                uint32_t &w = tab[(size_t)(j / 2) * nPx + i];
                w = (j & 1) ? ((w & 0xFFFFu) | ((uint32_t)px_cell[k] << 16)) : ((w & 0xFFFF0000u) | (uint32_t)px_cell[k]);
            }
    }
    int rc;
    if (m2 > 0 && (rc = dev_upload(&p->d_bp_tab, tab.data(), tab.size()))) return rc;
    p->bp_m2 = m2;
    if ((rc = dev_upload(&p->d_px_off, px_off.data(), px_off.size()))) return rc;
    if ((rc = dev_upload(&p->d_px_cell, px_cell.data(), px_cell.size()))) return rc;
    if ((rc = dev_upload(&p->d_icnt, icnt.data(), icnt.size()))) return rc;
    if ((rc = dev_upload(&p->d_pre_off, pre_off.data(), pre_off.size()))) return rc;
    if ((rc = dev_upload(&p->d_pre, pre.data(), pre.size()))) return rc;
    return 0;
}

static unsigned grid_for(int64_t total)
{
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536));
}

extern "C" int ebsd_band_grad(ebsd_plan *p, const float *d_radon, int64_t N, const float *d_maxloc,
                              const float *d_aveloc, const int8_t *d_valid, const float *d_grad_rho,
                              const float *d_grad_theta, float *d_grad_conv, void *stream)
{
    if (!p) return fail(-1, "ebsd_band_grad: null plan");
    if (N <= 0) return N == 0 ? 0 : fail(-1, "ebsd_band_grad: negative N");
    if (!d_radon || !d_maxloc || !d_aveloc || !d_valid || !d_grad_rho || !d_grad_theta || !d_grad_conv)
        return fail(-1, "ebsd_band_grad: null device pointer");
    BandGradArgs a;
    a.radon = d_radon;
    a.N = N;
    a.maxloc = d_maxloc;
    a.aveloc = d_aveloc;
    a.valid = d_valid;
    a.g_rho = d_grad_rho;
    a.g_theta = d_grad_theta;
    a.grid = p->d_grid;
    a.dconv = d_grad_conv;
    a.g = p->g;
    hipLaunchKernelGGL(k_band_grad, dim3((unsigned)N), dim3(256), 0, (hipStream_t)stream, a);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int ebsd_conv_transpose(ebsd_plan *p, const float *d_grad_conv, int64_t N, float *d_grad_radon,
                                   void *stream)
{
    if (!p) return fail(-1, "ebsd_conv_transpose: null plan");
    if (N <= 0) return N == 0 ? 0 : fail(-1, "ebsd_conv_transpose: negative N");
    if (!d_grad_conv || !d_grad_radon) return fail(-1, "ebsd_conv_transpose: null device pointer");
    int rc;
    if ((rc = ensure_adjoint(p))) return rc;
    const PeakGeo &g = p->g;
    if ((g.kr == 9 || g.kr == 5) && g.kt == 5) {
        const int64_t grid = std::min<int64_t>(N, 2 * (int64_t)p->ncu);
        const size_t need = (size_t)grid * g.nRp * g.nTp;
        if (p->tscr_cap < need) {
            if (p->d_tscr) {
                HIPCHK(hipDeviceSynchronize()); // earlier launches may still use the old buffer
                (void)hipFree(p->d_tscr);
            }
            p->d_tscr = nullptr;
            p->tscr_cap = 0;
            if (hipMalloc((void **)&p->d_tscr, need * sizeof(double)) != hipSuccess) {
                p->d_tscr = nullptr;
                return fail(-3, "ebsd_conv_transpose: device allocation failed");
            }
            p->tscr_cap = need;
        }
        auto kern = g.kr == 9 ? k_conv_transpose_sep<9, 5> : k_conv_transpose_sep<5, 5>;
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(1024), 0, (hipStream_t)stream, d_grad_conv, N, g,
                           p->d_pre_off, p->d_pre, p->d_tscr, d_grad_radon);
        HIPCHK(hipGetLastError());
        return 0;
    }
    const size_t lds = (size_t)p->g.nRp * p->g.nTp * 4;
    if ((int)lds <= p->lds_max) {
        HIPCHK(hipFuncSetAttribute((const void *)k_conv_transpose<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
        hipLaunchKernelGGL(k_conv_transpose<true>, dim3((unsigned)N), dim3(1024), lds, (hipStream_t)stream,
                           d_grad_conv, N, p->g, p->d_pre_off, p->d_pre, d_grad_radon);
    } else {
        hipLaunchKernelGGL(k_conv_transpose<false>, dim3(grid_for(N * (int64_t)p->nCells)), dim3(256), 0,
                           (hipStream_t)stream, d_grad_conv, N, p->g, p->d_pre_off, p->d_pre, d_grad_radon);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int ebsd_backproject(ebsd_plan *p, const float *d_grad_radon, int64_t N, float *d_grad_pats,
                                void *stream)
{
    if (!p) return fail(-1, "ebsd_backproject: null plan");
    if (N <= 0) return N == 0 ? 0 : fail(-1, "ebsd_backproject: negative N");
    if (!d_grad_radon || !d_grad_pats) return fail(-1, "ebsd_backproject: null device pointer");
    int rc;
    if ((rc = ensure_adjoint(p))) return rc;
    const size_t lds = (size_t)(p->nCells + 1) * 8;
    if (p->bp_m2 > 0 && (int)lds <= p->lds_max) {
        HIPCHK(hipFuncSetAttribute((const void *)k_backproject<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
        hipLaunchKernelGGL(k_backproject<true>, dim3((unsigned)N), dim3(1024), lds, (hipStream_t)stream, d_grad_radon,
                           N, p->nPx, p->nCells, p->d_px_off, p->d_px_cell, p->d_icnt, p->d_bp_tab, p->bp_m2,
                           d_grad_pats);
    } else {
        hipLaunchKernelGGL(k_backproject<false>, dim3(grid_for(N * (int64_t)p->nPx)), dim3(256), 0,
                           (hipStream_t)stream, d_grad_radon, N, p->nPx, p->nCells, p->d_px_off, p->d_px_cell,
                           p->d_icnt, (const uint32_t *)nullptr, 0, d_grad_pats);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int ebsd_gnomonic_bands(int64_t N, int M, const double *d_normals, const double *d_pc,
                                   int pc_per_pattern, int H, int W, int vendor, double *d_rho, double *d_theta,
                                   double *d_drho_dpc, void *stream)
{
    if (N <= 0 || M <= 0) return (N == 0 || M == 0) ? 0 : fail(-1, "ebsd_gnomonic_bands: negative size");
    if (!d_normals || !d_pc || !d_rho || !d_theta || !d_drho_dpc) return fail(-1, "ebsd_gnomonic_bands: null pointer");
    if (vendor < 0 || vendor > 2) return fail(-1, "ebsd_gnomonic_bands: vendor is 0 (EDAX), 1 (OXFORD) or 2 (BRUKER)");
    hipLaunchKernelGGL(k_gnomonic_bands, dim3(grid_for(N * (int64_t)M)), dim3(256), 0, (hipStream_t)stream, N, M,
                       d_normals, d_pc, pc_per_pattern, H, W, vendor, d_rho, d_theta, d_drho_dpc);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int ebsd_radon2pole(int64_t N, int nB, const float *d_theta, const float *d_rho, const double *d_pc,
                               int pc_per_pattern, int H, int W, int vendor, float *d_normals, double *d_dn_dpc,
                               void *stream)
{
    if (N <= 0 || nB <= 0) return (N == 0 || nB == 0) ? 0 : fail(-1, "ebsd_radon2pole: negative size");
    if (!d_theta || !d_rho || !d_pc || !d_normals) return fail(-1, "ebsd_radon2pole: null pointer");
    if (vendor < 0 || vendor > 2) return fail(-1, "ebsd_radon2pole: vendor is 0 (EDAX), 1 (OXFORD) or 2 (BRUKER)");
    hipLaunchKernelGGL(k_radon2pole, dim3(grid_for(N * (int64_t)nB)), dim3(256), 0, (hipStream_t)stream, N, nB,
                       d_theta, d_rho, d_pc, pc_per_pattern, H, W, vendor, d_normals, d_dn_dpc);
    HIPCHK(hipGetLastError());
    return 0;
}

// Band records in the byte layout of BandDetect.dataType (band_detect.py:110-115, numpy's
// packed layout: fields in declaration order, no padding): id @0, max @4, normmax @8,
// maxloc[2] @12, avemax @20, aveloc[2] @24, pqmax @32, width @36, theta @40, rho @44,
// valid @48, band_match_index[nPhases] @49, bandfit[nPhases] @49+8*nPhases; record size
// 49 + 16*nPhases.  One thread per output byte, grid-stride (the stores of a wavefront
// cover 64 consecutive bytes).  Fields that find_bands does not compute are zero, as in the
// reference's result.
__global__ void k_pack_band_records(ebsd_bands b, int64_t nbytes, int R, uint8_t *dst)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nbytes;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / R; // record = pattern * nB + band
        const int o = (int)(i - r * R);
        uint8_t v = 0;
        if (o == 48) {
            v = (uint8_t)b.d_valid[r];
        } else if (o >= 4 && o < 48 && (o < 32 || o >= 36)) {
            // the float field word: fi = 0 max, 1 normmax, 2-3 maxloc, 4 avemax, 5-6 aveloc,
            // 8 width, 9 theta, 10 rho -- chosen by selects, without branches (a switch in
            // their place gave rho's bytes theta's values on the GPU)
            const int fi = (o - 4) >> 2;
            const float *f = fi == 0 ? b.d_max : b.d_normmax;
            f = (fi == 2 || fi == 3) ? b.d_maxloc : f;
            f = fi == 4 ? b.d_avemax : f;
            f = (fi == 5 || fi == 6) ? b.d_aveloc : f;
            f = fi == 8 ? b.d_width : f;
            f = fi == 9 ? b.d_theta : f;
            f = fi == 10 ? b.d_rho : f;
            const int64_t k = (fi == 2 || fi == 3) ? 2 * r + (fi - 2) : ((fi == 5 || fi == 6) ? 2 * r + (fi - 5) : r);
            v = (uint8_t)(__float_as_uint(f[k]) >> (8 * (o & 3)));
        }
        dst[i] = v;
    }
}

extern "C" int ebsd_pack_band_records(const ebsd_bands *bands, int64_t N, int nB, int nPhases, void *d_records,
                                      void *stream)
{
    if (N < 0 || nB < 0 || nPhases < 0) return fail(-1, "ebsd_pack_band_records: negative size");
    if (N == 0 || nB == 0) return 0;
    if (!bands || !d_records || !bands->d_max || !bands->d_normmax || !bands->d_avemax || !bands->d_maxloc ||
        !bands->d_aveloc || !bands->d_width || !bands->d_theta || !bands->d_rho || !bands->d_valid)
        return fail(-1, "ebsd_pack_band_records: null pointer");
    const int R = 49 + 16 * nPhases;
    const int64_t nbytes = N * (int64_t)nB * R;
    hipLaunchKernelGGL(k_pack_band_records, dim3(grid_for(nbytes)), dim3(256), 0, (hipStream_t)stream, *bands,
                       nbytes, R, (uint8_t *)d_records);
    HIPCHK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------------------
// batches of mixed pattern sizes: one persistent gather launch and one persistent peak
// launch over heterogeneous work lists, optionally captured once into a hipGraph
// ---------------------------------------------------------------------------------------
struct VariantInfo {
    int id, IN, P, WB;
    bool FLT;
};

static const VariantInfo kVariants[] = {
#define EBSD_GV_INFO(ID, IN, FLT, P) {ID, IN, P, GatherCfg<IN, FLT, P>::WB, FLT},
    EBSD_GATHER_VARIANTS(EBSD_GV_INFO)
#undef EBSD_GV_INFO
};

// The variant the single-set launcher would pick: the largest pack (within the plan's
// tuning) whose pixel-major LDS words fit.
static int choose_variant(const ebsd_plan *p, int dtype, bool flt, VariantInfo *out)
{
    int pmax = flt ? p->gpack_f32 : (dtype == EBSD_U8 ? p->gpack_u8 : p->gpack_u16);
    int best = -1;
    for (const VariantInfo &v : kVariants) {
        if (v.IN != dtype || v.FLT != flt || v.P > pmax) continue;
        if ((size_t)(p->full.nwords + 1) * v.WB + kGatherLdsExtra > (size_t)p->lds_max) continue;
        if (best < 0 || v.P > kVariants[best].P) best = (int)(&v - kVariants);
    }
    if (best < 0) return fail(-2, "pattern too large for the Radon kernel's LDS staging");
    *out = kVariants[best];
    return 0;
}

// Sets whose staged patterns need similar LDS share one persistent gather launch, so small
// patterns keep the occupancy their own launches would get.
struct GatherClass {
    int threads = 1024;
    size_t lds = 0;
    int64_t item0 = 0, nitems = 0; // slice of the batch's gather work list
    int grid = 0;
};

// Sets whose Radon geometry is identical share one persistent peak launch; the geometry
// travels in that launch's kernel arguments.
struct PeakClass {
    PeakGeo g{};
    PeakLds L{};
    size_t lds = 0;
    int kr = 0, kt = 0; // 0 = run-time kernel sizes in the peak kernel
    int threads = 512;
    bool stream = false; // k_peaks_stream (one work-group per item) instead of k_peaks
    int64_t item0 = 0, nitems = 0; // slice of the batch's peak work list
    int grid = 0;
};

// patterns from which a set of a batch runs the per-size launchers (a gather round of the
// whole GPU at 4 patterns per CU: the set then fills the GPU by itself; measured on mixed
// batches of 1,024 patterns a set, the per-size launches beat the merged ones by 7%)
constexpr int64_t kSingleSetMin = 1024;

struct ebsd_batch {
    int nsets = 0, device = 0, ncu = 0;
    GatherArgs *d_gsets = nullptr;
    int *d_variant = nullptr;
    int2 *d_gitems = nullptr;
    int64_t ngitems = 0;
    std::vector<GatherClass> classes;
    PeakPtrs *d_psets = nullptr;
    int2 *d_pitems = nullptr;
    int64_t npitems = 0;
    std::vector<PeakClass> pclasses;
    // sets that fill the GPU by themselves: the per-size gather and peak launchers, one
    // set after the other, inside the batch
    struct Single {
        ebsd_plan *plan;
        const void *pats;
        int dtype;
        int64_t N;
        float *radon;
        ebsd_bands out;
    };
    std::vector<Single> singles;
    uint32_t *part = nullptr;        // the singles' row-segment partial sums (in the workspace;
    size_t part_bytes = 0;           // the singles run one after another, so they share it)
    std::vector<ebsd_plan *> pinned; // distinct plans of the batch, pinned while it lives
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static void batch_free(ebsd_batch *b)
{
    if (!b) return;
    for (ebsd_plan *p : b->pinned) --p->pins;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    (void)hipFree(b->d_gsets);
    (void)hipFree(b->d_variant);
    (void)hipFree(b->d_gitems);
    (void)hipFree(b->d_psets);
    (void)hipFree(b->d_pitems);
    delete b;
}

// Partial-sum bytes a large set may need when its per-size gather runs in row segments
// (float64 sums, the larger kind; at most one launch sequence's worth).
static size_t single_part_bytes(const ebsd_plan *p, int64_t N)
{
    if (N < kSingleSetMin) return 0;
    const size_t npos = (size_t)((p->nCells + 63) / 64) * 64;
    return std::min<size_t>((size_t)N * npos * 8, (size_t)kPartBytes);
}

extern "C" size_t ebsd_batch_workspace_bytes(int nsets, ebsd_plan *const *plans, const int64_t *N)
{
    if (nsets <= 0 || !plans || !N) return 0;
    size_t total = 0, part = 0;
    for (int i = 0; i < nsets; ++i) {
        if (!plans[i] || N[i] < 0) return 0;
        total += (((size_t)N[i] * plans[i]->nCells * sizeof(float) + 255) / 256) * 256;
        part = std::max(part, single_part_bytes(plans[i], N[i]));
    }
    return total + part;
}

extern "C" int ebsd_batch_create(int nsets, ebsd_plan *const *plans, const void *const *d_pats,
                                 const int *dtypes, const int64_t *N, const ebsd_bands *outs, void *d_work,
                                 size_t ws_bytes, ebsd_batch **out)
{
    if (!out || nsets <= 0 || !plans || !d_pats || !dtypes || !N || !outs || !d_work)
        return fail(-1, "ebsd_batch_create: null argument");
    *out = nullptr;
    if (ws_bytes < ebsd_batch_workspace_bytes(nsets, plans, N))
        return fail(-1, "ebsd_batch_create: workspace smaller than ebsd_batch_workspace_bytes");
    ebsd_batch *b = new ebsd_batch();
    b->nsets = nsets;
    b->device = plans[0]->device;
    b->ncu = plans[0]->ncu;
    std::vector<GatherArgs> gs(nsets);
    std::vector<int> var(nsets);
    std::vector<PeakPtrs> ps(nsets);
    char *work = (char *)d_work;
    int rc;
    // LDS occupancy class of each set: work-groups per CU its staged pack allows (1..4);
    // classes of 2+ work-groups per CU run 512-thread work-groups, class 1 runs 1024
    std::vector<int> cls(nsets);
    std::vector<char> fltv(nsets, 0);
    for (int i = 0; i < nsets; ++i) {
        ebsd_plan *p = plans[i];
        if (!p || N[i] < 0) { batch_free(b); return fail(-1, "ebsd_batch_create: bad set"); }
        if (p->device != b->device) { batch_free(b); return fail(-1, "ebsd_batch_create: plans on different devices"); }
        if (!bands_ok(&outs[i]) || (N[i] > 0 && !d_pats[i])) { batch_free(b); return fail(-1, "ebsd_batch_create: null set pointer"); }
        const int dtype = dtypes[i] & 0xff;
        if (dtype > EBSD_F32) { batch_free(b); return fail(-1, "ebsd_batch_create: bad dtype"); }
        const bool use_bg = !(dtypes[i] & EBSD_FLAG_NOBG) && p->bg_mode != 0;
        const bool flt = dtype == EBSD_F32 || (use_bg && p->bg_mode == 2);
        fltv[i] = flt;
        VariantInfo v;
        if ((rc = choose_variant(p, dtype, flt, &v))) { batch_free(b); return rc; }
        var[i] = v.id;
        const size_t lds = (size_t)(p->full.nwords + 1) * v.WB + kGatherLdsExtra;
        const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(4, (size_t)p->lds_max / std::max<size_t>(lds, 1)));
        cls[i] = per_cu;
        float *radon = (float *)work;
        work += (((size_t)N[i] * p->nCells * sizeof(float) + 255) / 256) * 256;
        GatherArgs &a = gs[i];
        a.pats = d_pats[i];
        a.N = N[i];
        a.nPx = p->nPx;
        a.W = p->W;
        a.pitch = p->spitch;
        a.px0 = 0;
        a.rows = p->H;
        a.segmode = 0;
        a.nbatch = p->full.nbatch;
        a.bhdr = p->full.d_bhdr;
        a.lane = p->full.d_lane;
        a.lrc = p->full.d_lrc;
        a.linebg = (use_bg && !flt && p->bg_mode == 1) ? p->d_linebg : nullptr;
        a.bg = (use_bg && flt) ? p->d_bg : nullptr;
        a.out = radon;
        a.nCells = p->nCells;
        a.npacks = (N[i] + v.P - 1) / v.P;
        PeakPtrs &S = ps[i];
        S.radon = radon;
        S.N = N[i];
        S.mask = p->d_mask;
        S.grid = p->d_grid;
        S.mlane = p->d_mlane;
        S.mrow = p->d_mrow;
        S.out = outs[i];
    }
    // a set that fills the GPU by itself (kSingleSetMin patterns or more) runs the per-size
    // launchers, gather (with the plan's tuned shape and row segments) then peaks, inside
    // the batch, as a plain call would: the peak kernel reads the Radon its gather has just
    // written, while the cache still holds much of it.  The partial sums of its row
    // segments go to the batch's own buffer (after the Radons in the workspace; the singles
    // run one after another and share it).  The warm-up below compiles every row
    // segmentation they use, and a plan keeps its segmentations (one per segment count)
    // until it is destroyed; the merged launches are for the others
    b->part = (uint32_t *)work;
    b->part_bytes = (size_t)((char *)d_work + ws_bytes - work);
    std::vector<char> single(nsets, 0);
    for (int i = 0; i < nsets; ++i) {
        if (N[i] < kSingleSetMin) continue;
        b->singles.push_back(ebsd_batch::Single{plans[i], d_pats[i], dtypes[i], N[i], gs[i].out, outs[i]});
        cls[i] = 0;
        single[i] = 1;
    }
    // gather items, class by class (4 work-groups per CU first); inside a class the largest
    // patterns first, so the static round-robin over persistent work-groups ends cheap
    std::vector<int2> gitems;
    std::vector<int> order(nsets);
    for (int i = 0; i < nsets; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return plans[x]->nPx > plans[y]->nPx; });
    // a class of c work-groups per CU runs 512 threads (c >= 2) or 1,024; each set's packs are
    // split into slices (work-groups sharing a pack's lines) so that the class has about two
    // resident rounds of work-groups
    for (int c = 4; c >= 1; --c) {
        GatherClass gc;
        gc.threads = c >= 2 ? 512 : 1024;
        gc.item0 = (int64_t)gitems.size();
        int64_t packs = 0;
        for (int i : order) {
            if (cls[i] != c) continue;
            gc.lds = std::max(gc.lds, (size_t)(plans[i]->full.nwords + 1) * kVariants[var[i]].WB + kGatherLdsExtra);
            packs += gs[i].npacks;
        }
        if (packs == 0) continue;
        int occ = 0;
        if (hipFuncSetAttribute((const void *)k_radon_gather_multi, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)gc.lds) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_radon_gather_multi, gc.threads, gc.lds) !=
                hipSuccess || occ < 1) {
            batch_free(b);
            return fail(-3, "batched gather kernel cannot be resident");
        }
        const int nw = gc.threads / 64;
        for (int i : order) {
            if (cls[i] != c) continue;
            GatherArgs &a = gs[i];
            a.nslice = choose_slices(plans[i], plans[i]->full, packs, nw, occ);
            ebsd_plan::Layout lay;
            if ((rc = get_layout(plans[i]->full, nw * a.nslice,
                                 kVariants[var[i]].FLT ? 0 : sched_kind(kVariants[var[i]].WB), &lay))) {
                batch_free(b);
                return rc;
            }
            a.gidx = lay.gidx;
            a.wofs = lay.wofs;
            a.wcnt = lay.wcnt;
            for (int64_t k = 0; k < a.npacks; ++k)
                for (int sl = 0; sl < a.nslice; ++sl) gitems.push_back(make_int2(i, (int)(k * a.nslice + sl)));
        }
        gc.nitems = (int64_t)gitems.size() - gc.item0;
        gc.grid = (int)std::min<int64_t>(gc.nitems, (int64_t)occ * b->ncu);
        b->classes.push_back(gc);
    }
    // peak items, one class per distinct geometry, in first-seen order (the streamed kernel
    // ignores the tiled kernel's tile shape, so sets that differ only there share a launch)
    auto geo_key = [](const ebsd_plan *p) {
        PeakGeo k = p->g;
        if (use_stream(p)) k.tile = k.ntiles = k.rh = k.apitch = k.bpitch = 0;
        return k;
    };
    std::vector<int2> pitems;
    std::vector<int> done(nsets, 0);
    for (int i = 0; i < nsets; ++i) done[i] = single[i];
    for (int i0 = 0; i0 < nsets; ++i0) {
        if (done[i0]) continue;
        const PeakGeo g = geo_key(plans[i0]);
        PeakClass pc;
        pc.g = g;
        pc.L = peak_lds_layout(g);
        pc.lds = (size_t)pc.L.total * 4;
        pc.threads = plans[i0]->pthreads;
        const bool fixed = g.hr == g.kr && g.ht == g.kt;
        pc.kr = fixed && g.kt == 5 && (g.kr == 9 || g.kr == 5) ? g.kr : 0;
        pc.kt = pc.kr ? 5 : 0;
        pc.stream = use_stream(plans[i0]);
        if (pc.stream) {
            pc.threads = 64 * stream_waves(g);
            pc.lds = (size_t)stream_lds_layout(stream_waves(g), g.nB).total * 4;
        }
        pc.item0 = (int64_t)pitems.size();
        for (int i : order) {
            if (done[i] || use_stream(plans[i]) != pc.stream) continue;
            const PeakGeo gi = geo_key(plans[i]);
            if (std::memcmp(&gi, &g, sizeof(PeakGeo)) != 0) continue;
            done[i] = 1;
            for (int64_t q = 0; q < N[i]; ++q) pitems.push_back(make_int2(i, (int)q));
        }
        pc.nitems = (int64_t)pitems.size() - pc.item0;
        if (pc.nitems > 0) b->pclasses.push_back(pc);
    }
    b->ngitems = (int64_t)gitems.size();
    b->npitems = (int64_t)pitems.size();
    if ((rc = dev_upload(&b->d_gsets, gs.data(), gs.size()))) { batch_free(b); return rc; }
    if ((rc = dev_upload(&b->d_variant, var.data(), var.size()))) { batch_free(b); return rc; }
    if ((rc = dev_upload(&b->d_gitems, gitems.data(), gitems.size()))) { batch_free(b); return rc; }
    if ((rc = dev_upload(&b->d_psets, ps.data(), ps.size()))) { batch_free(b); return rc; }
    if ((rc = dev_upload(&b->d_pitems, pitems.data(), pitems.size()))) { batch_free(b); return rc; }
    int occ = 0;
    // the dynamic-LDS limit of the gather kernel covers every class (set once, not per launch)
    size_t lds_all = 0;
    for (const GatherClass &gc : b->classes) lds_all = std::max(lds_all, gc.lds);
    if (lds_all > 0 && hipFuncSetAttribute((const void *)k_radon_gather_multi,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_all) != hipSuccess) {
        batch_free(b);
        return fail(-3, "batched gather kernel LDS limit");
    }
    size_t lds_pk[3] = {0, 0, 0}, lds_st[2] = {0, 0};
    for (PeakClass &pc : b->pclasses) {
        if (pc.stream) {
            pc.grid = (int)pc.nitems;
            const int k = pc.kr == 9 ? 0 : 1;
            lds_st[k] = std::max(lds_st[k], pc.lds);
            continue;
        }
        const int k = pc.kr == 9 ? 0 : pc.kr == 5 ? 1 : 2;
        const void *pk = k == 0 ? (const void *)k_peaks<9, 5, true>
                       : k == 1 ? (const void *)k_peaks<5, 5, true> : (const void *)k_peaks<0, 0, true>;
        occ = 0;
        if (hipFuncSetAttribute(pk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pc.lds) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, pk, pc.threads, pc.lds) != hipSuccess || occ < 1) {
            batch_free(b);
            return fail(-3, "batched peak kernel cannot be resident");
        }
        pc.grid = (int)std::min<int64_t>(pc.nitems, (int64_t)occ * b->ncu);
        lds_pk[k] = std::max(lds_pk[k], pc.lds);
    }
    // each instantiation's dynamic-LDS limit covers all of its classes
    const void *pks[3] = {(const void *)k_peaks<9, 5, true>, (const void *)k_peaks<5, 5, true>,
                          (const void *)k_peaks<0, 0, true>};
    for (int k = 0; k < 3; ++k)
        if (lds_pk[k] > 0 &&
            hipFuncSetAttribute(pks[k], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_pk[k]) != hipSuccess) {
            batch_free(b);
            return fail(-3, "batched peak kernel LDS limit");
        }
    const void *sks[2] = {(const void *)k_peaks_stream<9, 5, true>, (const void *)k_peaks_stream<5, 5, true>};
    for (int k = 0; k < 2; ++k)
        if (lds_st[k] > 0 &&
            hipFuncSetAttribute(sks[k], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_st[k]) != hipSuccess) {
            batch_free(b);
            return fail(-3, "batched peak kernel LDS limit");
        }
    // one launch of each per-size gather now: its plan structures and stream layouts exist
    // afterwards, so a later launch (or its graph capture) allocates nothing
    GatherOpts go;
    go.part = b->part;
    go.part_bytes = b->part_bytes;
    for (const ebsd_batch::Single &g : b->singles) {
        if ((rc = launch_gather(g.plan, g.pats, g.dtype, g.N, g.radon, nullptr, go)) ||
            (rc = launch_peaks(g.plan, g.radon, g.N, &g.out, nullptr))) {
            batch_free(b);
            return rc;
        }
    }
    if (!b->singles.empty() && hipDeviceSynchronize() != hipSuccess) {
        batch_free(b);
        return fail(-3, "batched per-size gather warm-up failed");
    }
    for (int i = 0; i < nsets; ++i)
        if (std::find(b->pinned.begin(), b->pinned.end(), plans[i]) == b->pinned.end()) {
            b->pinned.push_back(plans[i]);
            ++plans[i]->pins;
        }
    *out = b;
    return 0;
}

static int batch_launch(ebsd_batch *b, hipStream_t s)
{
    GatherOpts go;
    go.part = b->part;
    go.part_bytes = b->part_bytes;
    for (const ebsd_batch::Single &g : b->singles) {
        int rc = launch_gather(g.plan, g.pats, g.dtype, g.N, g.radon, s, go);
        if (!rc) rc = launch_peaks(g.plan, g.radon, g.N, &g.out, s);
        if (rc) return rc;
    }
    for (const GatherClass &gc : b->classes) {
        GatherMultiArgs gm;
        gm.sets = b->d_gsets;
        gm.variant = b->d_variant;
        gm.items = b->d_gitems + gc.item0;
        gm.nitems = gc.nitems;
        hipLaunchKernelGGL(k_radon_gather_multi, dim3(gc.grid), dim3(gc.threads), gc.lds, s, gm);
        HIPCHK(hipGetLastError());
    }
    for (const PeakClass &pc : b->pclasses) {
        PeakArgs a{};
        a.sets = b->d_psets;
        a.items = b->d_pitems + pc.item0;
        a.nitems = pc.nitems;
        a.L = pc.L;
        a.g = pc.g;
        if (pc.stream && pc.kr == 9)
            hipLaunchKernelGGL((k_peaks_stream<9, 5, true>), dim3(pc.grid), dim3(pc.threads), pc.lds, s, a);
        else if (pc.stream)
            hipLaunchKernelGGL((k_peaks_stream<5, 5, true>), dim3(pc.grid), dim3(pc.threads), pc.lds, s, a);
        else if (pc.kr == 9)
            hipLaunchKernelGGL((k_peaks<9, 5, true>), dim3(pc.grid), dim3(pc.threads), pc.lds, s, a);
        else if (pc.kr == 5)
            hipLaunchKernelGGL((k_peaks<5, 5, true>), dim3(pc.grid), dim3(pc.threads), pc.lds, s, a);
        else
            hipLaunchKernelGGL((k_peaks<0, 0, true>), dim3(pc.grid), dim3(pc.threads), pc.lds, s, a);
        HIPCHK(hipGetLastError());
    }
    return 0;
}

extern "C" int ebsd_batch_run(ebsd_batch *b, void *stream)
{
    if (!b) return fail(-1, "ebsd_batch_run: null batch");
    return batch_launch(b, (hipStream_t)stream);
}

extern "C" int ebsd_batch_graph_launch(ebsd_batch *b, void *stream)
{
    if (!b) return fail(-1, "ebsd_batch_graph_launch: null batch");
    hipStream_t s = (hipStream_t)stream;
    if (!s) return fail(-1, "ebsd_batch_graph_launch: graphs are captured on a non-default stream");
    if (!b->exec) {
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = batch_launch(b, s);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return fail(-3, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        b->graph = g;
        HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(b->exec, s));
    return 0;
}

extern "C" int ebsd_batch_destroy(ebsd_batch *b)
{
    batch_free(b);
    return 0;
}
