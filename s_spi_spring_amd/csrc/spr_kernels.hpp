// gfx950 kernels of the per-individual phase-retrieval iteration, batched over a
// population of P individuals that share one detector amplitude.
//
// HBM layout (one context = one pattern):
//   rho  [P][N][N] float2   real-space densities (the state between operations)
//   work [P][N][N] float2   half-transformed fields (row-transformed / column-processed)
//   supp [P][N][N] uint8    supports (0/1)
//   sbits[P][N*N/32] uint32 the supports one bit per pixel (bit i&31 of word i>>5),
//                           packed from supp at the start of every iteration block
//   amp  [N][N]    float    corner-origin amplitude with the mask encoding of the reference
//   tw   [N]       float2   exp(-2*pi*i*m/N)
//
// One iteration = two launches (the minimal number of HBM round trips for a field
// that does not fit in one CU's LDS), in one of two forms chosen by the row-pass
// variant:
//   state stored (rho_n read and written every iteration)
//     column pass  work <- colIFFT( P_M( colFFT(work) ) )                   16 B/px
//     row pass     rho' = rowIFFT(work)/N^2 ; rho <- update(rho', rho, S) ;
//                  work <- rowFFT(rho)                                32.125 B/px
//   state recomputed (work holds rowFFT(rho_n); rho is read and written only at
//   the ends of a block)
//     column pass  proj <- colIFFT( P_M( colFFT(work) ) )                   16 B/px
//     row pass     HIO: rho' = rowIFFT(proj)/N^2 ; rho_n = rowIFFT(work)/N ;
//                       work <- rowFFT(update(rho', rho_n, S))
//                  RAAR: q = rowIFFT(2 proj/N^2 - work/N) (= 2 rho' - rho_n) ;
//                       work <- proj/N - beta rowFFT(q on the rejected pixels)
//                                                                     24.125 B/px
// so HIO/RAAR move 48.125 or 40.125 B/px per individual-iteration (the second
// form saves the state's round trip for one more row transform in HIO and none
// in RAAR); ER, which needs no rho_n, moves 32.125 B/px in both. proj is the
// work2 buffer. (The support packing reads 1 and writes 0.125 B/px once per
// block.)
#pragma once

#include <cstdint>

#include <type_traits>

#include "spr_fft.hpp"

namespace spr {

// ---------------------------------------------------------------------------
// Launch geometry
// ---------------------------------------------------------------------------
constexpr int pow2_floor(int x) { return x >= 2 ? 2 * pow2_floor(x / 2) : 1; }

// Geometry of the row and column passes of size N with radix plan PL_, CCOLS
// columns per column-pass workgroup (0: the default width), RTHREADS threads per
// row-pass workgroup, row-pass prefetch mode RPF (-1: SPR_ROW_PREFETCH), PAD
// float2 of padding per row of the column tile, and TWG = 1 to read the column
// pass's twiddle table from global memory (through the caches) instead of
// building it in LDS. CSK > 0 lays the column tile out unpadded with one spare
// tile row after every 2^CSK rows (row i of the tile at (i + i/2^CSK) * C; PAD
// is then unused). CGR > 1 gives each column-pass workgroup CGR adjacent groups
// of C columns (a strip of C*CGR columns, so that C*CGR = 16 reads whole
// 128-byte row segments), transformed one group after another through the same
// C-column LDS tile; WPE > 0 sets the column pass's waves per SIMD (its
// register budget) instead of deriving it from the workgroups a CU holds.
// RST = 1 makes the row update passes recompute rho_n from its row transform
// instead of reading and writing it (the second iteration form above).
template <int N, class PL_ = RadixPlan<N>, int CCOLS = 0, int RTHREADS = 256, int RPF = -1, int PAD = 1, int TWG = 0,
          int RSH = 3, int CSK = 0, int CGR = 1, int WPE = 0, int RST = 0>
struct GeoT {
    using PL = PL_;
    static constexpr bool TW_GLOBAL = TWG != 0;
    static constexpr int ROW_PAD_SHIFT = RSH;            // row LDS: one float2 of padding per 2^RSH
    static constexpr int V = PL::V;                       // values per thread
    static constexpr int T = N / V;                       // threads per line
    static constexpr int ROW_THREADS = RTHREADS;
    static constexpr int ROW_PREFETCH = RPF;
    static constexpr bool ROW_RECOMPUTE = RST != 0;      // rho_n recomputed from work
    static constexpr int RW = ROW_THREADS / T;            // rows per row-pass workgroup
    static constexpr int RS = N + (N >> RSH);             // padded LDS row length (float2)
#ifndef SPR_COL_COLUMNS
#define SPR_COL_COLUMNS 8192   // columns x rows of a column-pass tile
#endif
#ifndef SPR_COL_WAVE
#define SPR_COL_WAVE 0         // 1: wavefront-owned columns when N/V <= 64
#endif
#ifndef SPR_COL_XCD
#define SPR_COL_XCD 1          // XCD-aware strip order for strips narrower than 128 bytes
#endif
#ifndef SPR_COL_AMP_PREFETCH
#define SPR_COL_AMP_PREFETCH 1 // load the amplitude with the field, before the forward transform
#endif
    // columns per column-pass workgroup: a power of two (it divides N)
    static constexpr int C = CCOLS ? CCOLS : (pow2_floor(SPR_COL_COLUMNS / N) < N ? pow2_floor(SPR_COL_COLUMNS / N) : N);
    static constexpr int COL_GROUPS = CGR;                // groups of C columns per column-pass workgroup
    static constexpr int COL_STRIP = C * CGR;             // columns per column-pass workgroup
    static constexpr bool COL_WAVE = SPR_COL_WAVE && T <= 64;
    static constexpr int COL_THREADS = C * T;
    static constexpr int CS = C + PAD;                    // padded LDS row length of a column tile
    static constexpr int COL_SKEW = CSK;
    // Column-major LDS tile of the wavefront-owned column pass (T <= 64): column
    // c, row i at c*CST + i + i/8. Within a column this is the row pass's padded
    // layout; the column stride CST = RPW (mod 2*RPW) float2, RPW = 64/C rows
    // per wavefront in the coalesced load layout, spreads a wavefront's C columns
    // x RPW rows evenly over the LDS banks when the tile is transposed.
    static constexpr int RPW = (C >= 64) ? 1 : 64 / C;
    static constexpr int CST = ((N + N / 8) / (2 * RPW) + 1) * (2 * RPW) + RPW;
    // LDS of a column-pass workgroup: the tile (float2), the twiddle table unless it
    // is read from global memory, and 512 bytes for the loss pass's reduction.
    static constexpr int COL_TILE = COL_WAVE ? C * CST : CSK ? (N + (N >> CSK)) * C : N * CS;   // float2 of the column tile
    static constexpr int COL_LDS = (COL_TILE + (TWG ? 0 : TwTableP<PL>::SIZE)) * 8 + 512;
#ifndef SPR_COL_WG_PER_CU
#define SPR_COL_WG_PER_CU 8
#endif
    // Workgroups per CU: as many as LDS and the 32 wave slots of a CU allow,
    // at most SPR_COL_WG_PER_CU; the register budget follows from it.
    static constexpr int COL_WG_LDS = (160 * 1024) / COL_LDS;
    static constexpr int COL_WG_SLOTS = 32 / (COL_THREADS / 64);
    static constexpr int COL_WG = COL_WG_LDS < COL_WG_SLOTS ? (COL_WG_LDS < SPR_COL_WG_PER_CU ? COL_WG_LDS : SPR_COL_WG_PER_CU)
                                                            : (COL_WG_SLOTS < SPR_COL_WG_PER_CU ? COL_WG_SLOTS : SPR_COL_WG_PER_CU);
    static constexpr int COL_WAVES_PER_EU = WPE > 0 ? WPE : (COL_WG > 0 ? COL_WG : 1) * COL_THREADS / 256;
    static_assert(RW >= 1 && ROW_THREADS % T == 0 && N % RW == 0, "row geometry");
    static_assert(COL_THREADS <= 1024 && N % COL_STRIP == 0, "column geometry");
    static_assert(CGR == 1 || !COL_WAVE, "column groups need the tile layout");
};
template <int N> using Geo = GeoT<N>;

// Variants of the two iteration passes that the autotuner chooses between
// (spr_autotune): variant 0 is the default geometry. All variants read and
// write the same HBM layout, so row and column variants combine freely; the
// column width and the row-pass threads and prefetch run the same operations
// (up to the compiler's contraction of multiply-adds), a different radix plan
// changes the rounding of the transforms.
template <int N, int CV> struct ColVariant { static constexpr bool valid = false; };
template <int N> struct ColVariant<N, 0> { static constexpr bool valid = true; using G = GeoT<N>; };
template <int N, int RV> struct RowVariant { static constexpr bool valid = false; };
template <int N> struct RowVariant<N, 0> { static constexpr bool valid = true; using G = GeoT<N>; };
#define SPR_COL_VARIANT(N_, I_, ...) \
    template <> struct ColVariant<N_, I_> { static constexpr bool valid = true; using G = GeoT<N_, __VA_ARGS__>; };
#define SPR_ROW_VARIANT(N_, I_, ...) \
    template <> struct RowVariant<N_, I_> { static constexpr bool valid = true; using G = GeoT<N_, __VA_ARGS__>; };
// column variants: plan, columns per workgroup
SPR_COL_VARIANT(256, 1, RadixPlan<256>, 16)
SPR_COL_VARIANT(256, 2, RadixPlan<256>, 8)
SPR_COL_VARIANT(384, 1, RadixPlan<384>, 8)
SPR_COL_VARIANT(384, 2, RadixPlan<384>, 32)
SPR_COL_VARIANT(384, 3, RadixPlan<384>, 4)
SPR_COL_VARIANT(512, 1, RadixPlan<512>, 8)
SPR_COL_VARIANT(512, 2, Plan<512, 16, 16, 4, 8>, 16)
SPR_COL_VARIANT(512, 3, RadixPlan<512>, 4)
SPR_COL_VARIANT(768, 1, RadixPlan<768>, 4, 256, -1, 1, 0, 3, 0, 4, 2)   // 4 groups of 4 columns
SPR_COL_VARIANT(768, 2, RadixPlan<768>, 4)
SPR_COL_VARIANT(1024, 1, Plan<1024, 16, 16, 8, 8>, 8, 256, -1, 1, 1, 3, 0, 2, 4)   // 2 groups of 8
SPR_COL_VARIANT(1024, 2, Plan<1024, 16, 16, 8, 8>, 8)
SPR_COL_VARIANT(1024, 3, Plan<1024, 16, 16, 8, 8>, 8, 256, -1, 0, 1, 3, 0, 2, 4)   // 2 groups of 8, unpadded
SPR_COL_VARIANT(1024, 4, Plan<1024, 16, 16, 8, 8>, 8, 256, -1, 0, 0, 3, 4)
// (plan, columns, -, -, padding, global twiddles): tiles under 80 KB, two workgroups per CU
SPR_COL_VARIANT(384, 4, RadixPlan<384>, 16, 256, -1, 1, 1)
SPR_COL_VARIANT(512, 4, RadixPlan<512>, 16, 256, -1, 1, 1)
SPR_COL_VARIANT(512, 5, RadixPlan<512>, 8, 256, -1, 1, 1)
SPR_COL_VARIANT(768, 3, RadixPlan<768>, 8, 256, -1, 1, 1)
SPR_COL_VARIANT(768, 4, RadixPlan<768>, 8, 256, -1, 0, 1)
SPR_COL_VARIANT(768, 5, RadixPlan<768>, 4, 256, -1, 1, 1)
SPR_COL_VARIANT(1024, 5, Plan<1024, 16, 16, 8, 8>, 8, 256, -1, 1, 1)
SPR_COL_VARIANT(1024, 6, Plan<1024, 16, 16, 8, 8>, 8, 256, -1, 0, 0)
SPR_COL_VARIANT(1024, 7, RadixPlan<1024>, 8, 256, -1, 1, 1)
SPR_COL_VARIANT(2048, 1, RadixPlan<2048>, 4, 256, -1, 0, 1)
SPR_COL_VARIANT(2048, 2, Plan<2048, 16, 16, 8, 16>, 8, 256, -1, 0, 1)
// (plan, columns, -, -, -, -, -, spare tile row every 2^k rows): tiles without
// the bank conflicts of the padded layouts. Measured on MI355X the conflicts
// cost little (the pass is bound by HBM and VALU): 768^2 4 columns 73.7 vs
// 76.7 us, 1024^2 radix 16 8 columns 68.7 vs 69.6 us, 512^2 8 columns 56.6 vs
// 57.4 us; at 256^2, 384^2 and 2048^2 the skewed tiles were slower.
SPR_COL_VARIANT(512, 6, RadixPlan<512>, 8, 256, -1, 0, 0, 3, 3)
SPR_COL_VARIANT(768, 6, RadixPlan<768>, 4, 256, -1, 0, 0, 3, 2)
// (plan, columns, -, -, padding, -, -, -, column groups, waves per SIMD): strips
// of 16 columns (whole 128-byte row segments) transformed as groups of 8 or 4
// columns through one tile, so that one group's transforms overlap the next
// group's loads and the previous group's stores
SPR_COL_VARIANT(512, 7, RadixPlan<512>, 8, 256, -1, 1, 0, 3, 0, 2, 0)
SPR_COL_VARIANT(768, 7, RadixPlan<768>, 8, 256, -1, 1, 0, 3, 0, 2, 4)
// row variants: plan, -, threads per workgroup, prefetch mode
SPR_ROW_VARIANT(256, 1, RadixPlan<256>, 0, 512)
SPR_ROW_VARIANT(384, 1, RadixPlan<384>, 0, 512)
SPR_ROW_VARIANT(512, 1, RadixPlan<512>, 0, 512)
SPR_ROW_VARIANT(768, 1, RadixPlan<768>, 0, 256, 1)
SPR_ROW_VARIANT(768, 2, RadixPlan<768>, 0, 512)
SPR_ROW_VARIANT(1024, 1, Plan<1024, 16, 16, 8, 8>)
SPR_ROW_VARIANT(1024, 2, Plan<1024, 16, 16, 8, 8>, 0, 256, 1)
SPR_ROW_VARIANT(1024, 3, Plan<1024, 16, 16, 8, 8>, 0, 512)
SPR_ROW_VARIANT(2048, 1, Plan<2048, 16, 16, 8, 16>)
// (plan, -, threads, prefetch, -, -, padding period 2^4)
SPR_ROW_VARIANT(1024, 4, Plan<1024, 16, 16, 8, 8>, 0, 256, -1, 1, 0, 4)
SPR_ROW_VARIANT(2048, 2, Plan<2048, 16, 16, 8, 16>, 0, 256, -1, 1, 0, 4)
// default plans with the padding period that a model of the LDS banks finds
// fewest conflicts for (2^4, or 2^5 for the radix-12 plans)
SPR_ROW_VARIANT(256, 2, RadixPlan<256>, 0, 256, -1, 1, 0, 4)
SPR_ROW_VARIANT(384, 2, RadixPlan<384>, 0, 256, -1, 1, 0, 5)
SPR_ROW_VARIANT(512, 2, RadixPlan<512>, 0, 256, -1, 1, 0, 4)
SPR_ROW_VARIANT(768, 3, RadixPlan<768>, 0, 256, -1, 1, 0, 5)
SPR_ROW_VARIANT(1024, 5, RadixPlan<1024>, 0, 256, -1, 1, 0, 4)
SPR_ROW_VARIANT(2048, 3, RadixPlan<2048>, 0, 256, -1, 1, 0, 4)
// (plan, -, threads, prefetch, -, -, padding period, -, -, -, recomputed state):
// the row update passes recompute rho_n from its row transform instead of
// reading and writing it (one more row transform for 8 fewer bytes per pixel);
// the column pass then writes its projection to a second field buffer
SPR_ROW_VARIANT(256, 3, RadixPlan<256>, 0, 256, -1, 1, 0, 4, 0, 1, 0, 1)
SPR_ROW_VARIANT(384, 3, RadixPlan<384>, 0, 256, -1, 1, 0, 5, 0, 1, 0, 1)
SPR_ROW_VARIANT(512, 3, RadixPlan<512>, 0, 256, -1, 1, 0, 3, 0, 1, 0, 1)
SPR_ROW_VARIANT(512, 4, RadixPlan<512>, 0, 256, -1, 1, 0, 4, 0, 1, 0, 1)
SPR_ROW_VARIANT(768, 4, RadixPlan<768>, 0, 256, -1, 1, 0, 5, 0, 1, 0, 1)
SPR_ROW_VARIANT(1024, 6, RadixPlan<1024>, 0, 256, -1, 1, 0, 4, 0, 1, 0, 1)
SPR_ROW_VARIANT(1024, 7, Plan<1024, 16, 16, 8, 8>, 0, 256, -1, 1, 0, 3, 0, 1, 0, 1)
SPR_ROW_VARIANT(2048, 4, RadixPlan<2048>, 0, 256, -1, 1, 0, 4, 0, 1, 0, 1)
SPR_ROW_VARIANT(2048, 5, Plan<2048, 16, 16, 8, 16>, 0, 256, -1, 1, 0, 4, 0, 1, 0, 1)
#undef SPR_COL_VARIANT
#undef SPR_ROW_VARIANT

constexpr int MAX_VARIANTS = 8;
template <int N> constexpr int col_variants() {
    int k = 1;
    if constexpr (ColVariant<N, 1>::valid) k = 2;
    if constexpr (ColVariant<N, 2>::valid) k = 3;
    if constexpr (ColVariant<N, 3>::valid) k = 4;
    if constexpr (ColVariant<N, 4>::valid) k = 5;
    if constexpr (ColVariant<N, 5>::valid) k = 6;
    if constexpr (ColVariant<N, 6>::valid) k = 7;
    if constexpr (ColVariant<N, 7>::valid) k = 8;
    return k;
}
template <int N> constexpr int row_variants() {
    int k = 1;
    if constexpr (RowVariant<N, 1>::valid) k = 2;
    if constexpr (RowVariant<N, 2>::valid) k = 3;
    if constexpr (RowVariant<N, 3>::valid) k = 4;
    if constexpr (RowVariant<N, 4>::valid) k = 5;
    if constexpr (RowVariant<N, 5>::valid) k = 6;
    if constexpr (RowVariant<N, 6>::valid) k = 7;
    if constexpr (RowVariant<N, 7>::valid) k = 8;
    return k;
}

// Loads and stores of the streamed fields (each value is read and written once
// per pass). With SPR_NT they carry the non-temporal hint.
#ifndef SPR_NT
#define SPR_NT 0
#endif
__device__ __forceinline__ float2 ld_stream(const float2* p) {
    if constexpr (SPR_NT) {
        using v2 = float __attribute__((ext_vector_type(2)));
        const v2 x = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
        return make_float2(x.x, x.y);
    } else {
        return *p;
    }
}
__device__ __forceinline__ void st_stream(float2* p, float2 v) {
    if constexpr (SPR_NT) {
        using v2 = float __attribute__((ext_vector_type(2)));
        v2 x;
        x.x = v.x;
        x.y = v.y;
        __builtin_nontemporal_store(x, reinterpret_cast<v2*>(p));
    } else {
        *p = v;
    }
}

// LDS row of a line transform with one float2 of padding after every 2^SH
// elements (SH = 3 by default; 4 suits plans whose first radix is 16, whose
// first-stage stores stride by 16 elements).
template <int SH> struct RowAccT {
    float2* base;
    __device__ __forceinline__ float2 ld(int i) const { return base[i + (i >> SH)]; }
    __device__ __forceinline__ void st(int i, float2 v) const { base[i + (i >> SH)] = v; }
    // Element i0 + k*STEP for STEP a multiple of 2^SH: one base register plus
    // constant offsets, since (i0 + k*STEP) >> SH = (i0 >> SH) + k*STEP/2^SH.
    template <int STEP> __device__ __forceinline__ float2 ldk(int i0, int k) const {
        static_assert(STEP % (1 << SH) == 0, "stride");
        return base[i0 + (i0 >> SH) + k * (STEP + (STEP >> SH))];
    }
    template <int STEP> __device__ __forceinline__ void stk(int i0, int k, float2 v) const {
        static_assert(STEP % (1 << SH) == 0, "stride");
        base[i0 + (i0 >> SH) + k * (STEP + (STEP >> SH))] = v;
    }
};
using RowAcc = RowAccT<3>;

// Column c of a tile of C columns stored unpadded, with one spare tile row after
// every 2^SH rows: element i at (i + (i >> SH)) * C. Two rows 2^SH apart then
// fall C float2 apart modulo 2C, so the rows that the lanes of one LDS access
// touch together spread over the banks.
template <int C, int SH> struct ColAccSkew {
    float2* base;
    __device__ __forceinline__ float2 ld(int i) const { return base[(i + (i >> SH)) * C]; }
    __device__ __forceinline__ void st(int i, float2 v) const { base[(i + (i >> SH)) * C] = v; }
    template <int STEP> __device__ __forceinline__ float2 ldk(int i0, int k) const {
        static_assert(STEP % (1 << SH) == 0, "stride");
        return base[(i0 + (i0 >> SH) + k * (STEP + (STEP >> SH))) * C];
    }
    template <int STEP> __device__ __forceinline__ void stk(int i0, int k, float2 v) const {
        static_assert(STEP % (1 << SH) == 0, "stride");
        base[(i0 + (i0 >> SH) + k * (STEP + (STEP >> SH))) * C] = v;
    }
};

template <int CS> struct ColAcc {
    float2* base;
    __device__ __forceinline__ float2 ld(int i) const { return base[i * CS]; }
    __device__ __forceinline__ void st(int i, float2 v) const { base[i * CS] = v; }
    template <int STEP> __device__ __forceinline__ float2 ldk(int i0, int k) const { return base[(i0 + k * STEP) * CS]; }
    template <int STEP> __device__ __forceinline__ void stk(int i0, int k, float2 v) const {
        base[(i0 + k * STEP) * CS] = v;
    }
};

// ---------------------------------------------------------------------------
// Per-pixel operators (src/gpu/gpu_context.cu:79-111, 519-593)
// ---------------------------------------------------------------------------

// Reality window: the phase fails when atan2f(y,x) >= hi_ge or <= lo_le. The
// float bounds are the nearest floats strictly beyond the reference's double
// thresholds, so the float comparison is the reference's double comparison.
// pi_only marks reality == 0, where only atan2f == +-pi_f can fail; the test
// then skips atan2f for every value that is not (nearly) a negative real.
struct PhaseCut {
    float hi_ge, lo_le;
    int pi_only;
};

__device__ __forceinline__ bool phase_fails(float x, float y, const PhaseCut& pc) {
    if (pc.pi_only) {
        if (!signbit(x) || fabsf(y) > 1e-6f * fabsf(x)) return false;
    }
    const float ph = atan2f(y, x);
    return ph >= pc.hi_ge || ph <= pc.lo_le;
}

// a * (cos phi, sin phi) with phi = atan2f(y, x) for a zero (or NaN) value:
// atan2f(+-0, +0) = +-0 and atan2f(+-0, -0) = +-pi_f, whose cosine rounds to -1
// and whose sine is -+8.742278e-8 (sin of the float nearest pi).
__device__ __forceinline__ float2 polar_of_zero(float2 F, float a) {
    if (F.x != F.x || F.y != F.y) return make_float2(F.x + F.y, F.x + F.y);
    if (!signbit(F.x)) return make_float2(a, copysignf(0.f, F.y) * a);
    const float s = 8.742278e-8f;
    return make_float2(-a, signbit(F.y) ? s * a : -s * a);
}

// Exact P_M for the rare cases: measured pixels whose |F|^2 is not a normal
// float (zero, tiny or NaN) and bounded pixels (a < -1.5).
__device__ __forceinline__ float2 project_modulus_exact(float2 F, float a, float m2) {
    const float m = sqrtf(m2);
    if (a >= 0.f) {
        if (m > 0.f) return make_float2(a * F.x / m, a * F.y / m);
        return polar_of_zero(F, a);
    }
    if (a < -1.5f && m > -a - 2.f) {
        const float nm = -a - 2.f;
        return make_float2(nm * F.x / m, nm * F.y / m);
    }
    return F;
}

// Fourier modulus projection P_M (sub_intensities). The measured-pixel case
// scales by a/|F| through the hardware reciprocal square root (within 2 ulp of
// the reference's a*F/|F|) whenever |F|^2 is a normal float, and free pixels
// (-1.5 <= a < 0) keep F (scale 1): both without a branch. Only a wavefront
// holding a tiny, zero or bounded value runs the exact formula.
__device__ __forceinline__ float2 project_modulus(float2 F, float a) {
    const float m2 = F.y * F.y + F.x * F.x;
    const bool fast = a >= 0.f && m2 > 1e-36f;
    const float s = fast ? a * rsqrtf(m2) : 1.f;
    float2 r = make_float2(F.x * s, F.y * s);
    if (!fast && !(a < 0.f && a >= -1.5f)) r = project_modulus_exact(F, a, m2);
    return r;
}

enum Alg { ALG_ER = 0, ALG_HIO = 1, ALG_RAAR = 2 };

// ---------------------------------------------------------------------------
// Row pass
// ---------------------------------------------------------------------------
enum RowMode {
    ROW_FWD = 0,      // work <- rowFFT+(rho)
    ROW_FINAL = 1,    // rho <- rowFFT-(work)/N^2
    ROW_HIO = 2,      // rho' = rowFFT-(work)/N^2; rho <- HIO; work <- rowFFT+(rho)
    ROW_RAAR = 3,
    ROW_ER = 4,       // as HIO without rho_n; rho is not written
    ROW_FWD_INPLACE = 5,   // rho <- rowFFT+(rho)   (2-D FT, first half)
    ROW_INV_INPLACE = 6,   // rho <- rowFFT-(rho)/N^2 (2-D IFT, second half)
};

struct RowArgs {
    float2* rho;
    float2* work;
    const uint8_t* supp;     // supports, one byte per pixel (on-chip blocks)
    const uint32_t* sbits;   // supports, one bit per pixel (row update passes)
    const float2* tw;
    float beta;
    float inv_nn;  // 1/N^2 (exact for power-of-two N; within 1 ulp of dividing by N^2 otherwise)
    PhaseCut pc;
    const float2* proj;  // column-projected fields, read by variants that recompute rho_n
    float inv_n;         // 1/N
};

// Support bit of pixel i of a field: bit i&31 of word i>>5 (N*N and the
// row-block offsets are multiples of 32).
__device__ __forceinline__ uint8_t support_bit(const uint32_t* __restrict__ bits, int i) {
    return (uint8_t)((bits[i >> 5] >> (i & 31)) & 1u);
}

// supp (bytes) -> bits: word w holds pixels [32w, 32w+32), bit k = (supp[32w+k] != 0).
// Each thread reads its 32 bytes as two 16-byte loads (supp 16-byte aligned).
__global__ void k_pack_support(const uint8_t* __restrict__ supp, uint32_t* __restrict__ bits, size_t nwords) {
    const size_t step = (size_t)gridDim.x * blockDim.x;
    for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < nwords; w += step) {
        const uint4 lo = *reinterpret_cast<const uint4*>(supp + 32 * w);
        const uint4 hi = *reinterpret_cast<const uint4*>(supp + 32 * w + 16);
        const unsigned x[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
        uint32_t m = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) m |= (uint32_t)(((x[k] >> (8 * j)) & 0xffu) != 0u) << (4 * k + j);
        bits[w] = m;
    }
}

// Loads of the state rho_n and the support issued with the field, ahead of the
// first transform: 0 none, 1 rho_n, 2 rho_n and the support. At 512^2 the RAAR
// row pass measured 85 us per 64-field launch with 2 (6 waves/SIMD) against
// 93 us with 0 (7 waves/SIMD).
#ifndef SPR_ROW_PREFETCH
#define SPR_ROW_PREFETCH 2
#endif

template <int N, int MODE, int RV = 0>
__global__ __launch_bounds__((RowVariant<N, RV>::G::ROW_THREADS)) void k_row(RowArgs a) {
    using G = typename RowVariant<N, RV>::G;
    using PL = typename G::PL;
    constexpr int V = G::V, T = G::T;
    constexpr int PREFETCH = G::ROW_PREFETCH >= 0 ? G::ROW_PREFETCH : SPR_ROW_PREFETCH;
    constexpr bool WV = T <= 64;   // every row inside one wavefront: no workgroup barriers
    constexpr bool RS = G::ROW_RECOMPUTE;
    static_assert(!RS || PREFETCH >= 1, "the recomputed state is transformed from its prefetched row");
    __shared__ float2 tw[TwTableP<PL>::SIZE];
    __shared__ float2 lds[G::RW * G::RS];
    const int tid = threadIdx.x;
    const int t = tid % T;
    const int rl = tid / T;
    tw_fill_p<PL>(tw, a.tw, tid, G::ROW_THREADS);
    // Block-uniform bases with 32-bit per-thread offsets (SGPR base addressing).
    const size_t blk = (size_t)blockIdx.x * G::RW * N;
    float2* __restrict__ rho = a.rho + blk;
    float2* __restrict__ work = a.work + blk;
    const uint32_t* __restrict__ sbits = a.sbits + blk / 32;
    const int off = rl * N + t;
    const RowAccT<G::ROW_PAD_SHIFT> acc{lds + rl * G::RS};
    float2 v[V];
    __syncthreads();

    if constexpr (MODE == ROW_FWD || MODE == ROW_FWD_INPLACE) {
#pragma unroll
        for (int r = 0; r < V; ++r) v[r] = rho[off + r * T];
        fft_line_p<PL, +1, WV>(v, acc, tw, t);
        float2* dst = (MODE == ROW_FWD) ? work : rho;
#pragma unroll
        for (int q = 0; q < V; ++q) dst[off + q * T] = v[q];
    } else if constexpr (MODE == ROW_FINAL || MODE == ROW_INV_INPLACE) {
        const float2* src = (MODE == ROW_FINAL) ? (RS ? a.proj + blk : work) : rho;
#pragma unroll
        for (int r = 0; r < V; ++r) v[r] = src[off + r * T];
        fft_line_p<PL, -1, WV>(v, acc, tw, t);
#pragma unroll
        for (int q = 0; q < V; ++q) rho[off + q * T] = make_float2(v[q].x * a.inv_nn, v[q].y * a.inv_nn);
    } else {
        // The state and the support may be loaded with the field, before the
        // transform, so that their latency hides behind it. With RS the field
        // comes from proj and the state is recomputed from its row transform
        // (work): rho_n = rowIFFT(rowFFT(rho_n)) / N, within float rounding.
        float2 bn[V];
        uint8_t sn[V];
        const float2* __restrict__ fin = RS ? a.proj + blk : work;
#pragma unroll
        for (int r = 0; r < V; ++r) {
            v[r] = ld_stream(fin + off + r * T);
            if constexpr (PREFETCH >= 2) sn[r] = support_bit(sbits, off + r * T);
            if constexpr (MODE != ROW_ER && PREFETCH >= 1) bn[r] = ld_stream((RS ? work : rho) + off + r * T);
        }
        if constexpr (RS && MODE == ROW_RAAR) {
            // RAAR needs rho' and rho_n only through q = 2 rho' - rho_n: its
            // rejected branch beta rho_n + (1 - 2 beta) rho' is rho' - beta q. The
            // transforms are linear, so q is one inverse transform of
            // 2 proj / N^2 - work / N, and the new row transform is
            // rowFFT(rho') - beta rowFFT(c) = proj / N - beta rowFFT(c), with c = q
            // on the rejected pixels and 0 elsewhere: two row transforms, as with
            // the state stored.
            const float two_nn = 2.f * a.inv_nn;
#pragma unroll
            for (int r = 0; r < V; ++r) {
                const float2 pr = v[r], wr = bn[r];
                v[r] = make_float2(two_nn * pr.x - a.inv_n * wr.x, two_nn * pr.y - a.inv_n * wr.y);
                bn[r] = make_float2(a.inv_n * pr.x, a.inv_n * pr.y);
            }
            fft_line_p<PL, -1, WV>(v, acc, tw, t);   // q
#pragma unroll
            for (int q = 0; q < V; ++q) {
                const uint8_t s = PREFETCH >= 2 ? sn[q] : support_bit(sbits, off + q * T);
                if (s != 0 && !phase_fails(v[q].x, v[q].y, a.pc)) v[q] = make_float2(0.f, 0.f);
            }
            fft_line_p<PL, +1, WV>(v, acc, tw, t);   // rowFFT(c)
#pragma unroll
            for (int q = 0; q < V; ++q)
                st_stream(work + off + q * T, make_float2(bn[q].x - a.beta * v[q].x, bn[q].y - a.beta * v[q].y));
            return;
        }
        if constexpr (RS && MODE != ROW_ER) {
            fft_line_p<PL, -1, WV>(bn, acc, tw, t);
#pragma unroll
            for (int r = 0; r < V; ++r) bn[r] = make_float2(bn[r].x * a.inv_n, bn[r].y * a.inv_n);
        }
        fft_line_p<PL, -1, WV>(v, acc, tw, t);
#pragma unroll
        for (int q = 0; q < V; ++q) {
            const int i = off + q * T;
            const float2 d = make_float2(v[q].x * a.inv_nn, v[q].y * a.inv_nn);  // rho'
            const uint8_t s = PREFETCH >= 2 ? sn[q] : support_bit(sbits, i);
            float2 o;
            if constexpr (MODE == ROW_ER) {
                const float ts = (s != 0 && !phase_fails(d.x, d.y, a.pc)) ? 1.f : 0.f;
                o = make_float2(d.x * ts, d.y * ts);
            } else if constexpr (MODE == ROW_HIO) {
                const float2 b = PREFETCH >= 1 ? bn[q] : rho[i];
                const float ts = (s != 0 && !phase_fails(d.x, d.y, a.pc)) ? 1.f : 0.f;
                o = make_float2((b.x - a.beta * d.x) * (1.f - ts) + d.x * ts,
                                (b.y - a.beta * d.y) * (1.f - ts) + d.y * ts);
            } else {  // RAAR
                const float2 b = PREFETCH >= 1 ? bn[q] : rho[i];
                const float qx = 2.f * d.x - b.x, qy = 2.f * d.y - b.y;
                o = d;
                if (s == 0 || phase_fails(qx, qy, a.pc)) {
                    const float c = 1.f - 2.f * a.beta;
                    o = make_float2(a.beta * b.x + c * d.x, a.beta * b.y + c * d.y);
                }
            }
            if constexpr (MODE != ROW_ER && !RS) st_stream(rho + i, o);
            v[q] = o;
        }
        fft_line_p<PL, +1, WV>(v, acc, tw, t);
#pragma unroll
        for (int q = 0; q < V; ++q) st_stream(work + off + q * T, v[q]);
    }
}

// ---------------------------------------------------------------------------
// Column pass
// ---------------------------------------------------------------------------
enum ColMode {
    COL_PROJ = 0,   // colIFFT(P_M(colFFT(.)))
    COL_FWD = 1,    // colFFT+(.)
    COL_INV = 2,    // colFFT-(.)
};

struct ColArgs {
    float2* f;          // field batch, transformed in place
    const float* amp;   // amplitudes [patterns][N][N]
    const float2* tw;
    const float* amp_t; // the amplitudes transposed, amp_t[x*N + y] = amp[y*N + x]
    const int* amp_of;  // field -> pattern (NULL: every field belongs to pattern 0)
    const float2* twtab;  // the plan's twiddle table in global memory (variants with TW_GLOBAL)
    float2* out;          // output fields (= f: in place)
};

// Offset of the amplitude of field p (block-uniform: one scalar load).
template <int N> __device__ __forceinline__ size_t amp_offset(const int* amp_of, int p) {
    return amp_of ? (size_t)amp_of[p] * N * N : 0;
}

// Column strip of workgroup `b`. Workgroups are dispatched round-robin over the
// 8 XCDs (b mod 8), each with its own L2. When a strip is narrower than a
// 128-byte line (C < 16, N >= 1024), the 16/C strips that share those lines
// are given to workgroups b, b+8, ... of the same XCD, so each line is fetched
// from HBM into one L2 only.
template <int N, int C = Geo<N>::C> __device__ __forceinline__ int col_strip(int b) {
    constexpr int GS = C < 16 ? 16 / C : 1;     // strips per 128-byte line
    if constexpr (GS == 1 || SPR_COL_XCD == 0) {
        return b;
    } else {
        constexpr int CH = 8 * GS;              // workgroups per chunk of lines
        static_assert((N / C) % CH == 0, "a field's strips must fill whole chunks (the remap stays in range)");
        const int j = b % CH;
        return (b - j) + (j % 8) * GS + j / 8;
    }
}

// Column pass over C adjacent columns of one field. Global memory is read and
// written as rows of C consecutive values (coalesced); the C columns' transforms
// exchange data through a padded LDS tile with workgroup barriers. With
// SPR_COL_WAVE and N/8 <= 64 the tile is instead transposed through LDS so that
// each wavefront owns whole columns and its transforms synchronise with
// wavefront fences, reading the amplitude from its transposed copy; at 512^2
// this form measured 71 us per 64-field launch against 66 us for the default.
template <int N, int MODE, int CV = 0>
__global__ __launch_bounds__((ColVariant<N, CV>::G::COL_THREADS))
__attribute__((amdgpu_waves_per_eu((ColVariant<N, CV>::G::COL_WAVES_PER_EU)))) void k_col(ColArgs a) {
    using G = typename ColVariant<N, CV>::G;
    using PL = typename G::PL;
    constexpr int V = G::V, T = G::T, C = G::C, CS = G::CS;
    constexpr bool WV = G::COL_WAVE;
    __shared__ float2 twl[G::TW_GLOBAL ? 1 : TwTableP<PL>::SIZE];
    __shared__ float2 lds[G::COL_TILE];
    const int tid = threadIdx.x;
    const int c = tid % C;
    const int t = tid / C;
    const float2* tw = twl;
    if constexpr (G::TW_GLOBAL) tw = a.twtab;
    else tw_fill_p<PL>(twl, a.tw, tid, G::COL_THREADS);
    constexpr int CG = G::COL_GROUPS;
    constexpr int TILES = N / G::COL_STRIP;
    const int strip = col_strip<N, G::COL_STRIP>(blockIdx.x);
    const int p = strip / TILES;
    const int x0 = (strip % TILES) * G::COL_STRIP;
    // Block-uniform bases with 32-bit per-thread offsets (SGPR base addressing).
    float2* __restrict__ f = a.f + (size_t)p * N * N + x0;
    float2* __restrict__ fo = a.out + (size_t)p * N * N + x0;
    const int off = t * N + c;
    // Group g of the strip is columns [g*C, (g+1)*C). The loads are issued
    // group after group, so that the wait for group 0's values leaves the later
    // groups' loads in flight during group 0's transforms (the other half of
    // each line comes from the L2).
    float2 vg[CG][V];
#pragma unroll
    for (int g = 0; g < CG; ++g)
#pragma unroll
        for (int r = 0; r < V; ++r) vg[g][r] = ld_stream(f + off + g * C + r * T * N);
    float2(&v)[V] = vg[0];
    if constexpr (WV) {
        constexpr int CST = G::CST;
        const int cc = tid / T;   // column owned by this thread's wavefront
        const int tl = tid % T;
        const RowAcc mine{lds + c * CST};
#pragma unroll
        for (int r = 0; r < V; ++r) mine.stk<T>(t, r, v[r]);
        __syncthreads();
        const RowAcc acc{lds + cc * CST};
#pragma unroll
        for (int r = 0; r < V; ++r) v[r] = acc.ldk<T>(tl, r);
        if constexpr (MODE == COL_PROJ || MODE == COL_FWD) fft_line_p<PL, +1, true>(v, acc, tw, tl);
        if constexpr (MODE == COL_PROJ) {
            const float* __restrict__ at = a.amp_t + amp_offset<N>(a.amp_of, p) + (x0 + cc) * N;
#pragma unroll
            for (int q = 0; q < V; ++q) v[q] = project_modulus(v[q], at[tl + q * T]);
        }
        if constexpr (MODE == COL_PROJ || MODE == COL_INV) fft_line_p<PL, -1, true>(v, acc, tw, tl);
#pragma unroll
        for (int q = 0; q < V; ++q) acc.stk<T>(tl, q, v[q]);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < V; ++r) v[r] = mine.ldk<T>(t, r);
    } else {
        const float* __restrict__ amp = a.amp + amp_offset<N>(a.amp_of, p) + x0;
        using Acc = std::conditional_t<(G::COL_SKEW > 0), ColAccSkew<C, (G::COL_SKEW > 0 ? G::COL_SKEW : 1)>, ColAcc<CS>>;
        const Acc acc{lds + c};
        // The amplitude of group 0 is loaded with the field; that of a later
        // group when its turn starts, so that its latency hides behind the
        // group's forward transform while only one group's values are live.
        float am[V];
        if constexpr (MODE == COL_PROJ && SPR_COL_AMP_PREFETCH) {
#pragma unroll
            for (int q = 0; q < V; ++q) am[q] = amp[off + q * T * N];
        }
        __syncthreads();
#pragma unroll
        for (int g = 0; g < CG; ++g) {
            if constexpr (MODE == COL_PROJ && SPR_COL_AMP_PREFETCH) {
                if (g > 0) {
#pragma unroll
                    for (int q = 0; q < V; ++q) am[q] = amp[off + g * C + q * T * N];
                }
            }
            if constexpr (MODE == COL_PROJ || MODE == COL_FWD) fft_line_p<PL, +1>(vg[g], acc, tw, t);
            if constexpr (MODE == COL_PROJ) {
#pragma unroll
                for (int q = 0; q < V; ++q)
                    vg[g][q] = project_modulus(vg[g][q], SPR_COL_AMP_PREFETCH ? am[q] : amp[off + g * C + q * T * N]);
            }
            if constexpr (MODE == COL_PROJ || MODE == COL_INV) fft_line_p<PL, -1>(vg[g], acc, tw, t);
            if constexpr (CG > 1) {
#pragma unroll
                for (int q = 0; q < V; ++q) st_stream(fo + off + g * C + q * T * N, vg[g][q]);
            }
        }
        if constexpr (CG > 1) return;
    }
#pragma unroll
    for (int q = 0; q < V; ++q) st_stream(fo + off + q * T * N, v[q]);
}

}  // namespace spr

namespace spr {

// The twiddle table of plan PL in global memory (one workgroup), for the
// column-pass variants that read it from there.
template <class PL> __global__ void k_build_twtab(float2* out, const float2* __restrict__ root) {
    tw_fill_p<PL>(out, root, threadIdx.x, blockDim.x);
}

}  // namespace spr
