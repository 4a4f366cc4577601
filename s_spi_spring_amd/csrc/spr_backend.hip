// Host side of libspring_hip.so: population-resident contexts and the C ABI
// declared in include/spring_hip.h. Kernels are in spr_kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "spring_hip.h"
#include "spr_fused.hpp"
#include "spr_grad.hpp"
#include "spr_tuned.hpp"

// Working set per chunk of an iteration block: a population up to this size
// runs whole; a larger one in the fewest equal chunks under it. On MI355X (256
// MB Infinity Cache) 300 MB measured best: 512^2 with 512 individuals runs at
// 0.72 of the HBM peak in chunks of 64 against 0.59 unchunked, while splitting a
// population that already fits (64 individuals, 285 MB) costs 8 %.
#ifndef SPR_CHUNK_MB_DEFAULT
#define SPR_CHUNK_MB_DEFAULT 300
#endif

// Populations from which iteration blocks of 64^2 and 128^2 fields run on chip
// (one workgroup per individual) unless spr_set_fused says otherwise. Measured
// on MI355X, RAAR blocks of 25 iterations: at 128^2 the on-chip form runs at
// 1.03x the two-pass form with 8 individuals, 1.64x with 64 and 2.7x with 256;
// at 64^2 at 1.27x with 8 and 2.5x with 256. Smaller populations, which occupy
// fewer CUs than the two-pass form spreads them over, were not measured.
#ifndef SPR_FUSED_MIN_P
#define SPR_FUSED_MIN_P 8
#endif

using namespace spr;

struct spr_ctx {
    int n = 0, cap = 0, P = 0, device = 0;
    hipStream_t stream = nullptr;
    // Second lane of the pipelined iteration: the population is split in two
    // halves whose column and row passes run on two streams half an iteration
    // apart, so that the VALU-bound column pass of one half overlaps the
    // HBM-bound row pass of the other.
    int lanes = 1;
    hipStream_t lane2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

    float2* rho = nullptr;
    float2* work = nullptr;
    float2* work2 = nullptr;      // second half-transformed field per individual (correlations, |b|)
    uint8_t* supp = nullptr;
    uint32_t* sbits = nullptr;    // supports one bit per pixel, packed at the start of every iteration block
    float* amp = nullptr;         // amplitudes [npat][n][n]
    float* amp_t = nullptr;       // amp transposed (read along columns by the column pass)
    int npat = 1;                 // patterns sharing this handle (one amplitude each)
    int colvar = 0, rowvar = 0;   // variants of the column and row iteration passes
    float2* twtab[MAX_VARIANTS] = {};  // global twiddle tables of the column variants that read one
    size_t chunk_bytes = 0;       // iteration blocks run in chunks of this working set (0: whole population)
    int fused = -1;               // on-chip iteration blocks (N <= 128): -1 by population size, 0 never, 1 always
    int* amp_of = nullptr;        // slot -> pattern when npat > 1, else NULL
    float2* tw = nullptr;
    // reduction scratch
    double* partial = nullptr;    // per workgroup partial sums
    double* sums = nullptr;       // [cap][SUMS_PER_ITEM]
    unsigned* maxbits = nullptr;  // [cap][2]
    unsigned long long* keys = nullptr;  // [cap][2]
    unsigned* radix = nullptr;    // [3][cap]: prefix, mask, remaining
    unsigned* hist = nullptr;     // [cap][2048]
    size_t partial_cap = 0;       // doubles
    // Alignment / averaging / crossover workspace, grown on demand.
    struct Buf {
        void* p = nullptr;
        size_t bytes = 0;
    };
    Buf spec, xf, xs, supp2, fieldbuf, ampavg, xidx, shiftp, xchild, xmaps, xphase, adam_m, adam_v, gidx, lrpart;
    Buf pinned;                   // pinned host staging of spr_put/fetch_individuals (hipHostMalloc)
    std::string err;
    // Opt-in per-launch timing: events bracket every launch on the stream.
    bool timing = false;
    int timing_stride = 1;             // bracket every timing_stride-th launch of each kind
    long kind_seen[SPR_TIMING_KINDS] = {};
    bool bracket_open = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_kind;          // kind of the launch bracketed by events 2i, 2i+1
    double kind_ms[SPR_TIMING_KINDS] = {};
    long kind_count[SPR_TIMING_KINDS] = {};
};

namespace {

thread_local std::string g_create_error;

int fail(spr_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

int harvest_timing(spr_ctx* c);

#define SPR_HIP(ctx, call)                                                                        \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail((ctx), SPR_ERR_HIP, std::string(#call " failed: ") + hipGetErrorString(e_)); \
    } while (0)

float next_up(double x) {
    float f = (float)x;
    if ((double)f <= x) f = std::nextafter(f, INFINITY);
    return f;
}
float next_down(double x) {
    float f = (float)x;
    if ((double)f >= x) f = std::nextafter(f, -INFINITY);
    return f;
}

// Event bracket of one launch (no-op unless timing is enabled).
int timing_begin(spr_ctx* c, int kind) {
    c->bracket_open = false;
    if (!c->timing) return SPR_OK;
    if (c->kind_seen[kind]++ % c->timing_stride != 0) return SPR_OK;
    c->bracket_open = true;
    const size_t i = c->ev_kind.size();
    while (c->ev_pool.size() < 2 * (i + 1)) {
        hipEvent_t e;
        SPR_HIP(c, hipEventCreate(&e));
        c->ev_pool.push_back(e);
    }
    c->ev_kind.push_back(kind);
    SPR_HIP(c, hipEventRecord(c->ev_pool[2 * i], c->stream));
    return SPR_OK;
}
int timing_end(spr_ctx* c) {
    if (!c->bracket_open) return SPR_OK;
    c->bracket_open = false;
    const size_t i = c->ev_kind.size() - 1;
    SPR_HIP(c, hipEventRecord(c->ev_pool[2 * i + 1], c->stream));
    return SPR_OK;
}
int harvest_timing(spr_ctx* c) {
    for (size_t i = 0; i < c->ev_kind.size(); ++i) {
        float ms = 0.f;
        SPR_HIP(c, hipEventElapsedTime(&ms, c->ev_pool[2 * i], c->ev_pool[2 * i + 1]));
        c->kind_ms[c->ev_kind[i]] += ms;
        c->kind_count[c->ev_kind[i]] += 1;
    }
    c->ev_kind.clear();
    return SPR_OK;
}

std::vector<float> transpose_amp(const float* amp, int n) {
    std::vector<float> t((size_t)n * n);
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) t[(size_t)x * n + y] = amp[(size_t)y * n + x];
    return t;
}

// Reality windows of src/gpu/gpu_context.cu:530, 553 (ER/HIO) and 582 (RAAR).
PhaseCut cut_er_hio(float r) {
    const double t = (double)(1.f - r) * M_PI;
    return PhaseCut{next_up(t), next_down(-t), r == 0.f ? 1 : 0};
}
PhaseCut cut_raar(float r) {
    return PhaseCut{next_up((1.0 - (double)r) * M_PI), next_down(-(double)(1.f - r) * M_PI), r == 0.f ? 1 : 0};
}

// Row / column passes over `cnt` individuals starting at the pointers in the
// arguments, on stream `s`, in the handle's variant.
template <int N, int RV> int launch_row_v(spr_ctx* c, int mode, const RowArgs& a, int cnt, hipStream_t s) {
    using G = typename RowVariant<N, RV>::G;
    const dim3 grid((unsigned)((size_t)cnt * N / G::RW)), block(G::ROW_THREADS);
    switch (mode) {
        case ROW_FWD: hipLaunchKernelGGL((k_row<N, ROW_FWD, RV>), grid, block, 0, s, a); break;
        case ROW_FINAL: hipLaunchKernelGGL((k_row<N, ROW_FINAL, RV>), grid, block, 0, s, a); break;
        case ROW_HIO: hipLaunchKernelGGL((k_row<N, ROW_HIO, RV>), grid, block, 0, s, a); break;
        case ROW_RAAR: hipLaunchKernelGGL((k_row<N, ROW_RAAR, RV>), grid, block, 0, s, a); break;
        case ROW_ER: hipLaunchKernelGGL((k_row<N, ROW_ER, RV>), grid, block, 0, s, a); break;
        case ROW_FWD_INPLACE: hipLaunchKernelGGL((k_row<N, ROW_FWD_INPLACE, RV>), grid, block, 0, s, a); break;
        case ROW_INV_INPLACE: hipLaunchKernelGGL((k_row<N, ROW_INV_INPLACE, RV>), grid, block, 0, s, a); break;
        default: return fail(c, SPR_ERR_ARG, "bad row mode");
    }
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}
template <int N> int launch_row(spr_ctx* c, int mode, const RowArgs& a, int cnt, hipStream_t s) {
    switch (c->rowvar) {
        case 0: return launch_row_v<N, 0>(c, mode, a, cnt, s);
        case 1:
            if constexpr (RowVariant<N, 1>::valid) return launch_row_v<N, 1>(c, mode, a, cnt, s);
            break;
        case 2:
            if constexpr (RowVariant<N, 2>::valid) return launch_row_v<N, 2>(c, mode, a, cnt, s);
            break;
        case 3:
            if constexpr (RowVariant<N, 3>::valid) return launch_row_v<N, 3>(c, mode, a, cnt, s);
            break;
        case 4:
            if constexpr (RowVariant<N, 4>::valid) return launch_row_v<N, 4>(c, mode, a, cnt, s);
            break;
        case 5:
            if constexpr (RowVariant<N, 5>::valid) return launch_row_v<N, 5>(c, mode, a, cnt, s);
            break;
        case 6:
            if constexpr (RowVariant<N, 6>::valid) return launch_row_v<N, 6>(c, mode, a, cnt, s);
            break;
        case 7:
            if constexpr (RowVariant<N, 7>::valid) return launch_row_v<N, 7>(c, mode, a, cnt, s);
            break;
    }
    return fail(c, SPR_ERR_ARG, "bad row variant");
}

// Column passes: the projection pass in every variant, the plain transforms
// (2-D FT / IFT, off the iteration path) in variant 0; in place, or into `out`.
template <int N, int CV>
int launch_col_v(spr_ctx* c, int mode, float2* f, int cnt, hipStream_t s, int first, float2* out = nullptr) {
    using G = typename ColVariant<N, CV>::G;
    if (G::TW_GLOBAL && !c->twtab[CV]) return fail(c, SPR_ERR_STATE, "twiddle table missing");
    const ColArgs a{f, c->amp, c->tw, c->amp_t, c->amp_of ? c->amp_of + first : nullptr, c->twtab[CV], out ? out : f};
    const dim3 grid((unsigned)((size_t)cnt * (N / G::COL_STRIP))), block(G::COL_THREADS);
    switch (mode) {
        case COL_PROJ: hipLaunchKernelGGL((k_col<N, COL_PROJ, CV>), grid, block, 0, s, a); break;
        case COL_FWD:
            if constexpr (CV == 0) hipLaunchKernelGGL((k_col<N, COL_FWD, 0>), grid, block, 0, s, a);
            break;
        case COL_INV:
            if constexpr (CV == 0) hipLaunchKernelGGL((k_col<N, COL_INV, 0>), grid, block, 0, s, a);
            break;
        default: return fail(c, SPR_ERR_ARG, "bad column mode");
    }
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}

// Global twiddle tables of every column variant of size N that reads one.
template <int N, int CV = 0> int build_twtabs(spr_ctx* c) {
    if constexpr (CV < MAX_VARIANTS) {
        if constexpr (ColVariant<N, CV>::valid) {
            using G = typename ColVariant<N, CV>::G;
            if constexpr (G::TW_GLOBAL) {
                constexpr int SIZE = TwTableP<typename G::PL>::SIZE;
                SPR_HIP(c, hipMalloc(&c->twtab[CV], SIZE * sizeof(float2)));
                hipLaunchKernelGGL((k_build_twtab<typename G::PL>), dim3(1), dim3(256), 0, c->stream, c->twtab[CV],
                                   (const float2*)c->tw);
                SPR_HIP(c, hipGetLastError());
            }
        }
        return build_twtabs<N, CV + 1>(c);
    }
    return SPR_OK;
}

template <int N> int launch_col(spr_ctx* c, int mode, float2* f, int cnt, hipStream_t s, int first, float2* out) {
    if (mode != COL_PROJ) return launch_col_v<N, 0>(c, mode, f, cnt, s, first, out);
    switch (c->colvar) {
        case 0: return launch_col_v<N, 0>(c, mode, f, cnt, s, first, out);
        case 1:
            if constexpr (ColVariant<N, 1>::valid) return launch_col_v<N, 1>(c, mode, f, cnt, s, first, out);
            break;
        case 2:
            if constexpr (ColVariant<N, 2>::valid) return launch_col_v<N, 2>(c, mode, f, cnt, s, first, out);
            break;
        case 3:
            if constexpr (ColVariant<N, 3>::valid) return launch_col_v<N, 3>(c, mode, f, cnt, s, first, out);
            break;
        case 4:
            if constexpr (ColVariant<N, 4>::valid) return launch_col_v<N, 4>(c, mode, f, cnt, s, first, out);
            break;
        case 5:
            if constexpr (ColVariant<N, 5>::valid) return launch_col_v<N, 5>(c, mode, f, cnt, s, first, out);
            break;
        case 6:
            if constexpr (ColVariant<N, 6>::valid) return launch_col_v<N, 6>(c, mode, f, cnt, s, first, out);
            break;
        case 7:
            if constexpr (ColVariant<N, 7>::valid) return launch_col_v<N, 7>(c, mode, f, cnt, s, first, out);
            break;
    }
    return fail(c, SPR_ERR_ARG, "bad column variant");
}

// Grid sizes with compiled kernels (RadixPlan in spr_fft.hpp).
#define SPR_SIZES(X) X(64) X(128) X(256) X(384) X(512) X(768) X(1024) X(2048)

#define SPR_DISPATCH_N(c, FN, ...)                                          \
    switch ((c)->n) {                                                       \
        case 64: return FN<64>(__VA_ARGS__);                                \
        case 128: return FN<128>(__VA_ARGS__);                              \
        case 256: return FN<256>(__VA_ARGS__);                              \
        case 384: return FN<384>(__VA_ARGS__);                              \
        case 512: return FN<512>(__VA_ARGS__);                              \
        case 768: return FN<768>(__VA_ARGS__);                              \
        case 1024: return FN<1024>(__VA_ARGS__);                            \
        case 2048: return FN<2048>(__VA_ARGS__);                            \
        default: return fail((c), SPR_ERR_ARG, "unsupported grid size");   \
    }

int row_n(spr_ctx* c, int mode, const RowArgs& a, int cnt, hipStream_t s) {
    SPR_DISPATCH_N(c, launch_row, c, mode, a, cnt, s);
}
// Column pass over fields f[0, cnt), the individuals [first, first + cnt),
// in place or into out[0, cnt).
int col_n(spr_ctx* c, int mode, float2* f, int cnt, hipStream_t s, int first = 0, float2* out = nullptr) {
    SPR_DISPATCH_N(c, launch_col, c, mode, f, cnt, s, first, out);
}

// Timed passes (spr_timing) over `cnt` individuals from the argument pointers
// (all resident individuals by default).
int row(spr_ctx* c, int mode, const RowArgs& a, int cnt = -1) {
    const int kind = (mode == ROW_HIO || mode == ROW_RAAR || mode == ROW_ER) ? SPR_KIND_ROW_UPDATE : SPR_KIND_ROW_OTHER;
    int rc = timing_begin(c, kind);
    if (rc != SPR_OK) return rc;
    rc = row_n(c, mode, a, cnt < 0 ? c->P : cnt, c->stream);
    if (rc != SPR_OK) return rc;
    return timing_end(c);
}
int col(spr_ctx* c, int mode, float2* f, int cnt = -1, int first = 0, float2* out = nullptr) {
    const int kind = (mode == COL_PROJ) ? SPR_KIND_COL_PROJ : SPR_KIND_COL_OTHER;
    int rc = timing_begin(c, kind);
    if (rc != SPR_OK) return rc;
    rc = col_n(c, mode, f, cnt < 0 ? c->P : cnt, c->stream, first, out);
    if (rc != SPR_OK) return rc;
    return timing_end(c);
}

RowArgs row_args(spr_ctx* c) {
    RowArgs a{};
    a.rho = c->rho;
    a.work = c->work;
    a.supp = c->supp;
    a.sbits = c->sbits;
    a.tw = c->tw;
    a.beta = 0.f;
    a.inv_nn = 1.0f / (float)((double)c->n * c->n);
    a.pc = cut_er_hio(0.f);
    a.proj = nullptr;
    a.inv_n = 1.0f / (float)c->n;
    return a;
}

// Whether row variant rv of size n recomputes rho_n from its row transform
// (RowVariant::G::ROW_RECOMPUTE); false for variants that do not exist.
template <int N, int RV = 0> bool row_recomputes_n(int rv) {
    if constexpr (RV < MAX_VARIANTS) {
        if constexpr (RowVariant<N, RV>::valid)
            if (rv == RV) return RowVariant<N, RV>::G::ROW_RECOMPUTE;
        return row_recomputes_n<N, RV + 1>(rv);
    }
    return false;
}
bool row_recomputes(int n, int rv) {
#define SPR_RS_CASE(N_) if (n == N_) return row_recomputes_n<N_>(rv);
    SPR_SIZES(SPR_RS_CASE)
#undef SPR_RS_CASE
    return false;
}

#define SPR_TRY(expr)                 \
    do {                              \
        int rc_ = (expr);             \
        if (rc_ != SPR_OK) return rc_; \
    } while (0)

// ---------------------------------------------------------------------------
// Auxiliary passes (shrink-wrap, fitness, symmetry, correlations)
// ---------------------------------------------------------------------------
constexpr int SUMS_PER_ITEM = 8;   // msum, err, nrm, rerr, rnrm
constexpr int RADIX_BINS = 2048;

template <int N> int launch_row_aux(spr_ctx* c, int mode, const AuxRowArgs& a, int items) {
    const dim3 grid((unsigned)((size_t)items * N / Geo<N>::RW)), block(Geo<N>::ROW_THREADS);
    switch (mode) {
        case AROW_ABS_FWD_WORK: hipLaunchKernelGGL((k_row_aux<N, AROW_ABS_FWD_WORK>), grid, block, 0, c->stream, a); break;
        case AROW_MAG_MAX: hipLaunchKernelGGL((k_row_aux<N, AROW_MAG_MAX>), grid, block, 0, c->stream, a); break;
        case AROW_ERR_FWD: hipLaunchKernelGGL((k_row_aux<N, AROW_ERR_FWD>), grid, block, 0, c->stream, a); break;
        case AROW_ERR_REAL: hipLaunchKernelGGL((k_row_aux<N, AROW_ERR_REAL>), grid, block, 0, c->stream, a); break;
        case AROW_ABS_FWD_RHO: hipLaunchKernelGGL((k_row_aux<N, AROW_ABS_FWD_RHO>), grid, block, 0, c->stream, a); break;
        case AROW_CORR: hipLaunchKernelGGL((k_row_aux<N, AROW_CORR>), grid, block, 0, c->stream, a); break;
        default: return fail(c, SPR_ERR_ARG, "bad aux row mode");
    }
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}

template <int N> int launch_col_aux(spr_ctx* c, int mode, const AuxColArgs& a, int items) {
    const dim3 grid((unsigned)((size_t)items * (N / Geo<N>::C))), block(Geo<N>::COL_THREADS);
    switch (mode) {
        case ACOL_BLUR: hipLaunchKernelGGL((k_col_aux<N, ACOL_BLUR>), grid, block, 0, c->stream, a); break;
        case ACOL_ERR: hipLaunchKernelGGL((k_col_aux<N, ACOL_ERR>), grid, block, 0, c->stream, a); break;
        case ACOL_CORR: hipLaunchKernelGGL((k_col_aux<N, ACOL_CORR>), grid, block, 0, c->stream, a); break;
        case ACOL_FWD_UNIT: hipLaunchKernelGGL((k_col_aux<N, ACOL_FWD_UNIT>), grid, block, 0, c->stream, a); break;
        default: return fail(c, SPR_ERR_ARG, "bad aux column mode");
    }
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}

int row_aux(spr_ctx* c, int mode, const AuxRowArgs& a, int items) {
    SPR_DISPATCH_N(c, launch_row_aux, c, mode, a, items);
}
int col_aux(spr_ctx* c, int mode, const AuxColArgs& a, int items) {
    SPR_DISPATCH_N(c, launch_col_aux, c, mode, a, items);
}

// Rows per row-pass workgroup and columns per column-pass workgroup of size n.
int rows_per_wg(int n) {
#define SPR_RW_CASE(N_) if (n == N_) return Geo<N_>::RW;
    SPR_SIZES(SPR_RW_CASE)
#undef SPR_RW_CASE
    return 1;
}
int cols_per_wg(int n) {
#define SPR_C_CASE(N_) if (n == N_) return Geo<N_>::C;
    SPR_SIZES(SPR_C_CASE)
#undef SPR_C_CASE
    return 1;
}

AuxRowArgs aux_row_args(spr_ctx* c) {
    AuxRowArgs a{};
    a.src = c->rho;
    a.work = c->work;
    a.work2 = c->work2;
    a.supp = c->supp;
    a.tw = c->tw;
    a.mag = reinterpret_cast<float*>(c->work2);
    a.maxbits = c->maxbits;
    a.key1 = c->keys;
    a.key2 = c->keys + c->cap;
    a.partial = c->partial;
    a.inv_nn = 1.0f / (float)((double)c->n * c->n);
    a.pc = cut_er_hio(0.f);
    a.src_index = nullptr;
    return a;
}
AuxColArgs aux_col_args(spr_ctx* c) {
    AuxColArgs a{};
    a.f = c->work;
    a.f2 = c->work2;
    a.amp = c->amp;
    a.amp_of = c->amp_of;
    a.tw = c->tw;
    a.partial = c->partial;
    a.maxbits = c->maxbits;
    return a;
}

// The reference forms the window constants in float/double mixes
// (src/gpu/gpu_context.cu:1171-1180); reproduced here on the host.
CorrWin make_corr_win(int n, float minr, float maxr) {
    CorrWin w;
    const float sigmatot = (float)((double)((float)n * minr) / 2. + 1e-12);
    const float sigma = (float)((double)sigmatot / 10.);
    const float sigmatotmax = (float)((double)((float)n * maxr) / 2. + 1e-12);
    const float sigmamax = (float)((double)sigmatotmax / 10.);
    w.sigtot_minus_sig = sigmatot - sigma;
    w.sig2 = sigma * sigma;
    w.sigtotmax_minus_sigmax = sigmatotmax - sigmamax;
    w.sigmax2 = sigmamax * sigmamax;
    return w;
}

// Grid of a grid-stride fill: enough workgroups to cover the chip, at most one
// per 256 elements.
dim3 fill_grid(size_t n) { return dim3((unsigned)std::min<size_t>((n + 255) / 256, 8192)); }
int fill_u32(spr_ctx* c, unsigned* p, unsigned v, size_t n) {
    if (n == 0) return SPR_OK;
    hipLaunchKernelGGL(k_fill_u32, fill_grid(n), dim3(256), 0, c->stream, p, v, n);
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}
int fill_u64(spr_ctx* c, unsigned long long* p, unsigned long long v, size_t n) {
    if (n == 0) return SPR_OK;
    hipLaunchKernelGGL(k_fill_u64, fill_grid(n), dim3(256), 0, c->stream, p, v, n);
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}
int reduce(spr_ctx* c, int per_item, int K, int out_off, int items) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(items), dim3(256), 0, c->stream, (const double*)c->partial, per_item,
                       K, c->sums, SUMS_PER_ITEM, out_off, items);
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}

#define SPR_TRY2(expr)                 \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ != SPR_OK) return rc_; \
    } while (0)

// |b| of shrink-wrap into mag = (float*)work2 and max|b| into maxbits[0..P)
// (src/gpu/gpu_context.cu:911-944).
int queue_blur_magnitude(spr_ctx* c, float sigma, bool want_max) {
    RowArgs ra = row_args(c);
    SPR_TRY2(row(c, ROW_FWD, ra));
    SPR_TRY2(col(c, COL_PROJ, c->work));
    AuxRowArgs a = aux_row_args(c);
    SPR_TRY2(row_aux(c, AROW_ABS_FWD_WORK, a, c->P));
    AuxColArgs ca = aux_col_args(c);
    ca.sigma_ft = (float)((double)((float)c->n / sigma) / M_PI);
    SPR_TRY2(col_aux(c, ACOL_BLUR, ca, c->P));
    if (want_max) SPR_TRY2(fill_u32(c, c->maxbits, 0u, c->P));
    SPR_TRY2(row_aux(c, AROW_MAG_MAX, a, c->P));
    return SPR_OK;
}

int queue_sw(spr_ctx* c, float sigma, float threshold) {
    if (c->P <= 0) return SPR_OK;
    SPR_TRY2(queue_blur_magnitude(c, sigma, true));
    const size_t total = (size_t)c->n * c->n * c->P;
    hipLaunchKernelGGL(k_cut_max, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->stream,
                       (const float*)c->work2, (const unsigned*)c->maxbits, threshold, c->supp, c->n * c->n, c->P);
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}

int queue_swa(spr_ctx* c, float sigma, int spixels) {
    if (c->P <= 0) return SPR_OK;
    const int npix = c->n * c->n;
    SPR_TRY2(queue_blur_magnitude(c, sigma, true));
    unsigned* prefix = c->radix;
    unsigned* mask = c->radix + c->cap;
    unsigned* rem = c->radix + 2 * c->cap;
    const unsigned want = spixels <= 0 ? 0u : (unsigned)(spixels > npix ? npix : spixels);
    SPR_TRY2(fill_u32(c, prefix, 0u, c->P));
    SPR_TRY2(fill_u32(c, mask, 0u, c->P));
    SPR_TRY2(fill_u32(c, rem, want, c->P));
    const int bpi = 16;   // histogram blocks per individual
    const int los[3] = {21, 10, 0};
    for (int pass = 0; pass < 3; ++pass) {
        SPR_TRY2(fill_u32(c, c->hist, 0u, (size_t)c->P * RADIX_BINS));
        const int lo = los[pass];
        if (pass < 2) {
            hipLaunchKernelGGL((k_radix_hist<11>), dim3(c->P * bpi), dim3(256), 0, c->stream, (const float*)c->work2,
                               npix, lo, (const unsigned*)prefix, (const unsigned*)mask, c->hist, bpi);
            SPR_HIP(c, hipGetLastError());
            hipLaunchKernelGGL((k_radix_pick<11>), dim3((c->P + 63) / 64), dim3(64), 0, c->stream,
                               (const unsigned*)c->hist, lo, prefix, mask, rem, c->P);
        } else {
            hipLaunchKernelGGL((k_radix_hist<10>), dim3(c->P * bpi), dim3(256), 0, c->stream, (const float*)c->work2,
                               npix, lo, (const unsigned*)prefix, (const unsigned*)mask, c->hist, bpi);
            SPR_HIP(c, hipGetLastError());
            hipLaunchKernelGGL((k_radix_pick<10>), dim3((c->P + 63) / 64), dim3(64), 0, c->stream,
                               (const unsigned*)c->hist, lo, prefix, mask, rem, c->P);
        }
        SPR_HIP(c, hipGetLastError());
    }
    const size_t total = (size_t)npix * c->P;
    hipLaunchKernelGGL(k_cut_bits, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->stream,
                       (const float*)c->work2, (const unsigned*)prefix, c->supp, npix, c->P);
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}

// get_error: sums[p] = {msum, err, nrm, rerr, rnrm}, maxbits[p] = max|F|.
int queue_error(spr_ctx* c, float reality) {
    const int n = c->n;
    AuxRowArgs a = aux_row_args(c);
    a.pc = cut_er_hio(reality);
    SPR_TRY2(row_aux(c, AROW_ERR_FWD, a, c->P));
    SPR_TRY2(reduce(c, n / rows_per_wg(n), 1, 0, c->P));
    AuxColArgs ca = aux_col_args(c);
    SPR_TRY2(fill_u32(c, c->maxbits, 0u, c->P));
    SPR_TRY2(col_aux(c, ACOL_ERR, ca, c->P));
    SPR_TRY2(reduce(c, n / cols_per_wg(n), 2, 1, c->P));
    SPR_TRY2(row_aux(c, AROW_ERR_REAL, a, c->P));
    SPR_TRY2(reduce(c, n / rows_per_wg(n), 2, 3, c->P));
    return SPR_OK;
}

// Symmetry correlation maxima into keys (key1: conj(a) a, key2: a a).
int queue_symmetry(spr_ctx* c, float symmin, float symmax) {
    AuxRowArgs a = aux_row_args(c);
    a.src = c->rho;
    SPR_TRY2(row_aux(c, AROW_ABS_FWD_RHO, a, c->P));
    AuxColArgs ca = aux_col_args(c);
    ca.win = make_corr_win(c->n, symmin, symmax);
    ca.unit_scale = 1;
    ca.ref = nullptr;
    ca.ref_index = nullptr;
    SPR_TRY2(col_aux(c, ACOL_CORR, ca, c->P));
    SPR_TRY2(fill_u64(c, c->keys, 0ull, 2 * (size_t)c->cap));
    SPR_TRY2(row_aux(c, AROW_CORR, a, c->P));
    return SPR_OK;
}

float key_value(unsigned long long k) {
    const unsigned b = (unsigned)(k >> 32);
    float f;
    std::memcpy(&f, &b, 4);
    return k ? f : 0.f;
}
int key_index(unsigned long long k) { return k ? (int)(0xFFFFFFFFu - (unsigned)(k & 0xFFFFFFFFu)) : 0; }

// supp -> sbits for individuals [0, cnt) from the pointers given (one bit per
// pixel, read by the row update passes).
int pack_support(spr_ctx* c, const uint8_t* supp, uint32_t* bits, int cnt, hipStream_t s) {
    const size_t nwords = (size_t)cnt * c->n * c->n / 32;
    if (nwords == 0) return SPR_OK;
    const dim3 grid((unsigned)std::min<size_t>((nwords + 255) / 256, 8192));
    hipLaunchKernelGGL(k_pack_support, grid, dim3(256), 0, s, supp, bits, nwords);
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}

// Pipelined form of the iteration over two halves of the population: lane 1
// (individuals [0, h), the context's stream) and lane 2 ([h, P), a second
// stream) issue the same pass sequence; lane 2 starts after lane 1's first row
// pass, so that the column pass of each half runs beside the row pass of the
// other. The halves share no memory (the twiddles and the amplitude are read
// only); the context's stream waits for lane 2 at the end.
int queue_iterate_lanes(spr_ctx* c, int mode, int niter, const RowArgs& a0) {
    const int h = (c->P + 1) / 2, h2 = c->P - h;
    const size_t off = (size_t)h * c->n * c->n;
    // With a row variant that recomputes rho_n the column passes project into work2.
    const bool rs = row_recomputes(c->n, c->rowvar);
    float2* pa = rs ? c->work2 : nullptr;
    float2* pb = rs ? c->work2 + off : nullptr;
    RowArgs a = a0;
    a.proj = pa;
    RowArgs b = a;
    b.rho += off;
    b.work += off;
    b.supp += off;
    b.sbits += off / 32;
    b.proj = pb;
    hipStream_t s1 = c->stream, s2 = c->lane2;
    SPR_TRY(pack_support(c, c->supp, c->sbits, c->P, s1));
    SPR_TRY(row_n(c, ROW_FWD, a, h, s1));
    SPR_HIP(c, hipEventRecord(c->ev_fork, s1));
    SPR_HIP(c, hipStreamWaitEvent(s2, c->ev_fork, 0));
    SPR_TRY(row_n(c, ROW_FWD, b, h2, s2));
    for (int i = 0; i < niter; ++i) {
        SPR_TRY(col_n(c, COL_PROJ, a.work, h, s1, 0, pa));
        SPR_TRY(col_n(c, COL_PROJ, b.work, h2, s2, h, pb));
        SPR_TRY(row_n(c, mode, a, h, s1));
        SPR_TRY(row_n(c, mode, b, h2, s2));
    }
    SPR_TRY(col_n(c, COL_PROJ, a.work, h, s1, 0, pa));
    SPR_TRY(col_n(c, COL_PROJ, b.work, h2, s2, h, pb));
    SPR_TRY(row_n(c, ROW_FINAL, a, h, s1));
    SPR_TRY(row_n(c, ROW_FINAL, b, h2, s2));
    SPR_HIP(c, hipEventRecord(c->ev_join, s2));
    SPR_HIP(c, hipStreamWaitEvent(s1, c->ev_join, 0));
    return SPR_OK;
}

// Individuals per chunk of an iteration block: the population split into the
// fewest equal chunks whose densities, fields and supports (17 bytes per pixel;
// with the recomputed state the iterations touch the two field buffers and the
// support bits instead, about as much) fit in c->chunk_bytes (at least one
// individual per chunk).
int chunk_individuals(const spr_ctx* c) {
    if (c->chunk_bytes == 0 || c->P <= 1) return std::max(c->P, 1);
    const size_t per = (size_t)17 * c->n * c->n;
    const int fit = (int)std::max<size_t>(1, c->chunk_bytes / per);
    const int chunks = (c->P + fit - 1) / fit;
    return (c->P + chunks - 1) / chunks;
}

// One iteration block on the context's stream: the first row transform, niter
// x [column pass, row update], the trailing projection and the last row pass,
// run chunk after chunk. Individuals are independent, so the chunks compute
// exactly what one pass over all of them would; a chunk's working set stays in
// the GPU's last-level cache (256 MB Infinity Cache on MI355X) across its
// iterations instead of streaming the whole population from HBM per pass. With
// a row variant that recomputes rho_n the column passes project into work2
// (proj) and rho is read by the first row pass and written by the last only.
int queue_iterate_block(spr_ctx* c, int mode, int niter, const RowArgs& a0) {
    const int chunk = chunk_individuals(c);
    const size_t np = (size_t)c->n * c->n;
    const bool rs = row_recomputes(c->n, c->rowvar);
    SPR_TRY(pack_support(c, a0.supp, const_cast<uint32_t*>(a0.sbits), c->P, c->stream));
    for (int first = 0; first < c->P; first += chunk) {
        const int cnt = std::min(chunk, c->P - first);
        RowArgs a = a0;
        a.rho += first * np;
        a.work += first * np;
        a.supp += first * np;
        a.sbits += first * np / 32;
        float2* pj = rs ? c->work2 + first * np : nullptr;
        a.proj = pj;
        SPR_TRY(row(c, ROW_FWD, a, cnt));
        for (int i = 0; i < niter; ++i) {
            SPR_TRY(col(c, COL_PROJ, a.work, cnt, first, pj));
            SPR_TRY(row(c, mode, a, cnt));
        }
        SPR_TRY(col(c, COL_PROJ, a.work, cnt, first, pj));
        SPR_TRY(row(c, ROW_FINAL, a, cnt));
    }
    return SPR_OK;
}

// One iteration block with each individual's whole field in one workgroup's
// LDS (spr_fused.hpp): one launch, HBM read at its start and written at its end.
template <int N> int launch_fused(spr_ctx* c, int mode, int niter, const RowArgs& a) {
    if constexpr (fused_size<N>()) {
        const FusedArgs f{a.rho, a.supp, c->amp_t, c->amp_of, c->tw, niter, a.beta, a.inv_nn, a.pc};
        const dim3 grid((unsigned)c->P), block(N * (N / 16));
        switch (mode) {
            case ROW_ER: hipLaunchKernelGGL((k_iter_fused<N, ROW_ER>), grid, block, 0, c->stream, f); break;
            case ROW_HIO: hipLaunchKernelGGL((k_iter_fused<N, ROW_HIO>), grid, block, 0, c->stream, f); break;
            case ROW_RAAR: hipLaunchKernelGGL((k_iter_fused<N, ROW_RAAR>), grid, block, 0, c->stream, f); break;
            default: return fail(c, SPR_ERR_ARG, "bad iteration mode");
        }
        SPR_HIP(c, hipGetLastError());
        return SPR_OK;
    } else {
        return fail(c, SPR_ERR_ARG, "no on-chip iteration block for this grid size");
    }
}

int fused_available(int n) { return n == 64 || n == 128; }

bool use_fused(const spr_ctx* c) {
    if (!fused_available(c->n) || c->fused == 0) return false;
    return c->fused == 1 || c->P >= SPR_FUSED_MIN_P;
}

int fused_n(spr_ctx* c, int mode, int niter, const RowArgs& a) { SPR_DISPATCH_N(c, launch_fused, c, mode, niter, a); }

int queue_iterate_fused(spr_ctx* c, int mode, int niter, const RowArgs& a) {
    SPR_TRY(timing_begin(c, SPR_KIND_BLOCK));
    SPR_TRY(fused_n(c, mode, niter, a));
    return timing_end(c);
}

int queue_iterate(spr_ctx* c, int alg, int niter, float beta, float reality) {
    if (alg < SPR_ER || alg > SPR_RAAR) return fail(c, SPR_ERR_ARG, "unknown algorithm");
    if (niter < 0) return fail(c, SPR_ERR_ARG, "negative iteration count");
    if (c->P <= 0) return SPR_OK;
    RowArgs a = row_args(c);
    a.beta = beta;
    a.pc = (alg == SPR_RAAR) ? cut_raar(reality) : cut_er_hio(reality);
    const int mode = (alg == SPR_HIO) ? ROW_HIO : (alg == SPR_RAAR) ? ROW_RAAR : ROW_ER;
    SPR_HIP(c, hipSetDevice(c->device));
    if (use_fused(c)) return queue_iterate_fused(c, mode, niter, a);
    if (c->lanes > 1 && c->P >= 2 && !c->timing) return queue_iterate_lanes(c, mode, niter, a);
    return queue_iterate_block(c, mode, niter, a);
}

// ---------------------------------------------------------------------------
// Workspace and batched helpers of alignment, averages and crossover
// ---------------------------------------------------------------------------
int ensure(spr_ctx* c, spr_ctx::Buf& b, size_t bytes) {
    if (b.bytes >= bytes) return SPR_OK;
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    if (b.p) SPR_HIP(c, hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
    SPR_HIP(c, hipMalloc(&b.p, bytes));
    b.bytes = bytes;
    return SPR_OK;
}

template <int N> int launch_fft_rows(spr_ctx* c, int sign, FieldBatch b, int items) {
    const dim3 grid((unsigned)((size_t)items * N / Geo<N>::RW)), block(Geo<N>::ROW_THREADS);
    if (sign > 0) hipLaunchKernelGGL((k_fft_rows<N, +1>), grid, block, 0, c->stream, b);
    else hipLaunchKernelGGL((k_fft_rows<N, -1>), grid, block, 0, c->stream, b);
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}
template <int N> int launch_fft_cols(spr_ctx* c, int sign, FieldBatch b, int items) {
    const dim3 grid((unsigned)((size_t)items * (N / Geo<N>::C))), block(Geo<N>::COL_THREADS);
    if (sign > 0) hipLaunchKernelGGL((k_fft_cols<N, +1>), grid, block, 0, c->stream, b);
    else hipLaunchKernelGGL((k_fft_cols<N, -1>), grid, block, 0, c->stream, b);
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}
template <int N> int launch_corr_spec(spr_ctx* c, CorrSpecArgs a, int items) {
    a.amp = c->amp;
    a.tw = c->tw;
    hipLaunchKernelGGL((k_corr_spec_cols<N>), dim3((unsigned)((size_t)items * (N / Geo<N>::C))),
                       dim3(Geo<N>::COL_THREADS), 0, c->stream, a);
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}
template <int N> int launch_reshift(spr_ctx* c, ReshiftArgs a, int items) {
    hipLaunchKernelGGL((k_reshift<N>), dim3((N * N + 255) / 256, items), dim3(256), 0, c->stream, a);
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}
template <int N> int launch_cross(spr_ctx* c, CrossArgs a, int items) {
    hipLaunchKernelGGL((k_cross<N>), dim3((N * N + 255) / 256, items), dim3(256), 0, c->stream, a);
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}
int fft_rows(spr_ctx* c, int sign, FieldBatch b, int items) { SPR_DISPATCH_N(c, launch_fft_rows, c, sign, b, items); }
int fft_cols(spr_ctx* c, int sign, FieldBatch b, int items) { SPR_DISPATCH_N(c, launch_fft_cols, c, sign, b, items); }
int corr_spec(spr_ctx* c, const CorrSpecArgs& a, int items) { SPR_DISPATCH_N(c, launch_corr_spec, c, a, items); }
int reshift_fields(spr_ctx* c, const ReshiftArgs& a, int items) { SPR_DISPATCH_N(c, launch_reshift, c, a, items); }
int cross_fields(spr_ctx* c, const CrossArgs& a, int items) { SPR_DISPATCH_N(c, launch_cross, c, a, items); }

// 2-D FT (sign +1, unnormalised) or IFT (sign -1, divided by N^2) of listed fields.
int fft2_fields(spr_ctx* c, int sign, float2* base, const int* index, int items) {
    if (items <= 0) return SPR_OK;
    FieldBatch b{base, index, c->tw, 1.f};
    if (sign > 0) {
        SPR_TRY2(fft_rows(c, +1, b, items));
        return fft_cols(c, +1, b, items);
    }
    SPR_TRY2(fft_cols(c, -1, b, items));
    b.scale = 1.0f / (float)((double)c->n * c->n);
    return fft_rows(c, -1, b, items);
}

// FT(|src field|) of listed fields into consecutive spectra at dst.
int modulus_spectra(spr_ctx* c, const float2* src, const int* src_index, float2* dst, int items) {
    if (items <= 0) return SPR_OK;
    AuxRowArgs a = aux_row_args(c);
    a.src = src;
    a.src_index = src_index;
    a.work = dst;
    SPR_TRY2(row_aux(c, AROW_ABS_FWD_RHO, a, items));
    FieldBatch b{dst, nullptr, c->tw, 1.f};
    return fft_cols(c, +1, b, items);
}

// Correlate spectra d_index[i] against r_index[i], pick the shifts and write
// them to shiftp (first index of the strict maximum, direct vs twin).
int correlate_and_pick(spr_ctx* c, const float2* spec, const int* d_index, const int* r_index, float minr,
                       float maxr, int unit_scale, const float* pdirect, const float* pflip, ShiftParam* shiftp,
                       int items) {
    CorrSpecArgs ca{};
    ca.spec = spec;
    ca.d_index = d_index;
    ca.r_index = r_index;
    ca.out1 = c->work;
    ca.out2 = c->work2;
    ca.win = make_corr_win(c->n, minr, maxr);
    ca.unit_scale = unit_scale;
    SPR_TRY2(corr_spec(c, ca, items));
    SPR_TRY2(fill_u64(c, c->keys, 0ull, 2 * (size_t)c->cap));
    AuxRowArgs a = aux_row_args(c);
    SPR_TRY2(row_aux(c, AROW_CORR, a, items));
    hipLaunchKernelGGL(k_pick_shift, dim3((items + 63) / 64), dim3(64), 0, c->stream,
                       (const unsigned long long*)c->keys, (const unsigned long long*)(c->keys + c->cap), pdirect,
                       pflip, c->n, shiftp, items);
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}

// ---------------------------------------------------------------------------
// Masked Fourier loss and gradient (spr_grad.hpp)
// ---------------------------------------------------------------------------
// Column loss pass in the geometry of column variant CV. want_loss: write one
// loss partial per workgroup.
template <int N, int CV> int launch_col_loss_v(spr_ctx* c, LossArgs a, int items, bool want_loss) {
    using G = typename ColVariant<N, CV>::G;
    if (G::TW_GLOBAL && !c->twtab[CV]) return fail(c, SPR_ERR_STATE, "twiddle table missing");
    a.twtab = c->twtab[CV];
    const dim3 grid((unsigned)((size_t)items * (N / G::COL_STRIP))), block(G::COL_THREADS);
    if (want_loss) hipLaunchKernelGGL((k_col_loss<N, CV, true>), grid, block, 0, c->stream, a);
    else hipLaunchKernelGGL((k_col_loss<N, CV, false>), grid, block, 0, c->stream, a);
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}
// Column variants whose tile holds the default radix plan compute the same
// arithmetic as variant 0 (only the LDS layout and the columns per workgroup
// differ).
template <int N, int CV> constexpr bool same_plan_variant() {
    if constexpr (ColVariant<N, CV>::valid) {
        using G = typename ColVariant<N, CV>::G;
        return G::COL_GROUPS == 1 && !G::COL_WAVE && std::is_same<typename G::PL, RadixPlan<N>>::value;
    }
    return false;
}
template <int N> int launch_col_loss(spr_ctx* c, LossArgs a, int items, bool want_loss) {
    if (!want_loss) {
        switch (c->colvar) {
            case 1:
                if constexpr (same_plan_variant<N, 1>()) return launch_col_loss_v<N, 1>(c, a, items, false);
                break;
            case 2:
                if constexpr (same_plan_variant<N, 2>()) return launch_col_loss_v<N, 2>(c, a, items, false);
                break;
            case 3:
                if constexpr (same_plan_variant<N, 3>()) return launch_col_loss_v<N, 3>(c, a, items, false);
                break;
            case 4:
                if constexpr (same_plan_variant<N, 4>()) return launch_col_loss_v<N, 4>(c, a, items, false);
                break;
            case 5:
                if constexpr (same_plan_variant<N, 5>()) return launch_col_loss_v<N, 5>(c, a, items, false);
                break;
            case 6:
                if constexpr (same_plan_variant<N, 6>()) return launch_col_loss_v<N, 6>(c, a, items, false);
                break;
            case 7:
                if constexpr (same_plan_variant<N, 7>()) return launch_col_loss_v<N, 7>(c, a, items, false);
                break;
        }
    }
    return launch_col_loss_v<N, 0>(c, a, items, want_loss);
}
// The loss pass with one loss partial per workgroup (n / cols_per_wg(n) per
// field) in the default geometry, and the refinement steps' gradient pass
// without partials, in the handle's column variant when that variant holds the
// default radix plan (the same gradients as spr_loss_grad, bit for bit), else
// in the default geometry too (at 1024^2 the radix-16 variants measured no
// faster for this pass: 0.577 against 0.575 ms per refinement step).
int col_loss(spr_ctx* c, const LossArgs& a, int items) { SPR_DISPATCH_N(c, launch_col_loss, c, a, items, true); }
int col_loss_step(spr_ctx* c, const LossArgs& a, int items) { SPR_DISPATCH_N(c, launch_col_loss, c, a, items, false); }

// Row pass of a refinement step in the geometry of row variant RV.
template <int N, int RV> int launch_row_adam_v(spr_ctx* c, const AdamRowArgs& a, bool masked, int items) {
    using G = typename RowVariant<N, RV>::G;
    const dim3 grid((unsigned)((size_t)items * N / G::RW)), block(G::ROW_THREADS);
    if (masked) hipLaunchKernelGGL((k_row_adam<N, RV, true>), grid, block, 0, c->stream, a);
    else hipLaunchKernelGGL((k_row_adam<N, RV, false>), grid, block, 0, c->stream, a);
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}
// In the default row geometry: the tuned row variants measured up to 2.5 %
// slower for this pass (1024^2, 32 fields).
template <int N> int launch_row_adam(spr_ctx* c, const AdamRowArgs& a, bool masked, int items) {
    return launch_row_adam_v<N, 0>(c, a, masked, items);
}
int row_adam(spr_ctx* c, const AdamRowArgs& a, bool masked, int items) {
    SPR_DISPATCH_N(c, launch_row_adam, c, a, masked, items);
}

// `steps` Adam steps of the Fourier loss on the resident densities, two passes
// per step: the column loss pass and the fused adjoint-row / Adam / row-FFT pass.
int queue_adam_steps(spr_ctx* c, int steps, float lr, const float* lrp, int kind, float eps) {
    const int P = c->P;
    const size_t n = (size_t)c->n * c->n * P;
    SPR_TRY2(ensure(c, c->adam_m, n * sizeof(float2)));
    SPR_TRY2(ensure(c, c->adam_v, n * sizeof(float2)));
    SPR_TRY2(fill_u32(c, (unsigned*)c->adam_m.p, 0u, 2 * n));
    SPR_TRY2(fill_u32(c, (unsigned*)c->adam_v.p, 0u, 2 * n));
    FieldBatch fwd{c->work, nullptr, c->tw, 1.f, c->rho};
    SPR_TRY2(fft_rows(c, +1, fwd, P));
    const float b1 = 0.9f, b2 = 0.999f;
    float p1 = 1.f, p2 = 1.f;
    for (int s = 1; s <= steps; ++s) {
        LossArgs la{c->work, c->amp, c->amp_of, c->tw, nullptr, eps, kind, nullptr};
        SPR_TRY2(col_loss_step(c, la, P));
        p1 *= b1;
        p2 *= b2;
        AdamRowArgs ar{c->rho, (float2*)c->adam_m.p, (float2*)c->adam_v.p, c->work, c->supp, lrp, c->tw,
                       lr, b1, b2, 1e-8f, 1.f - p1, 1.f - p2};
        SPR_TRY2(row_adam(c, ar, lrp != nullptr, P));
    }
    return SPR_OK;
}

// The launches of spr_refine_support on the handle's stream, without waiting
// for them (the GRAD operation of spr_run and spr_batch_run).
int queue_refine_support(spr_ctx* c, int steps, float lr_rel, int kind, float eps) {
    const int P = c->P;
    if (P <= 0 || steps == 0) return SPR_OK;
    const int np = c->n * c->n;
    SPR_TRY(ensure(c, c->xphase, (size_t)P * sizeof(float)));
    float* lrp = (float*)c->xphase.p;
    // Start from the projection onto the supports: after an ER/HIO/RAAR block
    // the densities are IFT(P_M(FT rho)), whose Fourier magnitudes already
    // minimise the loss, so descending from them without the support
    // constraint has nothing to gain.
    const size_t total = (size_t)P * np;
    hipLaunchKernelGGL(k_zero_outside, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->stream, c->rho,
                       (const uint8_t*)c->supp, total);
    SPR_HIP(c, hipGetLastError());
    // Every supported n is a multiple of 8, so SUPPORT_CHUNKS = 64 divides n^2.
    SPR_TRY(ensure(c, c->lrpart, (size_t)P * SUPPORT_CHUNKS * 2 * sizeof(double)));
    hipLaunchKernelGGL(k_support_partials, dim3((unsigned)(P * SUPPORT_CHUNKS)), dim3(256), 0, c->stream,
                       (const float2*)c->rho, (const uint8_t*)c->supp, np, (double*)c->lrpart.p);
    SPR_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_support_lr, dim3((unsigned)((P + 63) / 64)), dim3(64), 0, c->stream, (const double*)c->lrpart.p,
                       P, lr_rel, lrp);
    SPR_HIP(c, hipGetLastError());
    return queue_adam_steps(c, steps, 0.f, lrp, kind, eps);
}

// Runs the handle's launches on another stream for the lifetime of the guard.
struct StreamSwap {
    spr_ctx* c;
    hipStream_t saved;
    StreamSwap(spr_ctx* c_, hipStream_t s) : c(c_), saved(c_->stream) {
        if (s) c->stream = s;
    }
    ~StreamSwap() { c->stream = saved; }
};

// The passes of the loss and its gradient for P fields at device addresses:
// the row FFT of rho into `field`, the column loss pass on it (one loss
// partial per workgroup into c->partial) and, unless grad is NULL, the
// adjoint row pass grad <- 2 FT^H[f'(|F|^2) F]. grad may be `field` itself
// (the adjoint pass then runs in place, row by row).
int queue_loss_passes(spr_ctx* c, int P, const float2* rho, float2* field, float2* grad, int kind, float eps) {
    FieldBatch fwd{field, nullptr, c->tw, 1.f, rho};
    SPR_TRY2(fft_rows(c, +1, fwd, P));
    LossArgs la{field, c->amp, c->amp_of, c->tw, c->partial, eps, kind, nullptr};
    SPR_TRY2(col_loss(c, la, P));
    if (!grad) return SPR_OK;
    FieldBatch adj{grad, nullptr, c->tw, 2.f, grad == field ? nullptr : field};
    return fft_rows(c, -1, adj, P);
}

// grad <- 2 FT^H[f'(|F|^2) F], F = FT(rho) (grad NULL: the losses only);
// per-field losses into sums[p*SUMS_PER_ITEM + 7].
int queue_loss_grad(spr_ctx* c, int P, const float2* rho, float2* grad, int kind, float eps) {
    const int n = c->n;
    SPR_TRY2(queue_loss_passes(c, P, rho, c->work, grad, kind, eps));
    return reduce(c, n / cols_per_wg(n), 1, 7, P);
}

// Loss of every resident density after a refinement (host, P doubles).
int final_loss(spr_ctx* c, int kind, float eps, double* loss) {
    const int P = c->P;
    SPR_TRY2(queue_loss_grad(c, P, c->rho, nullptr, kind, eps));
    std::vector<double> sm((size_t)P * SUMS_PER_ITEM);
    SPR_HIP(c, hipMemcpyAsync(sm.data(), c->sums, sm.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    for (int p = 0; p < P; ++p) loss[p] = sm[(size_t)p * SUMS_PER_ITEM + 7];
    return SPR_OK;
}

// ---------------------------------------------------------------------------
// Variant choice and the autotuner
// ---------------------------------------------------------------------------
struct VariantPair {
    int col = 0, row = 0;
};
std::mutex g_variant_mu;
std::map<int, VariantPair> g_variant;   // sizes whose default differs from the compiled table

int variant_counts(int n, int* nc, int* nr) {
#define SPR_VAR_CASE(N_)                      \
    if (n == N_) {                            \
        *nc = col_variants<N_>();             \
        *nr = row_variants<N_>();             \
        return SPR_OK;                        \
    }
    SPR_SIZES(SPR_VAR_CASE)
#undef SPR_VAR_CASE
    return SPR_ERR_ARG;
}

// Working set per chunk of an iteration block for new handles (SPR_CHUNK_MB
// overrides; 0 = no chunking).
size_t default_chunk_bytes() {
    if (const char* e = std::getenv("SPR_CHUNK_MB")) return (size_t)std::max(0.0, std::atof(e)) * 1024 * 1024;
    return (size_t)SPR_CHUNK_MB_DEFAULT * 1024 * 1024;
}

// Variants a new handle of size n starts with: the last spr_autotune or
// spr_set_default_variants of this process for n, else the compiled table.
VariantPair default_variants(int n) {
    std::lock_guard<std::mutex> lk(g_variant_mu);
    auto it = g_variant.find(n);
    if (it != g_variant.end()) return it->second;
    VariantPair v;
    for (const TunedVariants& t : kTuned)
        if (t.n == n) {
            v.col = t.col;
            v.row = t.row;
        }
    int nc = 1, nr = 1;
    if (variant_counts(n, &nc, &nr) != SPR_OK || v.col >= nc || v.row >= nr) v = VariantPair{};
    return v;
}

// Mean duration (ms) of `reps` launches of one pass in the handle's current variant.
// Average time of one iteration (column pass, then RAAR row pass) with the
// handle's current variants. Timing whole iterations, not one pass repeated,
// keeps the caches as the iteration leaves them: a pass that reads a table from
// global memory finds it evicted by the other pass, as it does in a block.
// The densities are not changed: a row variant that stores rho_n updates a
// fresh copy of them in work2; one that recomputes it leaves them alone, and
// its column passes project into work2.
int time_iteration(spr_ctx* c, const RowArgs& ra0, int reps, hipEvent_t e0, hipEvent_t e1, float* ms) {
    RowArgs ra = ra0;
    float2* out = nullptr;
    if (row_recomputes(c->n, c->rowvar)) {
        ra.proj = c->work2;
        out = c->work2;
    } else {
        const int np = c->n * c->n;
        hipLaunchKernelGGL(k_copy_fields, dim3((np + 255) / 256, c->P), dim3(256), 0, c->stream, (const float2*)c->rho,
                           (const int*)nullptr, c->work2, (const int*)nullptr, np);
        SPR_HIP(c, hipGetLastError());
        ra.rho = c->work2;
    }
    auto once = [&]() {
        SPR_TRY(col_n(c, COL_PROJ, c->work, c->P, c->stream, 0, out));
        return row_n(c, ROW_RAAR, ra, c->P, c->stream);
    };
    SPR_TRY(once());   // warm-up (code load, caches)
    SPR_HIP(c, hipEventRecord(e0, c->stream));
    for (int i = 0; i < reps; ++i) SPR_TRY(once());
    SPR_HIP(c, hipEventRecord(e1, c->stream));
    SPR_HIP(c, hipEventSynchronize(e1));
    float t = 0.f;
    SPR_HIP(c, hipEventElapsedTime(&t, e0, e1));
    *ms = t / (float)reps;
    return SPR_OK;
}

// Times every column and row variant of the handle's size on its resident
// population and keeps the fastest: first the column variants, each in whole
// iterations with the handle's row variant, then the row variants with the
// chosen column variant. The densities and supports are not changed (see
// time_iteration); the passes transform `work` and `work2`, which every
// operation rewrites first.
int autotune(spr_ctx* c, int reps, float* col_ms, float* row_ms) {
    if (c->P <= 0) return fail(c, SPR_ERR_ARG, "autotune needs resident individuals");
    if (reps < 1) return fail(c, SPR_ERR_ARG, "reps must be >= 1");
    int nc = 1, nr = 1;
    SPR_TRY(variant_counts(c->n, &nc, &nr));
    SPR_HIP(c, hipSetDevice(c->device));
    RowArgs ra = row_args(c);
    ra.beta = 0.87f;
    ra.pc = cut_raar(0.f);
    const VariantPair keep{c->colvar, c->rowvar};
    c->rowvar = 0;
    SPR_TRY(row_n(c, ROW_FWD, ra, c->P, c->stream));   // realistic spectra in work
    SPR_TRY(pack_support(c, c->supp, c->sbits, c->P, c->stream));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SPR_HIP(c, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(c, SPR_ERR_HIP, "hipEventCreate failed");
    }
    // The lowest-numbered variant within TIE of the fastest: near-ties are timing
    // noise, and preferring the lower index keeps the choice stable between runs.
    constexpr float TIE = 1.015f;
    auto pick = [](const float* t, int cnt) {
        float m = t[0];
        for (int v = 1; v < cnt; ++v) m = std::min(m, t[v]);
        for (int v = 0; v < cnt; ++v)
            if (t[v] <= m * TIE) return v;
        return 0;
    };
    float tc[MAX_VARIANTS] = {}, tr[MAX_VARIANTS] = {};
    int rc = SPR_OK;
    VariantPair best;
    c->rowvar = keep.row;
    for (int v = 0; v < nc && rc == SPR_OK; ++v) {
        c->colvar = v;
        rc = time_iteration(c, ra, reps, e0, e1, &tc[v]);
        if (col_ms) col_ms[v] = tc[v];
    }
    if (rc == SPR_OK) best.col = pick(tc, nc);
    c->colvar = best.col;
    for (int v = 0; v < nr && rc == SPR_OK; ++v) {
        c->rowvar = v;
        rc = time_iteration(c, ra, reps, e0, e1, &tr[v]);
        if (row_ms) row_ms[v] = tr[v];
    }
    if (rc == SPR_OK) best.row = pick(tr, nr);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != SPR_OK) {
        c->colvar = keep.col;
        c->rowvar = keep.row;
        return rc;
    }
    c->colvar = best.col;
    c->rowvar = best.row;
    std::lock_guard<std::mutex> lk(g_variant_mu);
    g_variant[c->n] = best;
    return SPR_OK;
}

// Pinned host staging, grown on demand.
int ensure_pinned(spr_ctx* c, size_t bytes) {
    spr_ctx::Buf& b = c->pinned;
    if (b.bytes >= bytes) return SPR_OK;
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    if (b.p) SPR_HIP(c, hipHostFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
    SPR_HIP(c, hipHostMalloc(&b.p, bytes, hipHostMallocDefault));
    b.bytes = bytes;
    return SPR_OK;
}

// f(i) for i in [0, count) on up to 16 host threads (host-side copies of whole fields).
template <class F> void host_parallel(int count, F f) {
    const int nt = std::min(count, std::max(1, std::min(16, (int)std::thread::hardware_concurrency())));
    if (nt <= 1) {
        for (int i = 0; i < count; ++i) f(i);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&, t] {
            for (int i = t; i < count; i += nt) f(i);
        });
    for (std::thread& th : pool) th.join();
}

int check_indices(spr_ctx* c, const int32_t* idx, int count, int limit, const char* what) {
    for (int i = 0; i < count; ++i)
        if (idx[i] < 0 || idx[i] >= limit) return fail(c, SPR_ERR_ARG, std::string(what) + ": index out of range");
    return SPR_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Fixed device scratch of spr_loss_grad_noalloc, the entry point that
// allocates nothing. Layout: the twiddle roots exp(-2 pi i m/N) of every
// supported size (NOALLOC_TW_FLOAT2 float2, sizes in SPR_SIZES order), then
// NOALLOC_PARTIALS doubles of per-workgroup loss partials. Each GPU holds its
// own copy for the life of the process. The array's name is a host-side
// stand-in in host code: host code reaches the memory only through the address
// hipGetSymbolAddress returns for the current device (noalloc_setup).
// ---------------------------------------------------------------------------
constexpr int NOALLOC_TW_FLOAT2 = 64 + 128 + 256 + 384 + 512 + 768 + 1024 + 2048;
constexpr size_t NOALLOC_TW_BYTES = ((NOALLOC_TW_FLOAT2 * sizeof(float2) + 255) / 256) * 256;
constexpr size_t NOALLOC_PARTIALS = 131072;
constexpr size_t NOALLOC_BYTES = NOALLOC_TW_BYTES + NOALLOC_PARTIALS * sizeof(double);
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {

// Offset (float2) of size n's roots in the scratch twiddle region; -1 if n is
// not supported.
int noalloc_tw_offset(int n) {
    int off = 0;
#define SPR_TWOFF_CASE(N_)         \
    if (n == N_) return off;       \
    off += N_;
    SPR_SIZES(SPR_TWOFF_CASE)
#undef SPR_TWOFF_CASE
    return -1;
}

std::mutex g_noalloc_mu;
unsigned char* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};
float2 g_noalloc_tw_host[NOALLOC_TW_FLOAT2];   // staging of the roots (set-up only)

// Once per device, with that device current: the device address of `scratch`
// and the upload of every size's twiddle roots (computed in double and rounded
// to float, as for a handle). Synchronous; later calls only read the address.
int noalloc_setup(int dev, unsigned char** out) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(nullptr, SPR_ERR_ARG, "device index out of range");
    if (!g_noalloc_ptr[dev]) {
        unsigned char* p = nullptr;
        hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(nullptr, SPR_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        int off = 0;
#define SPR_TWFILL_CASE(N_)                                                                      \
    for (int k = 0; k < N_; ++k) {                                                               \
        const double ang = -2.0 * M_PI * (double)k / (double)N_;                                 \
        g_noalloc_tw_host[off + k] = make_float2((float)std::cos(ang), (float)std::sin(ang));    \
    }                                                                                            \
    off += N_;
        SPR_SIZES(SPR_TWFILL_CASE)
#undef SPR_TWFILL_CASE
        e = hipMemcpy(p, g_noalloc_tw_host, sizeof(g_noalloc_tw_host), hipMemcpyHostToDevice);
        if (e != hipSuccess)
            return fail(nullptr, SPR_ERR_HIP, std::string("upload of the twiddle roots failed: ") + hipGetErrorString(e));
        g_noalloc_ptr[dev] = p;
    }
    *out = g_noalloc_ptr[dev];
    return SPR_OK;
}

}  // namespace

extern "C" {

int spr_supported_size(int n) {
#define SPR_SUPPORTED_CASE(N_) if (n == N_) return 1;
    SPR_SIZES(SPR_SUPPORTED_CASE)
#undef SPR_SUPPORTED_CASE
    return 0;
}

// A handle for `npat` patterns of size n: amps[k] is pattern k's corner-origin
// amplitude and amp_of (capacity entries) gives every slot's pattern (NULL for a
// single pattern).
static int create_context(spr_ctx** out, int n, int capacity, int npat, const float* const* amps, const int* amp_of,
                          int device) {
    if (!out) return fail(nullptr, SPR_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!spr_supported_size(n)) return fail(nullptr, SPR_ERR_ARG, "unsupported grid size " + std::to_string(n));
    if (capacity < 1) return fail(nullptr, SPR_ERR_ARG, "capacity must be >= 1");
    if (npat < 1 || !amps) return fail(nullptr, SPR_ERR_ARG, "amp is NULL");
    for (int k = 0; k < npat; ++k)
        if (!amps[k]) return fail(nullptr, SPR_ERR_ARG, "amp is NULL");
    if (npat > 1) {
        if (!amp_of) return fail(nullptr, SPR_ERR_ARG, "pattern map is NULL");
        for (int i = 0; i < capacity; ++i)
            if (amp_of[i] < 0 || amp_of[i] >= npat) return fail(nullptr, SPR_ERR_ARG, "pattern map out of range");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(nullptr, SPR_ERR_HIP, "no HIP device");
    if (device < 0 || device >= ndev) return fail(nullptr, SPR_ERR_ARG, "bad device id");
    spr_ctx* c = new spr_ctx();
    c->n = n;
    c->cap = capacity;
    c->device = device;
    c->npat = npat;
    c->chunk_bytes = default_chunk_bytes();
    {
        const VariantPair v = default_variants(n);
        c->colvar = v.col;
        c->rowvar = v.row;
    }
    const size_t np = (size_t)n * n;
    auto cleanup_fail = [&](const char* what, hipError_t e) {
        g_create_error = std::string(what) + ": " + hipGetErrorString(e);
        spr_destroy(c);
        return SPR_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return cleanup_fail("hipSetDevice", e);
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return cleanup_fail("hipStreamCreate", e);
    if (const char* l = std::getenv("SPR_ITERATE_LANES")) c->lanes = std::atoi(l) == 2 ? 2 : 1;
    if (c->lanes == 2) {
        if ((e = hipStreamCreateWithFlags(&c->lane2, hipStreamNonBlocking)) != hipSuccess)
            return cleanup_fail("hipStreamCreate", e);
        if ((e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming)) != hipSuccess)
            return cleanup_fail("hipEventCreate", e);
    }
    if ((e = hipMalloc(&c->rho, np * capacity * sizeof(float2))) != hipSuccess) return cleanup_fail("hipMalloc rho", e);
    if ((e = hipMalloc(&c->work, np * capacity * sizeof(float2))) != hipSuccess)
        return cleanup_fail("hipMalloc work", e);
    if ((e = hipMalloc(&c->supp, np * capacity)) != hipSuccess) return cleanup_fail("hipMalloc supp", e);
    if ((e = hipMalloc(&c->sbits, np / 8 * capacity)) != hipSuccess) return cleanup_fail("hipMalloc sbits", e);
    if ((e = hipMalloc(&c->amp, np * npat * sizeof(float))) != hipSuccess) return cleanup_fail("hipMalloc amp", e);
    if ((e = hipMalloc(&c->amp_t, np * npat * sizeof(float))) != hipSuccess) return cleanup_fail("hipMalloc amp_t", e);
    if ((e = hipMalloc(&c->tw, n * sizeof(float2))) != hipSuccess) return cleanup_fail("hipMalloc tw", e);
    if ((e = hipMalloc(&c->work2, np * capacity * sizeof(float2))) != hipSuccess)
        return cleanup_fail("hipMalloc work2", e);
    {
        const int per = std::max(n / rows_per_wg(n), n / cols_per_wg(n));
        c->partial_cap = (size_t)capacity * per * 2;
        if ((e = hipMalloc(&c->partial, c->partial_cap * sizeof(double))) != hipSuccess)
            return cleanup_fail("hipMalloc partial", e);
    }
    if ((e = hipMalloc(&c->sums, (size_t)capacity * SUMS_PER_ITEM * sizeof(double))) != hipSuccess)
        return cleanup_fail("hipMalloc sums", e);
    if ((e = hipMalloc(&c->maxbits, (size_t)capacity * 2 * sizeof(unsigned))) != hipSuccess)
        return cleanup_fail("hipMalloc maxbits", e);
    if ((e = hipMalloc(&c->keys, (size_t)capacity * 2 * sizeof(unsigned long long))) != hipSuccess)
        return cleanup_fail("hipMalloc keys", e);
    if ((e = hipMalloc(&c->radix, (size_t)capacity * 3 * sizeof(unsigned))) != hipSuccess)
        return cleanup_fail("hipMalloc radix", e);
    if ((e = hipMalloc(&c->hist, (size_t)capacity * RADIX_BINS * sizeof(unsigned))) != hipSuccess)
        return cleanup_fail("hipMalloc hist", e);
    if (npat > 1) {
        if ((e = hipMalloc(&c->amp_of, (size_t)capacity * sizeof(int))) != hipSuccess)
            return cleanup_fail("hipMalloc pattern map", e);
        if ((e = hipMemcpy(c->amp_of, amp_of, (size_t)capacity * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
            return cleanup_fail("upload pattern map", e);
    }
    std::vector<float2> tw(n);
    for (int k = 0; k < n; ++k) {
        const double ang = -2.0 * M_PI * (double)k / (double)n;
        tw[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    if ((e = hipMemcpy(c->tw, tw.data(), n * sizeof(float2), hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup_fail("upload twiddles", e);
    {
        int rc = SPR_ERR_ARG;
        switch (n) {
#define SPR_TWTAB_CASE(N_) \
    case N_: rc = build_twtabs<N_>(c); break;
            SPR_SIZES(SPR_TWTAB_CASE)
#undef SPR_TWTAB_CASE
        }
        if (rc != SPR_OK || (e = hipStreamSynchronize(c->stream)) != hipSuccess) {
            g_create_error = "twiddle tables: " + c->err;
            spr_destroy(c);
            return SPR_ERR_HIP;
        }
    }
    for (int k = 0; k < npat; ++k) {
        if ((e = hipMemcpy(c->amp + k * np, amps[k], np * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess)
            return cleanup_fail("upload amplitude", e);
        const std::vector<float> at = transpose_amp(amps[k], n);
        if ((e = hipMemcpy(c->amp_t + k * np, at.data(), np * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess)
            return cleanup_fail("upload amplitude", e);
    }
    if ((e = hipMemset(c->supp, 0, np * capacity)) != hipSuccess) return cleanup_fail("clear supports", e);
    *out = c;
    return SPR_OK;
}

int spr_create(spr_ctx** out, int n, int capacity, const float* amp, int device) {
    const float* amps[1] = {amp};
    return create_context(out, n, capacity, 1, amps, nullptr, device);
}

int spr_destroy(spr_ctx* c) {
    if (!c) return SPR_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    (void)hipFree(c->rho);
    (void)hipFree(c->work);
    (void)hipFree(c->supp);
    (void)hipFree(c->sbits);
    (void)hipFree(c->amp);
    (void)hipFree(c->amp_t);
    (void)hipFree(c->amp_of);
    for (float2* t : c->twtab) (void)hipFree(t);
    (void)hipFree(c->tw);
    (void)hipFree(c->work2);
    (void)hipFree(c->partial);
    (void)hipFree(c->sums);
    (void)hipFree(c->maxbits);
    (void)hipFree(c->keys);
    (void)hipFree(c->radix);
    (void)hipFree(c->hist);
    for (spr_ctx::Buf* b : {&c->spec, &c->xf, &c->xs, &c->supp2, &c->fieldbuf, &c->ampavg, &c->xidx, &c->shiftp,
                            &c->xchild, &c->xmaps, &c->xphase, &c->adam_m, &c->adam_v, &c->gidx, &c->lrpart})
        (void)hipFree(b->p);
    if (c->pinned.p) (void)hipHostFree(c->pinned.p);
    if (c->lane2) (void)hipStreamSynchronize(c->lane2);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->lane2) (void)hipStreamDestroy(c->lane2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SPR_OK;
}

const char* spr_last_error(const spr_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int spr_size(const spr_ctx* c, int* n, int* capacity) {
    if (!c) return SPR_ERR_ARG;
    if (n) *n = c->n;
    if (capacity) *capacity = c->cap;
    return SPR_OK;
}

int spr_set_pattern(spr_ctx* c, const float* amp) {
    if (!c || !amp) return fail(c, SPR_ERR_ARG, "NULL argument");
    if (c->npat != 1) return fail(c, SPR_ERR_ARG, "the handle holds several patterns");
    SPR_HIP(c, hipSetDevice(c->device));
    const std::vector<float> at = transpose_amp(amp, c->n);
    SPR_HIP(c, hipMemcpyAsync(c->amp, amp, (size_t)c->n * c->n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    SPR_HIP(c, hipMemcpyAsync(c->amp_t, at.data(), (size_t)c->n * c->n * sizeof(float), hipMemcpyHostToDevice,
                              c->stream));
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    return SPR_OK;
}

int spr_get_pattern(spr_ctx* c, float* amp) {
    if (!c || !amp) return fail(c, SPR_ERR_ARG, "NULL argument");
    if (c->npat != 1) return fail(c, SPR_ERR_ARG, "the handle holds several patterns");
    SPR_HIP(c, hipSetDevice(c->device));
    SPR_HIP(c, hipMemcpyAsync(amp, c->amp, (size_t)c->n * c->n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    return SPR_OK;
}

int spr_set_population(spr_ctx* c, int P, const float* rho, const int32_t* support) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (P < 0 || P > c->cap) return fail(c, SPR_ERR_ARG, "population size exceeds capacity");
    if (P > 0 && !rho) return fail(c, SPR_ERR_ARG, "rho is NULL");
    const size_t np = (size_t)c->n * c->n;
    SPR_HIP(c, hipSetDevice(c->device));
    if (P > 0) {
        SPR_HIP(c, hipMemcpyAsync(c->rho, rho, np * P * sizeof(float2), hipMemcpyHostToDevice, c->stream));
        if (support) {
            std::vector<uint8_t> s(np * P);
            for (size_t i = 0; i < s.size(); ++i) s[i] = support[i] != 0 ? 1 : 0;
            SPR_HIP(c, hipMemcpyAsync(c->supp, s.data(), s.size(), hipMemcpyHostToDevice, c->stream));
            SPR_HIP(c, hipStreamSynchronize(c->stream));
        }
    }
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    c->P = P;
    return SPR_OK;
}

int spr_get_population(spr_ctx* c, int P, float* rho, int32_t* support) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (P < 0 || P > c->P) return fail(c, SPR_ERR_ARG, "more individuals requested than resident");
    const size_t np = (size_t)c->n * c->n;
    SPR_HIP(c, hipSetDevice(c->device));
    if (P == 0) return SPR_OK;
    if (rho) SPR_HIP(c, hipMemcpyAsync(rho, c->rho, np * P * sizeof(float2), hipMemcpyDeviceToHost, c->stream));
    if (support) {
        std::vector<uint8_t> s(np * P);
        SPR_HIP(c, hipMemcpyAsync(s.data(), c->supp, s.size(), hipMemcpyDeviceToHost, c->stream));
        SPR_HIP(c, hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < s.size(); ++i) support[i] = s[i];
    }
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    return SPR_OK;
}

int spr_set_active(spr_ctx* c, int P) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (P < 0 || P > c->cap) return fail(c, SPR_ERR_ARG, "population size exceeds capacity");
    c->P = P;
    return SPR_OK;
}

int spr_gather(spr_ctx* c, int count, const int* src) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (count < 0 || count > c->cap) return fail(c, SPR_ERR_ARG, "count exceeds capacity");
    if (count > 0 && !src) return fail(c, SPR_ERR_ARG, "src is NULL");
    std::vector<int> from, to;
    for (int i = 0; i < count; ++i) {
        if (src[i] < -1 || src[i] >= c->cap) return fail(c, SPR_ERR_ARG, "slot index out of range");
        if (src[i] >= 0 && src[i] != i) {
            from.push_back(src[i]);
            to.push_back(i);
        }
    }
    if (from.empty()) return SPR_OK;
    const int m = (int)from.size();
    const int np = c->n * c->n;
    SPR_HIP(c, hipSetDevice(c->device));
    SPR_TRY(ensure(c, c->gidx, 2 * m * sizeof(int)));
    SPR_TRY(ensure(c, c->supp2, (size_t)m * np));
    std::vector<int> idx(from);
    idx.insert(idx.end(), to.begin(), to.end());
    int* d_from = (int*)c->gidx.p;
    int* d_to = d_from + m;
    SPR_HIP(c, hipMemcpyAsync(d_from, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    // Every source is read into the staging fields before any slot is written.
    const dim3 gf((np + 255) / 256, m), gs((np / 16 + 255) / 256, m);
    hipLaunchKernelGGL(k_copy_fields, gf, dim3(256), 0, c->stream, (const float2*)c->rho, d_from, c->work,
                       (const int*)nullptr, np);
    hipLaunchKernelGGL(k_copy_supports, gs, dim3(256), 0, c->stream, (const uint4*)c->supp, d_from,
                       (uint4*)c->supp2.p, (const int*)nullptr, np / 16);
    hipLaunchKernelGGL(k_copy_fields, gf, dim3(256), 0, c->stream, (const float2*)c->work, (const int*)nullptr,
                       c->rho, d_to, np);
    hipLaunchKernelGGL(k_copy_supports, gs, dim3(256), 0, c->stream, (const uint4*)c->supp2.p, (const int*)nullptr,
                       (uint4*)c->supp, d_to, np / 16);
    SPR_HIP(c, hipGetLastError());
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    return SPR_OK;
}

int spr_put_individuals(spr_ctx* c, int count, const int* slots, const float* const* rho,
                        const int32_t* const* support) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (count < 0 || count > c->cap) return fail(c, SPR_ERR_ARG, "count exceeds capacity");
    if (count == 0) return SPR_OK;
    if (!slots) return fail(c, SPR_ERR_ARG, "slots is NULL");
    for (int i = 0; i < count; ++i)
        if (slots[i] < 0 || slots[i] >= c->cap) return fail(c, SPR_ERR_ARG, "slot index out of range");
    if (support)
        for (int i = 0; i < count; ++i)
            if (!support[i]) return fail(c, SPR_ERR_ARG, "support pointer is NULL");
    const size_t np = (size_t)c->n * c->n;
    SPR_HIP(c, hipSetDevice(c->device));
    // Host arrays -> pinned staging (densities as they are, supports narrowed to
    // uint8 0/1) on host threads, then one DMA copy per individual and part.
    const size_t fb = np * sizeof(float2);
    SPR_TRY(ensure_pinned(c, (size_t)count * (fb + np)));
    char* stage = (char*)c->pinned.p;
    host_parallel(count, [&](int i) {
        if (rho && rho[i]) std::memcpy(stage + (size_t)i * fb, rho[i], fb);
        if (support) {
            uint8_t* d = (uint8_t*)(stage + (size_t)count * fb + (size_t)i * np);
            const int32_t* src = support[i];
            for (size_t k = 0; k < np; ++k) d[k] = src[k] != 0 ? 1 : 0;
        }
    });
    for (int i = 0; i < count; ++i) {
        if (rho && rho[i])
            SPR_HIP(c, hipMemcpyAsync(c->rho + (size_t)slots[i] * np, stage + (size_t)i * fb, fb,
                                      hipMemcpyHostToDevice, c->stream));
        if (support)
            SPR_HIP(c, hipMemcpyAsync(c->supp + (size_t)slots[i] * np, stage + (size_t)count * fb + (size_t)i * np,
                                      np, hipMemcpyHostToDevice, c->stream));
    }
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    return SPR_OK;
}

int spr_fetch_individuals(spr_ctx* c, int count, const int* slots, float* const* rho, int32_t* const* support) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (count < 0 || count > c->cap) return fail(c, SPR_ERR_ARG, "count exceeds capacity");
    if (count == 0) return SPR_OK;
    if (!slots) return fail(c, SPR_ERR_ARG, "slots is NULL");
    for (int i = 0; i < count; ++i)
        if (slots[i] < 0 || slots[i] >= c->cap) return fail(c, SPR_ERR_ARG, "slot index out of range");
    const size_t np = (size_t)c->n * c->n;
    SPR_HIP(c, hipSetDevice(c->device));
    // One DMA copy per individual and part into pinned staging (supports as
    // uint8), then host threads copy the densities out and widen the supports.
    const size_t fb = np * sizeof(float2);
    SPR_TRY(ensure_pinned(c, (size_t)count * (fb + np)));
    char* stage = (char*)c->pinned.p;
    for (int i = 0; i < count; ++i) {
        if (rho && rho[i])
            SPR_HIP(c, hipMemcpyAsync(stage + (size_t)i * fb, c->rho + (size_t)slots[i] * np, fb,
                                      hipMemcpyDeviceToHost, c->stream));
        if (support && support[i])
            SPR_HIP(c, hipMemcpyAsync(stage + (size_t)count * fb + (size_t)i * np, c->supp + (size_t)slots[i] * np,
                                      np, hipMemcpyDeviceToHost, c->stream));
    }
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    host_parallel(count, [&](int i) {
        if (rho && rho[i]) std::memcpy(rho[i], stage + (size_t)i * fb, fb);
        if (support && support[i]) {
            const uint8_t* src = (const uint8_t*)(stage + (size_t)count * fb + (size_t)i * np);
            int32_t* d = support[i];
            for (size_t k = 0; k < np; ++k) d[k] = src[k];
        }
    });
    return SPR_OK;
}

int spr_ft(spr_ctx* c) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (c->P <= 0) return SPR_OK;
    SPR_HIP(c, hipSetDevice(c->device));
    RowArgs a = row_args(c);
    SPR_TRY(row(c, ROW_FWD_INPLACE, a));
    SPR_TRY(col(c, COL_FWD, c->rho));
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    return SPR_OK;
}

int spr_ift(spr_ctx* c) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (c->P <= 0) return SPR_OK;
    SPR_HIP(c, hipSetDevice(c->device));
    RowArgs a = row_args(c);
    SPR_TRY(col(c, COL_INV, c->rho));
    SPR_TRY(row(c, ROW_INV_INPLACE, a));
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    return SPR_OK;
}

int spr_iterate_async(spr_ctx* c, int alg, int niter, float beta, float reality) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    return queue_iterate(c, alg, niter, beta, reality);
}

int spr_iterate(spr_ctx* c, int alg, int niter, float beta, float reality) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    SPR_TRY(queue_iterate(c, alg, niter, beta, reality));
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    return harvest_timing(c);
}

int spr_sw(spr_ctx* c, float sigma, float threshold, float reality) {
    (void)reality;
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    SPR_HIP(c, hipSetDevice(c->device));
    SPR_TRY(queue_sw(c, sigma, threshold));
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    return harvest_timing(c);
}

int spr_swa(spr_ctx* c, float sigma, int spixels, float reality) {
    (void)reality;
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    SPR_HIP(c, hipSetDevice(c->device));
    SPR_TRY(queue_swa(c, sigma, spixels));
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    return harvest_timing(c);
}

int spr_evaluate(spr_ctx* c, float reality, float* error_fourier, float* error_real, float* complexsum,
                 float* modulesum, float* symmetry) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (c->P <= 0) return SPR_OK;
    SPR_HIP(c, hipSetDevice(c->device));
    SPR_TRY(queue_error(c, reality));
    std::vector<double> sums((size_t)c->P * SUMS_PER_ITEM);
    std::vector<unsigned> mx(c->P);
    SPR_HIP(c, hipMemcpyAsync(sums.data(), c->sums, sums.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SPR_HIP(c, hipMemcpyAsync(mx.data(), c->maxbits, mx.size() * sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    std::vector<unsigned long long> keys(2 * (size_t)c->cap);
    if (symmetry) {
        SPR_TRY(queue_symmetry(c, 0.f, 1.f));
        SPR_HIP(c, hipMemcpyAsync(keys.data(), c->keys, keys.size() * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, c->stream));
    }
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    for (int p = 0; p < c->P; ++p) {
        const double* s = &sums[(size_t)p * SUMS_PER_ITEM];
        float fm;
        std::memcpy(&fm, &mx[p], 4);
        if (error_fourier) error_fourier[p] = (float)std::sqrt(s[1] / s[2]);
        if (error_real) error_real[p] = (float)std::sqrt(s[3] / s[4]);
        if (complexsum) complexsum[p] = fm;
        if (modulesum) modulesum[p] = (float)s[0];
        if (symmetry) {
            const float m1 = key_value(keys[p]), m2 = key_value(keys[c->cap + p]);
            symmetry[p] = std::sqrt(m2 / m1);
        }
    }
    return harvest_timing(c);
}

int spr_symmetry(spr_ctx* c, float symmin, float symmax, float* symmetry) {
    if (!c || !symmetry) return fail(c, SPR_ERR_ARG, "NULL argument");
    if (c->P <= 0) return SPR_OK;
    SPR_HIP(c, hipSetDevice(c->device));
    SPR_TRY(queue_symmetry(c, symmin, symmax));
    std::vector<unsigned long long> keys(2 * (size_t)c->cap);
    SPR_HIP(c, hipMemcpyAsync(keys.data(), c->keys, keys.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              c->stream));
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    for (int p = 0; p < c->P; ++p) {
        const float m1 = key_value(keys[p]), m2 = key_value(keys[c->cap + p]);
        symmetry[p] = std::sqrt(m2 / m1);
    }
    return harvest_timing(c);
}

// Every operation of a list is checked before any is queued, so that a bad
// entry leaves the population untouched. NULL when the list is valid.
static const char* op_list_error(const spr_op* ops, int nops) {
    for (int i = 0; i < nops; ++i) {
        const spr_op& o = ops[i];
        if (o.kind < SPR_OP_ER || o.kind > SPR_OP_GRAD) return "unknown op kind";
        if (o.kind <= SPR_OP_RAAR && o.iterations < 0) return "negative iteration count";
        if (o.kind == SPR_OP_GRAD) {
            if (o.iterations < 0) return "negative GRAD step count";
            if (o.pixels != LOSS_POISSON && o.pixels != LOSS_AMPLITUDE) return "unknown GRAD loss kind";
        }
    }
    return nullptr;
}

int spr_run(spr_ctx* c, const spr_op* ops, int nops, int nrep, float reality) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (nops < 0 || nrep < 0 || (nops > 0 && !ops)) return fail(c, SPR_ERR_ARG, "bad op list");
    const char* bad = op_list_error(ops, nops);
    if (bad) return fail(c, SPR_ERR_ARG, bad);
    SPR_HIP(c, hipSetDevice(c->device));
    for (int r = 0; r < nrep; ++r) {
        for (int i = 0; i < nops; ++i) {
            const spr_op& o = ops[i];
            switch (o.kind) {
                case SPR_OP_ER: SPR_TRY(queue_iterate(c, SPR_ER, o.iterations, 0.f, reality)); break;
                case SPR_OP_HIO: SPR_TRY(queue_iterate(c, SPR_HIO, o.iterations, o.beta, reality)); break;
                case SPR_OP_RAAR: SPR_TRY(queue_iterate(c, SPR_RAAR, o.iterations, o.beta, reality)); break;
                case SPR_OP_SW: SPR_TRY(queue_sw(c, o.sigma, o.threshold)); break;
                case SPR_OP_SWA: SPR_TRY(queue_swa(c, o.sigma, o.pixels)); break;
                case SPR_OP_GRAD: SPR_TRY(queue_refine_support(c, o.iterations, o.beta, o.pixels, o.threshold)); break;
            }
        }
    }
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    return harvest_timing(c);
}

int spr_align(spr_ctx* c, const float* ref_rho, float minr, float maxr, const float* phase_direct,
              const float* phase_flip, int32_t* shifts) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (!ref_rho || !phase_direct || !phase_flip) return fail(c, SPR_ERR_ARG, "NULL argument");
    const int P = c->P;
    if (P <= 0) return SPR_OK;
    const size_t np = (size_t)c->n * c->n;
    SPR_HIP(c, hipSetDevice(c->device));
    SPR_TRY(ensure(c, c->spec, (P + 1) * np * sizeof(float2)));
    SPR_TRY(ensure(c, c->fieldbuf, np * sizeof(float2)));
    SPR_TRY(ensure(c, c->supp2, (size_t)c->cap * np));
    SPR_TRY(ensure(c, c->shiftp, (size_t)P * sizeof(ShiftParam)));
    SPR_TRY(ensure(c, c->xphase, 2 * (size_t)P * sizeof(float)));
    SPR_TRY(ensure(c, c->xidx, 2 * (size_t)P * sizeof(int)));
    std::vector<int> idx(2 * (size_t)P);
    for (int i = 0; i < P; ++i) {
        idx[i] = i;        // d: the individual's spectrum
        idx[P + i] = P;    // r: the reference spectrum
    }
    float2* spec = (float2*)c->spec.p;
    float* ph = (float*)c->xphase.p;
    int* di = (int*)c->xidx.p;
    ShiftParam* sp = (ShiftParam*)c->shiftp.p;
    SPR_HIP(c, hipMemcpyAsync(c->fieldbuf.p, ref_rho, np * sizeof(float2), hipMemcpyHostToDevice, c->stream));
    SPR_HIP(c, hipMemcpyAsync(ph, phase_direct, P * sizeof(float), hipMemcpyHostToDevice, c->stream));
    SPR_HIP(c, hipMemcpyAsync(ph + P, phase_flip, P * sizeof(float), hipMemcpyHostToDevice, c->stream));
    SPR_HIP(c, hipMemcpyAsync(di, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    // set_ref_density and get_shift both unit-scale their spectra (normalize_ft_shift)
    SPR_TRY(modulus_spectra(c, c->rho, nullptr, spec, P));
    SPR_TRY(modulus_spectra(c, (const float2*)c->fieldbuf.p, nullptr, spec + (size_t)P * np, 1));
    SPR_TRY(correlate_and_pick(c, spec, di, di + P, minr, maxr, 1, ph, ph + P, sp, P));
    ReshiftArgs ra{};
    ra.src = c->rho;
    ra.ssupp = c->supp;
    ra.dst = c->work;
    ra.dsupp = (uint8_t*)c->supp2.p;
    ra.shift = sp;
    SPR_TRY(reshift_fields(c, ra, P));
    SPR_HIP(c, hipMemcpyAsync(c->rho, c->work, P * np * sizeof(float2), hipMemcpyDeviceToDevice, c->stream));
    SPR_HIP(c, hipMemcpyAsync(c->supp, c->supp2.p, P * np, hipMemcpyDeviceToDevice, c->stream));
    std::vector<ShiftParam> hs(P);
    SPR_HIP(c, hipMemcpyAsync(hs.data(), sp, P * sizeof(ShiftParam), hipMemcpyDeviceToHost, c->stream));
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    if (shifts)
        for (int i = 0; i < P; ++i) {
            shifts[3 * i] = hs[i].sx;
            shifts[3 * i + 1] = hs[i].sy;
            shifts[3 * i + 2] = hs[i].flip;
        }
    return harvest_timing(c);
}

int spr_average(spr_ctx* c, const int32_t* idx, int navg, float* rho_out, int32_t* supp_out) {
    if (!c || !idx || !rho_out || !supp_out) return fail(c, SPR_ERR_ARG, "NULL argument");
    if (navg < 1) return fail(c, SPR_ERR_ARG, "navg must be >= 1");
    SPR_TRY(check_indices(c, idx, navg, c->P, "spr_average"));
    const size_t np = (size_t)c->n * c->n;
    SPR_HIP(c, hipSetDevice(c->device));
    SPR_TRY(ensure(c, c->xidx, navg * sizeof(int)));
    SPR_TRY(ensure(c, c->fieldbuf, np * sizeof(float2)));
    SPR_TRY(ensure(c, c->supp2, (size_t)c->cap * np));
    SPR_HIP(c, hipMemcpyAsync(c->xidx.p, idx, navg * sizeof(int), hipMemcpyHostToDevice, c->stream));
    const float norm = (float)(1. / (float)navg);
    hipLaunchKernelGGL(k_average, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->stream, (const float2*)c->rho,
                       (const uint8_t*)c->supp, (const int*)c->xidx.p, navg, (int)np, norm, (float2*)c->fieldbuf.p,
                       (uint8_t*)c->supp2.p);
    SPR_HIP(c, hipGetLastError());
    std::vector<uint8_t> s(np);
    SPR_HIP(c, hipMemcpyAsync(rho_out, c->fieldbuf.p, np * sizeof(float2), hipMemcpyDeviceToHost, c->stream));
    SPR_HIP(c, hipMemcpyAsync(s.data(), c->supp2.p, np, hipMemcpyDeviceToHost, c->stream));
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < np; ++i) supp_out[i] = s[i];
    return harvest_timing(c);
}

int spr_average_module(spr_ctx* c, const int32_t* idx, int navg, int ft, float* out) {
    if (!c || !idx || !out) return fail(c, SPR_ERR_ARG, "NULL argument");
    if (navg < 1) return fail(c, SPR_ERR_ARG, "navg must be >= 1");
    SPR_TRY(check_indices(c, idx, navg, c->P, "spr_average_module"));
    const size_t np = (size_t)c->n * c->n;
    SPR_HIP(c, hipSetDevice(c->device));
    SPR_TRY(ensure(c, c->xidx, navg * sizeof(int)));
    SPR_TRY(ensure(c, c->ampavg, np * sizeof(float)));
    SPR_HIP(c, hipMemcpyAsync(c->xidx.p, idx, navg * sizeof(int), hipMemcpyHostToDevice, c->stream));
    const float norm = (float)(1. / (float)navg);
    const dim3 g1((unsigned)((np + 255) / 256));
    if (ft) {
        SPR_TRY(ensure(c, c->spec, (size_t)navg * np * sizeof(float2)));
        hipLaunchKernelGGL(k_copy_fields, dim3((unsigned)((np + 255) / 256), navg), dim3(256), 0, c->stream,
                           (const float2*)c->rho, (const int*)c->xidx.p, (float2*)c->spec.p, (const int*)nullptr,
                           (int)np);
        SPR_HIP(c, hipGetLastError());
        SPR_TRY(fft2_fields(c, +1, (float2*)c->spec.p, nullptr, navg));
        hipLaunchKernelGGL(k_sum_modulus, g1, dim3(256), 0, c->stream, (const float2*)c->spec.p, (const int*)nullptr,
                           navg, (int)np, norm, (float*)c->ampavg.p);
    } else {
        hipLaunchKernelGGL(k_sum_modulus, g1, dim3(256), 0, c->stream, (const float2*)c->rho, (const int*)c->xidx.p,
                           navg, (int)np, norm, (float*)c->ampavg.p);
    }
    SPR_HIP(c, hipGetLastError());
    SPR_HIP(c, hipMemcpyAsync(out, c->ampavg.p, np * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    return harvest_timing(c);
}

int spr_crossover(spr_ctx* c, int nchild, const spr_cross_desc* desc, const int32_t* crossmaps, int crossmap_len,
                  const float* avg_rho, const float* amp_avg, float crossdiff, float use_average, float minr,
                  float maxr, float reality, int update_support, float* child_rho, int32_t* child_supp) {
    (void)update_support;
    if (!c || !desc || !crossmaps || !avg_rho || !amp_avg || !child_rho || !child_supp)
        return fail(c, SPR_ERR_ARG, "NULL argument");
    const int P = c->P, K = nchild, n = c->n;
    if (K < 1 || K > c->cap) return fail(c, SPR_ERR_ARG, "number of children exceeds the capacity");
    const size_t np = (size_t)n * n;
    for (int k = 0; k < K; ++k) {
        const spr_cross_desc& d = desc[k];
        const int pp[4] = {d.p0, d.p1, d.p2, d.p3};
        SPR_TRY(check_indices(c, pp, 4, P, "spr_crossover parent"));
        if (d.squaresizex < 1 || d.squaresizey < 1) return fail(c, SPR_ERR_ARG, "cell size must be >= 1");
        if (d.crosssizex != (n + d.squaresizex - 1) / d.squaresizex + 1 ||
            d.crosssizey != (n + d.squaresizey - 1) / d.squaresizey + 1)
            return fail(c, SPR_ERR_ARG, "crossmap grid does not match the cell sizes");
        if (d.offsetx < 0 || d.offsetx >= d.squaresizex + 1 || d.offsety < 0 || d.offsety >= d.squaresizey + 1)
            return fail(c, SPR_ERR_ARG, "cell offset out of range");
        if (d.crossmap_offset < 0 || d.crossmap_offset + d.crosssizex * d.crosssizey > crossmap_len)
            return fail(c, SPR_ERR_ARG, "crossmap out of range");
    }
    SPR_HIP(c, hipSetDevice(c->device));
    const bool with_avg = use_average != 0.f;
    // fields: child k -> 5k+0 p0, 5k+1 p1, 5k+2 p2, 5k+3 p3, 5k+4 avg; supports: 3k+0 p0, 3k+1 p2, 3k+2 p3
    SPR_TRY(ensure(c, c->xf, 5 * (size_t)K * np * sizeof(float2)));
    SPR_TRY(ensure(c, c->xs, 3 * (size_t)K * np));
    SPR_TRY(ensure(c, c->spec, (size_t)(P + 1) * np * sizeof(float2)));
    SPR_TRY(ensure(c, c->fieldbuf, np * sizeof(float2)));
    SPR_TRY(ensure(c, c->ampavg, np * sizeof(float)));
    SPR_TRY(ensure(c, c->shiftp, (size_t)K * sizeof(ShiftParam)));
    SPR_TRY(ensure(c, c->xchild, (size_t)K * sizeof(CrossChild)));
    SPR_TRY(ensure(c, c->xmaps, (size_t)std::max(crossmap_len, 1) * sizeof(int32_t)));
    SPR_TRY(ensure(c, c->xphase, 8 * (size_t)K * sizeof(float)));
    // host-side index lists, uploaded once
    enum { I_D = 0, I_R, I_SRC, I_DST, I_SSRC, I_SDST, NLIST };
    const int L = 4 * K;   // one list slot per (role, child)
    std::vector<int> lists((size_t)NLIST * L + 8 * (size_t)K, 0);
    auto list = [&](int which, int role, int k) -> int& { return lists[(size_t)which * L + role * K + k]; };
    std::vector<float> phases(8 * (size_t)K);
    std::vector<CrossChild> cc(K);
    int* ftl = &lists[(size_t)NLIST * L];            // fields to transform (crossft children), up to 5K
    int nft = 0;
    std::vector<int> child_fields, er_f, er_s;
    for (int k = 0; k < K; ++k) {
        const spr_cross_desc& d = desc[k];
        const int roles[4] = {d.p0, d.p2, d.p3, P};   // spectrum / source of each role (P = the average)
        for (int r = 0; r < 4; ++r) {
            list(I_D, r, k) = roles[r];
            list(I_R, r, k) = d.p1;
            list(I_SRC, r, k) = r < 3 ? roles[r] : 0;
            list(I_DST, r, k) = 5 * k + (r == 0 ? 0 : r == 1 ? 2 : r == 2 ? 3 : 4);
            list(I_SSRC, r, k) = r < 3 ? roles[r] : 0;
            list(I_SDST, r, k) = 3 * k + (r < 3 ? r : 0);
            phases[(size_t)r * 2 * K + k] = d.phase[2 * r];
            phases[(size_t)r * 2 * K + K + k] = d.phase[2 * r + 1];
        }
        CrossChild& x = cc[k];
        x.crossft = d.crossft;
        x.crosssym = d.crosssym;
        x.crosssizex = d.crosssizex;
        x.squaresizex = d.squaresizex;
        x.squaresizey = d.squaresizey;
        x.offsetx = d.offsetx;
        x.offsety = d.offsety;
        x.rotangle = d.rotangle;
        x.crossmap_offset = d.crossmap_offset;
        x.f0 = 5 * k;
        x.f1 = 5 * k + 1;
        x.f2 = 5 * k + 2;
        x.f3 = 5 * k + 3;
        x.favg = 5 * k + 4;
        x.s0 = 3 * k;
        x.s1 = d.p1;
        x.s2 = 3 * k + 1;
        x.s3 = 3 * k + 2;
        if (d.crossft) {
            for (int f = 0; f < (with_avg ? 5 : 4); ++f) ftl[nft++] = 5 * k + f;
            child_fields.push_back(5 * k);
        }
        for (int r = 0; r < 3; ++r) {
            er_f.push_back(5 * k + (r == 0 ? 0 : r + 1));
            er_s.push_back(3 * k + r);
        }
    }
    // p1 copies: field 5k+1 <- rho[p1]; and the ER lists of p1 with the population supports
    std::vector<int> p1_src(K), p1_dst(K);
    for (int k = 0; k < K; ++k) {
        p1_src[k] = desc[k].p1;
        p1_dst[k] = 5 * k + 1;
    }
    const int nchild_ft = (int)child_fields.size();
    std::vector<int> out_f(K), out_s(K);   // every child: its field and support slots
    for (int k = 0; k < K; ++k) {
        out_f[k] = 5 * k;
        out_s[k] = 3 * k;
    }
    std::vector<int> tail;
    tail.insert(tail.end(), p1_src.begin(), p1_src.end());
    tail.insert(tail.end(), p1_dst.begin(), p1_dst.end());
    tail.insert(tail.end(), er_f.begin(), er_f.end());
    tail.insert(tail.end(), er_s.begin(), er_s.end());
    tail.insert(tail.end(), child_fields.begin(), child_fields.end());
    tail.insert(tail.end(), out_f.begin(), out_f.end());
    tail.insert(tail.end(), out_s.begin(), out_s.end());
    const size_t tail_off = lists.size();
    lists.insert(lists.end(), tail.begin(), tail.end());
    SPR_TRY(ensure(c, c->xidx, lists.size() * sizeof(int)));
    int* dl = (int*)c->xidx.p;
    auto dlist = [&](int which, int role) { return dl + (size_t)which * L + role * K; };
    int* d_ft = dl + (size_t)NLIST * L;
    int* d_p1src = dl + tail_off;
    int* d_p1dst = d_p1src + K;
    int* d_erf = d_p1dst + K;
    int* d_ers = d_erf + 3 * K;
    int* d_child = d_ers + 3 * K;
    int* d_outf = d_child + nchild_ft;
    int* d_outs = d_outf + K;
    float* dph = (float*)c->xphase.p;
    float2* xf = (float2*)c->xf.p;
    uint8_t* xs = (uint8_t*)c->xs.p;
    float2* spec = (float2*)c->spec.p;
    ShiftParam* sp = (ShiftParam*)c->shiftp.p;
    SPR_HIP(c, hipMemcpyAsync(dl, lists.data(), lists.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    SPR_HIP(c, hipMemcpyAsync(dph, phases.data(), phases.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    SPR_HIP(c, hipMemcpyAsync(c->xchild.p, cc.data(), K * sizeof(CrossChild), hipMemcpyHostToDevice, c->stream));
    SPR_HIP(c, hipMemcpyAsync(c->xmaps.p, crossmaps, (size_t)crossmap_len * sizeof(int32_t), hipMemcpyHostToDevice,
                              c->stream));
    SPR_HIP(c, hipMemcpyAsync(c->fieldbuf.p, avg_rho, np * sizeof(float2), hipMemcpyHostToDevice, c->stream));
    SPR_HIP(c, hipMemcpyAsync(c->ampavg.p, amp_avg, np * sizeof(float), hipMemcpyHostToDevice, c->stream));

    // prepare_parents: spectra FT(|rho|) of the population and of the average (not unit-scaled)
    SPR_TRY(modulus_spectra(c, c->rho, nullptr, spec, P));
    if (with_avg) SPR_TRY(modulus_spectra(c, (const float2*)c->fieldbuf.p, nullptr, spec + (size_t)P * np, 1));
    // reshift_parent of p0, p2, p3 (and the average) against p1
    for (int r = 0; r < (with_avg ? 4 : 3); ++r) {
        SPR_TRY(correlate_and_pick(c, spec, dlist(I_D, r), dlist(I_R, r), minr, maxr, 0, dph + (size_t)r * 2 * K,
                                   dph + (size_t)r * 2 * K + K, sp, K));
        ReshiftArgs ra{};
        ra.src = r < 3 ? c->rho : (const float2*)c->fieldbuf.p;
        ra.src_index = dlist(I_SRC, r);
        ra.ssupp = r < 3 ? c->supp : nullptr;
        ra.ssupp_index = dlist(I_SSRC, r);
        ra.dst = xf;
        ra.dst_index = dlist(I_DST, r);
        ra.dsupp = r < 3 ? xs : nullptr;
        ra.dsupp_index = dlist(I_SDST, r);
        ra.shift = sp;
        SPR_TRY(reshift_fields(c, ra, K));
    }
    // get_new_individual: p1 copy, ER projections, transforms, cell mix, inverse, ER
    const dim3 gp((unsigned)((np + 255) / 256), K), gp3((unsigned)((np + 255) / 256), 3 * K);
    hipLaunchKernelGGL(k_copy_fields, gp, dim3(256), 0, c->stream, (const float2*)c->rho, (const int*)d_p1src, xf,
                       (const int*)d_p1dst, (int)np);
    SPR_HIP(c, hipGetLastError());
    const PhaseCut pc = cut_er_hio(reality);
    hipLaunchKernelGGL(k_er_project, gp, dim3(256), 0, c->stream, xf, (const int*)d_p1dst, (const uint8_t*)c->supp,
                       (const int*)d_p1src, (int)np, pc);
    SPR_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_er_project, gp3, dim3(256), 0, c->stream, xf, (const int*)d_erf, (const uint8_t*)xs,
                       (const int*)d_ers, (int)np, pc);
    SPR_HIP(c, hipGetLastError());
    SPR_TRY(fft2_fields(c, +1, xf, d_ft, nft));
    CrossArgs xa{};
    xa.fields = xf;
    xa.supps = xs;
    xa.pop_supp = c->supp;
    xa.child = (const CrossChild*)c->xchild.p;
    xa.crossmaps = (const int32_t*)c->xmaps.p;
    xa.amp = c->amp;
    xa.amp_avg = (const float*)c->ampavg.p;
    xa.crossdiff = crossdiff;
    xa.use_average = use_average;
    SPR_TRY(cross_fields(c, xa, K));
    SPR_TRY(fft2_fields(c, -1, xf, d_child, nchild_ft));
    hipLaunchKernelGGL(k_er_project, gp, dim3(256), 0, c->stream, xf, (const int*)d_outf, (const uint8_t*)xs,
                       (const int*)d_outs, (int)np, pc);
    SPR_HIP(c, hipGetLastError());
    // Children (field 5k, support 3k) through pinned staging; host threads copy
    // the fields out and widen the supports.
    const size_t fb = np * sizeof(float2);
    SPR_TRY(ensure_pinned(c, (size_t)K * (fb + np)));
    char* stage = (char*)c->pinned.p;
    for (int k = 0; k < K; ++k) {
        SPR_HIP(c, hipMemcpyAsync(stage + (size_t)k * fb, xf + (size_t)(5 * k) * np, fb, hipMemcpyDeviceToHost,
                                  c->stream));
        SPR_HIP(c, hipMemcpyAsync(stage + (size_t)K * fb + (size_t)k * np, xs + (size_t)(3 * k) * np, np,
                                  hipMemcpyDeviceToHost, c->stream));
    }
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    host_parallel(K, [&](int k) {
        std::memcpy(child_rho + 2 * (size_t)k * np, stage + (size_t)k * fb, fb);
        const uint8_t* src = (const uint8_t*)(stage + (size_t)K * fb + (size_t)k * np);
        int32_t* d = child_supp + (size_t)k * np;
        for (size_t i = 0; i < np; ++i) d[i] = src[i];
    });
    return harvest_timing(c);
}

int spr_loss_grad(spr_ctx* c, int P, const float* rho, float* grad, double* loss, int kind, float eps, void* stream) {
    if (!c || !rho) return fail(c, SPR_ERR_ARG, "NULL argument");
    if (P < 1 || P > c->cap) return fail(c, SPR_ERR_ARG, "field count exceeds the capacity");
    if (kind != LOSS_POISSON && kind != LOSS_AMPLITUDE) return fail(c, SPR_ERR_ARG, "unknown loss kind");
    SPR_HIP(c, hipSetDevice(c->device));
    StreamSwap guard(c, (hipStream_t)stream);
    SPR_TRY(queue_loss_grad(c, P, (const float2*)rho, (float2*)grad, kind, eps));
    if (loss) {
        std::vector<double> s((size_t)P * SUMS_PER_ITEM);
        SPR_HIP(c, hipMemcpyAsync(s.data(), c->sums, s.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        SPR_HIP(c, hipStreamSynchronize(c->stream));
        for (int p = 0; p < P; ++p) loss[p] = s[(size_t)p * SUMS_PER_ITEM + 7];
    }
    return harvest_timing(c);
}

int spr_loss_grad_noalloc(int n, int P, const float* amp, const float* rho, float* grad, double* loss, int kind,
                          float eps, int want_grad, void* stream) {
    if (!amp || !rho || !grad) return fail(nullptr, SPR_ERR_ARG, "NULL argument");
    if (!spr_supported_size(n)) return fail(nullptr, SPR_ERR_ARG, "unsupported pattern size");
    if (P < 1) return fail(nullptr, SPR_ERR_ARG, "field count must be >= 1");
    if (kind != LOSS_POISSON && kind != LOSS_AMPLITUDE) return fail(nullptr, SPR_ERR_ARG, "unknown loss kind");
    const size_t per_field = (size_t)(n / cols_per_wg(n));
    if ((size_t)P * per_field > NOALLOC_PARTIALS)
        return fail(nullptr, SPR_ERR_ARG,
                    "field count exceeds the fixed scratch (" + std::to_string(NOALLOC_PARTIALS / per_field) +
                        " fields of " + std::to_string(n) + "^2 at most)");
    const size_t bytes = (size_t)P * n * n * sizeof(float2);
    const char *r0 = (const char*)rho, *g0 = (const char*)grad;
    if (r0 < g0 + bytes && g0 < r0 + bytes) return fail(nullptr, SPR_ERR_ARG, "rho and grad overlap");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(nullptr, SPR_ERR_HIP, std::string("hipGetDevice failed: ") + hipGetErrorString(e));
    unsigned char* base = nullptr;
    SPR_TRY(noalloc_setup(dev, &base));
    // A handle on the stack that borrows the caller's buffers and the fixed
    // scratch; its launch helpers queue kernels and allocate nothing.
    spr_ctx c;
    c.n = n;
    c.cap = c.P = P;
    c.device = dev;
    c.stream = (hipStream_t)stream;
    c.amp = const_cast<float*>(amp);
    c.tw = reinterpret_cast<float2*>(base) + noalloc_tw_offset(n);
    c.partial = reinterpret_cast<double*>(base + NOALLOC_TW_BYTES);
    const float2* r = (const float2*)rho;
    float2* g = (float2*)grad;
    int rc = queue_loss_passes(&c, P, r, g, want_grad ? g : nullptr, kind, eps);
    if (rc == SPR_OK && loss) {
        hipLaunchKernelGGL(k_reduce_partials, dim3(P), dim3(256), 0, c.stream, (const double*)c.partial, (int)per_field,
                           1, loss, 1, 0, P);
        e = hipGetLastError();
        if (e != hipSuccess) rc = fail(&c, SPR_ERR_HIP, std::string("loss reduction launch failed: ") + hipGetErrorString(e));
    }
    if (rc != SPR_OK) g_create_error = c.err;
    return rc;
}

int spr_refine(spr_ctx* c, int steps, float lr, int kind, float eps, double* loss) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (steps < 0) return fail(c, SPR_ERR_ARG, "negative step count");
    if (kind != LOSS_POISSON && kind != LOSS_AMPLITUDE) return fail(c, SPR_ERR_ARG, "unknown loss kind");
    const int P = c->P;
    if (P <= 0 || steps == 0) return SPR_OK;
    SPR_HIP(c, hipSetDevice(c->device));
    SPR_TRY(queue_adam_steps(c, steps, lr, nullptr, kind, eps));
    if (loss) SPR_TRY(final_loss(c, kind, eps, loss));
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    return harvest_timing(c);
}

int spr_refine_support(spr_ctx* c, int steps, float lr_rel, int kind, float eps, double* loss) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (steps < 0) return fail(c, SPR_ERR_ARG, "negative step count");
    if (kind != LOSS_POISSON && kind != LOSS_AMPLITUDE) return fail(c, SPR_ERR_ARG, "unknown loss kind");
    if (c->P <= 0 || steps == 0) return SPR_OK;
    SPR_HIP(c, hipSetDevice(c->device));
    SPR_TRY(queue_refine_support(c, steps, lr_rel, kind, eps));
    if (loss) SPR_TRY(final_loss(c, kind, eps, loss));
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    return harvest_timing(c);
}

int spr_device_count(int* n) {
    if (!n) return fail(nullptr, SPR_ERR_ARG, "NULL argument");
    *n = 0;
    int k = 0;
    if (hipGetDeviceCount(&k) != hipSuccess) return SPR_OK;
    *n = k;
    return SPR_OK;
}

int spr_device_name(int device, char* buf, int len) {
    if (!buf || len <= 0) return fail(nullptr, SPR_ERR_ARG, "NULL argument");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(nullptr, SPR_ERR_HIP, "bad device");
    std::snprintf(buf, (size_t)len, "%s", prop.name);
    return SPR_OK;
}

int spr_sync(spr_ctx* c) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    SPR_HIP(c, hipSetDevice(c->device));
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    return harvest_timing(c);
}

int spr_timing(spr_ctx* c, int enable) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    SPR_TRY(harvest_timing(c));
    if (enable < 0) return fail(c, SPR_ERR_ARG, "negative timing stride");
    c->timing = enable != 0;
    c->timing_stride = enable > 0 ? enable : 1;
    for (int k = 0; k < SPR_TIMING_KINDS; ++k) {
        c->kind_ms[k] = 0;
        c->kind_count[k] = 0;
        c->kind_seen[k] = 0;
    }
    return SPR_OK;
}

int spr_timing_read(spr_ctx* c, int kind, double* total_ms, long* count) {
    if (!c || kind < 0 || kind >= SPR_TIMING_KINDS) return fail(c, SPR_ERR_ARG, "bad timing kind");
    if (total_ms) *total_ms = c->kind_ms[kind];
    if (count) *count = c->kind_count[kind];
    return SPR_OK;
}

int spr_device_buffers(spr_ctx* c, void** rho, void** support_u8, void** amp) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (rho) *rho = c->rho;
    if (support_u8) *support_u8 = c->supp;
    if (amp) *amp = c->amp;
    return SPR_OK;
}

int spr_stream(spr_ctx* c, void** stream) {
    if (!c || !stream) return fail(c, SPR_ERR_ARG, "NULL argument");
    *stream = (void*)c->stream;
    return SPR_OK;
}

int spr_variants(int n, int* col_count, int* row_count) {
    int nc = 0, nr = 0;
    if (variant_counts(n, &nc, &nr) != SPR_OK) return fail(nullptr, SPR_ERR_ARG, "unsupported grid size");
    if (col_count) *col_count = nc;
    if (row_count) *row_count = nr;
    return SPR_OK;
}

int spr_row_variant_info(int n, int variant, int* recomputes_state) {
    int nc = 0, nr = 0;
    if (variant_counts(n, &nc, &nr) != SPR_OK) return fail(nullptr, SPR_ERR_ARG, "unsupported grid size");
    if (variant < 0 || variant >= nr) return fail(nullptr, SPR_ERR_ARG, "variant out of range");
    if (recomputes_state) *recomputes_state = row_recomputes(n, variant) ? 1 : 0;
    return SPR_OK;
}

int spr_default_variants(int n, int* col, int* row) {
    int nc = 0, nr = 0;
    if (variant_counts(n, &nc, &nr) != SPR_OK) return fail(nullptr, SPR_ERR_ARG, "unsupported grid size");
    const VariantPair v = default_variants(n);
    if (col) *col = v.col;
    if (row) *row = v.row;
    return SPR_OK;
}

int spr_set_default_variants(int n, int col, int row) {
    int nc = 0, nr = 0;
    if (variant_counts(n, &nc, &nr) != SPR_OK) return fail(nullptr, SPR_ERR_ARG, "unsupported grid size");
    if (col < 0 || col >= nc || row < 0 || row >= nr) return fail(nullptr, SPR_ERR_ARG, "variant out of range");
    std::lock_guard<std::mutex> lk(g_variant_mu);
    g_variant[n] = VariantPair{col, row};
    return SPR_OK;
}

int spr_set_variants(spr_ctx* c, int col, int row) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    int nc = 0, nr = 0;
    SPR_TRY(variant_counts(c->n, &nc, &nr));
    if (col < 0 || col >= nc || row < 0 || row >= nr) return fail(c, SPR_ERR_ARG, "variant out of range");
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    c->colvar = col;
    c->rowvar = row;
    return SPR_OK;
}

int spr_get_variants(spr_ctx* c, int* col, int* row) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (col) *col = c->colvar;
    if (row) *row = c->rowvar;
    return SPR_OK;
}

int spr_set_chunk(spr_ctx* c, long long bytes) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (bytes < 0) return fail(c, SPR_ERR_ARG, "negative chunk size");
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    c->chunk_bytes = (size_t)bytes;
    return SPR_OK;
}

int spr_get_chunk(spr_ctx* c, long long* bytes, int* individuals) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (bytes) *bytes = (long long)c->chunk_bytes;
    if (individuals) *individuals = chunk_individuals(c);
    return SPR_OK;
}

int spr_set_fused(spr_ctx* c, int mode) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (mode < -1 || mode > 1) return fail(c, SPR_ERR_ARG, "fused mode must be -1, 0 or 1");
    if (mode == 1 && !fused_available(c->n)) return fail(c, SPR_ERR_ARG, "no on-chip iteration block for this grid size");
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    c->fused = mode;
    return SPR_OK;
}

int spr_get_fused(spr_ctx* c, int* mode, int* active) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    if (mode) *mode = c->fused;
    if (active) *active = use_fused(c) ? 1 : 0;
    return SPR_OK;
}

int spr_autotune(spr_ctx* c, int reps, float* col_ms, float* row_ms) {
    if (!c) return fail(c, SPR_ERR_ARG, "NULL handle");
    SPR_HIP(c, hipSetDevice(c->device));
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    return autotune(c, reps, col_ms, row_ms);
}

// ---------------------------------------------------------------------------
// Batches of patterns of several sizes (spring_hip.h, spr_batch_*)
// ---------------------------------------------------------------------------
struct spr_batch {
    int device = 0;
    struct Class {
        int n = 0;
        spr_ctx* ctx = nullptr;       // every individual of every pattern of size n; its stream is the class's
        hipEvent_t done = nullptr;
    };
    std::vector<Class> classes;
    std::vector<int> pat_class, pat_first, pat_pop;   // per pattern: class, first slot, population
    hipStream_t origin = nullptr;                     // stream the block graphs are launched on
    hipEvent_t fork = nullptr;
    bool graphs = true;
    bool concurrent = false;                          // size classes of a block side by side, else in turn
    std::map<std::tuple<int, int, unsigned, unsigned>, hipGraphExec_t> blocks;
    std::string err;
};

static int bfail(spr_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    else g_create_error = msg;
    return code;
}

#define SPR_BHIP(b, call)                                                                           \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return bfail((b), SPR_ERR_HIP, std::string(#call " failed: ") + hipGetErrorString(e_)); \
    } while (0)

// A class handle's error becomes the batch's.
static int bclass(spr_batch* b, spr_batch::Class& k, int rc) {
    if (rc != SPR_OK) b->err = "size " + std::to_string(k.n) + ": " + k.ctx->err;
    return rc;
}

static int batch_sync(spr_batch* b) {
    SPR_BHIP(b, hipSetDevice(b->device));
    for (auto& k : b->classes) SPR_BHIP(b, hipStreamSynchronize(k.ctx->stream));
    SPR_BHIP(b, hipStreamSynchronize(b->origin));
    return SPR_OK;
}

// Every class stream's pending work before the origin stream's next launch, and
// the reverse: `to_origin` true makes the origin wait for the classes.
static int batch_join(spr_batch* b, bool to_origin) {
    if (to_origin) {
        for (auto& k : b->classes) {
            SPR_BHIP(b, hipEventRecord(k.done, k.ctx->stream));
            SPR_BHIP(b, hipStreamWaitEvent(b->origin, k.done, 0));
        }
    } else {
        SPR_BHIP(b, hipEventRecord(b->fork, b->origin));
        for (auto& k : b->classes) SPR_BHIP(b, hipStreamWaitEvent(k.ctx->stream, b->fork, 0));
    }
    return SPR_OK;
}

// The iteration block (alg, niter, beta, reality) of every class as one graph:
// captured once from the origin stream, forking to the class streams (one
// branch per size) and joining back.
static int batch_block_graph(spr_batch* b, int alg, int niter, float beta, float reality, hipGraphExec_t* out) {
    unsigned bb, rb;
    std::memcpy(&bb, &beta, 4);
    std::memcpy(&rb, &reality, 4);
    const auto key = std::make_tuple(alg, niter, bb, rb);
    auto it = b->blocks.find(key);
    if (it != b->blocks.end()) {
        *out = it->second;
        return SPR_OK;
    }
    SPR_BHIP(b, hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
    int rc = SPR_OK;
    if (hipEventRecord(b->fork, b->origin) != hipSuccess) rc = bfail(b, SPR_ERR_HIP, "capture fork failed");
    for (size_t i = 0; rc == SPR_OK && i < b->classes.size(); ++i) {
        auto& k = b->classes[i];
        // concurrent: every class branches from the fork; in turn: from the previous class
        hipEvent_t from = (b->concurrent || i == 0) ? b->fork : b->classes[i - 1].done;
        if (hipStreamWaitEvent(k.ctx->stream, from, 0) != hipSuccess) {
            rc = bfail(b, SPR_ERR_HIP, "capture fork failed");
            break;
        }
        rc = bclass(b, k, queue_iterate(k.ctx, alg, niter, beta, reality));
        if (rc == SPR_OK && (hipEventRecord(k.done, k.ctx->stream) != hipSuccess ||
                             hipStreamWaitEvent(b->origin, k.done, 0) != hipSuccess))
            rc = bfail(b, SPR_ERR_HIP, "capture join failed");
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->origin, &g);
    if (rc != SPR_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return bfail(b, SPR_ERR_HIP, std::string("hipStreamEndCapture failed: ") + hipGetErrorString(e));
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess)
        return bfail(b, SPR_ERR_HIP, std::string("hipGraphInstantiate failed: ") + hipGetErrorString(ei));
    b->blocks[key] = x;
    *out = x;
    return SPR_OK;
}

// Class i's stream waits for class i-1's work queued so far (in-turn order).
static int batch_chain(spr_batch* b, size_t i) {
    if (i == 0 || b->concurrent) return SPR_OK;
    auto& prev = b->classes[i - 1];
    SPR_BHIP(b, hipEventRecord(prev.done, prev.ctx->stream));
    SPR_BHIP(b, hipStreamWaitEvent(b->classes[i].ctx->stream, prev.done, 0));
    return SPR_OK;
}

static int batch_iterate(spr_batch* b, int alg, int niter, float beta, float reality) {
    if (!b->graphs) {
        for (size_t i = 0; i < b->classes.size(); ++i) {
            SPR_TRY(batch_chain(b, i));
            auto& k = b->classes[i];
            SPR_TRY(bclass(b, k, queue_iterate(k.ctx, alg, niter, beta, reality)));
        }
        return SPR_OK;
    }
    hipGraphExec_t x = nullptr;
    SPR_TRY(batch_join(b, true));
    SPR_TRY(batch_block_graph(b, alg, niter, beta, reality, &x));
    SPR_BHIP(b, hipGraphLaunch(x, b->origin));
    return batch_join(b, false);
}

int spr_batch_create(spr_batch** out, int npatterns, const int* n, const int* population, const float* const* amp,
                     int device) {
    if (!out) return bfail(nullptr, SPR_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (npatterns < 1 || !n || !population || !amp) return bfail(nullptr, SPR_ERR_ARG, "NULL or empty pattern list");
    for (int i = 0; i < npatterns; ++i) {
        if (!spr_supported_size(n[i]))
            return bfail(nullptr, SPR_ERR_ARG, "pattern " + std::to_string(i) + ": unsupported grid size " +
                                                   std::to_string(n[i]));
        if (population[i] < 1) return bfail(nullptr, SPR_ERR_ARG, "pattern " + std::to_string(i) + ": empty population");
        if (!amp[i]) return bfail(nullptr, SPR_ERR_ARG, "pattern " + std::to_string(i) + ": amp is NULL");
    }
    spr_batch* b = new spr_batch();
    b->device = device;
    b->pat_class.assign(npatterns, -1);
    b->pat_first.assign(npatterns, 0);
    b->pat_pop.assign(population, population + npatterns);
    // Classes by size, in order of first appearance; patterns keep their order within a class.
    std::vector<int> sizes;
    for (int i = 0; i < npatterns; ++i)
        if (std::find(sizes.begin(), sizes.end(), n[i]) == sizes.end()) sizes.push_back(n[i]);
    auto undo = [&](int rc) {
        spr_batch_destroy(b);
        return rc;
    };
    for (int size : sizes) {
        std::vector<const float*> amps;
        std::vector<int> amp_of;
        for (int i = 0; i < npatterns; ++i) {
            if (n[i] != size) continue;
            b->pat_class[i] = (int)b->classes.size();
            b->pat_first[i] = (int)amp_of.size();
            for (int j = 0; j < population[i]; ++j) amp_of.push_back((int)amps.size());
            amps.push_back(amp[i]);
        }
        spr_batch::Class k;
        k.n = size;
        const int cap = (int)amp_of.size();
        const int rc = create_context(&k.ctx, size, cap, (int)amps.size(), amps.data(),
                                      amps.size() > 1 ? amp_of.data() : nullptr, device);
        if (rc != SPR_OK) {
            g_create_error = "size " + std::to_string(size) + ": " + g_create_error;
            return undo(rc);
        }
        b->classes.push_back(k);
        spr_batch::Class& kk = b->classes.back();
        kk.ctx->lanes = 1;
        kk.ctx->P = cap;
        if (hipMemset(kk.ctx->rho, 0, (size_t)size * size * cap * sizeof(float2)) != hipSuccess ||
            hipEventCreateWithFlags(&kk.done, hipEventDisableTiming) != hipSuccess) {
            g_create_error = "batch set-up failed";
            return undo(SPR_ERR_HIP);
        }
    }
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
        g_create_error = "batch stream set-up failed";
        return undo(SPR_ERR_HIP);
    }
    *out = b;
    return SPR_OK;
}

int spr_batch_destroy(spr_batch* b) {
    if (!b) return SPR_OK;
    (void)hipSetDevice(b->device);
    if (b->origin) (void)hipStreamSynchronize(b->origin);
    for (auto& kv : b->blocks) (void)hipGraphExecDestroy(kv.second);
    for (auto& k : b->classes) {
        if (k.done) (void)hipEventDestroy(k.done);
        spr_destroy(k.ctx);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->origin) (void)hipStreamDestroy(b->origin);
    delete b;
    return SPR_OK;
}

const char* spr_batch_last_error(const spr_batch* b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int spr_batch_graphs(spr_batch* b, int enable) {
    if (!b) return bfail(b, SPR_ERR_ARG, "NULL handle");
    b->graphs = enable != 0;
    return SPR_OK;
}

int spr_batch_concurrent(spr_batch* b, int enable) {
    if (!b) return bfail(b, SPR_ERR_ARG, "NULL handle");
    SPR_TRY(batch_sync(b));
    if (b->concurrent != (enable != 0)) {
        for (auto& kv : b->blocks) (void)hipGraphExecDestroy(kv.second);
        b->blocks.clear();
    }
    b->concurrent = enable != 0;
    return SPR_OK;
}

static int batch_pattern(spr_batch* b, int pattern) {
    if (!b) return bfail(b, SPR_ERR_ARG, "NULL handle");
    if (pattern < 0 || pattern >= (int)b->pat_class.size()) return bfail(b, SPR_ERR_ARG, "pattern index out of range");
    return SPR_OK;
}

int spr_batch_set_population(spr_batch* b, int pattern, const float* rho, const int32_t* support) {
    SPR_TRY(batch_pattern(b, pattern));
    if (!rho && !support) return bfail(b, SPR_ERR_ARG, "rho and support are NULL");
    auto& k = b->classes[b->pat_class[pattern]];
    spr_ctx* c = k.ctx;
    const size_t np = (size_t)k.n * k.n, P = (size_t)b->pat_pop[pattern], first = (size_t)b->pat_first[pattern];
    SPR_TRY(batch_sync(b));
    if (rho) SPR_BHIP(b, hipMemcpy(c->rho + first * np, rho, np * P * sizeof(float2), hipMemcpyHostToDevice));
    if (support) {
        std::vector<uint8_t> s(np * P);
        for (size_t i = 0; i < s.size(); ++i) s[i] = support[i] != 0 ? 1 : 0;
        SPR_BHIP(b, hipMemcpy(c->supp + first * np, s.data(), s.size(), hipMemcpyHostToDevice));
    }
    return SPR_OK;
}

int spr_batch_get_population(spr_batch* b, int pattern, float* rho, int32_t* support) {
    SPR_TRY(batch_pattern(b, pattern));
    auto& k = b->classes[b->pat_class[pattern]];
    spr_ctx* c = k.ctx;
    const size_t np = (size_t)k.n * k.n, P = (size_t)b->pat_pop[pattern], first = (size_t)b->pat_first[pattern];
    SPR_TRY(batch_sync(b));
    if (rho) SPR_BHIP(b, hipMemcpy(rho, c->rho + first * np, np * P * sizeof(float2), hipMemcpyDeviceToHost));
    if (support) {
        std::vector<uint8_t> s(np * P);
        SPR_BHIP(b, hipMemcpy(s.data(), c->supp + first * np, s.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < s.size(); ++i) support[i] = s[i];
    }
    return SPR_OK;
}

int spr_batch_run(spr_batch* b, const spr_op* ops, int nops, int nrep, float reality) {
    if (!b) return bfail(b, SPR_ERR_ARG, "NULL handle");
    if (nops < 0 || nrep < 0 || (nops > 0 && !ops)) return bfail(b, SPR_ERR_ARG, "bad op list");
    const char* bad = op_list_error(ops, nops);
    if (bad) return bfail(b, SPR_ERR_ARG, bad);
    SPR_BHIP(b, hipSetDevice(b->device));
    for (int r = 0; r < nrep; ++r) {
        for (int i = 0; i < nops; ++i) {
            const spr_op& o = ops[i];
            switch (o.kind) {
                case SPR_OP_ER: SPR_TRY(batch_iterate(b, SPR_ER, o.iterations, 0.f, reality)); break;
                case SPR_OP_HIO: SPR_TRY(batch_iterate(b, SPR_HIO, o.iterations, o.beta, reality)); break;
                case SPR_OP_RAAR: SPR_TRY(batch_iterate(b, SPR_RAAR, o.iterations, o.beta, reality)); break;
                case SPR_OP_SW:
                    for (auto& k : b->classes) SPR_TRY(bclass(b, k, queue_sw(k.ctx, o.sigma, o.threshold)));
                    break;
                case SPR_OP_SWA:
                    for (auto& k : b->classes) SPR_TRY(bclass(b, k, queue_swa(k.ctx, o.sigma, o.pixels)));
                    break;
                case SPR_OP_GRAD:
                    for (auto& k : b->classes)
                        SPR_TRY(bclass(b, k, queue_refine_support(k.ctx, o.iterations, o.beta, o.pixels, o.threshold)));
                    break;
            }
        }
    }
    return batch_sync(b);
}

int spr_batch_set_chunk(spr_batch* b, long long bytes) {
    if (!b) return bfail(b, SPR_ERR_ARG, "NULL handle");
    if (bytes < 0) return bfail(b, SPR_ERR_ARG, "negative chunk size");
    SPR_TRY(batch_sync(b));
    for (auto& k : b->classes) k.ctx->chunk_bytes = (size_t)bytes;
    // captured blocks hold the previous chunking's launches
    for (auto& kv : b->blocks) (void)hipGraphExecDestroy(kv.second);
    b->blocks.clear();
    return SPR_OK;
}

int spr_batch_autotune(spr_batch* b, int reps) {
    if (!b) return bfail(b, SPR_ERR_ARG, "NULL handle");
    SPR_TRY(batch_sync(b));
    for (auto& k : b->classes) SPR_TRY(bclass(b, k, autotune(k.ctx, reps, nullptr, nullptr)));
    // captured blocks hold the previous variants' kernels
    for (auto& kv : b->blocks) (void)hipGraphExecDestroy(kv.second);
    b->blocks.clear();
    return batch_sync(b);
}

int spr_batch_evaluate(spr_batch* b, float reality, float* error_fourier, float* error_real, float* complexsum,
                       float* modulesum) {
    if (!b) return bfail(b, SPR_ERR_ARG, "NULL handle");
    SPR_BHIP(b, hipSetDevice(b->device));
    SPR_TRY(batch_sync(b));
    for (auto& k : b->classes) SPR_TRY(bclass(b, k, queue_error(k.ctx, reality)));
    std::vector<std::vector<double>> sums(b->classes.size());
    std::vector<std::vector<unsigned>> mx(b->classes.size());
    for (size_t i = 0; i < b->classes.size(); ++i) {
        spr_ctx* c = b->classes[i].ctx;
        sums[i].resize((size_t)c->P * SUMS_PER_ITEM);
        mx[i].resize(c->P);
        SPR_BHIP(b, hipMemcpyAsync(sums[i].data(), c->sums, sums[i].size() * sizeof(double), hipMemcpyDeviceToHost,
                                   c->stream));
        SPR_BHIP(b, hipMemcpyAsync(mx[i].data(), c->maxbits, mx[i].size() * sizeof(unsigned), hipMemcpyDeviceToHost,
                                   c->stream));
    }
    SPR_TRY(batch_sync(b));
    size_t out = 0;
    for (size_t pat = 0; pat < b->pat_class.size(); ++pat) {
        const int ci = b->pat_class[pat];
        for (int j = 0; j < b->pat_pop[pat]; ++j, ++out) {
            const size_t slot = (size_t)b->pat_first[pat] + j;
            const double* s = &sums[ci][slot * SUMS_PER_ITEM];
            float fm;
            std::memcpy(&fm, &mx[ci][slot], 4);
            if (error_fourier) error_fourier[out] = (float)std::sqrt(s[1] / s[2]);
            if (error_real) error_real[out] = (float)std::sqrt(s[3] / s[4]);
            if (complexsum) complexsum[out] = fm;
            if (modulesum) modulesum[out] = (float)s[0];
        }
    }
    return SPR_OK;
}

}  // extern "C"
