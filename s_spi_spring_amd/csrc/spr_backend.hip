// Host side of libspring_hip.so: population-resident contexts and the C ABI
// declared in include/spring_hip.h. Kernels are in spr_kernels.hpp.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "spring_hip.h"
#include "spr_kernels.hpp"

using namespace spr;

struct spr_ctx {
    int n = 0, cap = 0, P = 0, device = 0;
    hipStream_t stream = nullptr;
    float2* rho = nullptr;
    float2* work = nullptr;
    uint8_t* supp = nullptr;
    float* amp = nullptr;
    float2* tw = nullptr;
    std::string err;
    // Opt-in per-launch timing: events bracket every launch on the stream.
    bool timing = false;
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
    if (!c->timing) return SPR_OK;
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
    if (!c->timing) return SPR_OK;
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

// Reality windows of src/gpu/gpu_context.cu:530, 553 (ER/HIO) and 582 (RAAR).
PhaseCut cut_er_hio(float r) {
    const double t = (double)(1.f - r) * M_PI;
    return PhaseCut{next_up(t), next_down(-t), r == 0.f ? 1 : 0};
}
PhaseCut cut_raar(float r) {
    return PhaseCut{next_up((1.0 - (double)r) * M_PI), next_down(-(double)(1.f - r) * M_PI), r == 0.f ? 1 : 0};
}

template <int N> int launch_row(spr_ctx* c, int mode, const RowArgs& a) {
    const dim3 grid((unsigned)((size_t)c->P * N / Geo<N>::RW)), block(Geo<N>::ROW_THREADS);
    switch (mode) {
        case ROW_FWD: hipLaunchKernelGGL((k_row<N, ROW_FWD>), grid, block, 0, c->stream, a); break;
        case ROW_FINAL: hipLaunchKernelGGL((k_row<N, ROW_FINAL>), grid, block, 0, c->stream, a); break;
        case ROW_HIO: hipLaunchKernelGGL((k_row<N, ROW_HIO>), grid, block, 0, c->stream, a); break;
        case ROW_RAAR: hipLaunchKernelGGL((k_row<N, ROW_RAAR>), grid, block, 0, c->stream, a); break;
        case ROW_ER: hipLaunchKernelGGL((k_row<N, ROW_ER>), grid, block, 0, c->stream, a); break;
        case ROW_FWD_INPLACE: hipLaunchKernelGGL((k_row<N, ROW_FWD_INPLACE>), grid, block, 0, c->stream, a); break;
        case ROW_INV_INPLACE: hipLaunchKernelGGL((k_row<N, ROW_INV_INPLACE>), grid, block, 0, c->stream, a); break;
        default: return fail(c, SPR_ERR_ARG, "bad row mode");
    }
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}

template <int N> int launch_col(spr_ctx* c, int mode, float2* f) {
    const ColArgs a{f, c->amp, c->tw};
    const dim3 grid((unsigned)((size_t)c->P * (N / Geo<N>::C))), block(Geo<N>::COL_THREADS);
    switch (mode) {
        case COL_PROJ: hipLaunchKernelGGL((k_col<N, COL_PROJ>), grid, block, 0, c->stream, a); break;
        case COL_FWD: hipLaunchKernelGGL((k_col<N, COL_FWD>), grid, block, 0, c->stream, a); break;
        case COL_INV: hipLaunchKernelGGL((k_col<N, COL_INV>), grid, block, 0, c->stream, a); break;
        default: return fail(c, SPR_ERR_ARG, "bad column mode");
    }
    SPR_HIP(c, hipGetLastError());
    return SPR_OK;
}

#define SPR_DISPATCH_N(c, FN, ...)                         \
    switch ((c)->n) {                                      \
        case 64: return FN<64>(__VA_ARGS__);               \
        case 128: return FN<128>(__VA_ARGS__);             \
        case 256: return FN<256>(__VA_ARGS__);             \
        case 512: return FN<512>(__VA_ARGS__);             \
        case 1024: return FN<1024>(__VA_ARGS__);           \
        case 2048: return FN<2048>(__VA_ARGS__);           \
        default: return fail((c), SPR_ERR_ARG, "unsupported grid size"); \
    }

int row_n(spr_ctx* c, int mode, const RowArgs& a) { SPR_DISPATCH_N(c, launch_row, c, mode, a); }
int col_n(spr_ctx* c, int mode, float2* f) { SPR_DISPATCH_N(c, launch_col, c, mode, f); }

int row(spr_ctx* c, int mode, const RowArgs& a) {
    const int kind = (mode == ROW_HIO || mode == ROW_RAAR || mode == ROW_ER) ? SPR_KIND_ROW_UPDATE : SPR_KIND_ROW_OTHER;
    int rc = timing_begin(c, kind);
    if (rc != SPR_OK) return rc;
    rc = row_n(c, mode, a);
    if (rc != SPR_OK) return rc;
    return timing_end(c);
}
int col(spr_ctx* c, int mode, float2* f) {
    const int kind = (mode == COL_PROJ) ? SPR_KIND_COL_PROJ : SPR_KIND_COL_OTHER;
    int rc = timing_begin(c, kind);
    if (rc != SPR_OK) return rc;
    rc = col_n(c, mode, f);
    if (rc != SPR_OK) return rc;
    return timing_end(c);
}

RowArgs row_args(spr_ctx* c) {
    RowArgs a{};
    a.rho = c->rho;
    a.work = c->work;
    a.supp = c->supp;
    a.tw = c->tw;
    a.beta = 0.f;
    a.inv_nn = 1.0f / (float)((double)c->n * c->n);
    a.pc = cut_er_hio(0.f);
    return a;
}

#define SPR_TRY(expr)                 \
    do {                              \
        int rc_ = (expr);             \
        if (rc_ != SPR_OK) return rc_; \
    } while (0)

int queue_iterate(spr_ctx* c, int alg, int niter, float beta, float reality) {
    if (alg < SPR_ER || alg > SPR_RAAR) return fail(c, SPR_ERR_ARG, "unknown algorithm");
    if (niter < 0) return fail(c, SPR_ERR_ARG, "negative iteration count");
    if (c->P <= 0) return SPR_OK;
    RowArgs a = row_args(c);
    a.beta = beta;
    a.pc = (alg == SPR_RAAR) ? cut_raar(reality) : cut_er_hio(reality);
    const int mode = (alg == SPR_HIO) ? ROW_HIO : (alg == SPR_RAAR) ? ROW_RAAR : ROW_ER;
    SPR_HIP(c, hipSetDevice(c->device));
    SPR_TRY(row(c, ROW_FWD, a));
    for (int i = 0; i < niter; ++i) {
        SPR_TRY(col(c, COL_PROJ, c->work));
        SPR_TRY(row(c, mode, a));
    }
    SPR_TRY(col(c, COL_PROJ, c->work));
    SPR_TRY(row(c, ROW_FINAL, a));
    return SPR_OK;
}

}  // namespace

extern "C" {

int spr_supported_size(int n) {
    return (n == 64 || n == 128 || n == 256 || n == 512 || n == 1024 || n == 2048) ? 1 : 0;
}

int spr_create(spr_ctx** out, int n, int capacity, const float* amp, int device) {
    if (!out) return fail(nullptr, SPR_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!spr_supported_size(n)) return fail(nullptr, SPR_ERR_ARG, "unsupported grid size " + std::to_string(n));
    if (capacity < 1) return fail(nullptr, SPR_ERR_ARG, "capacity must be >= 1");
    if (!amp) return fail(nullptr, SPR_ERR_ARG, "amp is NULL");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(nullptr, SPR_ERR_HIP, "no HIP device");
    if (device < 0 || device >= ndev) return fail(nullptr, SPR_ERR_ARG, "bad device id");
    spr_ctx* c = new spr_ctx();
    c->n = n;
    c->cap = capacity;
    c->device = device;
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
    if ((e = hipMalloc(&c->rho, np * capacity * sizeof(float2))) != hipSuccess) return cleanup_fail("hipMalloc rho", e);
    if ((e = hipMalloc(&c->work, np * capacity * sizeof(float2))) != hipSuccess)
        return cleanup_fail("hipMalloc work", e);
    if ((e = hipMalloc(&c->supp, np * capacity)) != hipSuccess) return cleanup_fail("hipMalloc supp", e);
    if ((e = hipMalloc(&c->amp, np * sizeof(float))) != hipSuccess) return cleanup_fail("hipMalloc amp", e);
    if ((e = hipMalloc(&c->tw, n * sizeof(float2))) != hipSuccess) return cleanup_fail("hipMalloc tw", e);
    std::vector<float2> tw(n);
    for (int k = 0; k < n; ++k) {
        const double ang = -2.0 * M_PI * (double)k / (double)n;
        tw[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    if ((e = hipMemcpy(c->tw, tw.data(), n * sizeof(float2), hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup_fail("upload twiddles", e);
    if ((e = hipMemcpy(c->amp, amp, np * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup_fail("upload amplitude", e);
    if ((e = hipMemset(c->supp, 0, np * capacity)) != hipSuccess) return cleanup_fail("clear supports", e);
    *out = c;
    return SPR_OK;
}

int spr_destroy(spr_ctx* c) {
    if (!c) return SPR_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    (void)hipFree(c->rho);
    (void)hipFree(c->work);
    (void)hipFree(c->supp);
    (void)hipFree(c->amp);
    (void)hipFree(c->tw);
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
    SPR_HIP(c, hipSetDevice(c->device));
    SPR_HIP(c, hipMemcpyAsync(c->amp, amp, (size_t)c->n * c->n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    SPR_HIP(c, hipStreamSynchronize(c->stream));
    return SPR_OK;
}

int spr_get_pattern(spr_ctx* c, float* amp) {
    if (!c || !amp) return fail(c, SPR_ERR_ARG, "NULL argument");
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
    c->timing = enable != 0;
    for (int k = 0; k < SPR_TIMING_KINDS; ++k) {
        c->kind_ms[k] = 0;
        c->kind_count[k] = 0;
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

}  // extern "C"
