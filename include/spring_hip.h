/*
 * spring_hip.h -- C ABI of the MI355X (gfx950) backend of the SPRING memetic
 * phase-retrieval engine (libspring_hip.so).
 *
 * The reference's backend interface is the per-individual C++ class `Context`
 * (src/context.h:11-70), implemented for CUDA by `GPU_context`
 * (src/gpu/gpu_context.h:15-269, src/gpu/gpu_context.cu). Its only caller is
 * `Solver` (src/solver.cpp). This ABI replaces that interface with plain C
 * entry points that act on a whole population resident in HBM: one handle per
 * (pattern, device) instead of one Context per OpenMP thread. Every population
 * operation applied with P = 1 is the reference's single-individual operation.
 *
 * Conventions (identical to the reference):
 *   - fields are N x N, row-major, pixel p = x + N*y;
 *   - complex values are interleaved float pairs (re, im), i.e. std::complex<float>;
 *   - supports are int32 0/1 (SUPPORTTYPE, src/datatypes.h:11);
 *   - FT = unnormalised DFT with exp(+2 pi i k.x/N); IFT = exp(-2 pi i k.x/N) / (N*N)
 *     (src/gpu/gpu_context.cu:748-774);
 *   - the amplitude array has its origin in the corner and the reference's
 *     encoding: A >= 0 measured, A = -1 free, A <= -2 upper bound -A-2
 *     (src/solver.cpp:173-206).
 *
 * Error model: every function returns 0 on success and a negative code on
 * failure; spr_last_error() returns the message of the last failure on that
 * handle (or of the last failed spr_create when called with NULL). The
 * reference propagates no errors (cuFFT failures are printed and ignored,
 * src/gpu/gpu_context.cu:750-752); this ABI reports them instead.
 *
 * Threading: a handle is not thread-safe; distinct handles may be used
 * concurrently. All work of a handle runs on its own HIP stream (default
 * priority); every call returns after its work has completed unless stated.
 */
#ifndef SPRING_HIP_H
#define SPRING_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct spr_ctx spr_ctx;

enum spr_status {
    SPR_OK = 0,
    SPR_ERR_ARG = -1,      /* invalid argument (size, count, unsupported N) */
    SPR_ERR_HIP = -2,      /* HIP runtime failure */
    SPR_ERR_STATE = -3,    /* call not valid in the handle's state */
};

enum spr_alg { SPR_ER = 0, SPR_HIO = 1, SPR_RAAR = 2 };

/* Supported grid sizes N (spr_supported_size(N) != 0). */
int spr_supported_size(int n);

/* Creates a backend for one n x n pattern on `device` with room for `capacity`
 * individuals; uploads the corner-origin amplitude `amp` (n*n floats).
 * Replaces GPU_context::GPU_context (src/gpu/gpu_context.h:17-84), which the
 * reference constructs once per OpenMP thread (src/solver.cpp:32-39). */
int spr_create(spr_ctx** out, int n, int capacity, const float* amp, int device);

/* Replaces GPU_context::~GPU_context (src/gpu/gpu_context.h:87-137). */
int spr_destroy(spr_ctx* ctx);

/* Message of the last failure (never NULL). */
const char* spr_last_error(const spr_ctx* ctx);

/* Grid size and capacity of a handle. */
int spr_size(const spr_ctx* ctx, int* n, int* capacity);

/* Context::set_pattern / get_pattern (src/context.h:39-40,
 * src/gpu/gpu_context.cu:610-620, 669-674): replace / read the amplitude. */
int spr_set_pattern(spr_ctx* ctx, const float* amp);
int spr_get_pattern(spr_ctx* ctx, float* amp);

/* Context::set_density + set_support for P individuals at once
 * (src/context.h:23-24, src/gpu/gpu_context.cu:598-630): rho holds P*n*n
 * complex values, support P*n*n int32 (may be NULL to keep the supports). */
int spr_set_population(spr_ctx* ctx, int P, const float* rho, const int32_t* support);

/* Context::get_density + get_support (src/context.h:25-26,
 * src/gpu/gpu_context.cu:676-692) for the first P individuals; either output
 * may be NULL. */
int spr_get_population(spr_ctx* ctx, int P, float* rho, int32_t* support);

/* Context::get_ft / get_ift (src/context.h:21-22, src/gpu/gpu_context.cu:748-764)
 * applied to every resident individual's density. */
int spr_ft(spr_ctx* ctx);
int spr_ift(spr_ctx* ctx);

/* Context::HIO / RAAR / ER (src/context.h:44-47, src/gpu/gpu_context.cu:778-902)
 * on every resident individual: niter iterations of [FT; P_M; IFT; update]
 * followed by one trailing [FT; P_M; IFT]. beta is ignored for SPR_ER. */
int spr_iterate(spr_ctx* ctx, int alg, int niter, float beta, float reality);

/* Same as spr_iterate but returns as soon as the work is queued on the
 * handle's stream (for timing loops); spr_sync waits for it. */
int spr_iterate_async(spr_ctx* ctx, int alg, int niter, float beta, float reality);
int spr_sync(spr_ctx* ctx);

/* Device addresses of the resident population: rho (P*n*n float2) and the
 * uint8 supports; for callers that fill or read HBM directly. */
int spr_device_buffers(spr_ctx* ctx, void** rho, void** support_u8, void** amp);

/* Opt-in kernel timing. With timing enabled, every launch of the handle is
 * bracketed by HIP events on its stream; after each synchronising call the
 * durations are added per kind. spr_timing() also resets the totals. */
enum spr_timing_kind {
    SPR_KIND_COL_PROJ = 0,     /* column pass colIFFT(P_M(colFFT)) */
    SPR_KIND_ROW_UPDATE = 1,   /* row pass rowIFFT -> HIO/RAAR/ER update -> rowFFT */
    SPR_KIND_ROW_OTHER = 2,    /* other row passes (first/last of an operation) */
    SPR_KIND_COL_OTHER = 3,    /* other column passes */
    SPR_TIMING_KINDS = 4
};
int spr_timing(spr_ctx* ctx, int enable);
int spr_timing_read(spr_ctx* ctx, int kind, double* total_ms, long* count);

/* HIP stream of the handle (a hipStream_t). */
int spr_stream(spr_ctx* ctx, void** stream);

#ifdef __cplusplus
}
#endif

#endif /* SPRING_HIP_H */
