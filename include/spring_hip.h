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
 * handle (or, called with NULL, of the last failed call without a handle). The
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

/* Supported grid sizes N (spr_supported_size(N) != 0): 64, 128, 256, 384, 512,
 * 768, 1024 and 2048. */
int spr_supported_size(int n);

/* Device enumeration, replacing GPU_utils::get_count / get_info
 * (src/gpu/gpu_utils.cu:13-111): number of visible devices (0 without a GPU)
 * and the name of one. */
int spr_device_count(int* n);
int spr_device_name(int device, char* buf, int len);

/* Creates a backend for one n x n pattern on `device` with room for `capacity`
 * individuals; uploads the corner-origin amplitude `amp` (n*n floats).
 * Replaces GPU_context::GPU_context (src/gpu/gpu_context.h:17-84), which the
 * reference constructs once per OpenMP thread (src/solver.cpp:32-39). */
int spr_create(spr_ctx** out, int n, int capacity, const float* amp, int device);

/* Replaces GPU_context::~GPU_context (src/gpu/gpu_context.h:87-137). */
int spr_destroy(spr_ctx* ctx);

/* Message of the last failure (never NULL). With ctx == NULL: the last failure
 * of this thread's calls that take no handle (spr_create,
 * spr_loss_grad_noalloc). */
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

/* Slot-level residency: the handle holds `capacity` slots; operations act on
 * the active slots [0, P). These replace the per-individual set_density /
 * set_support / get_density / get_support of one Context per OpenMP thread
 * (src/context.h:23-26, src/solver.cpp:360-411) so that a caller can keep
 * individuals resident in HBM between operations instead of copying every
 * individual in and out around each one.
 *
 * spr_set_active: the active population size P (no data moves).
 * spr_gather: slot i <- slot src[i] for 0 <= i < count (density and support);
 *   src[i] = -1 or i leaves slot i alone; all sources are read before any slot
 *   is written, so src may be any mapping.
 * spr_put_individuals: upload individual k to slot slots[k] from separate host
 *   arrays (rho[k]: n*n complex, support[k]: n*n int32); either pointer array,
 *   or an entry of rho, may be NULL to leave that part.
 * spr_fetch_individuals: download slot slots[k] into rho[k] / support[k]
 *   (NULL arrays or entries are skipped). */
int spr_set_active(spr_ctx* ctx, int P);
int spr_gather(spr_ctx* ctx, int count, const int* src);
int spr_put_individuals(spr_ctx* ctx, int count, const int* slots, const float* const* rho,
                        const int32_t* const* support);
int spr_fetch_individuals(spr_ctx* ctx, int count, const int* slots, float* const* rho, int32_t* const* support);

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

/* Context::SW (src/context.h:49, src/gpu/gpu_context.cu:906-953) on every
 * resident individual: S <- [|b| >= max|b| * threshold], with b the Gaussian
 * (sigma) blur of |IFT(P_M(FT rho))|. The densities are not changed. `reality`
 * is accepted for signature parity and unused, as in the reference. */
int spr_sw(spr_ctx* ctx, float sigma, float threshold, float reality);

/* Context::SWA (src/context.h:50, src/gpu/gpu_context.cu:957-1027): as
 * spr_sw, with the threshold set to the `spixels`-th largest |b| of each
 * individual (exact selection on the device; the reference sorts on the host).
 * spixels <= 0 yields empty supports (out of bounds in the reference). */
int spr_swa(spr_ctx* ctx, float sigma, int spixels, float reality);

/* Context::get_error + get_symmetry(sym, 1, 0, 1) for every resident
 * individual (src/gpu/gpu_context.cu:1030-1134, 1378-1463), as called by
 * Solver::Sequence / Evaluate (src/solver.cpp:474-475, 549-550). Outputs are
 * arrays of P floats; any may be NULL (symmetry == NULL skips its cost).
 * Sums are accumulated in FP64 in a fixed order. */
int spr_evaluate(spr_ctx* ctx, float reality, float* error_fourier, float* error_real, float* complexsum,
                 float* modulesum, float* symmetry);

/* Context::get_symmetry(sym, 1, symmin, symmax) for every resident individual. */
int spr_symmetry(spr_ctx* ctx, float symmin, float symmax, float* symmetry);

/* One operation of an algorithm sequence (the closures Solver::get_algorithm
 * builds, src/solver.cpp:84-159). */
enum spr_op_kind { SPR_OP_ER = 0, SPR_OP_HIO = 1, SPR_OP_RAAR = 2, SPR_OP_SW = 3, SPR_OP_SWA = 4, SPR_OP_GRAD = 5 };
typedef struct spr_op {
    int kind;          /* spr_op_kind */
    int iterations;    /* ER/HIO/RAAR: iterations; GRAD: Adam steps */
    float beta;        /* HIO/RAAR: beta; GRAD: step relative to the RMS density (spr_refine_support) */
    float sigma;       /* SW/SWA */
    float threshold;   /* SW: threshold; GRAD: loss eps */
    int pixels;        /* SWA: support pixels; GRAD: loss kind (spr_loss_grad) */
} spr_op;

/* The device half of Solver::Sequence (src/solver.cpp:454-487): nrep
 * repetitions of the op list on every resident individual, queued on the
 * handle's stream and synchronised once at the end. */
int spr_run(spr_ctx* ctx, const spr_op* ops, int nops, int nrep, float reality);

/* Solver::Reshift on the device (src/solver.cpp:511-538): for every resident
 * individual, set_ref_density(ref) + get_shift(minr, maxr) + reshift
 * (src/gpu/gpu_context.cu:632-667, 1226-1371). ref_rho: the reference
 * density (n*n complex, host). phase_direct[p] / phase_flip[p]: the shift
 * phases the reference forms from float sums of the host densities
 * (density_ref_phase -/+ density_phase, :1309-1321). shifts (optional, 3*P
 * int32): the chosen (shiftx, shifty, flip) per individual. */
int spr_align(spr_ctx* ctx, const float* ref_rho, float minr, float maxr, const float* phase_direct,
              const float* phase_flip, int32_t* shifts);

/* Solver::GetAverage on the device (src/solver.cpp:618-686): the mean density
 * of the resident individuals listed in idx (navg entries, summed in that
 * order) and the union of their supports. rho_out: n*n complex, supp_out: n*n int32. */
int spr_average(spr_ctx* ctx, const int32_t* idx, int navg, float* rho_out, int32_t* supp_out);

/* Solver::GetAverageModule (src/solver.cpp:564-614): mean of |FT rho| (ft != 0)
 * or of |rho| over the listed individuals. out: n*n floats. */
int spr_average_module(spr_ctx* ctx, const int32_t* idx, int navg, int ft, float* out);

/* One child of Solver::Crossover (src/solver.cpp:742-814): parents are
 * resident individuals; p0 is the individual the child replaces. The random
 * draws are the caller's (host RNG, as in the reference). */
typedef struct spr_cross_desc {
    int p0, p1, p2, p3;
    int crossft, crosssym;
    int crosssizex, crosssizey, squaresizex, squaresizey, offsetx, offsety;
    float rotangle;
    int crossmap_offset;   /* first entry of this child's crossmap in the concatenated array */
    /* alignment phases to p1 (reshift_parent, src/gpu/gpu_context.cu:1536-1654) of
     * the roles p0, p2, p3, avg: {direct, flip} each */
    float phase[8];
} spr_cross_desc;

/* prepare_parents + get_new_individual (src/gpu/gpu_context.cu:1661-1694,
 * 2044-2117) for nchild children at once. avg_rho: the population average
 * density (n*n complex); amp_avg: GetAverageModule(ft, 0.66) (n*n floats).
 * Outputs: child densities (nchild*n*n complex) and the crossed supports
 * (nchild*n*n int32). update_support is accepted for signature parity and not
 * used here: as in the reference (src/solver.cpp:810), the caller keeps p0's
 * support when it is 0. */
int spr_crossover(spr_ctx* ctx, int nchild, const spr_cross_desc* desc, const int32_t* crossmaps, int crossmap_len,
                  const float* avg_rho, const float* amp_avg, float crossdiff, float use_average, float minr,
                  float maxr, float reality, int update_support, float* child_rho, int32_t* child_supp);

/* Masked Fourier-magnitude loss and its gradient (no counterpart in the
 * reference; the gradient-refinement operator of the memetic loop). For
 * F = FT(rho) and the measured pixels (A >= 0, counts I = A^2), u = |F|^2:
 *   kind 0 (Poisson NLL)  L = sum u - I log(u + eps)
 *   kind 1 (amplitude)    L = sum (|F| - A)^2
 * grad = dL/dRe(rho) + i dL/dIm(rho) = 2 FT^H[f'(u) F] (PyTorch's convention for
 * complex tensors). rho and grad are device arrays of P fields (P <= capacity);
 * grad may be NULL for the losses alone (the adjoint pass is skipped); loss
 * (optional, host, P doubles, FP64 sums). Runs on `stream` (a hipStream_t;
 * NULL: the handle's stream) and returns after the work is queued unless loss
 * is requested. */
int spr_loss_grad(spr_ctx* ctx, int P, const float* rho, float* grad, double* loss, int kind, float eps,
                  void* stream);

/* spr_loss_grad without a handle and without any allocation. No counterpart in
 * the reference: this entry point is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). It calls no hipMalloc, hipFree or allocating library
 * routine and never synchronises.
 *   n, P   grid size (spr_supported_size) and number of fields;
 *   amp    device, n*n floats: the corner-origin amplitude with the mask
 *          encoding of spr_create;
 *   rho    device, P*n*n complex: the densities (read only);
 *   grad   device, P*n*n complex, not overlapping rho: the gradient of
 *          spr_loss_grad when want_grad != 0; it is also the intermediate
 *          field of the passes, so with want_grad == 0 it holds no gradient;
 *   loss   device, P doubles (FP64 sums), or NULL;
 *   stream the hipStream_t to queue on (NULL: the null stream); the current
 *          device is the one used.
 * Results are identical to spr_loss_grad's (same kernels). Its only scratch
 * (loss partials per column workgroup, twiddle roots) is a fixed device array
 * of 1.04 MiB per GPU, held until the process ends, whose address and roots
 * are set up by the first call on each device (synchronously; make that first
 * call outside graph capture). The partials limit P * n / C <= 131072 (C =
 * columns per column-pass workgroup): at most 5461 fields of 384^2, 4096 of
 * 512^2, 1365 of 768^2, 1024 of 1024^2, 256 of 2048^2; larger counts are
 * rejected with SPR_ERR_ARG. Every call on one GPU shares that GPU's array:
 * calls that overlap in time on one GPU (on different streams or host threads)
 * race on it and must be serialised by the caller. Errors: spr_last_error(NULL). */
int spr_loss_grad_noalloc(int n, int P, const float* amp, const float* rho, float* grad, double* loss, int kind,
                          float eps, int want_grad, void* stream);

/* `steps` Adam steps (lr; beta 0.9/0.999, eps 1e-8, moments reset per call)
 * of the loss above on every resident density. loss (optional, host, P
 * doubles): the loss after the last step. */
int spr_refine(spr_ctx* ctx, int steps, float lr, int kind, float eps, double* loss);

/* Support-projected refinement (the GRAD operation of spr_run): projected
 * Adam on the loss above over densities that vanish outside the support. Each
 * density is first set to 0 outside its individual's support, then refined as
 * by spr_refine with the gradient masked by the support (pixels outside stay
 * 0); the step size of individual p is lr_rel * sqrt(sum_S |rho_p|^2 / |S_p|),
 * the RMS density on its support when the call starts. No counterpart in the
 * reference (new operator). */
int spr_refine_support(spr_ctx* ctx, int steps, float lr_rel, int kind, float eps, double* loss);

/* Device addresses of the resident population: rho (P*n*n float2) and the
 * uint8 supports; for callers that fill or read HBM directly. The amplitude
 * (n*n float) is for reading only: the context also keeps a transposed copy,
 * so change it through spr_set_pattern. */
int spr_device_buffers(spr_ctx* ctx, void** rho, void** support_u8, void** amp);

/* Opt-in kernel timing. spr_timing(ctx, k) with k >= 1 brackets every k-th
 * launch of each kind by HIP events on the handle's stream (k = 1: every
 * launch; a larger k keeps the events' own cost out of the timed work); after
 * each synchronising call the durations are added per kind. k = 0 turns timing
 * off. spr_timing() also resets the totals. */
enum spr_timing_kind {
    SPR_KIND_COL_PROJ = 0,     /* column pass colIFFT(P_M(colFFT)) */
    SPR_KIND_ROW_UPDATE = 1,   /* row pass rowIFFT -> HIO/RAAR/ER update -> rowFFT */
    SPR_KIND_ROW_OTHER = 2,    /* other row passes (first/last of an operation) */
    SPR_KIND_COL_OTHER = 3,    /* other column passes */
    SPR_KIND_BLOCK = 4,        /* whole iteration blocks on chip (64^2, 128^2; spr_set_fused) */
    SPR_TIMING_KINDS = 5
};
int spr_timing(spr_ctx* ctx, int enable);
int spr_timing_read(spr_ctx* ctx, int kind, double* total_ms, long* count);

/* HIP stream of the handle (a hipStream_t). */
int spr_stream(spr_ctx* ctx, void** stream);

/* ---------------------------------------------------------------------------
 * Kernel variants and the autotuner. No counterpart in the reference, whose
 * transforms are cuFFT plans (src/gpu/gpu_context.h:29-32).
 *
 * The column pass (colFFT -> P_M -> colIFFT) and the row pass (rowIFFT ->
 * update -> rowFFT) of each grid size are compiled in several variants:
 * columns per workgroup, values per thread (radix plan), threads and prefetch
 * of the row pass. Variant 0 is the default geometry; all variants read and
 * write the same layout. Width, threads and prefetch run the same operations
 * (up to the compiler's contraction of multiply-adds); another radix plan
 * changes the rounding of the transforms. Every variant meets the same
 * accuracy bar against the oracle. Some row variants recompute the HIO/RAAR
 * state rho_n in each iteration from its row transform, which the iteration
 * keeps anyway, instead of reading and writing it (8 fewer bytes per pixel
 * and iteration, for one more row transform in HIO and none in RAAR); the
 * column pass then writes its projection to a second field buffer. The
 * recomputed state agrees with the stored one to float rounding.
 *
 * spr_variants: number of column and row variants of size n.
 * spr_row_variant_info: whether row variant `variant` of size n recomputes
 *   the state (1) or stores it (0).
 * spr_autotune: times `reps` iterations (a column pass, then a RAAR row
 *   pass) with every variant on the handle's resident population (densities
 *   and supports are left unchanged): each column variant with the handle's
 *   row variant, then each row variant with the chosen column variant. It
 *   keeps the fastest pair (per pass, the lowest-numbered variant within
 *   1.5 % of the fastest) for this handle and makes it the default of handles
 *   of that size created later in the process; col_ms / row_ms (NULL or one
 *   entry per variant) receive the mean iteration durations in ms.
 * spr_set_variants / spr_get_variants: one handle's pair.
 * spr_default_variants / spr_set_default_variants: the pair new handles of
 *   size n start with (initially the table compiled from the last autotuning
 *   run on the target GPU, spr_tuned.hpp).
 * spr_batch_autotune: spr_autotune on every size class of a batch. */
int spr_variants(int n, int* col_count, int* row_count);
int spr_row_variant_info(int n, int variant, int* recomputes_state);

/* Chunked iteration blocks. No counterpart in the reference. An ER/HIO/RAAR
 * block runs over the population in chunks whose densities, fields and
 * supports (17 bytes per pixel) fit in `bytes`, each chunk through all of the
 * block's iterations before the next, so that a chunk's working set stays in
 * the last-level cache (the 256 MB Infinity Cache of MI355X) instead of every
 * pass streaming the whole population from HBM. A population whose working
 * set fits runs whole; a larger one runs in the fewest equal chunks that fit.
 * The results are identical to an unchunked block. bytes = 0 runs the whole
 * population at once. New handles start with the compiled default (300 MB) or
 * SPR_CHUNK_MB from the environment; spr_get_chunk also returns the
 * individuals per chunk. */
int spr_set_chunk(spr_ctx* ctx, long long bytes);
int spr_get_chunk(spr_ctx* ctx, long long* bytes, int* individuals);
/* On-chip iteration blocks. No counterpart in the reference. For 64^2 and
 * 128^2 fields an ER/HIO/RAAR block can run as one launch with one workgroup
 * per individual holding the whole field in LDS, so that HBM is read at the
 * block's start and written at its end only. mode -1 (the default) uses it from
 * a population of 8 individuals, 0 never, 1 always (SPR_ERR_ARG for other
 * sizes). The results agree with the two-pass form within float rounding (the
 * transforms use other radix plans). spr_get_fused returns the mode and whether
 * the handle's next block would run on chip. */
int spr_set_fused(spr_ctx* ctx, int mode);
int spr_get_fused(spr_ctx* ctx, int* mode, int* active);
int spr_autotune(spr_ctx* ctx, int reps, float* col_ms, float* row_ms);
int spr_set_variants(spr_ctx* ctx, int col, int row);
int spr_get_variants(spr_ctx* ctx, int* col, int* row);
int spr_default_variants(int n, int* col, int* row);
int spr_set_default_variants(int n, int col, int row);

/* ---------------------------------------------------------------------------
 * Batches of independent patterns of several sizes on one device.
 *
 * No counterpart in the reference, which reconstructs one pattern per Solver
 * (src/solver.cpp:9-52) and runs several patterns as separate processes. A
 * batch holds `npatterns` patterns (sizes n[i], any supported size, mixed
 * freely), each with population[i] >= 1 individuals. Individuals of all
 * patterns of one size share one set of launches: every column and row pass
 * covers the whole size class, reading each individual's own amplitude.
 * spr_batch_run applies the op list of spr_run to every individual of every
 * pattern; an ER/HIO/RAAR block runs as one HIP graph launch whose branches
 * (one per size) run one after another, or side by side after
 * spr_batch_concurrent(b, 1), captured on first use of that block's
 * (algorithm, iterations, beta, reality) and replayed afterwards;
 * spr_batch_graphs(b, 0) issues the same launches directly instead.
 * Populations start at rho = 0 and empty supports; spr_batch_set_population
 * and spr_batch_get_population copy one pattern's population[i] individuals
 * (layouts as spr_set_population). spr_batch_evaluate fills the metrics of
 * spr_evaluate (no symmetry) for all individuals, pattern by pattern, in
 * arrays of sum(population) entries. */
typedef struct spr_batch spr_batch;
int spr_batch_create(spr_batch** out, int npatterns, const int* n, const int* population, const float* const* amp,
                     int device);
int spr_batch_destroy(spr_batch* b);
const char* spr_batch_last_error(const spr_batch* b);
int spr_batch_graphs(spr_batch* b, int enable);
int spr_batch_concurrent(spr_batch* b, int enable);
int spr_batch_set_population(spr_batch* b, int pattern, const float* rho, const int32_t* support);
int spr_batch_get_population(spr_batch* b, int pattern, float* rho, int32_t* support);
int spr_batch_run(spr_batch* b, const spr_op* ops, int nops, int nrep, float reality);
int spr_batch_evaluate(spr_batch* b, float reality, float* error_fourier, float* error_real, float* complexsum,
                       float* modulesum);
int spr_batch_autotune(spr_batch* b, int reps);
/* spr_set_chunk on every size class of a batch. */
int spr_batch_set_chunk(spr_batch* b, long long bytes);

#ifdef __cplusplus
}
#endif

#endif /* SPRING_HIP_H */
