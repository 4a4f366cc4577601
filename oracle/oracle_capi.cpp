// C entry points of the CPU oracle, for ctypes (oracle/oracle.py).
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py; never by the product.
//
// Complex fields are interleaved (re, im) arrays of n*n pairs; supports are int32.
// Suffix _f32 runs the FP32 oracle, _f64 the FP64 oracle (inputs/outputs in that type).
#include <omp.h>

#include <memory>

#include "fast_port.hpp"
#include "spring_oracle.hpp"

using namespace orc;

namespace {
template <class T> cplx<T>* C(T* p) { return reinterpret_cast<cplx<T>*>(p); }
template <class T> const cplx<T>* C(const T* p) { return reinterpret_cast<const cplx<T>*>(p); }
}  // namespace

#define ORC_DEFINE(T, SFX)                                                                                  \
    extern "C" int orc_fft2_##SFX(int n, T* f, int sign) {                                                  \
        Plan pl(n);                                                                                         \
        fft2_raw(pl, C(f), sign);                                                                           \
        return 0;                                                                                           \
    }                                                                                                       \
    extern "C" int orc_ft_##SFX(int n, T* f) {                                                              \
        Plan pl(n);                                                                                         \
        ft(pl, C(f));                                                                                       \
        return 0;                                                                                           \
    }                                                                                                       \
    extern "C" int orc_ift_##SFX(int n, T* f) {                                                             \
        Plan pl(n);                                                                                         \
        ift(pl, C(f));                                                                                      \
        return 0;                                                                                           \
    }                                                                                                       \
    extern "C" int orc_project_##SFX(int n, T* f, const T* amp) {                                           \
        project_modulus(C(f), amp, (size_t)n * n);                                                          \
        return 0;                                                                                           \
    }                                                                                                       \
    extern "C" int orc_apply_support_##SFX(int n, int alg, T* rho, const int32_t* S, T* buf, T beta,        \
                                           float reality) {                                                 \
        const size_t np = (size_t)n * n;                                                                    \
        if (alg == ALG_ER) apply_support_er(C(rho), S, np, reality);                                        \
        else if (alg == ALG_HIO) apply_support_hio(C(rho), S, C(buf), beta, np, reality);                   \
        else apply_support_raar(C(rho), S, C(buf), beta, np, reality);                                      \
        return 0;                                                                                           \
    }                                                                                                       \
    /* P individuals, each iterated independently (one per OpenMP thread, as the reference's Solver). */   \
    extern "C" int orc_iterate_##SFX(int n, int P, T* rho, const int32_t* S, const T* amp, int alg,         \
                                     int niter, T beta, float reality, int nthreads) {                      \
        Plan pl(n);                                                                                         \
        const size_t np = (size_t)n * n;                                                                    \
        if (nthreads <= 0) nthreads = omp_get_max_threads();                                                \
        _Pragma("omp parallel for schedule(dynamic) num_threads(nthreads)") for (int i = 0; i < P; ++i)     \
            iterate(pl, C(rho) + i * np, S + i * np, amp, alg, niter, beta, reality);                       \
        return 0;                                                                                           \
    }                                                                                                       \
    extern "C" int orc_blur_##SFX(int n, const T* rho, const T* amp, float sigma, T* b_out) {               \
        Plan pl(n);                                                                                         \
        std::vector<cplx<T>> b;                                                                             \
        shrinkwrap_blur(pl, C(rho), amp, sigma, b);                                                         \
        std::memcpy(b_out, b.data(), b.size() * sizeof(cplx<T>));                                           \
        return 0;                                                                                           \
    }                                                                                                       \
    extern "C" int orc_sw_##SFX(int n, const T* rho, int32_t* S, const T* amp, float sigma, float thr) {    \
        Plan pl(n);                                                                                         \
        sw(pl, C(rho), S, amp, sigma, thr);                                                                 \
        return 0;                                                                                           \
    }                                                                                                       \
    extern "C" int orc_swa_##SFX(int n, const T* rho, int32_t* S, const T* amp, float sigma, int spix) {    \
        Plan pl(n);                                                                                         \
        swa(pl, C(rho), S, amp, sigma, spix);                                                               \
        return 0;                                                                                           \
    }                                                                                                       \
    /* out: error_fourier, error_real, complexsum, modulesum */                                             \
    extern "C" int orc_get_error_##SFX(int n, const T* rho, const int32_t* S, const T* amp, float reality, \
                                       double* out) {                                                       \
        Plan pl(n);                                                                                         \
        ErrorOut<T> e = get_error(pl, C(rho), S, amp, reality);                                             \
        out[0] = e.error_fourier;                                                                           \
        out[1] = e.error_real;                                                                              \
        out[2] = e.complexsum;                                                                              \
        out[3] = e.modulesum;                                                                               \
        return 0;                                                                                           \
    }                                                                                                       \
    extern "C" int orc_get_symmetry_##SFX(int n, const T* rho, const T* amp, float symmin, float symmax,   \
                                          double* out) {                                                    \
        Plan pl(n);                                                                                         \
        out[0] = get_symmetry(pl, C(rho), amp, symmin, symmax);                                             \
        return 0;                                                                                           \
    }                                                                                                       \
    /* Population Reshift step: align rho to ref (get_shift with unit-scaled spectra). */                   \
    extern "C" int orc_get_shift_##SFX(int n, const T* rho, const T* ref, const T* amp, float minr,        \
                                       float maxr, int unit_scale_ft, int* out_i, float* out_phase) {        \
        Plan pl(n);                                                                                         \
        const size_t np = (size_t)n * n;                                                                    \
        std::vector<cplx<T>> rf(np);                                                                        \
        modulus_field(C(ref), rf.data(), np);                                                               \
        ft(pl, rf.data());                                                                                  \
        if (unit_scale_ft) unit_scale(rf.data(), np);                                                       \
        Shift s = get_shift(pl, C(rho), rf.data(), field_sum(C(ref), np), amp, minr, maxr,                  \
                            unit_scale_ft != 0);                                                            \
        out_i[0] = s.sx;                                                                                    \
        out_i[1] = s.sy;                                                                                    \
        out_i[2] = s.flip ? 1 : 0;                                                                          \
        out_phase[0] = s.phase;                                                                             \
        return 0;                                                                                           \
    }                                                                                                       \
    extern "C" int orc_reshift_##SFX(int n, T* rho, int32_t* S, int sx, int sy, float phase, int flip) {    \
        Shift s{sx, sy, phase, flip != 0};                                                                  \
        reshift(n, C(rho), S, s);                                                                           \
        return 0;                                                                                           \
    }                                                                                                       \
    extern "C" int orc_cross_child_##SFX(int n, const T* amp, const T* amp_avg, T* rho0, int32_t* S0,       \
                                         const T* rho1, const int32_t* S1, const T* rho2, const int32_t* S2, \
                                         const T* rho3, const int32_t* S3, const T* avg,                    \
                                         const int32_t* crossmap, const int* di, const float* df,           \
                                         float minr, float maxr, float reality, int update_support) {       \
        Plan pl(n);                                                                                         \
        CrossDesc d{di[0], di[1], di[2], di[3], di[4], di[5], df[0], df[1], df[2], di[6], di[7]};           \
        crossover_child(pl, amp, amp_avg, C(rho0), S0, C(rho1), S1, C(rho2), S2, C(rho3), S3, C(avg),       \
                        crossmap, d, minr, maxr, reality, update_support != 0);                             \
        return 0;                                                                                           \
    }

ORC_DEFINE(float, f32)
ORC_DEFINE(double, f64)

// Cell map of a crossover descriptor (int32 n*n), di = {csx, csy, sqx, sqy, offx, offy, crossft, crosssym},
// df = {rotangle, crossdiff, use_average}.
extern "C" int orc_cross_cells(int n, const int* di, const float* df, int32_t* cells) {
    CrossDesc d{di[0], di[1], di[2], di[3], di[4], di[5], df[0], df[1], df[2], di[6], di[7]};
    for (int iy = 0; iy < n; ++iy)
        for (int ix = 0; ix < n; ++ix) cells[ix + (size_t)iy * n] = cross_cell(n, ix, iy, d);
    return 0;
}

// Intensity -> corner-origin amplitude (Solver::set_boundaries + prepare).
extern "C" int orc_prepare_pattern(const float* intensity, int n, int cx, int cy, int use_bounds, float* amp) {
    prepare_pattern(intensity, n, cx, cy, use_bounds != 0, amp);
    return 0;
}

// Draws of `nchild` children, each from its own Rng seeded by seeds[i]
// (di_out: nchild x 11 ints {p1,p2,p3,csx,csy,sqx,sqy,offx,offy,crossft,crosssym};
//  df_out: nchild x 1 {rotangle}; crossmaps concatenated into cm_out, cm_len[i] entries each).
extern "C" int orc_crossover_draws(int nchild, const int* seeds, int P, const int* sorted_idx, float suppdim,
                                   int n, float crossexp, float crossftin, float crossbalance, int* di_out,
                                   float* df_out, int32_t* cm_out, int* cm_len, int cm_cap) {
    size_t off = 0;
    for (int c = 0; c < nchild; ++c) {
        Rng g(seeds[c]);
        int p1, p2, p3;
        CrossDesc d{};
        std::vector<int32_t> cm;
        crossover_draws(g, P, sorted_idx, suppdim, n, crossexp, crossftin, p1, p2, p3, d, cm, crossbalance);
        int* o = di_out + 11 * c;
        o[0] = p1; o[1] = p2; o[2] = p3; o[3] = d.crosssizex; o[4] = d.crosssizey; o[5] = d.squaresizex;
        o[6] = d.squaresizey; o[7] = d.offsetx; o[8] = d.offsety; o[9] = d.crossft; o[10] = d.crosssym;
        df_out[c] = d.rotangle;
        cm_len[c] = (int)cm.size();
        if (off + cm.size() > (size_t)cm_cap) return -1;
        std::copy(cm.begin(), cm.end(), cm_out + off);
        off += cm.size();
    }
    return 0;
}

// Draws of `nchild` consecutive children from ONE Rng seeded with `seed`, as the
// children that one reference context serves in turn (src/solver.cpp:742-801);
// outputs as orc_crossover_draws.
extern "C" int orc_crossover_draws_stream(int nchild, int seed, int P, const int* sorted_idx, float suppdim, int n,
                                          float crossexp, float crossftin, float crossbalance, int* di_out,
                                          float* df_out, int32_t* cm_out, int* cm_len, int cm_cap) {
    Rng g(seed);
    size_t off = 0;
    for (int c = 0; c < nchild; ++c) {
        int p1, p2, p3;
        CrossDesc d{};
        std::vector<int32_t> cm;
        crossover_draws(g, P, sorted_idx, suppdim, n, crossexp, crossftin, p1, p2, p3, d, cm, crossbalance);
        int* o = di_out + 11 * c;
        o[0] = p1; o[1] = p2; o[2] = p3; o[3] = d.crosssizex; o[4] = d.crosssizey; o[5] = d.squaresizex;
        o[6] = d.squaresizey; o[7] = d.offsetx; o[8] = d.offsety; o[9] = d.crossft; o[10] = d.crosssym;
        df_out[c] = d.rotangle;
        cm_len[c] = (int)cm.size();
        if (off + cm.size() > (size_t)cm_cap) return -1;
        std::copy(cm.begin(), cm.end(), cm_out + off);
        off += cm.size();
    }
    return 0;
}

// First `count` draws u ~ U(0,1) of a Context RNG seeded with `seed` (src/context.h:13-18).
extern "C" int orc_rng_uniform(int seed, int count, float* out) {
    Rng g(seed);
    for (int i = 0; i < count; ++i) out[i] = g.rand();
    return 0;
}

// Population initialisation: `count` consecutive individuals drawn from ONE Rng
// seeded `seed` (one reference context serving them in turn), fields in
// rho_out (count*n*n complex64) and S_out (count*n*n int32).
// fp = {itemsize0, itemsize1, exponent0, exponent1, scaling0, scaling1, phase}, ip = {nitems0, nitems1}.
extern "C" int orc_init_disk(int n, const float* amp, int seed, int count, int support_size, const float* fp,
                             const int* ip, float* rho_out, int32_t* S_out) {
    Plan pl(n);
    Rng g(seed);
    const size_t np = (size_t)n * n;
    for (int r = 0; r < count; ++r)
        init_disk(g, pl, amp, support_size, fp, ip, fp + 2, fp + 4, fp[6],
                  reinterpret_cast<cplx<float>*>(rho_out) + r * np, S_out + r * np);
    return 0;
}
extern "C" int orc_init_ref(int n, const float* ref_rho, const int32_t* ref_S, int seed, int count, const float* fp,
                            const int* ip, float* rho_out, int32_t* S_out) {
    Rng g(seed);
    const size_t np = (size_t)n * n;
    for (int r = 0; r < count; ++r)
        init_ref(g, n, reinterpret_cast<const cplx<float>*>(ref_rho), ref_S, fp, ip, fp + 2, fp + 4, fp[6],
                 reinterpret_cast<cplx<float>*>(rho_out) + r * np, S_out + r * np);
    return 0;
}
// fp = {initfrac, scaling0, scaling1, phase}
extern "C" int orc_init_supp(int n, const float* amp, const int32_t* support, int seed, int count, const float* fp,
                             float* rho_out, int32_t* S_out) {
    Plan pl(n);
    Rng g(seed);
    const size_t np = (size_t)n * n;
    for (int r = 0; r < count; ++r)
        init_supp(g, pl, amp, support, fp[0], fp + 1, fp[3], reinterpret_cast<cplx<float>*>(rho_out) + r * np,
                  S_out + r * np);
    return 0;
}

// Stateful context RNG handles: one stream serving several calls in turn, as a
// reference context serves Initialize_rec and then Crossover draws.
extern "C" void* orc_rng_create(int seed) { return new Rng(seed); }
extern "C" void orc_rng_destroy(void* g) { delete static_cast<Rng*>(g); }
extern "C" int orc_init_disk_rng(void* g, int n, const float* amp, int support_size, const float* fp, const int* ip,
                                 float* rho_out, int32_t* S_out) {
    Plan pl(n);
    init_disk(*static_cast<Rng*>(g), pl, amp, support_size, fp, ip, fp + 2, fp + 4, fp[6],
              reinterpret_cast<cplx<float>*>(rho_out), S_out);
    return 0;
}
extern "C" int orc_crossover_draws_rng(void* g, int P, const int* sorted_idx, float suppdim, int n, float crossexp,
                                       float crossftin, float crossbalance, int* di_out, float* df_out,
                                       int32_t* cm_out, int* cm_len, int cm_cap) {
    int p1, p2, p3;
    CrossDesc d{};
    std::vector<int32_t> cm;
    crossover_draws(*static_cast<Rng*>(g), P, sorted_idx, suppdim, n, crossexp, crossftin, p1, p2, p3, d, cm,
                    crossbalance);
    di_out[0] = p1; di_out[1] = p2; di_out[2] = p3; di_out[3] = d.crosssizex; di_out[4] = d.crosssizey;
    di_out[5] = d.squaresizex; di_out[6] = d.squaresizey; di_out[7] = d.offsetx; di_out[8] = d.offsety;
    di_out[9] = d.crossft; di_out[10] = d.crosssym;
    df_out[0] = d.rotangle;
    *cm_len = (int)cm.size();
    if (cm.size() > (size_t)cm_cap) return -1;
    std::copy(cm.begin(), cm.end(), cm_out);
    return 0;
}

// Throughput port of the FP32 driver (fast_port.hpp), for the CPU baseline.
extern "C" int orc_fft2_fast(int n, float* f, int sign) {
    try {
        fastport::LinePlan pl(n);
        fastport::Work w(n);
        fastport::fft2(pl, C(f), sign, w);
    } catch (const std::exception&) {
        return -1;
    }
    return 0;
}
extern "C" int orc_apply_support_fast(int n, int alg, float* rho, const int32_t* S, float* buf, float beta,
                                      float reality) {
    std::vector<uint8_t> undecided((size_t)n * n);
    fastport::update(alg, C(rho), S, alg == ALG_ER ? nullptr : C(buf), beta, (size_t)n * n, reality,
                     undecided.data());
    return 0;
}
extern "C" int orc_iterate_fast(int n, int P, float* rho, const int32_t* S, const float* amp, int alg, int niter,
                                float beta, float reality, int nthreads) {
    std::unique_ptr<fastport::LinePlan> pl;
    try {
        pl.reset(new fastport::LinePlan(n));
    } catch (const std::exception&) {
        return -1;
    }
    const size_t np = (size_t)n * n;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
    {
        fastport::Work w(n);
#pragma omp for schedule(dynamic)
        for (int i = 0; i < P; ++i) fastport::iterate(*pl, C(rho) + i * np, S + i * np, amp, alg, niter, beta, reality, w);
    }
    return 0;
}
