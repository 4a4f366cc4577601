// gfx950 kernels of the population-level operations around the iteration:
// shrink-wrap (SW/SWA), fitness (get_error), symmetry and the Fourier
// cross-correlations of alignment. They reuse the row/column pass structure of
// spr_kernels.hpp; reductions are deterministic (per-workgroup partials reduced
// in a fixed order, FP64 sums) and maxima use order-independent integer atomics
// on the bit patterns of non-negative floats.
#pragma once

#include "spr_kernels.hpp"

namespace spr {

// ---------------------------------------------------------------------------
// Wave64 / workgroup reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// Max of non-negative values; NaN inputs are ignored (a > b is false for NaN).
__device__ __forceinline__ float vmax(float a, float b) { return b > a ? b : a; }
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = vmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// Sum of K doubles over the workgroup; result valid in thread 0. red has
// K * (blockDim/64) slots. Fixed reduction order.
template <int K, int NT>
__device__ __forceinline__ void block_sum(double (&v)[K], double* red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) red[wave * K + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0;
            for (int w = 0; w < NT / 64; ++w) s += red[w * K + k];
            v[k] = s;
        }
    }
    __syncthreads();
}

template <int NT>
__device__ __forceinline__ float block_max(float v, float* red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    v = wave_max(v);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < NT / 64; ++w) v = vmax(v, red[w]);
    __syncthreads();
    return v;
}

template <int NT>
__device__ __forceinline__ unsigned long long block_max_u64(unsigned long long v, unsigned long long* red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    v = wave_max_u64(v);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < NT / 64; ++w) v = red[w] > v ? red[w] : v;
    __syncthreads();
    return v;
}

// ---------------------------------------------------------------------------
// Spectral windows
// ---------------------------------------------------------------------------

// Shrink-wrap Gaussian g(k) = exp(-(kx^2+ky^2)/sigma_ft^2), k wrapped to
// [-N/2, N/2) (smooth_intensities, src/gpu/gpu_context.cu:199-217).
template <int N> __device__ __forceinline__ float sw_gauss(int ix, int iy, float sigma_ft) {
    int xx = ix, yy = iy;
    if (xx >= N / 2) xx -= N;
    if (yy >= N / 2) yy -= N;
    const float r2 = (float)(xx * xx + yy * yy);
    return expf(-r2 / (sigma_ft * sigma_ft));
}

// Radial band window of the cross-correlations (multiply(..., minr, maxr, sign),
// src/gpu/gpu_context.cu:1157-1200); the constants are precomputed on the host
// exactly as the reference forms them.
struct CorrWin {
    float sigtot_minus_sig;      // sigmatot - sigma      (minr side)
    float sig2;                  // sigma * sigma
    float sigtotmax_minus_sigmax;
    float sigmax2;
};
template <int N> __device__ __forceinline__ float corr_weight(int ix, int iy, const CorrWin& w) {
    float x = (float)ix, y = (float)iy;
    if (x >= N / 2) x -= N;
    if (y >= N / 2) y -= N;
    const float radius = sqrtf(x * x + y * y);
    const float eff = radius - w.sigtot_minus_sig;
    float coeff = eff > 0 ? (float)(1. - (double)expf(-eff * eff / w.sig2)) : 0.f;
    const float effmax = radius - w.sigtotmax_minus_sigmax;
    const float coeffmax = effmax > 0 ? expf(-effmax * effmax / w.sigmax2) : 1.f;
    coeff *= coeffmax;
    return coeff * coeff;
}

// ---------------------------------------------------------------------------
// Auxiliary row passes
// ---------------------------------------------------------------------------
enum AuxRowMode {
    AROW_ABS_FWD_WORK = 0,   // work <- rowFFT+( |rowFFT-(work)/N^2| )            (SW: a = |rho'|, start of FT(a))
    AROW_MAG_MAX = 1,        // mag <- |rowFFT-(work)| ; maxbits[p] = max         (SW: |b| and its maximum)
    AROW_ERR_FWD = 2,        // rho_S = S'(rho) rho; msum partial; work <- rowFFT+(rho_S)   (get_error)
    AROW_ERR_REAL = 3,       // rho'' = rowFFT-(work)/N^2; real-space error partials       (get_error)
    AROW_ABS_FWD_RHO = 4,    // work <- rowFFT+(|src|)                               (moduli spectra)
    AROW_CORR = 5,           // corr{1,2}: key[p] = max(|rowFFT-(c)|^2, first index) (alignment, symmetry)
};

struct AuxRowArgs {
    const float2* src;       // AROW_ABS_FWD_RHO / AROW_ERR_FWD input fields (base of a batch)
    float2* work;
    float2* work2;
    const uint8_t* supp;
    const float2* tw;
    float* mag;
    unsigned* maxbits;                   // per individual
    unsigned long long* key1;            // per item (AROW_CORR)
    unsigned long long* key2;
    double* partial;                     // per workgroup
    float inv_nn;
    PhaseCut pc;
    const int* src_index;                // optional indirection item -> source field (AROW_ABS_FWD_RHO)
};

template <int N, int MODE>
__global__ __launch_bounds__(Geo<N>::ROW_THREADS) void k_row_aux(AuxRowArgs a) {
    using G = Geo<N>;
    constexpr int V = G::V, T = G::T, NT = G::ROW_THREADS;
    constexpr int WG_PER_FIELD = N / G::RW;
    __shared__ float2 tw[TwTable<N>::SIZE];
    __shared__ float2 lds[G::RW * G::RS];
    __shared__ double red[4 * (NT / 64)];
    const int tid = threadIdx.x;
    const int t = tid % T;
    const int rl = tid / T;
    tw_fill<N>(tw, a.tw, tid, NT);
    const int item = blockIdx.x / WG_PER_FIELD;
    const int y = (blockIdx.x % WG_PER_FIELD) * G::RW + rl;
    const size_t base = (size_t)item * N * N + (size_t)y * N;
    const RowAcc acc{lds + rl * G::RS};
    float2 v[V];
    __syncthreads();

    if constexpr (MODE == AROW_ABS_FWD_WORK) {
#pragma unroll
        for (int r = 0; r < V; ++r) v[r] = a.work[base + t + r * T];
        fft_line<N, -1, (G::T <= 64)>(v, acc, tw, t);
#pragma unroll
        for (int q = 0; q < V; ++q) {
            const float x = v[q].x * a.inv_nn, yv = v[q].y * a.inv_nn;
            v[q] = make_float2(sqrtf(x * x + yv * yv), 0.f);
        }
        fft_line<N, +1, (G::T <= 64)>(v, acc, tw, t);
#pragma unroll
        for (int q = 0; q < V; ++q) a.work[base + t + q * T] = v[q];
    } else if constexpr (MODE == AROW_MAG_MAX) {
#pragma unroll
        for (int r = 0; r < V; ++r) v[r] = a.work[base + t + r * T];
        fft_line<N, -1, (G::T <= 64)>(v, acc, tw, t);
        float m = 0.f;
#pragma unroll
        for (int q = 0; q < V; ++q) {
            const float mm = sqrtf(v[q].y * v[q].y + v[q].x * v[q].x);
            a.mag[base + t + q * T] = mm;
            m = vmax(m, mm);
        }
        m = block_max<NT>(m, reinterpret_cast<float*>(red));
        if (tid == 0) atomicMax(&a.maxbits[item], __float_as_uint(m));
    } else if constexpr (MODE == AROW_ERR_FWD) {
        double ms[1] = {0.0};
#pragma unroll
        for (int r = 0; r < V; ++r) {
            const size_t i = base + t + r * T;
            float2 d = a.src[i];
            const float ts = (a.supp[i] != 0 && !phase_fails(d.x, d.y, a.pc)) ? 1.f : 0.f;
            d = make_float2(d.x * ts, d.y * ts);
            const float m = sqrtf(d.y * d.y + d.x * d.x);
            if (m > 0.f) ms[0] += (double)m;
            v[r] = d;
        }
        block_sum<1, NT>(ms, red);
        if (tid == 0) a.partial[blockIdx.x] = ms[0];
        fft_line<N, +1, (G::T <= 64)>(v, acc, tw, t);
#pragma unroll
        for (int q = 0; q < V; ++q) a.work[base + t + q * T] = v[q];
    } else if constexpr (MODE == AROW_ERR_REAL) {
#pragma unroll
        for (int r = 0; r < V; ++r) v[r] = a.work[base + t + r * T];
        fft_line<N, -1, (G::T <= 64)>(v, acc, tw, t);
        double s[2] = {0.0, 0.0};  // error, norm
#pragma unroll
        for (int q = 0; q < V; ++q) {
            const size_t i = base + t + q * T;
            const float dx = v[q].x * a.inv_nn, dy = v[q].y * a.inv_nn;
            const uint8_t sp = a.supp[i];
            const float ts = (sp != 0 && !phase_fails(dx, dy, a.pc)) ? 1.f : 0.f;
            const float sx = dx * ts, sy = dy * ts;
            s[1] += (double)(dy * dy + dx * dx);
            if (sp == 0) {
                const float e1 = dy - sy, e2 = dx * sx;
                s[0] += (double)(e1 * e1 + e2 * e2);
            }
        }
        block_sum<2, NT>(s, red);
        if (tid == 0) {
            a.partial[2 * blockIdx.x] = s[0];
            a.partial[2 * blockIdx.x + 1] = s[1];
        }
    } else if constexpr (MODE == AROW_ABS_FWD_RHO) {
        const size_t sbase = (size_t)(a.src_index ? a.src_index[item] : item) * N * N + (size_t)y * N;
#pragma unroll
        for (int r = 0; r < V; ++r) {
            const float2 d = a.src[sbase + t + r * T];
            v[r] = make_float2(sqrtf(d.x * d.x + d.y * d.y), 0.f);
        }
        fft_line<N, +1, (G::T <= 64)>(v, acc, tw, t);
#pragma unroll
        for (int q = 0; q < V; ++q) a.work[base + t + q * T] = v[q];
    } else if constexpr (MODE == AROW_CORR) {
        // two fields (work, work2) per item; first index of the strict maximum of |c|^2
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            const float2* src = f == 0 ? a.work : a.work2;
#pragma unroll
            for (int r = 0; r < V; ++r) v[r] = src[base + t + r * T];
            fft_line<N, -1, (G::T <= 64)>(v, acc, tw, t);
            unsigned long long best = 0;
#pragma unroll
            for (int q = 0; q < V; ++q) {
                const float val = v[q].x * v[q].x + v[q].y * v[q].y;
                if (val > 0.f) {
                    const unsigned idx = (unsigned)(y * N + t + q * T);
                    const unsigned long long k =
                        ((unsigned long long)__float_as_uint(val) << 32) | (unsigned long long)(0xFFFFFFFFu - idx);
                    best = k > best ? k : best;
                }
            }
            best = block_max_u64<NT>(best, reinterpret_cast<unsigned long long*>(red));
            if (tid == 0 && best) atomicMax(f == 0 ? &a.key1[item] : &a.key2[item], best);
        }
    }
}

// ---------------------------------------------------------------------------
// Auxiliary column passes
// ---------------------------------------------------------------------------
enum AuxColMode {
    ACOL_BLUR = 0,       // colFFT-( g * colFFT+(.) )                         (SW)
    ACOL_ERR = 1,        // F = colFFT+(.); error partials, max|F|; colFFT-(P_M(F))  (get_error)
    ACOL_CORR = 2,       // d = colFFT+(work[item]); r = ref spectrum;
                         // work <- colFFT-(win conj(d) r), work2 <- colFFT-(win d r)  (alignment, symmetry)
    ACOL_FWD_UNIT = 3,   // colFFT+(.) then z *= |z|/|z|  (normalize_ft_shift with e = 1)
};

struct AuxColArgs {
    float2* f;
    float2* f2;
    const float* amp;           // amplitudes [patterns][N][N]
    const int* amp_of;          // item -> pattern (NULL: pattern 0)
    const float2* tw;
    float sigma_ft;
    double* partial;
    unsigned* maxbits;
    const float2* ref;          // ACOL_CORR reference spectra (base of a batch)
    const int* ref_index;       // item -> reference spectrum (NULL: the item's own transformed field)
    CorrWin win;
    int unit_scale;             // ACOL_CORR: apply normalize_ft_shift to d first
};

template <int N, int MODE>
__global__ __launch_bounds__(Geo<N>::COL_THREADS) void k_col_aux(AuxColArgs a) {
    using G = Geo<N>;
    constexpr int V = G::V, T = G::T, C = G::C, NT = G::COL_THREADS;
    __shared__ float2 tw[TwTable<N>::SIZE];
    __shared__ float2 lds[N * G::CS];
    __shared__ double red[4 * (NT / 64)];
    const int tid = threadIdx.x;
    const int c = tid % C;
    const int t = tid / C;
    tw_fill<N>(tw, a.tw, tid, NT);
    constexpr int TILES = N / C;
    const int strip = col_strip<N>(blockIdx.x);
    const int item = strip / TILES;
    const int x0 = (strip % TILES) * C;
    const size_t base = (size_t)item * N * N + x0 + c;
    const float* __restrict__ amp = a.amp + amp_offset<N>(a.amp_of, item);
    const ColAcc<G::CS> acc{lds + c};
    float2 v[V];
#pragma unroll
    for (int r = 0; r < V; ++r) v[r] = a.f[base + (size_t)(t + r * T) * N];
    __syncthreads();
    fft_line<N, +1>(v, acc, tw, t);

    if constexpr (MODE == ACOL_BLUR) {
#pragma unroll
        for (int q = 0; q < V; ++q) {
            const float g = sw_gauss<N>(x0 + c, t + q * T, a.sigma_ft);
            v[q] = make_float2(v[q].x * g, v[q].y * g);
        }
        fft_line<N, -1>(v, acc, tw, t);
#pragma unroll
        for (int q = 0; q < V; ++q) a.f[base + (size_t)(t + q * T) * N] = v[q];
    } else if constexpr (MODE == ACOL_ERR) {
        double s[2] = {0.0, 0.0};  // sum (A-|F|)^2, sum A^2 over A >= 0
        float fmax = 0.f;
#pragma unroll
        for (int q = 0; q < V; ++q) {
            const float A = amp[(size_t)(t + q * T) * N + x0 + c];
            const float m = sqrtf(v[q].y * v[q].y + v[q].x * v[q].x);
            fmax = vmax(fmax, m);
            if (A >= 0.f) {
                s[1] += (double)(A * A);
                const float d = A - m;
                s[0] += (double)(d * d);
            }
            v[q] = project_modulus(v[q], A);
        }
        block_sum<2, NT>(s, red);
        if (tid == 0) {
            a.partial[2 * strip] = s[0];
            a.partial[2 * strip + 1] = s[1];
        }
        fmax = block_max<NT>(fmax, reinterpret_cast<float*>(red));
        if (tid == 0) atomicMax(&a.maxbits[item], __float_as_uint(fmax));
        fft_line<N, -1>(v, acc, tw, t);
#pragma unroll
        for (int q = 0; q < V; ++q) a.f[base + (size_t)(t + q * T) * N] = v[q];
    } else if constexpr (MODE == ACOL_FWD_UNIT) {
#pragma unroll
        for (int q = 0; q < V; ++q) {
            const float m = sqrtf(v[q].x * v[q].x + v[q].y * v[q].y);
            const float s = m / m;
            a.f[base + (size_t)(t + q * T) * N] = make_float2(v[q].x * s, v[q].y * s);
        }
    } else if constexpr (MODE == ACOL_CORR) {
        float2 d[V];
#pragma unroll
        for (int q = 0; q < V; ++q) {
            d[q] = v[q];
            if (a.unit_scale) {
                const float m = sqrtf(d[q].x * d[q].x + d[q].y * d[q].y);
                const float s = m / m;
                d[q] = make_float2(d[q].x * s, d[q].y * s);
            }
        }
        // sign -1: (conj d) * r ; sign +1: d * r ; both weighted and masked
        // (one loop body: unrolling it would hold both products' operands at once)
#pragma unroll 1
        for (int sgn = 0; sgn < 2; ++sgn) {
            const float sg = sgn == 0 ? -1.f : 1.f;
#pragma unroll
            for (int q = 0; q < V; ++q) {
                const int ky = t + q * T;
                float2 r;
                if (a.ref_index) r = a.ref[(size_t)a.ref_index[item] * N * N + (size_t)ky * N + x0 + c];
                else r = d[q];   // symmetry: correlate the spectrum with itself
                float cx = d[q].x * r.x - (sg * d[q].y) * r.y;
                float cy = d[q].x * r.y + (sg * d[q].y) * r.x;
                const float w = corr_weight<N>(x0 + c, ky, a.win);
                cx *= w;
                cy *= w;
                if (amp[(size_t)ky * N + x0 + c] < 0.f) {
                    cx = 0.f;
                    cy = 0.f;
                }
                v[q] = make_float2(cx, cy);
            }
            fft_line<N, -1>(v, acc, tw, t);
            float2* dst = sgn == 0 ? a.f : a.f2;
#pragma unroll
            for (int q = 0; q < V; ++q) dst[base + (size_t)(t + q * T) * N] = v[q];
        }
    }
}

// ---------------------------------------------------------------------------
// Small kernels
// ---------------------------------------------------------------------------
// p[0..n) <- v, grid-stride (any grid size covers any n).
__global__ void k_fill_u32(unsigned* p, unsigned v, size_t n) {
    const size_t step = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, size_t n) {
    const size_t step = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) p[i] = v;
}

// out[item*out_stride + out_off + k] = sum_j partial[(item*per_item + j)*K + k], fixed order.
__global__ __launch_bounds__(256) void k_reduce_partials(const double* partial, int per_item, int K, double* out,
                                                         int out_stride, int out_off, int items) {
    const int item = blockIdx.x;
    if (item >= items) return;
    __shared__ double red[4 * 4];
    for (int k = 0; k < K; ++k) {
        double s = 0;
        for (int j = threadIdx.x; j < per_item; j += 256) s += partial[((size_t)item * per_item + j) * K + k];
        double v[1] = {s};
        block_sum<1, 256>(v, red);
        if (threadIdx.x == 0) out[(size_t)item * out_stride + out_off + k] = v[0];
    }
}

// SW cut (cut_support with the maximum, src/gpu/gpu_context.cu:466-481): S = [|b| >= max*threshold].
__global__ void k_cut_max(const float* mag, const unsigned* maxbits, float threshold, uint8_t* supp, int npix,
                          int items) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)npix * items) return;
    const int item = (int)(i / npix);
    const float thr = __uint_as_float(maxbits[item]) * threshold;
    supp[i] = mag[i] >= thr ? 1 : 0;
}

// SWA cut (cut_support with an absolute threshold, :484-499); thr_bits per item,
// 0xFFFFFFFF meaning "select nothing".
__global__ void k_cut_bits(const float* mag, const unsigned* thr_bits, uint8_t* supp, int npix, int items) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)npix * items) return;
    const int item = (int)(i / npix);
    const unsigned tb = thr_bits[item];
    supp[i] = (tb != 0xFFFFFFFFu && __float_as_uint(mag[i]) >= tb) ? 1 : 0;
}

// Radix-select histogram of the digit [lo, lo+w) of the bit patterns of mag
// (non-negative floats), restricted to the values whose higher bits equal prefix.
template <int W>
__global__ __launch_bounds__(256) void k_radix_hist(const float* mag, int npix, int lo, const unsigned* prefix,
                                                     const unsigned* maskhi, unsigned* hist, int blocks_per_item) {
    __shared__ unsigned h[1 << W];
    for (int i = threadIdx.x; i < (1 << W); i += 256) h[i] = 0;
    __syncthreads();
    const int item = blockIdx.x / blocks_per_item;
    const int blk = blockIdx.x % blocks_per_item;
    const unsigned pf = prefix[item], mk = maskhi[item];
    const float* m = mag + (size_t)item * npix;
    for (int i = blk * 256 + threadIdx.x; i < npix; i += blocks_per_item * 256) {
        const unsigned b = __float_as_uint(m[i]);
        if ((b & mk) == pf) atomicAdd(&h[(b >> lo) & ((1u << W) - 1)], 1u);
    }
    __syncthreads();
    unsigned* gh = hist + (size_t)item * (1 << W);
    for (int i = threadIdx.x; i < (1 << W); i += 256)
        if (h[i]) atomicAdd(&gh[i], h[i]);
}

// One thread per item: pick the digit that holds the `remaining`-th largest value.
template <int W>
__global__ void k_radix_pick(const unsigned* hist, int lo, unsigned* prefix, unsigned* maskhi, unsigned* remaining,
                             int items) {
    const int item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= items) return;
    const unsigned* h = hist + (size_t)item * (1 << W);
    unsigned rem = remaining[item];
    if (rem == 0) {   // empty selection (spixels <= 0): a threshold that selects nothing
        prefix[item] = 0xFFFFFFFFu;
        maskhi[item] = 0xFFFFFFFFu;
        return;
    }
    unsigned cum = 0;
    int bsel = 0;
    for (int b = (1 << W) - 1; b >= 0; --b) {
        if (cum + h[b] >= rem) {
            bsel = b;
            break;
        }
        cum += h[b];
    }
    remaining[item] = rem - cum;
    prefix[item] |= ((unsigned)bsel) << lo;
    maskhi[item] |= ((1u << W) - 1) << lo;
}

}  // namespace spr

namespace spr {

// ---------------------------------------------------------------------------
// Generic batched line transforms over listed fields
// ---------------------------------------------------------------------------
// Fields live at base + field * N*N; item i of the launch processes field
// index[i] (or i when index is NULL).
struct FieldBatch {
    float2* base;
    const int* index;
    const float2* tw;
    float scale;      // applied after the transform (1 or 1/N^2)
    const float2* src = nullptr;   // row transforms: read from src (same field index) instead of base
};

__device__ __forceinline__ size_t field_of(const FieldBatch& b, int item) {
    return (size_t)(b.index ? b.index[item] : item);
}

// Row transforms (all rows of each listed field), in place.
template <int N, int SIGN>
__global__ __launch_bounds__(Geo<N>::ROW_THREADS) void k_fft_rows(FieldBatch b) {
    using G = Geo<N>;
    constexpr int V = G::V, T = G::T, NT = G::ROW_THREADS, WPF = N / G::RW;
    __shared__ float2 tw[TwTable<N>::SIZE];
    __shared__ float2 lds[G::RW * G::RS];
    const int tid = threadIdx.x, t = tid % T, rl = tid / T;
    tw_fill<N>(tw, b.tw, tid, NT);
    const int item = blockIdx.x / WPF;
    const int y = (blockIdx.x % WPF) * G::RW + rl;
    const size_t fo = field_of(b, item) * N * N + (size_t)y * N;
    float2* __restrict__ f = b.base + fo;
    const float2* __restrict__ s = b.src ? b.src + fo : f;
    const RowAcc acc{lds + rl * G::RS};
    float2 v[V];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < V; ++r) v[r] = s[t + r * T];
    fft_line<N, SIGN, (G::T <= 64)>(v, acc, tw, t);
#pragma unroll
    for (int q = 0; q < V; ++q) f[t + q * T] = make_float2(v[q].x * b.scale, v[q].y * b.scale);
}

// Column transforms (all columns of each listed field), in place.
template <int N, int SIGN>
__global__ __launch_bounds__(Geo<N>::COL_THREADS) void k_fft_cols(FieldBatch b) {
    using G = Geo<N>;
    constexpr int V = G::V, T = G::T, C = G::C, NT = G::COL_THREADS, TILES = N / C;
    __shared__ float2 tw[TwTable<N>::SIZE];
    __shared__ float2 lds[N * G::CS];
    const int tid = threadIdx.x, c = tid % C, t = tid / C;
    tw_fill<N>(tw, b.tw, tid, NT);
    const int strip = col_strip<N>(blockIdx.x);
    const int item = strip / TILES;
    const int x0 = (strip % TILES) * C;
    float2* __restrict__ f = b.base + field_of(b, item) * N * N + x0;
    const int off = t * N + c;
    const ColAcc<G::CS> acc{lds + c};
    float2 v[V];
#pragma unroll
    for (int r = 0; r < V; ++r) v[r] = f[off + r * T * N];
    __syncthreads();
    fft_line<N, SIGN>(v, acc, tw, t);
#pragma unroll
    for (int q = 0; q < V; ++q) f[off + q * T * N] = make_float2(v[q].x * b.scale, v[q].y * b.scale);
}

// ---------------------------------------------------------------------------
// Correlation of two listed spectra (both already fully transformed):
// out1 = colIFFT(win conj(d) r), out2 = colIFFT(win d r); rows follow with AROW_CORR.
// ---------------------------------------------------------------------------
struct CorrSpecArgs {
    const float2* spec;     // spectra batch
    const int* d_index;     // item -> spectrum of the aligned field
    const int* r_index;     // item -> spectrum of the reference
    float2* out1;           // [items][N][N]
    float2* out2;
    const float* amp;
    const float2* tw;
    CorrWin win;
    int unit_scale;         // apply normalize_ft_shift (z *= |z|/|z|) to d and r first
};

__device__ __forceinline__ float2 unit_scaled(float2 z) {
    const float m = sqrtf(z.x * z.x + z.y * z.y);
    const float s = m / m;
    return make_float2(z.x * s, z.y * s);
}

template <int N>
__global__ __launch_bounds__(Geo<N>::COL_THREADS) void k_corr_spec_cols(CorrSpecArgs a) {
    using G = Geo<N>;
    constexpr int V = G::V, T = G::T, C = G::C, NT = G::COL_THREADS, TILES = N / C;
    __shared__ float2 tw[TwTable<N>::SIZE];
    __shared__ float2 lds[N * G::CS];
    const int tid = threadIdx.x, c = tid % C, t = tid / C;
    tw_fill<N>(tw, a.tw, tid, NT);
    const int strip = col_strip<N>(blockIdx.x);
    const int item = strip / TILES;
    const int x0 = (strip % TILES) * C;
    const float2* __restrict__ d = a.spec + (size_t)a.d_index[item] * N * N + x0;
    const float2* __restrict__ r = a.spec + (size_t)a.r_index[item] * N * N + x0;
    const float* __restrict__ amp = a.amp + x0;
    const int off = t * N + c;
    const ColAcc<G::CS> acc{lds + c};
    __syncthreads();
#pragma unroll
    for (int sgn = 0; sgn < 2; ++sgn) {
        const float sg = sgn == 0 ? -1.f : 1.f;
        float2 v[V];
#pragma unroll
        for (int q = 0; q < V; ++q) {
            const int ky = t + q * T;
            float2 dv = d[off + q * T * N], rv = r[off + q * T * N];
            if (a.unit_scale) {
                dv = unit_scaled(dv);
                rv = unit_scaled(rv);
            }
            float cx = dv.x * rv.x - (sg * dv.y) * rv.y;
            float cy = dv.x * rv.y + (sg * dv.y) * rv.x;
            const float w = corr_weight<N>(x0 + c, ky, a.win);
            cx *= w;
            cy *= w;
            if (amp[off + q * T * N] < 0.f) {
                cx = 0.f;
                cy = 0.f;
            }
            v[q] = make_float2(cx, cy);
        }
        fft_line<N, -1>(v, acc, tw, t);
        float2* __restrict__ o = (sgn == 0 ? a.out1 : a.out2) + (size_t)item * N * N + x0;
#pragma unroll
        for (int q = 0; q < V; ++q) o[off + q * T * N] = v[q];
    }
}

// ---------------------------------------------------------------------------
// Shift choice and application (get_shift tail, reshift/shift kernel,
// src/gpu/gpu_context.cu:1309-1321, 1329-1357)
// ---------------------------------------------------------------------------
struct ShiftParam {
    int sx, sy, flip;
    float phase;
};

// keys: first index of the strict maximum of |c|^2 packed as (value bits << 32 | ~index).
__global__ void k_pick_shift(const unsigned long long* key1, const unsigned long long* key2,
                             const float* phase_direct, const float* phase_flip, int n, ShiftParam* out, int items) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= items) return;
    const unsigned long long k1 = key1[i], k2 = key2[i];
    const float m1 = k1 ? __uint_as_float((unsigned)(k1 >> 32)) : 0.f;
    const float m2 = k2 ? __uint_as_float((unsigned)(k2 >> 32)) : 0.f;
    const int i1 = k1 ? (int)(0xFFFFFFFFu - (unsigned)(k1 & 0xFFFFFFFFu)) : 0;
    const int i2 = k2 ? (int)(0xFFFFFFFFu - (unsigned)(k2 & 0xFFFFFFFFu)) : 0;
    ShiftParam s;
    if (m1 >= m2) {
        s.sx = i1 % n;
        s.sy = i1 / n;
        s.phase = phase_direct[i];
        s.flip = 0;
    } else {
        s.sx = -i2 % n;
        s.sy = -i2 / n;
        s.phase = phase_flip[i];
        s.flip = 1;
    }
    out[i] = s;
}

struct ReshiftArgs {
    const float2* src;        // source fields batch
    const int* src_index;
    const uint8_t* ssupp;     // source supports batch (NULL: no support)
    const int* ssupp_index;
    float2* dst;
    const int* dst_index;
    uint8_t* dsupp;
    const int* dsupp_index;
    const ShiftParam* shift;
};

// One thread per source pixel; the map old -> new is a bijection.
template <int N>
__global__ __launch_bounds__(256) void k_reshift(ReshiftArgs a) {
    constexpr int NP = N * N;
    const int item = blockIdx.y;
    const int old = blockIdx.x * 256 + threadIdx.x;
    if (old >= NP) return;
    const ShiftParam s = a.shift[item];
    const int ix = old % N, iy = old / N;
    int nx = ix + s.sx, ny = iy + s.sy;
    if (nx < 0) nx += N;
    if (nx >= N) nx -= N;
    if (ny < 0) ny += N;
    if (ny >= N) ny -= N;
    int ni = nx + ny * N;
    if (s.flip) ni = NP - 1 - ni;
    const float2 z = a.src[(size_t)(a.src_index ? a.src_index[item] : item) * NP + old];
    const float mod = sqrtf(z.x * z.x + z.y * z.y);
    float ph = atan2f(z.y, z.x);
    if (s.flip) ph = -ph;
    ph += s.phase;
    float sn, cs;
    sincosf(ph, &sn, &cs);
    a.dst[(size_t)(a.dst_index ? a.dst_index[item] : item) * NP + ni] = make_float2(mod * cs, mod * sn);
    if (a.ssupp && a.dsupp)
        a.dsupp[(size_t)(a.dsupp_index ? a.dsupp_index[item] : item) * NP + ni] =
            a.ssupp[(size_t)(a.ssupp_index ? a.ssupp_index[item] : item) * NP + old];
}

// rho <- S'(rho) rho (apply_support_er, src/gpu/gpu_context.cu:519-537) on listed fields.
__global__ __launch_bounds__(256) void k_er_project(float2* f, const int* f_index, const uint8_t* s,
                                                    const int* s_index, int np, PhaseCut pc) {
    const int item = blockIdx.y;
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= np) return;
    float2* z = f + (size_t)(f_index ? f_index[item] : item) * np + p;
    const uint8_t sp = s[(size_t)(s_index ? s_index[item] : item) * np + p];
    const float2 d = *z;
    const float ts = (sp != 0 && !phase_fails(d.x, d.y, pc)) ? 1.f : 0.f;
    *z = make_float2(d.x * ts, d.y * ts);
}

// ---------------------------------------------------------------------------
// Crossover (cross_density / cross_density_real / cross_support,
// src/gpu/gpu_context.cu:1699-1883)
// ---------------------------------------------------------------------------
struct CrossChild {
    int crossft, crosssym;
    int crosssizex, squaresizex, squaresizey, offsetx, offsety;
    float rotangle;
    int crossmap_offset;
    int f0, f1, f2, f3, favg;   // fields of the (aligned) roles in the crossover workspace
    int s0, s1, s2, s3;         // their supports (s1: population support of p1)
};

// Cell of pixel (ix, iy) in the rotated square grid; the polar geometry is
// evaluated in double so that host and device agree on every integer cell.
template <int N> __device__ __forceinline__ int cross_cell(int ix, int iy, const CrossChild& d) {
    int x = ix, y = iy;
    if (x >= N / 2) x -= N;
    if (y >= N / 2) y -= N;
    double radius = sqrt((double)(x * x + y * y));
    if (radius >= N / 2.) radius = (double)(N / 2 - 1);
    double angle = atan2((double)y, (double)x);
    if (d.crosssym && angle < 0) angle += M_PI;
    angle += (double)d.rotangle;
    double sn, cs;
    sincos(angle, &sn, &cs);
    x = (int)(cs * radius) + N / 2;
    y = (int)(sn * radius) + N / 2;
    return (x + d.offsetx) / d.squaresizex + ((y + d.offsety) / d.squaresizey) * d.crosssizex;
}

struct CrossArgs {
    float2* fields;              // crossover workspace fields
    uint8_t* supps;              // crossover workspace supports
    const uint8_t* pop_supp;     // population supports (s1 indexes it)
    const CrossChild* child;
    const int32_t* crossmaps;
    const float* amp;
    const float* amp_avg;
    float crossdiff, use_average;
};

template <int N>
__global__ __launch_bounds__(256) void k_cross(CrossArgs a) {
    constexpr int NP = N * N;
    const int c = blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= NP) return;
    const CrossChild d = a.child[c];
    float2* f0 = a.fields + (size_t)d.f0 * NP;
    const int cell = cross_cell<N>(i % N, i / N, d);
    if (a.crossmaps[d.crossmap_offset + cell] == 1) {
        const float2 r1 = a.fields[(size_t)d.f1 * NP + i];
        const float2 r2 = a.fields[(size_t)d.f2 * NP + i];
        const float2 r3 = a.fields[(size_t)d.f3 * NP + i];
        float2 d1 = r1;
        if (d.crossft && a.amp[i] < 0.f) {
            const float mod1 = sqrtf(r1.x * r1.x + r1.y * r1.y);
            const float2 av = a.fields[(size_t)d.favg * NP + i];
            if (a.use_average > 0) {
                const float avgmod = (a.amp_avg[i] + sqrtf(av.x * av.x + av.y * av.y)) * 0.5f;
                const float newmod = (float)((double)avgmod * a.use_average + (1. - a.use_average) * (double)mod1);
                if (mod1 > 0.f) d1 = make_float2(d1.x * (newmod / mod1), d1.y * (newmod / mod1));
            }
            if (a.use_average < 0 && mod1 > 0.f) {
                const float locua = fabsf(a.use_average);
                const float maxmod = a.amp_avg[i];
                const float minmod = sqrtf(av.x * av.x + av.y * av.y);
                float newmod = mod1;
                if (newmod < minmod) newmod = minmod * locua + newmod * (1.f - locua);
                else if (newmod > maxmod) newmod = maxmod * locua + newmod * (1.f - locua);
                d1 = make_float2(d1.x * (newmod / mod1), d1.y * (newmod / mod1));
            }
        }
        f0[i] = make_float2(d1.x + a.crossdiff * (r2.x - r3.x), d1.y + a.crossdiff * (r2.y - r3.y));
    }
    const int v = a.pop_supp[(size_t)d.s1 * NP + i] +
                  a.supps[(size_t)d.s2 * NP + i] * a.supps[(size_t)d.s3 * NP + i] * a.supps[(size_t)d.s0 * NP + i];
    a.supps[(size_t)d.s0 * NP + i] = v > 0.5 ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Averages (GetAverage / GetAverageModule, src/solver.cpp:564-686)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_average(const float2* rho, const uint8_t* supp, const int* idx, int navg,
                                                 int np, float norm, float2* out, uint8_t* out_supp) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= np) return;
    float sx = 0.f, sy = 0.f;
    int s = 0;
    for (int k = 0; k < navg; ++k) {
        const size_t o = (size_t)idx[k] * np + p;
        const float2 z = rho[o];
        sx += z.x;
        sy += z.y;
        s += supp[o];
    }
    out[p] = make_float2(sx * norm, sy * norm);
    out_supp[p] = s > 0 ? 1 : 0;
}

// out[p] = norm * sum_k |f[k][p]| (f either the listed fields or a contiguous batch).
__global__ __launch_bounds__(256) void k_sum_modulus(const float2* f, const int* idx, int count, int np, float norm,
                                                     float* out) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= np) return;
    float s = 0.f;
    for (int k = 0; k < count; ++k) {
        const float2 z = f[(size_t)(idx ? idx[k] : k) * np + p];
        s += sqrtf(z.x * z.x + z.y * z.y);
    }
    out[p] = s * norm;
}

// dst[didx[item]] <- src[sidx[item]] for listed fields (NULL index: identity).
__global__ __launch_bounds__(256) void k_copy_fields(const float2* src, const int* sidx, float2* dst, const int* didx,
                                                     int np) {
    const int item = blockIdx.y;
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= np) return;
    dst[(size_t)(didx ? didx[item] : item) * np + p] = src[(size_t)(sidx ? sidx[item] : item) * np + p];
}

// Support fields (uint8, np bytes each) in 16-byte words: dst[didx[k]] = src[sidx[k]].
__global__ __launch_bounds__(256) void k_copy_supports(const uint4* src, const int* sidx, uint4* dst, const int* didx,
                                                       int nw) {
    const int item = blockIdx.y;
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= nw) return;
    dst[(size_t)(didx ? didx[item] : item) * nw + p] = src[(size_t)(sidx ? sidx[item] : item) * nw + p];
}

}  // namespace spr
