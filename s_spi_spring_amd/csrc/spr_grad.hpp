// Differentiable masked Fourier-magnitude loss and its gradient (gfx950).
//
// For F = FT(rho) (unnormalised, exp(+2 pi i k.x/N)) and the measured pixels
// (A >= 0, counts I = A^2), with u = |F|^2:
//   Poisson  L = sum_k u_k - I_k log(u_k + eps)        f'(u) = 1 - I/(u + eps)
//   amplitude L = sum_k (|F_k| - A_k)^2                f'(u) = 1 - A/|F|   (0 where |F| = 0)
// The gradient in PyTorch's convention for complex inputs (dL/dRe + i dL/dIm =
// 2 dL/d conj(rho)) is 2 FT^H[f'(u) F], with FT^H the adjoint transform
// (exp(-2 pi i k.x/N), unnormalised). Loss and gradient come from one pass
// pair: the column pass computes F, the loss and the weighted spectrum and
// transforms it back along columns; the row pass finishes the adjoint.
#pragma once

#include "spr_aux.hpp"

namespace spr {

enum LossKind { LOSS_POISSON = 0, LOSS_AMPLITUDE = 1 };

struct LossArgs {
    float2* f;           // row-transformed fields (in), adjoint-column-transformed weights (out)
    const float* amp;    // amplitudes [patterns][N][N]
    const int* amp_of;   // field -> pattern (NULL: pattern 0)
    const float2* tw;
    double* partial;     // one loss partial per workgroup
    float eps;
    int kind;
    const float2* twtab; // the plan's twiddle table in global memory (column variants with TW_GLOBAL)
};

// Column pass of the loss in the geometry of column variant CV (the tile
// layouts of the projection pass). WANT_LOSS: one loss partial per workgroup
// into a.partial[strip] (summed in double); without it the pass computes only
// the weighted spectrum (the refinement steps, whose losses are not read).
template <int N, int CV = 0, bool WANT_LOSS = true>
__global__ __launch_bounds__((ColVariant<N, CV>::G::COL_THREADS))
__attribute__((amdgpu_waves_per_eu((ColVariant<N, CV>::G::COL_WAVES_PER_EU)))) void k_col_loss(LossArgs a) {
    using G = typename ColVariant<N, CV>::G;
    using PL = typename G::PL;
    constexpr int V = G::V, T = G::T, C = G::C, NT = G::COL_THREADS, TILES = N / G::COL_STRIP;
    static_assert(G::COL_GROUPS == 1 && !G::COL_WAVE, "loss passes use single-group LDS tiles");
    __shared__ float2 twl[G::TW_GLOBAL ? 1 : TwTableP<PL>::SIZE];
    __shared__ float2 lds[G::COL_TILE];
    __shared__ double red[WANT_LOSS ? 4 * (NT / 64) : 1];
    const int tid = threadIdx.x, c = tid % C, t = tid / C;
    const float2* tw = twl;
    if constexpr (G::TW_GLOBAL) tw = a.twtab;
    else tw_fill_p<PL>(twl, a.tw, tid, NT);
    const int strip = col_strip<N, G::COL_STRIP>(blockIdx.x);
    const int item = strip / TILES;
    const int x0 = (strip % TILES) * C;
    float2* __restrict__ f = a.f + (size_t)item * N * N + x0;
    const float* __restrict__ amp = a.amp + amp_offset<N>(a.amp_of, item) + x0;
    const int off = t * N + c;
    using Acc = std::conditional_t<(G::COL_SKEW > 0), ColAccSkew<C, (G::COL_SKEW > 0 ? G::COL_SKEW : 1)>, ColAcc<G::CS>>;
    const Acc acc{lds + c};
    float2 v[V];
    float Aq[V];
#pragma unroll
    for (int r = 0; r < V; ++r) {
        v[r] = f[off + r * T * N];
        Aq[r] = amp[off + r * T * N];
    }
    __syncthreads();
    fft_line_p<PL, +1>(v, acc, tw, t);
    double l[1] = {0.0};
#pragma unroll
    for (int q = 0; q < V; ++q) {
        const float A = Aq[q];
        const float u = v[q].x * v[q].x + v[q].y * v[q].y;
        float w = 0.f;
        if (A >= 0.f) {
            if (a.kind == LOSS_POISSON) {
                const float I = A * A;
                if constexpr (WANT_LOSS) l[0] += (double)u - (double)I * (double)logf(u + a.eps);
                w = 1.f - I / (u + a.eps);
            } else {
                const float m = sqrtf(u);
                const float d = m - A;
                if constexpr (WANT_LOSS) l[0] += (double)d * (double)d;
                w = m > 0.f ? 1.f - A / m : 0.f;
            }
        }
        v[q] = make_float2(v[q].x * w, v[q].y * w);
    }
    if constexpr (WANT_LOSS) {
        block_sum<1, NT>(l, red);
        if (tid == 0) a.partial[strip] = l[0];
    }
    fft_line_p<PL, -1>(v, acc, tw, t);
#pragma unroll
    for (int q = 0; q < V; ++q) f[off + q * T * N] = v[q];
}

// Projection onto the supports: rho <- 0 outside each individual's support (the
// start of the support-projected refinement). total = P * N^2.
__global__ __launch_bounds__(256) void k_zero_outside(float2* rho, const uint8_t* supp, size_t total) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < total && supp[i] == 0) rho[i] = make_float2(0.f, 0.f);
}

// Per-individual step size of the support-projected refinement:
// lr_p = lr_rel * sqrt(sum_{S} |rho|^2 / |S|) (0 for an empty support), summed
// in double in a fixed order: SUPPORT_CHUNKS workgroups per individual each sum
// a contiguous chunk (k_support_partials), then one thread per individual adds
// the chunks' sums in order (k_support_lr).
constexpr int SUPPORT_CHUNKS = 64;
__global__ __launch_bounds__(256) void k_support_partials(const float2* __restrict__ rho,
                                                          const uint8_t* __restrict__ supp, int np,
                                                          double* __restrict__ part) {
    const int p = blockIdx.x / SUPPORT_CHUNKS, j = blockIdx.x % SUPPORT_CHUNKS;
    const int len = np / SUPPORT_CHUNKS;
    const float2* r = rho + (size_t)p * np + (size_t)j * len;
    const uint8_t* s = supp + (size_t)p * np + (size_t)j * len;
    double e = 0, cnt = 0;
    for (int i = threadIdx.x; i < len; i += 256)
        if (s[i]) {
            e += (double)r[i].x * r[i].x + (double)r[i].y * r[i].y;
            cnt += 1;
        }
    __shared__ double se[256], sc[256];
    se[threadIdx.x] = e;
    sc[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            se[threadIdx.x] += se[threadIdx.x + w];
            sc[threadIdx.x] += sc[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * (size_t)blockIdx.x] = se[0];
        part[2 * (size_t)blockIdx.x + 1] = sc[0];
    }
}
__global__ __launch_bounds__(64) void k_support_lr(const double* __restrict__ part, int P, float lr_rel,
                                                   float* __restrict__ lrp) {
    const int p = blockIdx.x * 64 + threadIdx.x;
    if (p >= P) return;
    double e = 0, cnt = 0;
    for (int j = 0; j < SUPPORT_CHUNKS; ++j) {
        e += part[2 * ((size_t)p * SUPPORT_CHUNKS + j)];
        cnt += part[2 * ((size_t)p * SUPPORT_CHUNKS + j) + 1];
    }
    lrp[p] = cnt > 0 ? (float)(lr_rel * sqrt(e / cnt)) : 0.f;
}

// Row pass of one refinement step, fused: the adjoint row transform of the
// column pass's output gives the gradient g = 2 FT^H[f'(u) F] of each row; an
// Adam step (torch.optim.Adam semantics, no weight decay: m = b1 m + (1-b1) g;
// v = b2 v + (1-b2) g^2; x -= lr (m/(1-b1^t)) / (sqrt(v/(1-b2^t)) + eps), on
// the real and imaginary parts) updates the densities and the moments; the row FFT of the updated densities goes back to
// `work` for the next step's column pass. MASKED: only support pixels move and
// the step of field p is lrp[p]; otherwise every pixel moves by lr.
struct AdamRowArgs {
    float2* x;
    float2* m;
    float2* v;
    float2* work;
    const uint8_t* supp;
    const float* lrp;
    const float2* tw;
    float lr, b1, b2, eps, bc1, bc2;
};

// In the geometry of row variant RV; the densities are loaded with the field,
// before the adjoint transform, so that their latency hides behind it. MASKED:
// the densities outside the support are 0 (the support refinement zeroes them
// before its first step and never moves them), so only support pixels are read.
template <int N, int RV, bool MASKED>
__global__ __launch_bounds__((RowVariant<N, RV>::G::ROW_THREADS)) void k_row_adam(AdamRowArgs a) {
    using G = typename RowVariant<N, RV>::G;
    using PL = typename G::PL;
    constexpr int V = G::V, T = G::T;
    constexpr bool WV = T <= 64;
    __shared__ float2 tw[TwTableP<PL>::SIZE];
    __shared__ float2 lds[G::RW * G::RS];
    const int tid = threadIdx.x;
    const int t = tid % T;
    const int rl = tid / T;
    tw_fill_p<PL>(tw, a.tw, tid, G::ROW_THREADS);
    const size_t blk = (size_t)blockIdx.x * G::RW * N;
    const int field = (int)(blk / ((size_t)N * N));
    const int off = rl * N + t;
    float2* __restrict__ x = a.x + blk;
    float2* __restrict__ m = a.m + blk;
    float2* __restrict__ v2 = a.v + blk;
    float2* __restrict__ work = a.work + blk;
    const RowAccT<G::ROW_PAD_SHIFT> acc{lds + rl * G::RS};
    const float lr = MASKED ? a.lrp[field] : a.lr;
    float2 v[V], xs[V];
    uint8_t sb[V];
#pragma unroll
    for (int r = 0; r < V; ++r) {
        v[r] = work[off + r * T];
        if constexpr (MASKED) {
            sb[r] = a.supp[blk + off + r * T];
            xs[r] = sb[r] ? x[off + r * T] : make_float2(0.f, 0.f);
        } else {
            sb[r] = 1;
            xs[r] = x[off + r * T];
        }
    }
    __syncthreads();
    fft_line_p<PL, -1, WV>(v, acc, tw, t);
#pragma unroll
    for (int q = 0; q < V; ++q) {
        const int i = off + q * T;
        const float2 gi = make_float2(v[q].x * 2.f, v[q].y * 2.f);
        float2 xi = xs[q];
        if (!MASKED || sb[q]) {
            float2 mi = m[i], vi = v2[i];
            mi = make_float2(a.b1 * mi.x + (1.f - a.b1) * gi.x, a.b1 * mi.y + (1.f - a.b1) * gi.y);
            vi = make_float2(a.b2 * vi.x + (1.f - a.b2) * gi.x * gi.x, a.b2 * vi.y + (1.f - a.b2) * gi.y * gi.y);
            xi.x -= lr * (mi.x / a.bc1) / (sqrtf(vi.x / a.bc2) + a.eps);
            xi.y -= lr * (mi.y / a.bc1) / (sqrtf(vi.y / a.bc2) + a.eps);
            m[i] = mi;
            v2[i] = vi;
            x[i] = xi;
        }
        v[q] = xi;
    }
    fft_line_p<PL, +1, WV>(v, acc, tw, t);
#pragma unroll
    for (int q = 0; q < V; ++q) work[off + q * T] = v[q];
}

}  // namespace spr
