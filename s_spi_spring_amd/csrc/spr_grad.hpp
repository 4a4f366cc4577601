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
    const float* amp;
    const float2* tw;
    double* partial;     // one loss partial per workgroup
    float eps;
    int kind;
};

template <int N>
__global__ __launch_bounds__(Geo<N>::COL_THREADS) void k_col_loss(LossArgs a) {
    using G = Geo<N>;
    constexpr int T = G::T, C = G::C, NT = G::COL_THREADS, TILES = N / C;
    __shared__ float2 tw[TwTable<N>::SIZE];
    __shared__ float2 lds[N * G::CS];
    __shared__ double red[4 * (NT / 64)];
    const int tid = threadIdx.x, c = tid % C, t = tid / C;
    tw_fill<N>(tw, a.tw, tid, NT);
    const int item = blockIdx.x / TILES;
    const int x0 = (blockIdx.x % TILES) * C;
    float2* __restrict__ f = a.f + (size_t)item * N * N + x0;
    const float* __restrict__ amp = a.amp + x0;
    const int off = t * N + c;
    const ColAcc<G::CS> acc{lds + c};
    float2 v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = f[off + r * T * N];
    __syncthreads();
    fft_line<N, +1>(v, acc, tw, t);
    double l[1] = {0.0};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const float A = amp[off + q * T * N];
        const float u = v[q].x * v[q].x + v[q].y * v[q].y;
        float w = 0.f;
        if (A >= 0.f) {
            if (a.kind == LOSS_POISSON) {
                const float I = A * A;
                l[0] += (double)u - (double)I * (double)logf(u + a.eps);
                w = 1.f - I / (u + a.eps);
            } else {
                const float m = sqrtf(u);
                const float d = m - A;
                l[0] += (double)d * (double)d;
                w = m > 0.f ? 1.f - A / m : 0.f;
            }
        }
        v[q] = make_float2(v[q].x * w, v[q].y * w);
    }
    block_sum<1, NT>(l, red);
    if (tid == 0) a.partial[blockIdx.x] = l[0];
    fft_line<N, -1>(v, acc, tw, t);
#pragma unroll
    for (int q = 0; q < 8; ++q) f[off + q * T * N] = v[q];
}

// Adam step on complex fields viewed as real pairs (torch.optim.Adam semantics,
// no weight decay): m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
// x -= lr * (m / (1-b1^t)) / (sqrt(v / (1-b2^t)) + eps).
__global__ __launch_bounds__(256) void k_adam(float2* x, const float2* g, float2* m, float2* v, size_t n, float lr,
                                              float b1, float b2, float eps, float bc1, float bc2) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float2 gi = g[i];
    float2 mi = m[i], vi = v[i], xi = x[i];
    mi = make_float2(b1 * mi.x + (1.f - b1) * gi.x, b1 * mi.y + (1.f - b1) * gi.y);
    vi = make_float2(b2 * vi.x + (1.f - b2) * gi.x * gi.x, b2 * vi.y + (1.f - b2) * gi.y * gi.y);
    xi.x -= lr * (mi.x / bc1) / (sqrtf(vi.x / bc2) + eps);
    xi.y -= lr * (mi.y / bc1) / (sqrtf(vi.y / bc2) + eps);
    m[i] = mi;
    v[i] = vi;
    x[i] = xi;
}

// Per-individual step size of the support-projected refinement:
// lr_p = lr_rel * sqrt(sum_{S} |rho|^2 / |S|) (0 for an empty support), summed
// in double in a fixed order (one workgroup per individual).
__global__ __launch_bounds__(256) void k_support_scale(const float2* rho, const uint8_t* supp, int np, float lr_rel,
                                                       float* lrp) {
    const int p = blockIdx.x;
    const float2* r = rho + (size_t)p * np;
    const uint8_t* s = supp + (size_t)p * np;
    double e = 0, cnt = 0;
    for (int i = threadIdx.x; i < np; i += 256)
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
    if (threadIdx.x == 0) lrp[p] = sc[0] > 0 ? (float)(lr_rel * sqrt(se[0] / sc[0])) : 0.f;
}

// Adam step restricted to the support: pixels outside it keep their value (and
// zero moments), as torch.optim.Adam does for a gradient masked by the support;
// the step size of individual p is lrp[p].
__global__ __launch_bounds__(256) void k_adam_support(float2* x, const float2* g, float2* m, float2* v,
                                                      const uint8_t* supp, const float* lrp, int np, size_t n,
                                                      float b1, float b2, float eps, float bc1, float bc2) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n || !supp[i]) return;
    const float lr = lrp[i / np];
    const float2 gi = g[i];
    float2 mi = m[i], vi = v[i], xi = x[i];
    mi = make_float2(b1 * mi.x + (1.f - b1) * gi.x, b1 * mi.y + (1.f - b1) * gi.y);
    vi = make_float2(b2 * vi.x + (1.f - b2) * gi.x * gi.x, b2 * vi.y + (1.f - b2) * gi.y * gi.y);
    xi.x -= lr * (mi.x / bc1) / (sqrtf(vi.x / bc2) + eps);
    xi.y -= lr * (mi.y / bc1) / (sqrtf(vi.y / bc2) + eps);
    m[i] = mi;
    v[i] = vi;
    x[i] = xi;
}

}  // namespace spr
