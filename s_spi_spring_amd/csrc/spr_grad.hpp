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

}  // namespace spr
