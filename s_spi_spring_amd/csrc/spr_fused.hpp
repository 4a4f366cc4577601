// Whole iteration blocks of small fields on one CU (gfx950).
//
// A 128 x 128 complex field (128 KB) fits in the 160 KB of LDS of one CU, so
// for N <= 128 a whole block -- the first row transform, niter x [column pass,
// row update] and the trailing projection -- runs inside one workgroup per
// individual: the field stays in LDS, the state rho_n and the support bits stay
// in registers, the amplitude (N^2 floats per pattern, shared by its
// individuals) is read from L2 by every column phase, and HBM is touched only to
// read rho_0 and S at the start and to write rho at the end (17 B/px per block
// instead of 49 B/px per iteration). The operations are those of the two-pass iteration
// (spr_kernels.hpp); the transforms use 16 values per thread, so that every
// line's threads lie in one wavefront and synchronise with wavefront fences,
// with a workgroup barrier between the row and the column phases.
//
// LDS: element (y, x) of the field at lds[y*RS + x + (x >> 4)], RS = N + N/16.
// The row phase gives each wavefront 64/T whole rows (lane = row*T + t); the
// column phase gives it 64/T whole columns with the lanes ordered t-major
// (lane = t*(64/T) + column), so that the strided column accesses of lanes
// served together fall in distinct banks. Both layouts and lane orders were
// chosen with a bank model of the transforms' access patterns (one padding
// float2 per 16 elements, rows of N + N/16 float2).
#pragma once

#include "spr_kernels.hpp"

namespace spr {

template <int N> struct FusedGeo;
template <> struct FusedGeo<128> {
    using PL = Plan<128, 16, 16, 8>;
};
template <> struct FusedGeo<64> {
    using PL = Plan<64, 16, 16, 4>;
};

template <int N> constexpr bool fused_size() { return N == 64 || N == 128; }

struct FusedArgs {
    float2* rho;            // [P][N][N], read at the start and written at the end of the block
    const uint8_t* supp;
    const float* amp_t;     // amplitudes transposed, [patterns][N][N] with amp_t[x*N + y] = amp[y*N + x]
    const int* amp_of;      // individual -> pattern (NULL: pattern 0)
    const float2* tw;
    int niter;
    float beta;
    float inv_nn;
    PhaseCut pc;
};

template <int N, int MODE>
__global__ __launch_bounds__(N * (N / 16)) void k_iter_fused(FusedArgs a) {
    using PL = typename FusedGeo<N>::PL;
    constexpr int V = 16, T = N / V, SH = 4;
    constexpr int RS = N + (N >> SH);
    constexpr int LPW = 64 / T;   // lines per wavefront
    constexpr int THREADS = N * T;
    static_assert((1 << SH) % T == 0, "row offsets below assume T divides 2^SH");
    __shared__ float2 tw[TwTableP<PL>::SIZE];
    __shared__ float2 lds[N * RS];
    const int tid = threadIdx.x;
    tw_fill_p<PL>(tw, a.tw, tid, THREADS);
    const int p = blockIdx.x;
    const size_t base = (size_t)p * N * N;
    // Row phase: row `row`, elements x = t + T*j.
    const int t = tid % T;
    const int row = tid / T;
    const RowAccT<SH> racc{lds + row * RS};
    // Element t + T*j of the row: since t < T and T divides 2^SH,
    // (t + T*j) >> SH = (T*j) >> SH, a compile-time offset.
    float2* const rp = lds + row * RS + t;
    // Column phase: column `col`, elements y = tc + T*j.
    const int lane = tid % 64;
    const int tc = lane / LPW;
    const int col = (tid / 64) * LPW + lane % LPW;
    const ColAcc<RS> cacc{lds + col + (col >> SH)};
    // Amplitude at (y, x) = (tc + T j, col), from the transposed copy
    // (coalesced; N^2 floats per pattern, resident in L2).
    const float* __restrict__ at = a.amp_t + amp_offset<N>(a.amp_of, p);   // block-uniform base
    const int aoff = col * N + tc;

    float2 rn[V], v[V];
    unsigned sbits = 0;
    {
        const float2* __restrict__ r0 = a.rho + base + (size_t)row * N;
        const uint8_t* __restrict__ s0 = a.supp + base + (size_t)row * N;
#pragma unroll
        for (int j = 0; j < V; ++j) {
            rn[j] = r0[t + T * j];
            sbits |= (s0[t + T * j] != 0 ? 1u : 0u) << j;
        }
    }
    __syncthreads();   // twiddle table
#pragma unroll
    for (int j = 0; j < V; ++j) v[j] = rn[j];
    fft_line_p<PL, +1, true>(v, racc, tw, t);
#pragma unroll
    for (int j = 0; j < V; ++j) rp[T * j + ((T * j) >> SH)] = v[j];

    for (int it = 0;; ++it) {
        __syncthreads();
        // column phase: colIFFT(P_M(colFFT(.)))
#pragma unroll
        for (int j = 0; j < V; ++j) v[j] = cacc.template ldk<T>(tc, j);
        fft_line_p<PL, +1, true>(v, cacc, tw, tc);
#pragma unroll
        for (int j = 0; j < V; ++j) v[j] = project_modulus(v[j], at[aoff + T * j]);
        fft_line_p<PL, -1, true>(v, cacc, tw, tc);
#pragma unroll
        for (int j = 0; j < V; ++j) cacc.template stk<T>(tc, j, v[j]);
        __syncthreads();
        // row phase: rho' = rowIFFT(.)/N^2
#pragma unroll
        for (int j = 0; j < V; ++j) v[j] = rp[T * j + ((T * j) >> SH)];
        fft_line_p<PL, -1, true>(v, racc, tw, t);
        if (it == a.niter) {   // the trailing projection: rho <- rho'
            float2* __restrict__ r0 = a.rho + base + (size_t)row * N;
#pragma unroll
            for (int j = 0; j < V; ++j) r0[t + T * j] = make_float2(v[j].x * a.inv_nn, v[j].y * a.inv_nn);
            break;
        }
#pragma unroll
        for (int j = 0; j < V; ++j) {
            const float2 d = make_float2(v[j].x * a.inv_nn, v[j].y * a.inv_nn);
            const bool s = (sbits >> j) & 1u;
            float2 o;
            if constexpr (MODE == ROW_ER) {
                const float ts = (s && !phase_fails(d.x, d.y, a.pc)) ? 1.f : 0.f;
                o = make_float2(d.x * ts, d.y * ts);
            } else if constexpr (MODE == ROW_HIO) {
                const float2 b = rn[j];
                const float ts = (s && !phase_fails(d.x, d.y, a.pc)) ? 1.f : 0.f;
                o = make_float2((b.x - a.beta * d.x) * (1.f - ts) + d.x * ts,
                                (b.y - a.beta * d.y) * (1.f - ts) + d.y * ts);
            } else {   // RAAR
                const float2 b = rn[j];
                const float qx = 2.f * d.x - b.x, qy = 2.f * d.y - b.y;
                o = d;
                if (!s || phase_fails(qx, qy, a.pc)) {
                    const float c = 1.f - 2.f * a.beta;
                    o = make_float2(a.beta * b.x + c * d.x, a.beta * b.y + c * d.y);
                }
            }
            rn[j] = o;
            v[j] = o;
        }
        fft_line_p<PL, +1, true>(v, racc, tw, t);
#pragma unroll
        for (int j = 0; j < V; ++j) rp[T * j + ((T * j) >> SH)] = v[j];
    }
}

}  // namespace spr
