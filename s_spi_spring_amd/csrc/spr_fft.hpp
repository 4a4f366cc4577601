// Device-side FFT building blocks for gfx950 (CDNA4).
//
// One "line" (a row or a column of an N x N field) is transformed by T = N/8
// threads that each hold 8 complex values in registers. The transform is a
// Stockham autosort FFT: stage s (radix R, span Ns = product of the earlier
// radices) takes x[b + r*N/R] for butterfly b, multiplies input r by
// w^(r*(b mod Ns)) with w = exp(SIGN*2*pi*i/(Ns*R)), applies an R-point DFT and
// writes y[(b/Ns)*Ns*R + (b mod Ns) + q*Ns]. Stages exchange data through LDS;
// the first and the last stage are radix 8 so that thread t holds
// x[t + r*T] on entry and X[t + q*T] on exit. A forward and an inverse transform
// can therefore be chained in registers, with a per-element operation between
// them, without touching LDS.
//
// Twiddles come from a per-workgroup LDS table laid out per stage
// (TwTable<N>): for stage s with span Ns and radix R, row k in [0, Ns) holds
// w^(r*k), r = 1..R-1, densely at the odd stride R-1 so that the 64 lanes of a
// wave, which read consecutive k, spread over the LDS banks. The table is built from
// root[m] = exp(-2*pi*i*m/N), computed in double on the host and rounded to
// float, so every twiddle equals the float rounding of the exact root.
#pragma once

#include <hip/hip_runtime.h>

namespace spr {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * (S*i), S = +1 or -1
template <int S> __device__ __forceinline__ float2 mul_si(float2 a) {
    return S > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int S> __device__ __forceinline__ void dft2(float2& a0, float2& a1) {
    const float2 t = a0;
    a0 = cadd(t, a1);
    a1 = csub(t, a1);
}

// X[k] = sum_n a[n] exp(S*2*pi*i*n*k/4)
template <int S> __device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
    const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_si<S>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

// 8-point DFT as two 4-point DFTs (even/odd inputs) combined with w8^k.
template <int S> __device__ __forceinline__ void dft8(float2* v) {
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<S>(e0, e1, e2, e3);
    dft4<S>(o0, o1, o2, o3);
    const float h = 0.70710678118654752440f;
    // w8 = (1 + S i)/sqrt2, w8^2 = S i, w8^3 = (-1 + S i)/sqrt2
    const float2 p1 = make_float2((o1.x - S * o1.y) * h, (o1.y + S * o1.x) * h);
    const float2 p2 = mul_si<S>(o2);
    const float2 p3 = make_float2((-o3.x - S * o3.y) * h, (-o3.y + S * o3.x) * h);
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, p1);
    v[5] = csub(e1, p1);
    v[2] = cadd(e2, p2);
    v[6] = csub(e2, p2);
    v[3] = cadd(e3, p3);
    v[7] = csub(e3, p3);
}

template <int R, int S> __device__ __forceinline__ void dftR(float2* v) {
    if constexpr (R == 8) dft8<S>(v);
    else if constexpr (R == 4) dft4<S>(v[0], v[1], v[2], v[3]);
    else dft2<S>(v[0], v[1]);
}

// Radix plan: N = 8 * middle * 8 with middle in {1, 2, 4, 8, 16(=2x8), ...}.
// Stage radices (first and last are 8).
template <int N> struct RadixPlan;
template <> struct RadixPlan<64>   { static constexpr int ns = 2; static constexpr int r[4] = {8, 8, 0, 0}; };
template <> struct RadixPlan<128>  { static constexpr int ns = 3; static constexpr int r[4] = {8, 2, 8, 0}; };
template <> struct RadixPlan<256>  { static constexpr int ns = 3; static constexpr int r[4] = {8, 4, 8, 0}; };
template <> struct RadixPlan<512>  { static constexpr int ns = 3; static constexpr int r[4] = {8, 8, 8, 0}; };
template <> struct RadixPlan<1024> { static constexpr int ns = 4; static constexpr int r[4] = {8, 2, 8, 8}; };
template <> struct RadixPlan<2048> { static constexpr int ns = 4; static constexpr int r[4] = {8, 4, 8, 8}; };

template <int N> constexpr int span_before(int s) {
    int ns = 1;
    for (int i = 0; i < s; ++i) ns *= RadixPlan<N>::r[i];
    return ns;
}

// Per-stage twiddle tables in LDS. Stage s >= 1 (radix R, span NS) needs
// w^(r*k) = root[r*k*N/(NS*R)] for k < NS, 1 <= r < R; they are stored densely
// at tws[off(s) + k*(R-1) + r-1]. R is a power of two, so the row stride R-1 is
// odd and the 8-byte reads of 16 consecutive k fall in distinct LDS bank pairs;
// one base address serves all r.
template <int N> struct TwTable {
    static constexpr int off(int s) {
        int o = 0;
        for (int i = 1; i < s; ++i) o += span_before<N>(i) * (RadixPlan<N>::r[i] - 1);
        return o;
    }
    static constexpr int SIZE = off(RadixPlan<N>::ns);
};

template <int N, int S>
__device__ __forceinline__ void tw_fill_stage(float2* tws, const float2* __restrict__ root, int tid, int nt) {
    if constexpr (S < RadixPlan<N>::ns) {
        constexpr int R = RadixPlan<N>::r[S];
        constexpr int NS = span_before<N>(S);
        constexpr int STEP = N / (NS * R);
        constexpr int OFF = TwTable<N>::off(S);
        for (int e = tid; e < NS * (R - 1); e += nt) {
            const int k = e / (R - 1), r = e % (R - 1) + 1;
            tws[OFF + k * (R - 1) + r - 1] = root[r * k * STEP];
        }
        tw_fill_stage<N, S + 1>(tws, root, tid, nt);
    }
}

// Fill the per-stage tables from the global root table exp(-2 pi i m/N) (the
// caller synchronises before the first transform).
template <int N>
__device__ __forceinline__ void tw_fill(float2* tws, const float2* __restrict__ root, int tid, int nt) {
    tw_fill_stage<N, 1>(tws, root, tid, nt);
}

// Twiddle and butterfly of stage STAGE on registers v (8 values = 8/R butterflies).
template <int N, int STAGE, int SIGN>
__device__ __forceinline__ void stage_compute(float2 (&v)[8], const float2* __restrict__ tws, int t) {
    constexpr int R = RadixPlan<N>::r[STAGE];
    constexpr int NS = span_before<N>(STAGE);
    constexpr int T = N / 8;
    constexpr int NB = 8 / R;
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        if constexpr (NS > 1) {
            const int b = t + T * u;
            const int k = b & (NS - 1);
            const float2* w = tws + TwTable<N>::off(STAGE) + k * (R - 1);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 wr = w[r - 1];
                if (SIGN > 0) wr.y = -wr.y;
                v[u * R + r] = cmul(v[u * R + r], wr);
            }
        }
        dftR<R, SIGN>(&v[u * R]);
    }
}

// Stage inputs of stage STAGE from LDS: v[u*R + r] = x[b + r*N/R], b = t + T*u.
template <int N, int STAGE, class Acc>
__device__ __forceinline__ void stage_load(float2 (&v)[8], const Acc& acc, int t) {
    constexpr int R = RadixPlan<N>::r[STAGE];
    constexpr int T = N / 8;
    constexpr int NB = 8 / R;
    constexpr int M = N / R;
#pragma unroll
    for (int u = 0; u < NB; ++u)
#pragma unroll
        for (int r = 0; r < R; ++r) v[u * R + r] = acc.template ldk<M>(t + T * u, r);
}

// Stage outputs of stage STAGE to LDS: y[(b/Ns)*Ns*R + b%Ns + q*Ns] = v[u*R + q].
template <int N, int STAGE, class Acc>
__device__ __forceinline__ void stage_store(const float2 (&v)[8], const Acc& acc, int t) {
    constexpr int R = RadixPlan<N>::r[STAGE];
    constexpr int NS = span_before<N>(STAGE);
    constexpr int T = N / 8;
    constexpr int NB = 8 / R;
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        const int b = t + T * u;
        const int base = (b / NS) * NS * R + (b & (NS - 1));
#pragma unroll
        for (int q = 0; q < R; ++q) acc.st(base + q * NS, v[u * R + q]);
    }
}

// Synchronisation between the LDS exchanges of a line transform: a workgroup
// barrier, or, when every line lives inside one wavefront (WAVE), a wavefront
// fence. LDS operations of one wavefront execute in program order, so the
// fence only has to keep the compiler from moving LDS accesses across it.
template <bool WAVE> __device__ __forceinline__ void line_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// Full line transform. On entry v[r] = x[t + r*N/8]; on exit v[q] = X[t + q*N/8]
// with X[k] = sum_j x[j] exp(SIGN*2*pi*i*j*k/N) (unnormalised). With WAVE = false
// every thread of the workgroup must call it (it contains workgroup barriers);
// with WAVE = true the N/8 threads of each line must lie in one wavefront and
// the line's LDS must be touched by that wavefront only.
template <int N, int SIGN, bool WAVE = false, class Acc>
__device__ __forceinline__ void fft_line(float2 (&v)[8], const Acc& acc, const float2* __restrict__ tw, int t) {
    constexpr int NSTAGE = RadixPlan<N>::ns;
    stage_compute<N, 0, SIGN>(v, tw, t);
    stage_store<N, 0>(v, acc, t);
    line_sync<WAVE>();
    stage_load<N, 1>(v, acc, t);
    stage_compute<N, 1, SIGN>(v, tw, t);
    if constexpr (NSTAGE > 2) {
        line_sync<WAVE>();
        stage_store<N, 1>(v, acc, t);
        line_sync<WAVE>();
        stage_load<N, 2>(v, acc, t);
        stage_compute<N, 2, SIGN>(v, tw, t);
    }
    if constexpr (NSTAGE > 3) {
        line_sync<WAVE>();
        stage_store<N, 2>(v, acc, t);
        line_sync<WAVE>();
        stage_load<N, 3>(v, acc, t);
        stage_compute<N, 3, SIGN>(v, tw, t);
    }
    // Leave LDS free for the next user of this buffer.
    line_sync<WAVE>();
}

}  // namespace spr
