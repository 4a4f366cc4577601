// Device-side FFT building blocks for gfx950 (CDNA4).
//
// One "line" (a row or a column of an N x N field) is transformed by
// T = N/V threads that each hold V complex values in registers (V = 8 for
// power-of-two N, V = 12 for N = 384 and 768). The transform is a Stockham
// autosort FFT: stage s (radix R, span Ns = product of the earlier radices)
// takes x[b + r*N/R] for butterfly b, multiplies input r by w^(r*(b mod Ns))
// with w = exp(SIGN*2*pi*i/(Ns*R)), applies an R-point DFT and writes
// y[(b/Ns)*Ns*R + (b mod Ns) + q*Ns]. Thread t owns the butterflies t + T*u,
// u < V/R, of every stage. Stages exchange data through LDS; thread t holds
// x[t + j*T] (j < V) on entry and X[t + j*T] on exit (the first and the last
// stage see a compile-time permutation of these registers when their radix
// is below V). A forward and an inverse transform can therefore be chained in
// registers, with a per-element operation between them, without touching LDS.
//
// Twiddles come from a per-workgroup LDS table laid out per stage
// (TwTable<N>): for stage s with span Ns and radix R, row k in [0, Ns) holds
// w^(r*k), r = 1..R-1, densely at the stride R-1 (odd for every radix used),
// so that the 64 lanes of a wave, which read consecutive k, spread over the LDS
// banks. The table is built from root[m] = exp(-2*pi*i*m/N), computed in double
// on the host and rounded to float, so every twiddle equals the float rounding
// of the exact root.
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

constexpr float kH3 = 0.86602540378443864676f;   // sqrt(3)/2

// X[k] = sum_n a[n] exp(S*2*pi*i*n*k/3)
template <int S> __device__ __forceinline__ void dft3(float2& a0, float2& a1, float2& a2) {
    const float2 t1 = cadd(a1, a2), t2 = csub(a1, a2);
    const float2 m1 = make_float2(a0.x - 0.5f * t1.x, a0.y - 0.5f * t1.y);
    const float2 m2 = mul_si<S>(make_float2(kH3 * t2.x, kH3 * t2.y));
    a0 = cadd(a0, t1);
    a1 = cadd(m1, m2);
    a2 = csub(m1, m2);
}

// z * exp(S*2*pi*i*M/12), M a compile-time multiple of 30 degrees.
template <int M, int S> __device__ __forceinline__ float2 rot12(float2 z) {
    constexpr int m = M % 12;
    if constexpr (m == 0) {
        return z;
    } else if constexpr (m == 3) {
        return mul_si<S>(z);
    } else if constexpr (m == 6) {
        return make_float2(-z.x, -z.y);
    } else {
        static_assert(m == 1 || m == 2 || m == 4, "twiddle");
        // cos and sin of m*30 degrees
        constexpr float c = m == 1 ? kH3 : m == 2 ? 0.5f : -0.5f;
        constexpr float s = m == 1 ? 0.5f : kH3;
        return cmul(z, make_float2(c, S * s));
    }
}

// 6-point DFT as 2 x 3 (Cooley-Tukey, n = n1 + 2*n2, k = k2 + 3*k1).
template <int S> __device__ __forceinline__ void dft6(float2* v) {
    float2 a0 = v[0], a1 = v[2], a2 = v[4];
    float2 b0 = v[1], b1 = v[3], b2 = v[5];
    dft3<S>(a0, a1, a2);
    dft3<S>(b0, b1, b2);
    b1 = rot12<2, S>(b1);
    b2 = rot12<4, S>(b2);
    v[0] = cadd(a0, b0);
    v[3] = csub(a0, b0);
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[5] = csub(a2, b2);
}

// 12-point DFT as 4 x 3 (Cooley-Tukey, n = n1 + 4*n2, k = k2 + 3*k1): four
// 3-point DFTs, the twiddles w12^(n1*k2), three 4-point DFTs.
template <int S> __device__ __forceinline__ void dft12(float2* v) {
    float2 y0[3] = {v[0], v[4], v[8]}, y1[3] = {v[1], v[5], v[9]};
    float2 y2[3] = {v[2], v[6], v[10]}, y3[3] = {v[3], v[7], v[11]};
    dft3<S>(y0[0], y0[1], y0[2]);
    dft3<S>(y1[0], y1[1], y1[2]);
    dft3<S>(y2[0], y2[1], y2[2]);
    dft3<S>(y3[0], y3[1], y3[2]);
    y1[1] = rot12<1, S>(y1[1]);
    y1[2] = rot12<2, S>(y1[2]);
    y2[1] = rot12<2, S>(y2[1]);
    y2[2] = rot12<4, S>(y2[2]);
    y3[1] = rot12<3, S>(y3[1]);
    y3[2] = rot12<6, S>(y3[2]);
#pragma unroll
    for (int k2 = 0; k2 < 3; ++k2) {
        float2 a0 = y0[k2], a1 = y1[k2], a2 = y2[k2], a3 = y3[k2];
        dft4<S>(a0, a1, a2, a3);
        v[k2] = a0;
        v[k2 + 3] = a1;
        v[k2 + 6] = a2;
        v[k2 + 9] = a3;
    }
}

// z * exp(S*2*pi*i*M/16), M a compile-time multiple of 22.5 degrees.
template <int M, int S> __device__ __forceinline__ float2 rot16(float2 z) {
    constexpr int m = M % 16;
    if constexpr (m == 0) {
        return z;
    } else if constexpr (m == 4) {
        return mul_si<S>(z);
    } else if constexpr (m == 8) {
        return make_float2(-z.x, -z.y);
    } else if constexpr (m == 12) {
        return mul_si<-S>(z);
    } else {
        static_assert(m == 1 || m == 2 || m == 3 || m == 6 || m == 9, "twiddle");
        constexpr float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f, h = 0.70710678118654752440f;
        // cos and sin of m*22.5 degrees
        constexpr float c = m == 1 ? c1 : m == 2 ? h : m == 3 ? s1 : m == 6 ? -h : -c1;
        constexpr float s = m == 1 ? s1 : m == 2 ? h : m == 3 ? c1 : m == 6 ? h : -s1;
        return cmul(z, make_float2(c, S * s));
    }
}

// 16-point DFT as 4 x 4 (Cooley-Tukey, n = n1 + 4*n2, k = k2 + 4*k1).
template <int S> __device__ __forceinline__ void dft16(float2* v) {
    float2 y[4][4];
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) {
        y[n1][0] = v[n1];
        y[n1][1] = v[n1 + 4];
        y[n1][2] = v[n1 + 8];
        y[n1][3] = v[n1 + 12];
        dft4<S>(y[n1][0], y[n1][1], y[n1][2], y[n1][3]);
    }
    y[1][1] = rot16<1, S>(y[1][1]);
    y[1][2] = rot16<2, S>(y[1][2]);
    y[1][3] = rot16<3, S>(y[1][3]);
    y[2][1] = rot16<2, S>(y[2][1]);
    y[2][2] = rot16<4, S>(y[2][2]);
    y[2][3] = rot16<6, S>(y[2][3]);
    y[3][1] = rot16<3, S>(y[3][1]);
    y[3][2] = rot16<6, S>(y[3][2]);
    y[3][3] = rot16<9, S>(y[3][3]);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
        float2 a0 = y[0][k2], a1 = y[1][k2], a2 = y[2][k2], a3 = y[3][k2];
        dft4<S>(a0, a1, a2, a3);
        v[k2] = a0;
        v[k2 + 4] = a1;
        v[k2 + 8] = a2;
        v[k2 + 12] = a3;
    }
}

template <int R, int S> __device__ __forceinline__ void dftR(float2* v) {
    if constexpr (R == 16) dft16<S>(v);
    else if constexpr (R == 12) dft12<S>(v);
    else if constexpr (R == 8) dft8<S>(v);
    else if constexpr (R == 6) dft6<S>(v);
    else if constexpr (R == 4) dft4<S>(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 3) dft3<S>(v[0], v[1], v[2]);
    else dft2<S>(v[0], v[1]);
}

// A radix plan: the N-point transform with V values per thread in ns stages of
// radices r[] (each dividing V; unused entries 0).
template <int N_, int V_, int R0, int R1, int R2 = 0, int R3 = 0> struct Plan {
    static constexpr int N = N_;
    static constexpr int V = V_;
    static constexpr int ns = R2 == 0 ? 2 : R3 == 0 ? 3 : 4;
    static constexpr int r[4] = {R0, R1, R2, R3};
    static_assert(R0 * R1 * (R2 ? R2 : 1) * (R3 ? R3 : 1) == N_, "radices must multiply to N");
    static_assert(V_ % R0 == 0 && V_ % R1 == 0 && (R2 == 0 || V_ % R2 == 0) && (R3 == 0 || V_ % R3 == 0),
                  "every radix must divide the values per thread");
};

// Default plan of each supported size.
template <int N> struct RadixPlan;
template <> struct RadixPlan<64> : Plan<64, 8, 8, 8> {};
template <> struct RadixPlan<128> : Plan<128, 8, 8, 2, 8> {};
template <> struct RadixPlan<256> : Plan<256, 8, 8, 4, 8> {};
template <> struct RadixPlan<384> : Plan<384, 12, 12, 4, 4, 2> {};
template <> struct RadixPlan<512> : Plan<512, 8, 8, 8, 8> {};
template <> struct RadixPlan<768> : Plan<768, 12, 12, 4, 4, 4> {};
template <> struct RadixPlan<1024> : Plan<1024, 8, 8, 2, 8, 8> {};
template <> struct RadixPlan<2048> : Plan<2048, 8, 8, 4, 8, 8> {};

template <class PL> constexpr int span_before(int s) {
    int ns = 1;
    for (int i = 0; i < s; ++i) ns *= PL::r[i];
    return ns;
}

// Per-stage twiddle tables in LDS. Stage s >= 1 (radix R, span NS) needs
// w^(r*k) = root[r*k*N/(NS*R)] for k < NS, 1 <= r < R; they are stored densely
// at tws[off(s) + k*(R-1) + r-1]. The row stride R-1 is odd (R in {2, 4, 8, 12, 16}),
// so the 8-byte reads of 16 consecutive k fall in distinct LDS bank pairs; one
// base address serves all r.
template <class PL> struct TwTableP {
    static constexpr int off(int s) {
        int o = 0;
        for (int i = 1; i < s; ++i) o += span_before<PL>(i) * (PL::r[i] - 1);
        return o;
    }
    static constexpr int SIZE = off(PL::ns);
};
template <int N> using TwTable = TwTableP<RadixPlan<N>>;

template <class PL, int S>
__device__ __forceinline__ void tw_fill_stage(float2* tws, const float2* __restrict__ root, int tid, int nt) {
    if constexpr (S < PL::ns) {
        constexpr int R = PL::r[S];
        constexpr int NS = span_before<PL>(S);
        constexpr int STEP = PL::N / (NS * R);
        constexpr int OFF = TwTableP<PL>::off(S);
        for (int e = tid; e < NS * (R - 1); e += nt) {
            const int k = e / (R - 1), r = e % (R - 1) + 1;
            tws[OFF + k * (R - 1) + r - 1] = root[r * k * STEP];
        }
        tw_fill_stage<PL, S + 1>(tws, root, tid, nt);
    }
}

// Fill the per-stage tables from the global root table exp(-2 pi i m/N) (the
// caller synchronises before the first transform).
template <class PL>
__device__ __forceinline__ void tw_fill_p(float2* tws, const float2* __restrict__ root, int tid, int nt) {
    tw_fill_stage<PL, 1>(tws, root, tid, nt);
}
template <int N>
__device__ __forceinline__ void tw_fill(float2* tws, const float2* __restrict__ root, int tid, int nt) {
    tw_fill_p<RadixPlan<N>>(tws, root, tid, nt);
}

// Twiddle and butterfly of stage STAGE on registers v (V values = V/R butterflies).
template <class PL, int STAGE, int SIGN>
__device__ __forceinline__ void stage_compute(float2 (&v)[PL::V], const float2* __restrict__ tws, int t) {
    constexpr int R = PL::r[STAGE];
    constexpr int NS = span_before<PL>(STAGE);
    constexpr int T = PL::N / PL::V;
    constexpr int NB = PL::V / R;
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        if constexpr (NS > 1) {
            const int b = t + T * u;
            const int k = b % NS;
            const float2* w = tws + TwTableP<PL>::off(STAGE) + k * (R - 1);
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
template <class PL, int STAGE, class Acc>
__device__ __forceinline__ void stage_load(float2 (&v)[PL::V], const Acc& acc, int t) {
    constexpr int R = PL::r[STAGE];
    constexpr int T = PL::N / PL::V;
    constexpr int NB = PL::V / R;
    constexpr int M = PL::N / R;
#pragma unroll
    for (int u = 0; u < NB; ++u)
#pragma unroll
        for (int r = 0; r < R; ++r) v[u * R + r] = acc.template ldk<M>(t + T * u, r);
}

// Stage outputs of stage STAGE to LDS: y[(b/Ns)*Ns*R + b%Ns + q*Ns] = v[u*R + q].
template <class PL, int STAGE, class Acc>
__device__ __forceinline__ void stage_store(const float2 (&v)[PL::V], const Acc& acc, int t) {
    constexpr int R = PL::r[STAGE];
    constexpr int NS = span_before<PL>(STAGE);
    constexpr int T = PL::N / PL::V;
    constexpr int NB = PL::V / R;
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        const int b = t + T * u;
        const int base = (b / NS) * NS * R + b % NS;
#pragma unroll
        for (int q = 0; q < R; ++q) acc.st(base + q * NS, v[u * R + q]);
    }
}

// Register order of the first stage (radix R0, NB0 = V/R0 butterflies): it reads
// x[t + T*(u + r*NB0)] as v[u*R0 + r], while the caller supplies x[t + T*j] as
// v[j]. The last stage (radix RL, NBL = V/RL) leaves X[t + T*(u + q*NBL)] in
// v[u*RL + q]. Both reorderings are compile-time register renames.
template <int V, int R> __device__ __forceinline__ void regs_to_stage(float2 (&v)[V]) {
    if constexpr (R < V) {
        constexpr int NB = V / R;
        float2 w[V];
#pragma unroll
        for (int u = 0; u < NB; ++u)
#pragma unroll
            for (int r = 0; r < R; ++r) w[u * R + r] = v[u + r * NB];
#pragma unroll
        for (int j = 0; j < V; ++j) v[j] = w[j];
    }
}
template <int V, int R> __device__ __forceinline__ void stage_to_regs(float2 (&v)[V]) {
    if constexpr (R < V) {
        constexpr int NB = V / R;
        float2 w[V];
#pragma unroll
        for (int u = 0; u < NB; ++u)
#pragma unroll
            for (int q = 0; q < R; ++q) w[u + q * NB] = v[u * R + q];
#pragma unroll
        for (int j = 0; j < V; ++j) v[j] = w[j];
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

// Full line transform with plan PL. On entry v[j] = x[t + j*T]; on exit
// v[j] = X[t + j*T] (T = N/V) with X[k] = sum_n x[n] exp(SIGN*2*pi*i*n*k/N)
// (unnormalised). With WAVE = false every thread of the workgroup must call it
// (it contains workgroup barriers); with WAVE = true the T threads of each line
// must lie in one wavefront and the line's LDS must be touched by that
// wavefront only.
template <class PL, int SIGN, bool WAVE = false, class Acc>
__device__ __forceinline__ void fft_line_p(float2 (&v)[PL::V], const Acc& acc, const float2* __restrict__ tw, int t) {
    constexpr int NSTAGE = PL::ns;
    constexpr int V = PL::V;
    regs_to_stage<V, PL::r[0]>(v);
    stage_compute<PL, 0, SIGN>(v, tw, t);
    stage_store<PL, 0>(v, acc, t);
    line_sync<WAVE>();
    stage_load<PL, 1>(v, acc, t);
    stage_compute<PL, 1, SIGN>(v, tw, t);
    if constexpr (NSTAGE > 2) {
        line_sync<WAVE>();
        stage_store<PL, 1>(v, acc, t);
        line_sync<WAVE>();
        stage_load<PL, 2>(v, acc, t);
        stage_compute<PL, 2, SIGN>(v, tw, t);
    }
    if constexpr (NSTAGE > 3) {
        line_sync<WAVE>();
        stage_store<PL, 2>(v, acc, t);
        line_sync<WAVE>();
        stage_load<PL, 3>(v, acc, t);
        stage_compute<PL, 3, SIGN>(v, tw, t);
    }
    stage_to_regs<V, PL::r[NSTAGE - 1]>(v);
    // Leave LDS free for the next user of this buffer.
    line_sync<WAVE>();
}

// The default plan of size N.
template <int N, int SIGN, bool WAVE = false, class Acc>
__device__ __forceinline__ void fft_line(float2 (&v)[RadixPlan<N>::V], const Acc& acc, const float2* __restrict__ tw,
                                         int t) {
    fft_line_p<RadixPlan<N>, SIGN, WAVE>(v, acc, tw, t);
}

}  // namespace spr
