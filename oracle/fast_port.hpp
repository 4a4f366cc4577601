// Throughput port of the FP32 oracle's HIO/RAAR/ER driver, for the CPU baseline.
//
// TEST INFRASTRUCTURE ONLY, like the rest of oracle/: it is compiled into
// oracle/liboracle.so and called by the cpu_baseline leg of bench.py and by
// tests/test_oracle_fast_port.py, never by the product.
//
// It computes what orc::iterate<float> computes (GPU_context::HIO / RAAR / ER,
// src/gpu/gpu_context.cu:778-902: n iterations of [FT; P_M; IFT; support
// update], then one trailing [FT; P_M; IFT]) with the same per-pixel operators
// (the arithmetic of orc::project_modulus and of the updates, the same phase
// window), but with
// a transform written for speed on the host's vector units instead of for
// clarity:
//   * radix-8/4/2/3 Stockham stages (the index map of orc::fft_line) with the
//     twiddles of every stage tabulated once per plan (computed in double,
//     rounded to float, as orc::Plan does);
//   * 16 lines transformed at once, held split (real and imaginary planes) and
//     interleaved line-minor, so that every butterfly is a loop over 16 lanes
//     that the compiler turns into AVX2 vector instructions;
//   * the updates and the projection run as branch-free loops that the
//     compiler vectorises; the phase-window test calls atan2 only where the
//     window can fail: at reality 0 only for values with a negative real part
//     and |Im| <= 1e-6 |Re| (every other value has |atan2| < pi - 1e-6 and
//     passes), and the measured pixels of modulus 0 go through
//     orc::project_modulus itself.
// The transforms round differently from orc::fft_line (radix-8 and radix-4
// butterflies built from exact +-i rotations instead of direct R-point sums), so the results agree with the
// FP32 oracle to FP32 rounding, not bit for bit:
// tests/test_oracle_fast_port.py pins the transform against numpy FP64 and the
// driver against orc::iterate<float> (relative L2 <= 1e-5), and the update step
// bit for bit against orc::apply_support_* on values at the window's edges.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "spring_oracle.hpp"

namespace orc {
namespace fastport {

constexpr int LANES = 16;   // lines per transform block

struct Stage {
    int R, Ns;
    std::vector<float> wr, wi;   // w^(r*k), w = exp(-2 pi i/(Ns R)): index (r-1)*Ns + k
};

struct LinePlan {
    int n;
    std::vector<Stage> stages;
    explicit LinePlan(int n_) : n(n_) {
        if (n < 1) throw std::invalid_argument("FFT length must be >= 1");
        std::vector<int> rad;
        int m = n;
        while (m % 8 == 0) { rad.push_back(8); m /= 8; }
        while (m % 4 == 0) { rad.push_back(4); m /= 4; }
        while (m % 2 == 0) { rad.push_back(2); m /= 2; }
        while (m % 3 == 0) { rad.push_back(3); m /= 3; }
        if (m != 1) throw std::invalid_argument("fast port: lengths 2^a 3^b only");
        int Ns = 1;
        for (int R : rad) {
            Stage s{R, Ns, std::vector<float>((size_t)(R - 1) * Ns), std::vector<float>((size_t)(R - 1) * Ns)};
            for (int r = 1; r < R; ++r)
                for (int k = 0; k < Ns; ++k) {
                    const double a = -2.0 * M_PI * (double)(r * k) / (double)(Ns * R);
                    s.wr[(size_t)(r - 1) * Ns + k] = (float)std::cos(a);
                    s.wi[(size_t)(r - 1) * Ns + k] = (float)std::sin(a);
                }
            stages.push_back(std::move(s));
            Ns *= R;
        }
    }
};

// Per-thread work planes: two split-complex buffers of n x LANES floats and the
// update's mask.
struct Work {
    std::vector<float> ar, ai, br, bi;
    std::vector<uint8_t> undecided;   // n x n: pixels whose phase window needs atan2
    explicit Work(int n)
        : ar((size_t)n * LANES), ai((size_t)n * LANES), br((size_t)n * LANES), bi((size_t)n * LANES),
          undecided((size_t)n * n) {}
};

// One Stockham stage over LANES lines: x[j + r*m] (j < m = n/R), twiddled by
// w^(r*(j mod Ns)), R-point DFT over r, into y[(j/Ns)*Ns*R + j mod Ns + q*Ns].
// sign = +1 uses the conjugate twiddles (kernel exp(+2 pi i ...)).
inline void stage(const Stage& s, int n, int sign, const float* __restrict__ xr, const float* __restrict__ xi,
                  float* __restrict__ yr, float* __restrict__ yi) {
    constexpr int L = LANES;
    const int R = s.R, Ns = s.Ns, m = n / R;
    const float sg = sign > 0 ? 1.f : -1.f;
    for (int j = 0; j < m; ++j) {
        const int k = j % Ns;
        const int base = (j / Ns) * Ns * R + k;
        float wr[7] = {1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f}, wi[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int r = 1; r < R; ++r) {
            wr[r - 1] = s.wr[(size_t)(r - 1) * Ns + k];
            wi[r - 1] = sg > 0 ? -s.wi[(size_t)(r - 1) * Ns + k] : s.wi[(size_t)(r - 1) * Ns + k];
        }
        const float* x0r = xr + (size_t)j * L;
        const float* x0i = xi + (size_t)j * L;
        if (R == 8) {
            // Two 4-point DFTs (even and odd inputs), combined with the powers of
            // W = exp(sign 2 pi i / 8): y_q = E_q + W^q O_q, y_(q+4) = E_q - W^q O_q.
            const float c = 0.707106781186547524400844f;
            // offsets into the restrict-qualified planes (so that the lane loop vectorises)
            size_t xo[8], yo[8];
            for (int r = 0; r < 8; ++r) {
                xo[r] = (size_t)(j + r * m) * L;
                yo[r] = (size_t)(base + r * Ns) * L;
            }
#pragma GCC ivdep   // x and y are distinct planes
            for (int b = 0; b < L; ++b) {
                float vr[8], vi[8];
                vr[0] = xr[xo[0] + b];
                vi[0] = xi[xo[0] + b];
                for (int r = 1; r < 8; ++r) {
                    const float a = xr[xo[r] + b], bb = xi[xo[r] + b];
                    vr[r] = a * wr[r - 1] - bb * wi[r - 1];
                    vi[r] = a * wi[r - 1] + bb * wr[r - 1];
                }
                float er[4], ei[4], orr[4], oi[4];
                for (int h = 0; h < 2; ++h) {   // h = 0: inputs 0,2,4,6; h = 1: inputs 1,3,5,7
                    const float a0r = vr[h], a0i = vi[h], a1r = vr[h + 2], a1i = vi[h + 2];
                    const float a2r = vr[h + 4], a2i = vi[h + 4], a3r = vr[h + 6], a3i = vi[h + 6];
                    const float t0r = a0r + a2r, t0i = a0i + a2i, t1r = a0r - a2r, t1i = a0i - a2i;
                    const float t2r = a1r + a3r, t2i = a1i + a3i;
                    const float dr = a1r - a3r, di = a1i - a3i;
                    const float t3r = -sg * di, t3i = sg * dr;   // (sign i)(a1 - a3)
                    float* pr = h ? orr : er;
                    float* pi = h ? oi : ei;
                    pr[0] = t0r + t2r; pi[0] = t0i + t2i;
                    pr[1] = t1r + t3r; pi[1] = t1i + t3i;
                    pr[2] = t0r - t2r; pi[2] = t0i - t2i;
                    pr[3] = t1r - t3r; pi[3] = t1i - t3i;
                }
                // W^1 = c(1 + sign i), W^2 = sign i, W^3 = c(-1 + sign i)
                const float o1r = c * (orr[1] - sg * oi[1]), o1i = c * (oi[1] + sg * orr[1]);
                const float o2r = -sg * oi[2], o2i = sg * orr[2];
                const float o3r = -c * (orr[3] + sg * oi[3]), o3i = c * (sg * orr[3] - oi[3]);
                yr[yo[0] + b] = er[0] + orr[0]; yi[yo[0] + b] = ei[0] + oi[0];
                yr[yo[4] + b] = er[0] - orr[0]; yi[yo[4] + b] = ei[0] - oi[0];
                yr[yo[1] + b] = er[1] + o1r;    yi[yo[1] + b] = ei[1] + o1i;
                yr[yo[5] + b] = er[1] - o1r;    yi[yo[5] + b] = ei[1] - o1i;
                yr[yo[2] + b] = er[2] + o2r;    yi[yo[2] + b] = ei[2] + o2i;
                yr[yo[6] + b] = er[2] - o2r;    yi[yo[6] + b] = ei[2] - o2i;
                yr[yo[3] + b] = er[3] + o3r;    yi[yo[3] + b] = ei[3] + o3i;
                yr[yo[7] + b] = er[3] - o3r;    yi[yo[7] + b] = ei[3] - o3i;
            }
        } else if (R == 4) {
            const float *x1r = x0r + (size_t)m * L, *x1i = x0i + (size_t)m * L;
            const float *x2r = x1r + (size_t)m * L, *x2i = x1i + (size_t)m * L;
            const float *x3r = x2r + (size_t)m * L, *x3i = x2i + (size_t)m * L;
            float *y0r = yr + (size_t)base * L, *y0i = yi + (size_t)base * L;
            float *y1r = y0r + (size_t)Ns * L, *y1i = y0i + (size_t)Ns * L;
            float *y2r = y1r + (size_t)Ns * L, *y2i = y1i + (size_t)Ns * L;
            float *y3r = y2r + (size_t)Ns * L, *y3i = y2i + (size_t)Ns * L;
#pragma GCC ivdep   // x and y are distinct planes
            for (int b = 0; b < L; ++b) {
                const float ar = x0r[b], ai = x0i[b];
                const float br = x1r[b] * wr[0] - x1i[b] * wi[0], bi = x1r[b] * wi[0] + x1i[b] * wr[0];
                const float cr = x2r[b] * wr[1] - x2i[b] * wi[1], ci = x2r[b] * wi[1] + x2i[b] * wr[1];
                const float dr = x3r[b] * wr[2] - x3i[b] * wi[2], di = x3r[b] * wi[2] + x3i[b] * wr[2];
                const float t0r = ar + cr, t0i = ai + ci, t1r = ar - cr, t1i = ai - ci;
                const float t2r = br + dr, t2i = bi + di, t3r = br - dr, t3i = bi - di;
                // (sign i) * t3
                const float ur = -sg * t3i, ui = sg * t3r;
                y0r[b] = t0r + t2r; y0i[b] = t0i + t2i;
                y2r[b] = t0r - t2r; y2i[b] = t0i - t2i;
                y1r[b] = t1r + ur;  y1i[b] = t1i + ui;
                y3r[b] = t1r - ur;  y3i[b] = t1i - ui;
            }
        } else if (R == 2) {
            const float *x1r = x0r + (size_t)m * L, *x1i = x0i + (size_t)m * L;
            float *y0r = yr + (size_t)base * L, *y0i = yi + (size_t)base * L;
            float *y1r = y0r + (size_t)Ns * L, *y1i = y0i + (size_t)Ns * L;
#pragma GCC ivdep   // x and y are distinct planes
            for (int b = 0; b < L; ++b) {
                const float br = x1r[b] * wr[0] - x1i[b] * wi[0], bi = x1r[b] * wi[0] + x1i[b] * wr[0];
                y0r[b] = x0r[b] + br; y0i[b] = x0i[b] + bi;
                y1r[b] = x0r[b] - br; y1i[b] = x0i[b] - bi;
            }
        } else {   // R == 3: y_q = sum_r x_r e^(sign 2 pi i r q / 3)
            const float h = 0.866025403784438646763723f * sg;   // sign * sin(2 pi / 3)
            const float *x1r = x0r + (size_t)m * L, *x1i = x0i + (size_t)m * L;
            const float *x2r = x1r + (size_t)m * L, *x2i = x1i + (size_t)m * L;
            float *y0r = yr + (size_t)base * L, *y0i = yi + (size_t)base * L;
            float *y1r = y0r + (size_t)Ns * L, *y1i = y0i + (size_t)Ns * L;
            float *y2r = y1r + (size_t)Ns * L, *y2i = y1i + (size_t)Ns * L;
#pragma GCC ivdep   // x and y are distinct planes
            for (int b = 0; b < L; ++b) {
                const float br = x1r[b] * wr[0] - x1i[b] * wi[0], bi = x1r[b] * wi[0] + x1i[b] * wr[0];
                const float cr = x2r[b] * wr[1] - x2i[b] * wi[1], ci = x2r[b] * wi[1] + x2i[b] * wr[1];
                const float sr = br + cr, si = bi + ci, dr = br - cr, di = bi - ci;
                const float mr = x0r[b] - 0.5f * sr, mi = x0i[b] - 0.5f * si;
                y0r[b] = x0r[b] + sr; y0i[b] = x0i[b] + si;
                // i * h * d = (-h di, h dr)
                y1r[b] = mr - h * di; y1i[b] = mi + h * dr;
                y2r[b] = mr + h * di; y2i[b] = mi - h * dr;
            }
        }
    }
}

// All stages of LANES lines held in w.ar/w.ai; returns true when the result
// ended in w.br/w.bi.
inline bool transform_block(const LinePlan& pl, int sign, Work& w) {
    float *xr = w.ar.data(), *xi = w.ai.data(), *yr = w.br.data(), *yi = w.bi.data();
    bool in_b = false;
    for (const Stage& s : pl.stages) {
        stage(s, pl.n, sign, xr, xi, yr, yi);
        std::swap(xr, yr);
        std::swap(xi, yi);
        in_b = !in_b;
    }
    return in_b;
}

// Unnormalised 2-D transform of an n x n row-major field: rows, then columns
// (the order of orc::fft2_raw).
inline void fft2(const LinePlan& pl, std::complex<float>* f, int sign, Work& w) {
    constexpr int L = LANES;
    const int n = pl.n;
    for (int y0 = 0; y0 < n; y0 += L) {
        const int nb = std::min(L, n - y0);
        for (int b = 0; b < L; ++b)
            for (int i = 0; i < n; ++i) {
                const std::complex<float> v = b < nb ? f[(size_t)(y0 + b) * n + i] : std::complex<float>(0.f, 0.f);
                w.ar[(size_t)i * L + b] = v.real();
                w.ai[(size_t)i * L + b] = v.imag();
            }
        const bool in_b = transform_block(pl, sign, w);
        const float* rr = in_b ? w.br.data() : w.ar.data();
        const float* ri = in_b ? w.bi.data() : w.ai.data();
        for (int b = 0; b < nb; ++b)
            for (int i = 0; i < n; ++i)
                f[(size_t)(y0 + b) * n + i] = std::complex<float>(rr[(size_t)i * L + b], ri[(size_t)i * L + b]);
    }
    for (int x0 = 0; x0 < n; x0 += L) {
        const int nb = std::min(L, n - x0);
        for (int i = 0; i < n; ++i)
#pragma GCC ivdep   // x and y are distinct planes
            for (int b = 0; b < L; ++b) {
                const std::complex<float> v = b < nb ? f[(size_t)i * n + x0 + b] : std::complex<float>(0.f, 0.f);
                w.ar[(size_t)i * L + b] = v.real();
                w.ai[(size_t)i * L + b] = v.imag();
            }
        const bool in_b = transform_block(pl, sign, w);
        const float* rr = in_b ? w.br.data() : w.ar.data();
        const float* ri = in_b ? w.bi.data() : w.ai.data();
        for (int i = 0; i < n; ++i)
            for (int b = 0; b < nb; ++b)
                f[(size_t)i * n + x0 + b] = std::complex<float>(rr[(size_t)i * L + b], ri[(size_t)i * L + b]);
    }
}

// One pixel of orc::apply_support_er / _hio / _raar, with the window test of
// PhaseWindow::ok (w: PhaseWindow::er_hio for ER and HIO, ::raar for RAAR).
inline void update_pixel(int alg, std::complex<float>& r, int32_t s, std::complex<float>* b, float beta,
                         const PhaseWindow& w) {
    if (alg == ALG_ER) {
        float ts = (float)s;
        if (!w.ok(r)) ts = 0.f;
        r = std::complex<float>(r.real() * ts, r.imag() * ts);
    } else if (alg == ALG_HIO) {
        float ts = (float)s;
        if (!w.ok(r)) ts = 0.f;
        const float dx = r.real(), dy = r.imag();
        const float nx = (b->real() - beta * dx) * (1.f - ts) + dx * ts;
        const float ny = (b->imag() - beta * dy) * (1.f - ts) + dy * ts;
        r = std::complex<float>(nx, ny);
        *b = r;
    } else {
        const std::complex<float> q(2.f * r.real() - b->real(), 2.f * r.imag() - b->imag());
        if (!w.ok(q) || s == 0) {
            r = std::complex<float>(beta * b->real() + (1.f - 2.f * beta) * r.real(),
                                    beta * b->imag() + (1.f - 2.f * beta) * r.imag());
        }
        *b = r;
    }
}

// At reality 0 the window is [-pi, pi] in double and atan2f returns values in
// [-pi_f, pi_f], pi_f > pi, so a value fails it only when atan2f returns +-pi_f.
// A value with a non-negative real part (sign bit clear) or with
// |Im| > 1e-6 |Re| has |atan2f| <= pi - 1e-6 + rounding < pi and passes; only
// the others need atan2.
inline bool passes_without_atan2(float x, float y) { return !std::signbit(x) | (std::fabs(y) > 1e-6f * std::fabs(x)); }

// Branch-free passes of the support update at reality 0 over interleaved
// (re, im) planes: pixels whose outcome is known without atan2 are updated, the
// others are left as they are and marked in und. Returns whether any was marked.
inline int pass_er(float* __restrict__ r, const int32_t* __restrict__ S, uint8_t* __restrict__ und, size_t np) {
    int rare = 0;
    for (size_t p = 0; p < np; ++p) {
        const float x = r[2 * p], y = r[2 * p + 1];
        const bool out = S[p] == 0;
        const bool dec = out | passes_without_atan2(x, y);
        const float ts = dec ? (out ? 0.f : (float)S[p]) : 1.f;   // undecided: x * 1 = x
        r[2 * p] = x * ts;
        r[2 * p + 1] = y * ts;
        und[p] = (uint8_t)!dec;
        rare |= (int)!dec;
    }
    return rare;
}
inline int pass_hio(float* __restrict__ r, float* __restrict__ b, const int32_t* __restrict__ S,
                    uint8_t* __restrict__ und, float beta, size_t np) {
    int rare = 0;
    for (size_t p = 0; p < np; ++p) {
        const float dx = r[2 * p], dy = r[2 * p + 1], bx = b[2 * p], by = b[2 * p + 1];
        const bool out = S[p] == 0;
        const bool dec = out | passes_without_atan2(dx, dy);
        const float ts = out ? 0.f : (float)S[p];
        const float nx = (bx - beta * dx) * (1.f - ts) + dx * ts;
        const float ny = (by - beta * dy) * (1.f - ts) + dy * ts;
        r[2 * p] = dec ? nx : dx;
        r[2 * p + 1] = dec ? ny : dy;
        b[2 * p] = dec ? nx : bx;
        b[2 * p + 1] = dec ? ny : by;
        und[p] = (uint8_t)!dec;
        rare |= (int)!dec;
    }
    return rare;
}
inline int pass_raar(float* __restrict__ r, float* __restrict__ b, const int32_t* __restrict__ S,
                     uint8_t* __restrict__ und, float beta, size_t np) {
    int rare = 0;
    const float c = 1.f - 2.f * beta;
    for (size_t p = 0; p < np; ++p) {
        const float dx = r[2 * p], dy = r[2 * p + 1], bx = b[2 * p], by = b[2 * p + 1];
        const bool out = S[p] == 0;
        const bool dec = out | passes_without_atan2(2.f * dx - bx, 2.f * dy - by);
        // out of the support: the RAAR step; in it and passing: unchanged
        const float nx = out ? beta * bx + c * dx : dx;
        const float ny = out ? beta * by + c * dy : dy;
        r[2 * p] = nx;
        r[2 * p + 1] = ny;
        b[2 * p] = dec ? nx : bx;
        b[2 * p + 1] = dec ? ny : by;
        und[p] = (uint8_t)!dec;
        rare |= (int)!dec;
    }
    return rare;
}

// The support update of orc::apply_support_* over a field. At reality 0 one
// branch-free pass updates every pixel whose outcome is known without atan2 (out
// of the support, or passing the window by the test above) and marks the rest
// in `undecided` (np bytes of scratch), which update_pixel then finishes; at
// any other reality every pixel goes through update_pixel.
inline void update(int alg, std::complex<float>* rho, const int32_t* S, std::complex<float>* buf, float beta,
                   size_t np, float reality, uint8_t* undecided) {
    const PhaseWindow w = alg == ALG_RAAR ? PhaseWindow::raar(reality) : PhaseWindow::er_hio(reality);
    if (reality != 0.f) {
        for (size_t p = 0; p < np; ++p) update_pixel(alg, rho[p], S[p], buf ? buf + p : nullptr, beta, w);
        return;
    }
    float* r = reinterpret_cast<float*>(rho);
    float* b = reinterpret_cast<float*>(buf);
    const int rare = alg == ALG_ER    ? pass_er(r, S, undecided, np)
                     : alg == ALG_HIO ? pass_hio(r, b, S, undecided, beta, np)
                                      : pass_raar(r, b, S, undecided, beta, np);
    if (rare)
        for (size_t p = 0; p < np; ++p)
            if (undecided[p]) update_pixel(alg, rho[p], S[p], buf ? buf + p : nullptr, beta, w);
}

// orc::project_modulus with the same arithmetic per pixel, in a loop the
// compiler vectorises: measured pixels with m > 0 and bounded pixels in one
// pass, then the measured pixels with m == 0 (whose phase needs atan2) through
// orc::project_modulus itself.
inline void project(std::complex<float>* F, const float* amp, size_t np) {
    float* f = reinterpret_cast<float*>(F);
    int rare = 0;
    for (size_t p = 0; p < np; ++p) {
        const float re = f[2 * p], im = f[2 * p + 1];
        const float m = std::sqrt(im * im + re * re);
        const float a = amp[p];
        const float nm = -a - 2.f;
        const bool meas = a >= 0.f;
        const bool apply = (meas & (m > 0.f)) | ((!meas) & (a < -1.5f) & (m > nm));
        const float s = meas ? a : nm;
        const float nr = s * re / m, ni = s * im / m;
        f[2 * p] = apply ? nr : re;
        f[2 * p + 1] = apply ? ni : im;
        rare |= (int)(meas & !(m > 0.f));
    }
    if (rare) {
        for (size_t p = 0; p < np; ++p) {
            const float re = f[2 * p], im = f[2 * p + 1];
            if (amp[p] >= 0.f && !(std::sqrt(im * im + re * re) > 0.f)) project_modulus(F + p, amp + p, 1);
        }
    }
}

// rho <- IFT(P_M(FT(rho))): FT unnormalised with kernel exp(+2 pi i ...), IFT
// with exp(-2 pi i ...) and each component divided by float(n*n).
inline void projection(const LinePlan& pl, std::complex<float>* rho, const float* amp, Work& w) {
    const size_t np = (size_t)pl.n * pl.n;
    fft2(pl, rho, +1, w);
    project(rho, amp, np);
    fft2(pl, rho, -1, w);
    const float d = (float)(pl.n * pl.n);
    for (size_t p = 0; p < np; ++p) rho[p] = std::complex<float>(rho[p].real() / d, rho[p].imag() / d);
}

// The driver of orc::iterate<float>.
inline void iterate(const LinePlan& pl, std::complex<float>* rho, const int32_t* S, const float* amp, int alg,
                    int niter, float beta, float reality, Work& w) {
    const size_t np = (size_t)pl.n * pl.n;
    std::vector<std::complex<float>> buf;
    if (alg != ALG_ER) buf.assign(rho, rho + np);
    for (int it = 0; it < niter; ++it) {
        projection(pl, rho, amp, w);
        update(alg, rho, S, alg != ALG_ER ? buf.data() : nullptr, beta, np, reality, w.undecided.data());
    }
    projection(pl, rho, amp, w);
}

}  // namespace fastport
}  // namespace orc
