// CPU oracle for the SPRING memetic phase-retrieval hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP backend is
// compared against; it is compiled into oracle/liboracle.so and loaded only by
// tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py. The
// product (s_spi_spring_amd) never links, loads or calls it.
//
// It restates, per pixel and per driver, the semantics of the reference's GPU
// backend (the reference ships no working CPU backend: src/cpu/cpu_context.h:14-29
// is an empty stub). Every function cites the reference file:line it follows.
// Arithmetic is templated on the real type T (float = the FP32 oracle, double =
// the FP64 oracle). The FFT is this file's own mixed-radix Stockham autosort
// transform (radices 8/4/2/3/5, direct DFT for any other prime factor), written
// from the DFT definition; twiddles are computed in double and rounded to T.
//
// Conventions (src/gpu/gpu_context.cu:748-774, 65-77):
//   FT  = unnormalised 2-D DFT with kernel exp(+2*pi*i*k.x/N)   (cuFFT INVERSE)
//   IFT = 2-D DFT with kernel exp(-2*pi*i*k.x/N), then / T(N*N)  (cuFFT FORWARD + normalize)
//   fields are row-major, pixel p = x + N*y; amplitudes have the origin in the corner.
//
// How each part is pinned (tests/test_oracle_*.py):
//   fft2            numpy FP64 fft2/ifft2, delta/plane-wave closed forms, Parseval, linearity
//   project_modulus per-pixel closed form in numpy, idempotence, free pixels untouched, bound clamp
//   er/hio/raar     per-pixel closed forms in numpy; ER idempotence; RAAR(beta=1)==HIO(beta=1) at r=0;
//                   noise-free convergence with the true support; brute force on 8x8 in FP64
//   sw / swa        blur = exact Gaussian convolution (closed form, numpy); |S| >= n for SWA;
//                   threshold selection vs numpy.sort
//   get_error       closed form in numpy; E_F = 0 for an exact model
//   symmetry/shift  known cyclic shifts / point flips / global phases recovered exactly
//   crossover       cell-map closed form, cross_support truth table, numpy restatement
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <stdexcept>
#include <vector>

namespace orc {

template <class T> using cplx = std::complex<T>;

// ----------------------------------------------------------------------------
// FFT: mixed-radix Stockham autosort, one line at a time.
// Stage s with radix R and span Ns (product of the earlier radices) maps
//   x[j + r*N/R]  (j < N/R, r < R)  ->  y[(j/Ns)*Ns*R + (j%Ns) + q*Ns],
// after twiddling input r by w^(r*(j%Ns)) with w = exp(sign*2*pi*i/(Ns*R))
// and an R-point DFT over r. After the last stage y holds the DFT in natural
// order (C. Van Loan, Computational Frameworks for the FFT, SIAM 1992, sec. 1.7).
// ----------------------------------------------------------------------------
struct Plan {
    int n = 0;
    std::vector<int> radices;
    std::vector<std::complex<double>> root;  // root[m] = exp(-2*pi*i*m/n)

    explicit Plan(int n_) : n(n_) {
        if (n < 1) throw std::invalid_argument("FFT length must be >= 1");
        int m = n;
        while (m % 8 == 0) { radices.push_back(8); m /= 8; }
        while (m % 4 == 0) { radices.push_back(4); m /= 4; }
        while (m % 2 == 0) { radices.push_back(2); m /= 2; }
        for (int p = 3; m > 1; p += 2) {
            while (m % p == 0) { radices.push_back(p); m /= p; }
            if ((long long)p * p > m && m > 1) { radices.push_back(m); m = 1; }
        }
        root.resize(n);
        for (int k = 0; k < n; ++k) {
            double a = -2.0 * M_PI * (double)k / (double)n;
            root[k] = std::complex<double>(std::cos(a), std::sin(a));
        }
    }
};

// One line: data (length n, unit stride) transformed in place; scratch has n slots.
// sign = +1 uses exp(+2 pi i ...), sign = -1 uses exp(-2 pi i ...). Unnormalised.
template <class T>
void fft_line(const Plan& pl, cplx<T>* data, cplx<T>* scratch, int sign) {
    const int n = pl.n;
    if (n == 1) return;
    cplx<T>* x = data;
    cplx<T>* y = scratch;
    int Ns = 1;
    std::vector<cplx<T>> v(16), dft;
    for (int R : pl.radices) {
        if ((int)v.size() < R) v.resize(R);
        const int m = n / R;
        const int tw_step = n / (Ns * R);   // twiddle w^(r*k) = root[r*k*tw_step]
        const int dft_step = n / R;         // R-th roots: root[(r*q mod R)*dft_step]
        for (int j = 0; j < m; ++j) {
            const int k = j % Ns;
            for (int r = 0; r < R; ++r) {
                std::complex<double> w = pl.root[(size_t)r * k * tw_step];
                if (sign > 0) w = std::conj(w);
                v[r] = x[j + r * m] * cplx<T>((T)w.real(), (T)w.imag());
            }
            const int base = (j / Ns) * Ns * R + k;
            for (int q = 0; q < R; ++q) {
                cplx<T> acc(0, 0);
                for (int r = 0; r < R; ++r) {
                    std::complex<double> w = pl.root[(size_t)((r * q) % R) * dft_step];
                    if (sign > 0) w = std::conj(w);
                    acc += v[r] * cplx<T>((T)w.real(), (T)w.imag());
                }
                y[base + q * Ns] = acc;
            }
        }
        std::swap(x, y);
        Ns *= R;
    }
    if (x != data) std::memcpy(data, x, sizeof(cplx<T>) * n);
}

// 2-D unnormalised transform of an n x n row-major field: rows, then columns.
template <class T>
void fft2_raw(const Plan& pl, cplx<T>* f, int sign) {
    const int n = pl.n;
    std::vector<cplx<T>> line(n), scratch(n);
    for (int y = 0; y < n; ++y) fft_line(pl, f + (size_t)y * n, scratch.data(), sign);
    for (int x = 0; x < n; ++x) {
        for (int y = 0; y < n; ++y) line[y] = f[x + (size_t)y * n];
        fft_line(pl, line.data(), scratch.data(), sign);
        for (int y = 0; y < n; ++y) f[x + (size_t)y * n] = line[y];
    }
}

// GPU_context::get_ft (src/gpu/gpu_context.cu:748-754): cuFFT INVERSE, no scaling.
template <class T> void ft(const Plan& pl, cplx<T>* f) { fft2_raw(pl, f, +1); }

// GPU_context::get_ift + normalize (src/gpu/gpu_context.cu:757-764, 65-77):
// cuFFT FORWARD, then each component divided by FLOAT(npixels*npixels).
template <class T> void ift(const Plan& pl, cplx<T>* f) {
    fft2_raw(pl, f, -1);
    const T d = (T)(pl.n * pl.n);
    const size_t np = (size_t)pl.n * pl.n;
    for (size_t p = 0; p < np; ++p) f[p] = cplx<T>(f[p].real() / d, f[p].imag() / d);
}

// ----------------------------------------------------------------------------
// Per-pixel operators
// ----------------------------------------------------------------------------

// Modulus of a complex value in the reference's operand order: sqrt(y*y + x*x).
template <class T> inline T modulus_yx(const cplx<T>& z) {
    return std::sqrt(z.imag() * z.imag() + z.real() * z.real());
}
template <class T> inline T modulus_xy(const cplx<T>& z) {
    return std::sqrt(z.real() * z.real() + z.imag() * z.imag());
}

// Fourier modulus projection P_M, sub_intensities (src/gpu/gpu_context.cu:79-111).
// amp >= 0: measured; amp in (-1.5, 0): free; amp <= -1.5: upper bound -amp-2.
template <class T>
void project_modulus(cplx<T>* F, const T* amp, size_t np) {
    for (size_t p = 0; p < np; ++p) {
        const T re = F[p].real(), im = F[p].imag();
        const T m = modulus_yx(F[p]);
        const T a = amp[p];
        if (a >= T(0)) {
            if (m > T(0)) {
                F[p] = cplx<T>(a * re / m, a * im / m);
            } else {
                const T ph = std::atan2(im, re);
                F[p] = cplx<T>(a * std::cos(ph), a * std::sin(ph));
            }
        } else if (a < T(-1.5) && m > -a - T(2)) {
            const T nm = -a - T(2);
            F[p] = cplx<T>(nm * re / m, nm * im / m);
        }
    }
}

// Phase window of the reality constraint. The reference compares the FP32
// atan2 against a threshold formed in double:
//   ER/HIO/get_error (src/gpu/gpu_context.cu:530, 553):  phase > (1.f-r)*M_PI || phase < -(1.f-r)*M_PI
//   RAAR            (src/gpu/gpu_context.cu:582):        phase > (1.-r)*M_PI  || phase < -(1.f-r)*M_PI
struct PhaseWindow {
    double hi, lo;
    static PhaseWindow er_hio(float r) {
        double t = (double)(1.f - r) * M_PI;
        return {t, -t};
    }
    static PhaseWindow raar(float r) {
        return {(1.0 - (double)r) * M_PI, -(double)(1.f - r) * M_PI};
    }
    template <class T> bool ok(const cplx<T>& z) const {
        const double ph = (double)std::atan2(z.imag(), z.real());
        return !(ph > hi || ph < lo);
    }
};

// apply_support_er (src/gpu/gpu_context.cu:519-537): rho <- S'(rho) * rho.
template <class T>
void apply_support_er(cplx<T>* rho, const int32_t* S, size_t np, float reality) {
    const PhaseWindow w = PhaseWindow::er_hio(reality);
    for (size_t p = 0; p < np; ++p) {
        T ts = (T)S[p];
        if (!w.ok(rho[p])) ts = T(0);
        rho[p] = cplx<T>(rho[p].real() * ts, rho[p].imag() * ts);
    }
}

// apply_support_hio (src/gpu/gpu_context.cu:540-565).
template <class T>
void apply_support_hio(cplx<T>* rho, const int32_t* S, cplx<T>* buf, T beta, size_t np, float reality) {
    const PhaseWindow w = PhaseWindow::er_hio(reality);
    for (size_t p = 0; p < np; ++p) {
        T ts = (T)S[p];
        if (!w.ok(rho[p])) ts = T(0);
        const T dx = rho[p].real(), dy = rho[p].imag();
        const T nx = (buf[p].real() - beta * dx) * (T(1) - ts) + dx * ts;
        const T ny = (buf[p].imag() - beta * dy) * (T(1) - ts) + dy * ts;
        rho[p] = cplx<T>(nx, ny);
        buf[p] = rho[p];
    }
}

// apply_support_raar (src/gpu/gpu_context.cu:568-593).
template <class T>
void apply_support_raar(cplx<T>* rho, const int32_t* S, cplx<T>* buf, T beta, size_t np, float reality) {
    const PhaseWindow w = PhaseWindow::raar(reality);
    for (size_t p = 0; p < np; ++p) {
        const cplx<T> q(T(2) * rho[p].real() - buf[p].real(), T(2) * rho[p].imag() - buf[p].imag());
        if (!w.ok(q) || S[p] == 0) {
            rho[p] = cplx<T>(beta * buf[p].real() + (T(1) - T(2) * beta) * rho[p].real(),
                             beta * buf[p].imag() + (T(1) - T(2) * beta) * rho[p].imag());
        }
        buf[p] = rho[p];
    }
}

// ----------------------------------------------------------------------------
// Drivers (one individual)
// ----------------------------------------------------------------------------
enum Alg { ALG_ER = 0, ALG_HIO = 1, ALG_RAAR = 2 };

template <class T>
struct Individual {
    const Plan* pl;
    const T* amp;  // corner-origin amplitude with the mask encoding
    int n;
    size_t np;
};

// rho <- IFT(P_M(FT(rho)))
template <class T>
void projection(const Plan& pl, cplx<T>* rho, const T* amp) {
    ft(pl, rho);
    project_modulus(rho, amp, (size_t)pl.n * pl.n);
    ift(pl, rho);
}

// GPU_context::HIO / RAAR / ER (src/gpu/gpu_context.cu:778-902): n iterations of
// [FT; P_M; IFT; support update], then one trailing [FT; P_M; IFT]. HIO and RAAR
// start by copying rho into the buffer (rho_n); ER has no buffer.
template <class T>
void iterate(const Plan& pl, cplx<T>* rho, const int32_t* S, const T* amp, int alg, int niter,
             T beta, float reality) {
    const size_t np = (size_t)pl.n * pl.n;
    std::vector<cplx<T>> buf;
    if (alg != ALG_ER) buf.assign(rho, rho + np);
    for (int it = 0; it < niter; ++it) {
        projection(pl, rho, amp);
        if (alg == ALG_ER) apply_support_er(rho, S, np, reality);
        else if (alg == ALG_HIO) apply_support_hio(rho, S, buf.data(), beta, np, reality);
        else apply_support_raar(rho, S, buf.data(), beta, np, reality);
    }
    projection(pl, rho, amp);
}

// The blurred modulus b of shrink-wrap, shared by SW and SWA
// (src/gpu/gpu_context.cu:911-940, 962-999): a = |IFT(P_M(FT rho))|,
// b = FWD(g * FT(a)) with no 1/N^2, g(k) = exp(-(kx^2+ky^2)/sigma_ft^2),
// k wrapped to [-N/2, N/2), sigma_ft = float(N)/sigma/M_PI (src/gpu/gpu_context.cu:199-217, 935).
template <class T>
void shrinkwrap_blur(const Plan& pl, const cplx<T>* rho, const T* amp, float sigma, std::vector<cplx<T>>& b) {
    const int n = pl.n;
    const size_t np = (size_t)n * n;
    b.assign(rho, rho + np);
    projection(pl, b.data(), amp);
    for (size_t p = 0; p < np; ++p) b[p] = cplx<T>(modulus_xy(b[p]), T(0));
    ft(pl, b.data());
    const float sigma_ft = (float)((double)((float)n / sigma) / M_PI);
    const T s = (T)sigma_ft;
    for (int iy = 0; iy < n; ++iy) {
        for (int ix = 0; ix < n; ++ix) {
            int xx = ix, yy = iy;
            if (xx >= n / 2) xx -= n;
            if (yy >= n / 2) yy -= n;
            const T r2 = (T)(xx * xx + yy * yy);
            const T c = std::exp(-r2 / (s * s));
            cplx<T>& z = b[ix + (size_t)iy * n];
            z = cplx<T>(z.real() * c, z.imag() * c);
        }
    }
    fft2_raw(pl, b.data(), -1);
}

// GPU_context::SW (src/gpu/gpu_context.cu:906-953): S <- [|b| >= max|b| * threshold].
template <class T>
void sw(const Plan& pl, const cplx<T>* rho, int32_t* S, const T* amp, float sigma, float threshold) {
    std::vector<cplx<T>> b;
    shrinkwrap_blur(pl, rho, amp, sigma, b);
    const size_t np = b.size();
    T mx = T(0);
    for (size_t p = 0; p < np; ++p) {
        const T m = modulus_yx(b[p]);
        if (m > mx) mx = m;
    }
    const T thr = mx * (T)threshold;
    for (size_t p = 0; p < np; ++p) S[p] = (modulus_yx(b[p]) >= thr) ? 1 : 0;
}

// GPU_context::SWA (src/gpu/gpu_context.cu:957-1027): threshold = the spixels-th
// largest |b| (ascending sort, element N^2 - spixels); S <- [|b| >= threshold].
// spixels <= 0 reads past the end of the sorted array in the reference; here it
// yields an empty support.
template <class T>
void swa(const Plan& pl, const cplx<T>* rho, int32_t* S, const T* amp, float sigma, int spixels) {
    std::vector<cplx<T>> b;
    shrinkwrap_blur(pl, rho, amp, sigma, b);
    const size_t np = b.size();
    std::vector<T> v(np);
    for (size_t p = 0; p < np; ++p) v[p] = modulus_yx(b[p]);
    if (spixels <= 0) {
        std::fill(S, S + np, 0);
        return;
    }
    if ((size_t)spixels > np) spixels = (int)np;
    std::vector<T> sorted(v);
    std::nth_element(sorted.begin(), sorted.begin() + (np - spixels), sorted.end());
    const T thr = sorted[np - spixels];
    for (size_t p = 0; p < np; ++p) S[p] = (v[p] >= thr) ? 1 : 0;
}

// GPU_context::get_error (src/gpu/gpu_context.cu:1030-1134) with the kernels
// get_error_gpu (352-385), get_max_module (435-461), get_module_sum (219-246)
// and get_real_error_gpu (386-434). Sums are accumulated in double.
//   rho_S = S'(rho) rho ; F = FT(rho_S)
//   error_fourier = sqrt( sum_{A>=0} (A-|F|)^2 / sum_{A>=0} A^2 )
//   complexsum    = max_k |F|
//   modulesum     = sum_{|rho_S|>0} |rho_S|
//   rho2 = IFT(P_M(F));  error_real = sqrt( sum_{S==0} [ (Im rho2 - Im S'rho2)^2 + (Re rho2 * Re S'rho2)^2 ] / sum |rho2|^2 )
// (the product in the last term is the reference's, src/gpu/gpu_context.cu:417-418).
template <class T>
struct ErrorOut { double error_fourier, error_real, complexsum, modulesum; };

template <class T>
ErrorOut<T> get_error(const Plan& pl, const cplx<T>* rho, const int32_t* S, const T* amp, float reality) {
    const size_t np = (size_t)pl.n * pl.n;
    std::vector<cplx<T>> f(rho, rho + np);
    apply_support_er(f.data(), S, np, reality);
    double msum = 0;
    for (size_t p = 0; p < np; ++p) {
        const T m = modulus_yx(f[p]);
        if (m > T(0)) msum += (double)m;
    }
    ft(pl, f.data());
    double err = 0, nrm = 0;
    T fmax = T(0);
    for (size_t p = 0; p < np; ++p) {
        const T m = modulus_yx(f[p]);
        if (m > fmax) fmax = m;
        if (amp[p] >= T(0)) {
            nrm += (double)(amp[p] * amp[p]);
            const T d = amp[p] - m;
            err += (double)(d * d);
        }
    }
    project_modulus(f.data(), amp, np);
    ift(pl, f.data());
    const PhaseWindow w = PhaseWindow::er_hio(reality);
    double rerr = 0, rnrm = 0;
    for (size_t p = 0; p < np; ++p) {
        const cplx<T> z = f[p];
        T ts = (T)S[p];
        if (!w.ok(z)) ts = T(0);
        const T sx = z.real() * ts, sy = z.imag() * ts;
        rnrm += (double)(z.imag() * z.imag() + z.real() * z.real());
        if (S[p] == 0) {
            const T a = z.imag() - sy, b = z.real() * sx;
            rerr += (double)(a * a + b * b);
        }
    }
    ErrorOut<T> o;
    o.error_fourier = std::sqrt(err / nrm);
    o.error_real = std::sqrt(rerr / rnrm);
    o.complexsum = (double)fmax;
    o.modulesum = msum;
    return o;
}

// Spectral window of the alignment/symmetry cross-correlation,
// multiply(..., intensities, minr, maxr, sign) (src/gpu/gpu_context.cu:1157-1200):
//   c = (sign-conjugated d) * r * coeff^2, and c = 0 where A < 0.
template <class T>
void windowed_product(int n, const cplx<T>* d, const cplx<T>* r, const T* amp, float minr, float maxr,
                      float sign, cplx<T>* out) {
    const float sigmatot = (float)((double)n * minr / 2. + 1e-12);
    const float sigma = (float)(sigmatot / 10.);
    const float sigmatotmax = (float)((double)n * maxr / 2. + 1e-12);
    const float sigmamax = (float)(sigmatotmax / 10.);
    for (int iy = 0; iy < n; ++iy) {
        for (int ix = 0; ix < n; ++ix) {
            const size_t i = ix + (size_t)iy * n;
            float x = (float)ix, y = (float)iy;
            if (x >= n / 2) x -= n;
            if (y >= n / 2) y -= n;
            T cx = d[i].real() * r[i].real() - ((T)sign * d[i].imag()) * r[i].imag();
            T cy = d[i].real() * r[i].imag() + ((T)sign * d[i].imag()) * r[i].real();
            const float radius = std::sqrt(x * x + y * y);
            const float eff = radius - (sigmatot - sigma);
            float coeff = eff > 0 ? (float)(1. - (double)std::exp(-eff * eff / (sigma * sigma))) : 0.f;
            const float effmax = radius - (sigmatotmax - sigmamax);
            const float coeffmax = effmax > 0 ? (float)std::exp(-effmax * effmax / (sigmamax * sigmamax)) : 1.f;
            coeff *= coeffmax;
            cx *= (T)(coeff * coeff);
            cy *= (T)(coeff * coeff);
            if (amp[i] < T(0)) { cx = T(0); cy = T(0); }
            out[i] = cplx<T>(cx, cy);
        }
    }
}

// First index of the strict maximum of |c|^2, starting from 0 (src/gpu/gpu_context.cu:1271-1280).
template <class T>
void argmax_sq(const cplx<T>* c, size_t np, T& mval, int& midx) {
    mval = T(0);
    midx = 0;
    for (size_t i = 0; i < np; ++i) {
        const T v = c[i].real() * c[i].real() + c[i].imag() * c[i].imag();
        if (v > mval) { mval = v; midx = (int)i; }
    }
}

// |rho| as a real field (normalize_shift, e=1; src/gpu/gpu_context.cu:21-34).
template <class T>
void modulus_field(const cplx<T>* rho, cplx<T>* out, size_t np) {
    for (size_t p = 0; p < np; ++p) out[p] = cplx<T>(modulus_xy(rho[p]), T(0));
}

// normalize_ft_shift with e=1 (src/gpu/gpu_context.cu:37-50): z *= |z|/|z|
// (identity, except NaN where |z| == 0).
template <class T>
void unit_scale(cplx<T>* f, size_t np) {
    for (size_t p = 0; p < np; ++p) {
        const T m = modulus_xy(f[p]);
        const T s = m / m;
        f[p] = cplx<T>(f[p].real() * s, f[p].imag() * s);
    }
}

// Sequential float sum of a field (host loops at src/gpu/gpu_context.cu:638-641, 1236-1240).
template <class T>
cplx<T> field_sum(const cplx<T>* f, size_t np) {
    T sx = 0, sy = 0;
    for (size_t p = 0; p < np; ++p) { sx += f[p].real(); sy += f[p].imag(); }
    return cplx<T>(sx, sy);
}

struct Shift { int sx, sy; float phase; bool flip; };

// Alignment of rho to a reference by cross-correlating Fourier moduli, direct and
// twin: GPU_context::get_shift (src/gpu/gpu_context.cu:1226-1326) when
// unit_scale_ft is true (population Reshift, with set_ref_density at :632-667),
// reshift_parent (:1536-1654) when false (crossover parents, ref FT from
// prepare_parents :1661-1684). ref_ft is FT(|rho_ref|), already unit-scaled when
// unit_scale_ft is true; ref_sum is the float sum of the reference density.
template <class T>
Shift get_shift(const Plan& pl, const cplx<T>* rho, const cplx<T>* ref_ft, cplx<T> ref_sum, const T* amp,
                float minr, float maxr, bool unit_scale_ft) {
    const int n = pl.n;
    const size_t np = (size_t)n * n;
    const cplx<T> s = field_sum(rho, np);
    const T ph = std::atan2(s.imag(), s.real());
    const T ph_ref = std::atan2(ref_sum.imag(), ref_sum.real());
    std::vector<cplx<T>> a(np), c(np);
    modulus_field(rho, a.data(), np);
    ft(pl, a.data());
    if (unit_scale_ft) unit_scale(a.data(), np);
    T m1, m2;
    int i1, i2;
    windowed_product(n, a.data(), ref_ft, amp, minr, maxr, -1.f, c.data());
    fft2_raw(pl, c.data(), -1);
    argmax_sq(c.data(), np, m1, i1);
    windowed_product(n, a.data(), ref_ft, amp, minr, maxr, +1.f, c.data());
    fft2_raw(pl, c.data(), -1);
    argmax_sq(c.data(), np, m2, i2);
    Shift out;
    if (m1 >= m2) {
        out.sx = i1 % n; out.sy = i1 / n; out.phase = (float)(ph_ref - ph); out.flip = false;
    } else {
        out.sx = -i2 % n; out.sy = -i2 / n; out.phase = (float)(ph_ref + ph); out.flip = true;
    }
    return out;
}

// shift kernel (src/gpu/gpu_context.cu:1329-1357): cyclic shift (wrapped once),
// optional point flip p -> N^2-1-p, phase negated on flip, global phase added.
template <class T>
void reshift(int n, cplx<T>* rho, int32_t* S, const Shift& sh) {
    const size_t np = (size_t)n * n;
    std::vector<cplx<T>> r0(rho, rho + np);
    std::vector<int32_t> s0(S, S + np);
    for (int iy = 0; iy < n; ++iy) {
        for (int ix = 0; ix < n; ++ix) {
            const size_t old = ix + (size_t)iy * n;
            int nx = ix + sh.sx, ny = iy + sh.sy;
            if (nx < 0) nx += n;
            if (nx >= n) nx -= n;
            if (ny < 0) ny += n;
            if (ny >= n) ny -= n;
            size_t ni = nx + (size_t)ny * n;
            if (sh.flip) ni = np - 1 - ni;
            const T mod = modulus_xy(r0[old]);
            T phase = std::atan2(r0[old].imag(), r0[old].real());
            if (sh.flip) phase = -phase;
            phase += (T)sh.phase;
            rho[ni] = cplx<T>(mod * std::cos(phase), mod * std::sin(phase));
            S[ni] = s0[old];
        }
    }
}

// GPU_context::get_symmetry(sym, 1, symmin, symmax) (src/gpu/gpu_context.cu:1378-1463):
// a_hat = unit_scale(FT(|rho|)); c1 = FWD(win(conj(a_hat) a_hat)), c2 = FWD(win(a_hat a_hat));
// symmetry = sqrt(max|c2|^2 / max|c1|^2).
template <class T>
double get_symmetry(const Plan& pl, const cplx<T>* rho, const T* amp, float symmin, float symmax) {
    const int n = pl.n;
    const size_t np = (size_t)n * n;
    std::vector<cplx<T>> a(np), c(np);
    modulus_field(rho, a.data(), np);
    ft(pl, a.data());
    unit_scale(a.data(), np);
    T m1, m2;
    int i1, i2;
    windowed_product(n, a.data(), a.data(), amp, symmin, symmax, -1.f, c.data());
    fft2_raw(pl, c.data(), -1);
    argmax_sq(c.data(), np, m1, i1);
    windowed_product(n, a.data(), a.data(), amp, symmin, symmax, +1.f, c.data());
    fft2_raw(pl, c.data(), -1);
    argmax_sq(c.data(), np, m2, i2);
    return (double)std::sqrt((T)(m2 / m1));
}

}  // namespace orc

namespace orc {

// ----------------------------------------------------------------------------
// Crossover (src/solver.cpp:716-817, src/gpu/gpu_context.cu:1661-1694, 1699-1883, 2044-2117)
// ----------------------------------------------------------------------------
struct CrossDesc {
    int crosssizex, crosssizey, squaresizex, squaresizey, offsetx, offsety;
    float rotangle, crossdiff, use_average;
    int crossft, crosssym;
};

// Cell of pixel (ix, iy) in the rotated square grid of the crossover map
// (src/gpu/gpu_context.cu:1717-1739). The polar geometry (radius, angle,
// cos, sin) is evaluated in double here and on the GPU, so that the integer cell
// index is identical on both sides; the reference evaluates it in float with
// approximate device intrinsics, which decides differently only for pixels whose
// rotated coordinate lies within float rounding of an integer.
inline int cross_cell(int n, int ix, int iy, const CrossDesc& d) {
    int x = ix, y = iy;
    if (x >= n / 2) x -= n;
    if (y >= n / 2) y -= n;
    double radius = std::sqrt((double)(x * x + y * y));
    if (radius >= n / 2.) radius = (double)(n / 2 - 1);
    double angle = std::atan2((double)y, (double)x);
    if (d.crosssym && angle < 0) angle += M_PI;
    angle += (double)d.rotangle;
    x = (int)(std::cos(angle) * radius) + n / 2;
    y = (int)(std::sin(angle) * radius) + n / 2;
    return (x + d.offsetx) / d.squaresizex + ((y + d.offsety) / d.squaresizey) * d.crosssizex;
}

// cross_density (src/gpu/gpu_context.cu:1699-1796) when crossft, else
// cross_density_real (:1799-1853). rho0 is overwritten on selected cells.
template <class T>
void cross_density(int n, cplx<T>* rho0, const cplx<T>* rho1, const cplx<T>* rho2, const cplx<T>* rho3,
                   const cplx<T>* avg, const int32_t* crossmap, const T* amp, const T* amp_avg,
                   const CrossDesc& d) {
    for (int iy = 0; iy < n; ++iy) {
        for (int ix = 0; ix < n; ++ix) {
            const size_t i = ix + (size_t)iy * n;
            if (crossmap[cross_cell(n, ix, iy, d)] != 1) continue;
            cplx<T> d1 = rho1[i];
            if (d.crossft) {
                const T mod1 = modulus_xy(rho1[i]);
                if (amp[i] < T(0)) {
                    if (d.use_average > 0) {
                        const T avgmod = (amp_avg[i] + modulus_xy(avg[i])) * T(0.5);
                        const T newmod = (T)((double)avgmod * d.use_average + (1. - d.use_average) * (double)mod1);
                        if (mod1 > T(0)) d1 = cplx<T>(d1.real() * (newmod / mod1), d1.imag() * (newmod / mod1));
                    }
                    if (d.use_average < 0) {
                        if (mod1 > T(0)) {
                            const T locua = std::abs((T)d.use_average);
                            const T maxmod = amp_avg[i];
                            const T minmod = modulus_xy(avg[i]);
                            T newmod = mod1;
                            if (newmod < minmod) newmod = minmod * locua + newmod * (T(1) - locua);
                            else if (newmod > maxmod) newmod = maxmod * locua + newmod * (T(1) - locua);
                            d1 = cplx<T>(d1.real() * (newmod / mod1), d1.imag() * (newmod / mod1));
                        }
                    }
                }
            }
            const T cd = (T)d.crossdiff;
            rho0[i] = cplx<T>(d1.real() + cd * (rho2[i].real() - rho3[i].real()),
                              d1.imag() + cd * (rho2[i].imag() - rho3[i].imag()));
        }
    }
}

// cross_support (src/gpu/gpu_context.cu:1856-1883): S0 <- [S1 + S2*S3*S0 > 0.5].
inline void cross_support(int32_t* S0, const int32_t* S1, const int32_t* S2, const int32_t* S3, size_t np) {
    for (size_t i = 0; i < np; ++i) {
        const int v = S1[i] + S2[i] * S3[i] * S0[i];
        S0[i] = (v > 0.5) ? 1 : 0;
    }
}

// One child of Solver::Crossover given the parents and the host draws
// (prepare_parents + get_new_individual). rho0/S0 hold parent 0 on entry and the
// child on exit (support updated only if update_support). avg is the population
// average density; amp_avg the population-average Fourier modulus
// (GetAverageModule(ft=true, 0.66)). The best individual's alignment
// (reshift_parent(densitybest...), src/gpu/gpu_context.cu:1692) is omitted: its
// result is never read by cross_density.
template <class T>
void crossover_child(const Plan& pl, const T* amp, const T* amp_avg, cplx<T>* rho0, int32_t* S0,
                     const cplx<T>* rho1_in, const int32_t* S1_in, const cplx<T>* rho2_in, const int32_t* S2_in,
                     const cplx<T>* rho3_in, const int32_t* S3_in, const cplx<T>* avg_in,
                     const int32_t* crossmap, const CrossDesc& d, float minr, float maxr, float reality,
                     bool update_support) {
    const int n = pl.n;
    const size_t np = (size_t)n * n;
    std::vector<cplx<T>> r1(rho1_in, rho1_in + np), r2(rho2_in, rho2_in + np), r3(rho3_in, rho3_in + np),
        av(avg_in, avg_in + np), r0(rho0, rho0 + np);
    std::vector<int32_t> s1(S1_in, S1_in + np), s2(S2_in, S2_in + np), s3(S3_in, S3_in + np), s0(S0, S0 + np),
        sdummy(np, 0);
    // prepare_parents (src/gpu/gpu_context.cu:1661-1694)
    const cplx<T> ref_sum = field_sum(r1.data(), np);
    std::vector<cplx<T>> ref(np);
    modulus_field(r1.data(), ref.data(), np);
    ft(pl, ref.data());
    Shift sh = get_shift(pl, r0.data(), ref.data(), ref_sum, amp, minr, maxr, false);
    reshift(n, r0.data(), s0.data(), sh);
    sh = get_shift(pl, r2.data(), ref.data(), ref_sum, amp, minr, maxr, false);
    reshift(n, r2.data(), s2.data(), sh);
    sh = get_shift(pl, r3.data(), ref.data(), ref_sum, amp, minr, maxr, false);
    reshift(n, r3.data(), s3.data(), sh);
    sh = get_shift(pl, av.data(), ref.data(), ref_sum, amp, minr, maxr, false);
    reshift(n, av.data(), sdummy.data(), sh);
    // get_new_individual (src/gpu/gpu_context.cu:2044-2117)
    apply_support_er(r1.data(), s1.data(), np, reality);
    apply_support_er(r0.data(), s0.data(), np, reality);
    apply_support_er(r2.data(), s2.data(), np, reality);
    apply_support_er(r3.data(), s3.data(), np, reality);
    if (d.crossft) {
        ft(pl, r0.data());
        ft(pl, r1.data());
        ft(pl, r2.data());
        ft(pl, r3.data());
        ft(pl, av.data());
    }
    cross_density(n, r0.data(), r1.data(), r2.data(), r3.data(), av.data(), crossmap, amp, amp_avg, d);
    cross_support(s0.data(), s1.data(), s2.data(), s3.data(), np);
    if (d.crossft) ift(pl, r0.data());
    apply_support_er(r0.data(), s0.data(), np, reality);
    std::copy(r0.begin(), r0.end(), rho0);
    if (update_support) std::copy(s0.begin(), s0.end(), S0);
}

// ----------------------------------------------------------------------------
// Host-side steps of the Solver that the oracle also restates
// ----------------------------------------------------------------------------

// Context RNG (src/context.h:13-18, 68-69): rand(a,b) = u*(b-a)+a with
// u ~ std::uniform_real_distribution<float>(0,1) over std::mt19937.
struct Rng {
    std::mt19937 gen;
    std::uniform_real_distribution<float> dist{0.f, 1.f};
    explicit Rng(int seed) : gen((std::mt19937::result_type)seed) {}
    float rand(float a, float b) { return dist(gen) * (b - a) + a; }
    float rand() { return dist(gen); }
};

// Random draws of one child in Solver::Crossover (src/solver.cpp:750-801).
// sorted_idx: population indices sorted by ascending error. Outputs parent
// indices p1..p3, the cell grid and the crossmap (resized).
inline void crossover_draws(Rng& g, int P, const int* sorted_idx, float suppdim, int n, float crossexp,
                            float crossftin, int& p1, int& p2, int& p3, CrossDesc& d,
                            std::vector<int32_t>& crossmap, float crossbalance) {
    const int extracted = (int)(std::pow(g.rand(0, 1), crossexp) * (float)P);
    p1 = sorted_idx[extracted];
    p2 = (int)g.rand(0, (float)(P - 0.0001));
    p3 = (int)g.rand(0, (float)(P - 0.0001));
    d.crossft = g.rand(0, 1) < crossftin ? 1 : 0;
    if (d.crossft) {
        d.squaresizex = (int)g.rand((float)std::ceil((float)n / suppdim), (float)std::ceil((float)n / 4.));
        d.squaresizey = (int)g.rand((float)std::ceil((float)n / suppdim), (float)std::ceil((float)n / 4.));
    } else {
        d.squaresizex = (int)g.rand((float)(suppdim / 10. + 1.), (float)(suppdim / 2.));
        d.squaresizey = (int)g.rand((float)(suppdim / 10. + 1.), (float)(suppdim / 2.));
    }
    d.crosssizex = (n + d.squaresizex - 1) / d.squaresizex + 1;
    d.crosssizey = (n + d.squaresizey - 1) / d.squaresizey + 1;
    d.offsetx = (int)g.rand(0, (float)d.squaresizex);
    d.offsety = (int)g.rand(0, (float)d.squaresizey);
    d.rotangle = g.rand(0, (float)M_PI);
    crossmap.assign((size_t)d.crosssizex * d.crosssizey, 0);
    for (size_t i = 0; i < crossmap.size(); ++i)
        if (g.rand(0, 1) < crossbalance) crossmap[i] = 1;
}

// ----------------------------------------------------------------------------
// Population initialisation (Solver::Initialize_rec, src/solver.cpp:870-1174).
// Host arithmetic in the reference's types: FLOAT = float, COMPLEX =
// std::complex<float>; expressions that the reference forms in double (a float
// times a double literal, std::pow(float, int)) are formed in double here and
// rounded where the reference assigns to a FLOAT. The FT/IFT the reference runs
// on the device are this file's FP32 transforms. Each call draws one
// individual from `g` in the reference's order.
// ----------------------------------------------------------------------------

// Mean-preserving scale of an initial density: the reference's norm loop
// (src/solver.cpp:943-954, 1155-1166) over the measured pixels of amp.
inline float init_norm_factor(const cplx<float>* F, const float* amp, size_t np) {
    float norm = 0, norm_new = 0;
    for (size_t i = 0; i < np; ++i)
        if (amp[i] >= 0) {
            norm += std::pow(std::abs(F[i]), 2);   // std::pow(float, int): double, added in double
            norm_new += amp[i] * amp[i];
        }
    return std::sqrt(norm_new / norm);
}

// Initialize_rec with a disk support of diameter support_size centred at
// (n/2, n/2) and nitems random projected spheres (src/solver.cpp:870-957).
inline void init_disk(Rng& g, const Plan& pl, const float* amp, int support_size, const float itemsize[2],
                      const int nitems_in[2], const float exponent_in[2], const float scaling_in[2], float phase_in,
                      cplx<float>* rho, int32_t* S) {
    const int n = pl.n;
    const size_t np = (size_t)n * n;
    const int nitems = (int)g.rand((float)nitems_in[0], (float)(nitems_in[1] + 0.99));
    const float exponent = g.rand(exponent_in[0], exponent_in[1]);
    const float scaling = g.rand(scaling_in[0], scaling_in[1]);
    std::vector<float> sig(nitems), cx(nitems), cy(nitems), ph(nitems);
    for (int i = 0; i < nitems; ++i) {
        const float s2 = g.rand((float)std::pow(itemsize[0], 2), (float)std::pow(itemsize[1], 2));
        sig[i] = (float)(std::sqrt(s2) / 2.);
        cx[i] = g.rand(sig[i], (float)support_size - sig[i]);
        cy[i] = g.rand(sig[i], (float)support_size - sig[i]);
        ph[i] = g.rand((float)(-phase_in * M_PI), (float)(phase_in * M_PI));
    }
    std::fill(rho, rho + np, cplx<float>(0, 0));
    std::fill(S, S + np, 0);
    for (int iy = 0; iy < support_size; ++iy)
        for (int ix = 0; ix < support_size; ++ix) {
            const double dx = ix - support_size / 2., dy = iy - support_size / 2.;
            const float supprad = (float)std::sqrt(dx * dx + dy * dy);
            if (!(supprad <= support_size / 2.)) continue;
            const size_t idx = (size_t)(iy + n / 2 - support_size / 2) * n + (ix + n / 2 - support_size / 2);
            S[idx] = 1;
            for (int i = 0; i < nitems; ++i) {
                const float rad = std::sqrt((ix - cx[i]) * (ix - cx[i]) + (iy - cy[i]) * (iy - cy[i]));
                if (rad < sig[i]) {
                    const float val = (float)(2. * sig[i] * std::sin(std::acos(rad / sig[i])));
                    rho[idx] += std::polar(val, ph[i]);
                }
            }
            const float mod = std::abs(rho[idx]);
            if (mod > 0) rho[idx] = rho[idx] / mod * std::pow(mod, exponent);
        }
    std::vector<cplx<float>> F(rho, rho + np);
    ft(pl, F.data());
    const float f = init_norm_factor(F.data(), amp, np) * scaling;
    for (size_t i = 0; i < np; ++i) rho[i] *= f;
}

// Initialize_rec from a reference reconstruction: spheres centred on support
// pixels picked by rank, the reference's support, the sum's modulus scaled
// (src/solver.cpp:961-1058). The pixel of rank k is found as the reference
// does: the first index j at which k support pixels precede j.
inline void init_ref(Rng& g, int n, const cplx<float>* ref_rho, const int32_t* ref_S, const float itemsize[2],
                     const int nitems_in[2], const float exponent_in[2], const float scaling_in[2], float phase_in,
                     cplx<float>* rho, int32_t* S) {
    const size_t np = (size_t)n * n;
    const int nitems = (int)g.rand((float)nitems_in[0], (float)(nitems_in[1] + 0.99));
    const float exponent = g.rand(exponent_in[0], exponent_in[1]);
    const float scaling = g.rand(scaling_in[0], scaling_in[1]);
    int supppix = 0;
    cplx<float> sum = 0;
    for (size_t i = 0; i < np; ++i) {
        sum += ref_rho[i];
        if (ref_S[i] > 0) supppix += 1;
    }
    const float newnorm = std::abs(sum);
    std::vector<float> sig(nitems), cx(nitems), cy(nitems), ph(nitems);
    for (int i = 0; i < nitems; ++i) {
        const float s2 = g.rand((float)std::pow(itemsize[0], 2), (float)std::pow(itemsize[1], 2));
        sig[i] = (float)(std::sqrt(s2) / 2.);
        const int target = (int)g.rand(0.f, (float)supppix);
        int counter = 0;
        size_t index = 0;
        for (size_t j = 0; j < np; ++j) {
            if (counter == target) {
                index = j;
                break;
            }
            if (ref_S[j] > 0) counter += 1;
        }
        cx[i] = (float)(index % n);
        cy[i] = (float)(index / n);
        ph[i] = g.rand((float)(-phase_in * M_PI), (float)(phase_in * M_PI));
    }
    std::fill(rho, rho + np, cplx<float>(0, 0));
    for (int iy = 0; iy < n; ++iy)
        for (int ix = 0; ix < n; ++ix) {
            const size_t idx = (size_t)iy * n + ix;
            if (ref_S[idx] <= 0) continue;
            for (int i = 0; i < nitems; ++i) {
                const float rad = std::sqrt((ix - cx[i]) * (ix - cx[i]) + (iy - cy[i]) * (iy - cy[i]));
                if (rad < sig[i]) {
                    const float val = (float)(2. * sig[i] * std::sin(std::acos(rad / sig[i])));
                    rho[idx] += std::polar(val, ph[i]);
                }
            }
            const float mod = std::abs(rho[idx]);
            if (mod > 0) rho[idx] = rho[idx] / mod * std::pow(mod, exponent);
        }
    cplx<float> oldsum = 0;
    for (size_t i = 0; i < np; ++i) oldsum += rho[i];
    const float oldnorm = std::abs(oldsum);
    for (size_t i = 0; i < np; ++i) rho[i] *= newnorm / oldnorm * scaling;
    std::copy(ref_S, ref_S + np, S);
}

// Initialize_rec from a support with random Fourier cells: FT of the support,
// each selected cell of a rotated square grid scaled by U(0.5, 2) and phased
// by U(-pi phase, pi phase), IFT, mean-preserving scale (src/solver.cpp:1061-1174).
// The cell index is evaluated in float as the reference's host code does.
inline void init_supp(Rng& g, const Plan& pl, const float* amp, const int32_t* support, float initfrac,
                      const float scaling_in[2], float phase_in, cplx<float>* rho, int32_t* S) {
    const int n = pl.n;
    const size_t np = (size_t)n * n;
    const float normfactor = g.rand(scaling_in[0], scaling_in[1]);
    const int sqx = (int)g.rand(1.f, (float)((float)n / 16.));
    const int sqy = (int)g.rand(1.f, (float)((float)n / 16.));
    const int csx = (n + sqx - 1) / sqx + 1;
    const int csy = (n + sqy - 1) / sqy + 1;
    const int offx = (int)g.rand(0.f, (float)sqx);
    const int offy = (int)g.rand(0.f, (float)sqy);
    const float rot = g.rand(0.f, (float)M_PI);
    const size_t cells = (size_t)csx * csy;
    std::vector<int32_t> map(cells, 0);
    std::vector<float> scale(cells, 0), phase(cells, 0);
    for (size_t i = 0; i < cells; ++i)
        if (g.rand(0.f, 1.f) < initfrac) {
            map[i] = 1;
            scale[i] = g.rand(0.5f, 2.f);
            phase[i] = g.rand((float)(-M_PI * phase_in), (float)(M_PI * phase_in));
        }
    std::vector<cplx<float>> d(np);
    for (size_t i = 0; i < np; ++i) d[i] = cplx<float>((float)support[i], 0.f);
    ft(pl, d.data());
    const cplx<float> imag(0.0f, 1.0f);
    for (size_t i = 0; i < np; ++i) {
        int x = (int)(i % n), y = (int)(i / n);
        if (x >= n / 2) x -= n;
        if (y >= n / 2) y -= n;
        float radius = std::sqrt((float)(x * x + y * y));
        if (radius >= n / 2.) radius = (float)(n / 2 - 1);
        const float angle = std::atan2((float)y, (float)x) + rot;
        x = (int)(std::cos(angle) * radius) + n / 2;
        y = (int)(std::sin(angle) * radius) + n / 2;
        const int ci = (x + offx) / sqx + ((y + offy) / sqy) * csx;
        if (map[ci] == 1) d[i] = d[i] * scale[ci] * std::exp(imag * phase[ci]);
    }
    ift(pl, d.data());
    std::vector<cplx<float>> F(d);
    ft(pl, F.data());
    const float f = init_norm_factor(F.data(), amp, np) * normfactor;
    for (size_t i = 0; i < np; ++i) rho[i] = d[i] * f;
    std::copy(support, support + np, S);
}

// Solver::set_boundaries + Solver::prepare (src/solver.cpp:209-289, 173-206):
// intensity (centre at `center`, -1 masked, -2 saturated) -> corner-origin amplitude.
inline void prepare_pattern(const float* intensity, int n, int cx, int cy, bool use_bounds, float* amp_out) {
    const size_t np = (size_t)n * n;
    std::vector<float> pat(intensity, intensity + np);
    // set_boundaries
    const int binsize = 1, span = 1;
    const int profsize = (int)(std::sqrt((float)(cx * cx + cy * cy)) / (float)binsize) + 10;
    std::vector<float> radprof(profsize, 0), radprofstd(profsize, 0), counts(profsize, 0),
        totcounts(profsize, 0), totsat(profsize, 0);
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < n; ++i) {
            const float val = pat[i + (size_t)n * j];
            const float dist = std::sqrt((float)((j - cy) * (j - cy) + (i - cx) * (i - cx)));
            const int dc = (int)std::round(dist / (float)binsize);
            for (int di = dc - span; di <= dc + span; ++di) {
                if (di < profsize && di >= 0) {
                    if (val >= 0) {
                        radprof[di] += val;
                        radprofstd[di] += val * val;
                        counts[di] += 1.f;
                    }
                    if (val < -1.5f) totsat[di] += 1.f;
                    totcounts[di] += 1.f;
                }
            }
        }
    }
    for (int i = 0; i < profsize; ++i) {
        if (counts[i] > totcounts[i] * 0.66 && totsat[i] <= totcounts[i] * 0.05 && totcounts[i] > 40) {
            radprof[i] /= counts[i];
            radprofstd[i] = std::sqrt(std::abs(radprofstd[i] / counts[i] - radprof[i] * radprof[i]));
        } else {
            radprof[i] = -1;
        }
    }
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < n; ++i) {
            const size_t index = i + (size_t)n * j;
            if (pat[index] < 0) {
                const float dist = std::sqrt((float)((j - cy) * (j - cy) + (i - cx) * (i - cx)));
                const int di = (int)std::round(dist / (float)binsize);
                if (di < profsize && di > 0) {
                    if (radprof[di] >= 0 && pat[index] < -0.5f && pat[index] > -1.5f) {
                        const float bound = radprof[di] + 1.f * radprofstd[di];
                        pat[index] = -2.f - std::sqrt(bound);
                    } else {
                        pat[index] = -1;
                    }
                }
            }
        }
    }
    // prepare
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < n; ++i) {
            int kx = i - cx, ky = j - cy;
            if (kx < 0) kx += n;
            if (ky < 0) ky += n;
            const float t = pat[i + (size_t)n * j];
            float& o = amp_out[kx + (size_t)n * ky];
            if (t >= 0) o = std::sqrt(t);
            else o = use_bounds ? t : -1.f;
        }
    }
}

}  // namespace orc
