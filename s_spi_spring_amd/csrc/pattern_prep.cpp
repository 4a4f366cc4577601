// Detector-image preprocessing (see pattern_prep.h for the behaviour and the
// reference lines each function follows).
#include "pattern_prep.h"

#include <algorithm>
#include <cmath>

namespace spring {
namespace prep {

namespace {

constexpr double kKeep = 0.1;   // minimum weight fraction for an output pixel (pattern.cpp:12)

// Output grid accumulating bilinearly splatted values and their weights.
struct Grid {
    std::size_t nx, ny;
    std::vector<double> sum, weight;
    Grid(std::size_t nx_, std::size_t ny_) : nx(nx_), ny(ny_), sum(nx_ * ny_, 0.0), weight(nx_ * ny_, 0.0) {}

    // Spread `val` over the four pixels around (fx, fy); corners outside the
    // grid are dropped. Corner order: (0,0), (1,0), (0,1), (1,1).
    void splat(double fx, double fy, double val) {
        const int ix = (int)std::floor(fx), iy = (int)std::floor(fy);
        const double ax = fx - (double)ix, ay = fy - (double)iy;
        const double wx[2] = {1. - ax, ax};
        const double wy[2] = {1. - ay, ay};
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 2; ++i) {
                const long X = (long)ix + i, Y = (long)iy + j;
                if (X < 0 || Y < 0 || X >= (long)nx || Y >= (long)ny) continue;
                const double w = wx[i] * wy[j];
                sum[(std::size_t)X + (std::size_t)Y * nx] += val * w;
                weight[(std::size_t)X + (std::size_t)Y * nx] += w;
            }
    }

    double max_weight() const {
        double m = 0;
        for (double w : weight) m = std::max(m, w);
        return m;
    }

    // Weighted mean where weight > kKeep * ref[i], else -1.
    Image mean_where(const std::vector<double>& ref) const {
        Image out(sum.size());
        for (std::size_t i = 0; i < out.size(); ++i) out[i] = weight[i] > kKeep * ref[i] ? sum[i] / weight[i] : -1.;
        return out;
    }
    Image mean_where(double ref) const { return mean_where(std::vector<double>(sum.size(), ref)); }
};

// Flat detector -> Ewald-sphere projection about `center`: a pixel at distance r
// (physical units dx per pixel, detector distance R) moves to the radius that
// keeps its scattering angle on a grid whose outermost pixel stays in place.
struct Ewald {
    double R, dx, dx_new;
    std::array<double, 2> c;
    Ewald(std::size_t nx, std::size_t ny, std::array<double, 2> center, double R_, double dx_) : R(R_), dx(dx_), c(center) {
        const double far = std::max({c[0], c[1], (double)nx - c[0], (double)ny - c[1]});
        const double theta_max = std::atan2(dx * far, R);
        dx_new = R * 2. * std::sin(theta_max / 2.) / far;
    }
    // Displacement from the centre on the new grid, and the 1/cos^3 intensity factor.
    void map(std::size_t x, std::size_t y, double& ox, double& oy, double& gain) const {
        const double rx = (double)x - c[0], ry = (double)y - c[1];
        const double phi = std::atan2(ry, rx);
        const double theta = std::atan2(std::sqrt(rx * rx + ry * ry) * dx, R);
        const double rp = R * 2. * std::sin(theta / 2.) / dx_new;
        ox = rp * std::cos(phi);
        oy = rp * std::sin(phi);
        gain = 1. / std::pow(std::cos(theta), 3);
    }
};

}  // namespace

Image downscale(const double* data, std::size_t nx, std::size_t ny, std::array<std::size_t, 2> newsize) {
    Grid g(newsize[0], newsize[1]);
    const double s = (double)nx / (double)newsize[0];
    for (std::size_t y = 0; y < ny; ++y)
        for (std::size_t x = 0; x < nx; ++x) {
            const double v = data[x + y * nx];
            if (v >= 0) g.splat((double)x / s, (double)y / s, v);
        }
    // Only valid pixels are splatted, so "valid weight > 0.1 x all weight" holds
    // exactly where any weight arrived.
    return g.mean_where(g.weight);
}

Image downscale_and_correct(const double* data, const double* mask, std::size_t nx, std::size_t ny,
                            std::array<double, 2> center, double R, double dx, std::array<std::size_t, 2> newsize) {
    Grid g(newsize[0], newsize[1]);
    const Ewald e(nx, ny, center, R, dx);
    const double s0 = (double)nx / (double)newsize[0], s1 = (double)ny / (double)newsize[1];
    for (std::size_t y = 0; y < ny; ++y)
        for (std::size_t x = 0; x < nx; ++x) {
            const std::size_t i = x + y * nx;
            if (!(mask[i] > 0.5)) continue;
            double ox, oy, gain;
            e.map(x, y, ox, oy, gain);
            g.splat((ox + center[0]) / s0, (oy + center[1]) / s1, data[i] * gain);
        }
    return g.mean_where(g.max_weight());
}

std::tuple<Image, Image> cut_downscale_and_correct(const double* data, const double* mask, std::size_t nx,
                                                   std::size_t ny, std::array<double, 2> center, double R, double dx,
                                                   std::array<std::size_t, 2> cutsize,
                                                   std::array<std::size_t, 2> newsize) {
    Grid g(newsize[0], newsize[1]), m(newsize[0], newsize[1]);
    const Ewald e(nx, ny, center, R, dx);
    const double s0 = (double)cutsize[0] / (double)newsize[0], s1 = (double)cutsize[1] / (double)newsize[1];
    const double h0 = (double)(newsize[0] / 2), h1 = (double)(newsize[1] / 2);
    for (std::size_t y = 0; y < ny; ++y)
        for (std::size_t x = 0; x < nx; ++x) {
            const std::size_t i = x + y * nx;
            double ox, oy, gain;
            e.map(x, y, ox, oy, gain);
            const double fx = ox / s0 + h0, fy = oy / s1 + h1;
            if (mask[i] > 0.5) g.splat(fx, fy, data[i] * gain);
            m.splat(fx, fy, mask[i]);
        }
    return std::make_tuple(g.mean_where(g.max_weight()), m.mean_where(m.max_weight()));
}

std::tuple<Image, Image> cut_and_downscale(const double* data, const double* mask, std::size_t nx, std::size_t ny,
                                           std::array<double, 2> center, std::array<std::size_t, 2> cutsize,
                                           std::array<std::size_t, 2> newsize) {
    Grid g(newsize[0], newsize[1]), m(newsize[0], newsize[1]);
    const double s0 = (double)cutsize[0] / (double)newsize[0], s1 = (double)cutsize[1] / (double)newsize[1];
    const double h0 = (double)(newsize[0] / 2), h1 = (double)(newsize[1] / 2);
    for (std::size_t y = 0; y < ny; ++y)
        for (std::size_t x = 0; x < nx; ++x) {
            const std::size_t i = x + y * nx;
            const double fx = ((double)x - center[0]) / s0 + h0, fy = ((double)y - center[1]) / s1 + h1;
            if (mask[i] > 0.5) g.splat(fx, fy, data[i]);
            m.splat(fx, fy, mask[i]);
        }
    return std::make_tuple(g.mean_where(m.weight), m.mean_where(m.max_weight()));
}

std::vector<double> radial_profile(const double* data, std::size_t nx, std::size_t ny, double dtheta,
                                   std::array<double, 2> center, const std::vector<double>& angles,
                                   std::array<double, 2> phi) {
    const std::size_t na = angles.size();
    std::vector<double> prof(na, 0.0), norm(na, 0.0);
    for (std::size_t y = 0; y < ny && na >= 2; ++y)
        for (std::size_t x = 0; x < nx; ++x) {
            const double rx = (double)x - center[0], ry = (double)y - center[1];
            double ang = std::atan2(ry, rx) / M_PI * 180.;
            while (ang + 360. <= phi[1]) ang += 360.;
            if (!(ang <= phi[1] && ang >= phi[0])) continue;
            const double radius = std::sqrt(rx * rx + ry * ry) * dtheta;
            // first bracket angles[k] <= radius < angles[k+1]
            std::size_t k = 0;
            while (k + 1 < na && !(angles[k] <= radius && angles[k + 1] > radius)) ++k;
            if (k + 1 >= na) continue;
            const double v = data[x + y * nx];
            if (v < 0) continue;
            const double f = (angles[k + 1] - radius) / (angles[k + 1] - angles[k]);
            prof[k] += v * f;
            norm[k] += f;
            prof[k + 1] += v * (1. - f);
            norm[k + 1] += 1. - f;
        }
    for (std::size_t k = 0; k < na; ++k) prof[k] = norm[k] > 4 ? prof[k] / norm[k] : -1.;
    return prof;
}

}  // namespace prep
}  // namespace spring
