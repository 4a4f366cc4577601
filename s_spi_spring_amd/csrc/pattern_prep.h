// Detector-image preprocessing of spring_core.pattern (host, double precision):
// rescaling of a diffraction image by bilinear splatting of every input pixel
// onto the output grid, optionally after mapping the flat detector onto the
// Ewald sphere, and the radial profile. Behaviour follows the reference's
// Pattern class (src/pattern/pattern.h:8-30, src/pattern/pattern.cpp:16-516)
// as bound in src/wrapper.cpp:318-420; images are row-major [ny][nx], points
// are (x, y).
#pragma once

#include <array>
#include <cstddef>
#include <tuple>
#include <vector>

namespace spring {
namespace prep {

using Image = std::vector<double>;

// pattern.cpp:16-81 (downscale): pixels >= 0 are splatted at (x, y)/s with
// s = nx/newnx (used for both axes); an output pixel keeps the weighted mean when
// its weight from valid pixels exceeds 0.1 of its weight from all pixels, else -1.
Image downscale(const double* data, std::size_t nx, std::size_t ny, std::array<std::size_t, 2> newsize);

// pattern.cpp:85-183 (downscale_and_correct): pixels with mask > 0.5 are moved
// from the flat detector (distance R, pixel size dx, centre c) to the same
// scattering angle on the Ewald-sphere projection, weighted by 1/cos^3(theta),
// and splatted at (c + shifted position)/s per axis. Output pixels whose weight
// exceeds 0.1 of the largest weight keep the weighted mean, the others are -1.
Image downscale_and_correct(const double* data, const double* mask, std::size_t nx, std::size_t ny,
                            std::array<double, 2> center, double R, double dx, std::array<std::size_t, 2> newsize);

// pattern.cpp:190-318 (cut_downscale_and_correct): as downscale_and_correct, the
// output centred at newsize/2 with scale cutsize/newsize; the mask (every pixel)
// is splatted the same way and normalised against its own largest weight.
std::tuple<Image, Image> cut_downscale_and_correct(const double* data, const double* mask, std::size_t nx,
                                                   std::size_t ny, std::array<double, 2> center, double R, double dx,
                                                   std::array<std::size_t, 2> cutsize,
                                                   std::array<std::size_t, 2> newsize);

// pattern.cpp:323-438 (cut_and_downscale): no Ewald mapping; an output pixel
// keeps its data mean when the data weight exceeds 0.1 of its mask weight.
std::tuple<Image, Image> cut_and_downscale(const double* data, const double* mask, std::size_t nx, std::size_t ny,
                                           std::array<double, 2> center, std::array<std::size_t, 2> cutsize,
                                           std::array<std::size_t, 2> newsize);

// pattern.cpp:455-516 (get_radial_profile): mean of the pixels >= 0 whose
// direction (degrees, wrapped upward into the interval) lies in phi, linearly
// shared between the two bracketing scattering angles of `angles` (sorted);
// entries with total weight <= 4 are -1.
std::vector<double> radial_profile(const double* data, std::size_t nx, std::size_t ny, double dtheta,
                                   std::array<double, 2> center, const std::vector<double>& angles,
                                   std::array<double, 2> phi);

}  // namespace prep
}  // namespace spring
