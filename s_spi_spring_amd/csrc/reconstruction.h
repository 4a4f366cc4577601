// Reconstruction: one individual of the population, host side.
// Same members and accessors as the reference's type (src/reconstruction.h:15-110):
// density (complex64, N*N), support (int32, N*N) and seven float metrics.
#pragma once

#include <array>
#include <atomic>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

namespace spring {

// Process-wide counter for Reconstruction::version.
inline uint64_t next_version() {
    static std::atomic<uint64_t> v{1};
    return v.fetch_add(1, std::memory_order_relaxed);
}

using FLOAT = float;
using COMPLEX = std::complex<float>;
using SUPPORTTYPE = int32_t;

class Reconstruction {
public:
    Reconstruction(const COMPLEX* density_in, const SUPPORTTYPE* support_in, std::array<std::size_t, 2> dims_in)
        : dims(dims_in), error(-1), error_real(-1) {
        const std::size_t np = dims[0] * dims[1];
        density.assign(density_in, density_in + np);
        support.assign(support_in, support_in + np);
    }

    explicit Reconstruction(std::array<std::size_t, 2> dims_in) : dims(dims_in), error(-1), error_real(-1) {
        density.resize(dims[0] * dims[1]);
        support.resize(dims[0] * dims[1]);
    }

    // As the reference's copy constructor (src/reconstruction.h:35-45), only the
    // field values, `error` and `error_real` are copied; the other metrics are
    // zero here (the reference leaves them uninitialised).
    Reconstruction(const Reconstruction& o)
        : density(o.density), support(o.support), dims(o.dims), error(o.error), error_real(o.error_real) {}

    std::vector<COMPLEX> get_density() const { return density; }
    std::vector<SUPPORTTYPE> get_support() const { return support; }
    void set_support(const SUPPORTTYPE* s) {
        std::memcpy(support.data(), s, support.size() * sizeof(SUPPORTTYPE));
        version = next_version();
    }
    void set_density(const COMPLEX* d) {
        std::memcpy(density.data(), d, density.size() * sizeof(COMPLEX));
        version = next_version();
    }

    FLOAT get_error() const { return error; }
    FLOAT get_fourier_error() const { return error_fourier; }
    FLOAT get_real_error() const { return error_real; }
    FLOAT get_reality() const { return reality; }
    FLOAT get_symmetry() const { return symmetry; }
    FLOAT get_module_sum() const { return modulesum; }
    FLOAT get_complex_sum() const { return complexsum; }

    std::vector<COMPLEX> density;
    std::vector<SUPPORTTYPE> support;
    std::array<std::size_t, 2> dims;
    FLOAT error = -1;
    FLOAT error_fourier = 0;
    FLOAT error_real = -1;
    FLOAT reality = 0;
    FLOAT modulesum = 0;
    FLOAT complexsum = 0;
    FLOAT symmetry = 0;
    // Changes whenever density or support change (a device copy of this
    // individual is current only while it carries the same version).
    uint64_t version = next_version();
};

}  // namespace spring
