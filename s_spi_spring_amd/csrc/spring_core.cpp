// pybind11 module `spring_core` with the Python surface of the reference's
// binding (src/wrapper.cpp:24-420, 459-514): submodules `phase_retrieval`
// (Reconstruction, Solver), `pattern` (detector-image preprocessing) and
// `GPUtils`. Argument names, defaults and the
// single-or-list convention of every method are the reference's.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <variant>

#include "pattern_prep.h"
#include "solver.h"

namespace py = pybind11;
using namespace spring;

namespace {

using OneOrMany = std::variant<RecPtr, RecVec>;

// Calls f on a list; a single reconstruction is wrapped and unwrapped, as in
// every population method of src/wrapper.cpp.
template <class F> OneOrMany one_or_many(const OneOrMany& rec, F&& f) {
    if (const RecPtr* one = std::get_if<RecPtr>(&rec)) {
        RecVec v{*one};
        return f(v)[0];
    }
    return f(std::get<RecVec>(rec));
}

RecVec as_vec(const OneOrMany& rec) {
    if (const RecPtr* one = std::get_if<RecPtr>(&rec)) return RecVec{*one};
    return std::get<RecVec>(rec);
}

template <class T> py::array_t<T> square_array(const std::vector<T>& v, std::size_t n0, std::size_t n1) {
    py::array_t<T> a({n0, n1});
    std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(T));
    return a;
}

void require_square(const py::buffer_info& b, std::size_t n0, std::size_t n1, const char* what) {
    if (b.ndim != 2 || (std::size_t)b.shape[0] != n0 || (std::size_t)b.shape[1] != n1)
        throw std::invalid_argument(std::string(what) + " has the wrong shape");
}

}  // namespace

PYBIND11_MODULE(spring_core, sp) {
    sp.doc() = "MI355X backend of the SPRING memetic phase-retrieval engine";
    py::module pr = sp.def_submodule("phase_retrieval");

    py::class_<Reconstruction, RecPtr>(pr, "Reconstruction")
        .def(py::init([](py::array_t<std::complex<float>, py::array::c_style | py::array::forcecast> density,
                         py::array_t<SUPPORTTYPE, py::array::c_style | py::array::forcecast> support) {
                 auto d = density.request();
                 auto s = support.request();
                 if (d.ndim != 2) throw std::invalid_argument("density must be 2-D");
                 require_square(s, d.shape[0], d.shape[1], "support");
                 return std::make_shared<Reconstruction>(
                     (const COMPLEX*)d.ptr, (const SUPPORTTYPE*)s.ptr,
                     std::array<std::size_t, 2>{(std::size_t)d.shape[0], (std::size_t)d.shape[1]});
             }),
             "define the reconstruction object", py::arg("density"), py::arg("support"))
        .def("get_density", [](Reconstruction& r) { return square_array(r.density, r.dims[0], r.dims[1]); })
        .def("get_support", [](Reconstruction& r) { return square_array(r.support, r.dims[0], r.dims[1]); })
        .def("set_support",
             [](Reconstruction& r, py::array_t<SUPPORTTYPE, py::array::c_style | py::array::forcecast> s) {
                 auto b = s.request();
                 require_square(b, r.dims[0], r.dims[1], "support");
                 r.set_support((const SUPPORTTYPE*)b.ptr);
             })
        .def("set_density",
             [](Reconstruction& r, py::array_t<std::complex<float>, py::array::c_style | py::array::forcecast> d) {
                 auto b = d.request();
                 require_square(b, r.dims[0], r.dims[1], "density");
                 r.set_density((const COMPLEX*)b.ptr);
             })
        .def("get_error", &Reconstruction::get_error)
        .def("get_real_error", &Reconstruction::get_real_error)
        .def("get_fourier_error", &Reconstruction::get_fourier_error)
        .def("get_reality", &Reconstruction::get_reality)
        .def("get_symmetry", &Reconstruction::get_symmetry)
        .def("get_module_sum", &Reconstruction::get_module_sum)
        .def("get_complex_sum", &Reconstruction::get_complex_sum);

    py::class_<Solver, std::shared_ptr<Solver>>(pr, "Solver")
        .def(py::init([](py::array_t<FLOAT, py::array::c_style | py::array::forcecast> pattern,
                         std::array<int, 2> center, FLOAT reality, FLOAT realpen, int nthreads, bool use_bounds,
                         std::vector<int> gpus, int seed) {
                 auto b = pattern.request();
                 if (b.ndim != 2) throw std::invalid_argument("pattern must be 2-D");
                 return std::make_shared<Solver>(
                     (const FLOAT*)b.ptr, std::array<std::size_t, 2>{(std::size_t)b.shape[0], (std::size_t)b.shape[1]},
                     center, reality, realpen, nthreads, use_bounds, gpus, seed);
             }),
             "define the solver object", py::arg("pattern"), py::arg("center") = std::array<int, 2>({-1, -1}),
             py::arg("reality") = 0, py::arg("realpen") = 0, py::arg("nthreads") = 0, py::arg("use_bounds") = true,
             py::arg("gpus") = std::vector<int>({}), py::arg("seed") = 17)
        .def("get_pattern", [](Solver& s) { return square_array(s.get_pattern(), s.dims[0], s.dims[1]); })
        .def("set_pattern",
             [](Solver& s, py::array_t<FLOAT, py::array::c_style | py::array::forcecast> p) {
                 auto b = p.request();
                 require_square(b, s.dims[0], s.dims[1], "pattern");
                 s.set_pattern((const FLOAT*)b.ptr);
             })
        .def("set_shift_param", [](Solver& s, FLOAT a, FLOAT b) { s.set_shift_param(a, b); },
             py::arg("shiftmin") = 0., py::arg("shiftmax") = 2.)
        .def("Sequence",
             [](Solver& s, OneOrMany rec, int nrep) {
                 return one_or_many(rec, [&](RecVec v) { return s.Sequence(v, nrep); });
             },
             py::arg("reconstruction"), py::arg("nrep") = 1, py::call_guard<py::gil_scoped_release>())
        .def("Reshift",
             [](Solver& s, OneOrMany rec, RecPtr ref) {
                 return one_or_many(rec, [&](RecVec v) { return s.Reshift(v, ref); });
             },
             py::arg("reconstruction"), py::arg("reference"))
        .def("Crossover",
             [](Solver& s, RecVec rec, RecPtr best, RecPtr avg, FLOAT crossbalance, FLOAT crossdiff, FLOAT crossexp,
                FLOAT crossft, bool crosssym, FLOAT use_best, bool update_support) {
                 return s.Crossover(rec, best, avg, crossbalance, crossdiff, crossexp, crossft, crosssym, use_best,
                                    update_support);
             },
             py::arg("reconstruction"), py::arg("best"), py::arg("average"), py::arg("crossbalance"),
             py::arg("crossdiff"), py::arg("crossexp") = 1, py::arg("crossft") = true, py::arg("crosssym") = false,
             py::arg("usebest") = 0, py::arg("update_support") = true, py::call_guard<py::gil_scoped_release>())
        .def("HIO",
             [](Solver& s, OneOrMany rec, int iterations, FLOAT beta) {
                 return one_or_many(rec, [&](RecVec v) { return s.HIO(v, iterations, beta); });
             },
             py::arg("reconstruction"), py::arg("iterations"), py::arg("beta"))
        .def("RAAR",
             [](Solver& s, OneOrMany rec, int iterations, FLOAT beta) {
                 return one_or_many(rec, [&](RecVec v) { return s.RAAR(v, iterations, beta); });
             },
             py::arg("reconstruction"), py::arg("iterations"), py::arg("beta"))
        .def("ER",
             [](Solver& s, OneOrMany rec, int iterations) {
                 return one_or_many(rec, [&](RecVec v) { return s.ER(v, iterations); });
             },
             py::arg("reconstruction"), py::arg("iterations"))
        .def("SW",
             [](Solver& s, OneOrMany rec, FLOAT sigma, FLOAT threshold) {
                 return one_or_many(rec, [&](RecVec v) { return s.SW(v, sigma, threshold); });
             },
             py::arg("reconstruction"), py::arg("sigma"), py::arg("threshold"))
        .def("SWA",
             [](Solver& s, OneOrMany rec, FLOAT sigma, int pixels) {
                 return one_or_many(rec, [&](RecVec v) { return s.SWA(v, sigma, pixels); });
             },
             py::arg("reconstruction"), py::arg("sigma"), py::arg("pixels"))
        .def("Evaluate", [](Solver& s, OneOrMany rec) { s.Evaluate(as_vec(rec)); })
        .def("set_sequence", [](Solver& s, std::vector<SeqItem> seq) { s.set_sequence(seq); })
        .def("set_reality", [](Solver& s, FLOAT r) { s.set_reality(r); })
        .def("set_max_batch", [](Solver& s, int k) { s.set_max_batch(k); }, py::arg("k"))
        .def("get_max_batch", [](Solver& s) { return s.get_max_batch(); })
        .def("InitializePop",
             [](Solver& s, int nrec, int support_size, std::array<FLOAT, 2> itemsize, std::array<int, 2> nitems,
                std::array<FLOAT, 2> exponent, std::array<FLOAT, 2> scaling, FLOAT phase, FLOAT suppcut) {
                 return s.Initialize_pop(nrec, support_size, itemsize, nitems, exponent, scaling, phase, suppcut);
             },
             py::arg("popsize"), py::arg("suppsize"), py::arg("itemsize"), py::arg("nitems") = 8,
             py::arg("exponent") = 0, py::arg("scaling") = 0, py::arg("phaserange") = 0, py::arg("suppcut") = 0)
        .def("InitializePopRef",
             [](Solver& s, int nrec, RecPtr ref, std::array<FLOAT, 2> itemsize, std::array<int, 2> nitems,
                std::array<FLOAT, 2> exponent, std::array<FLOAT, 2> scaling, FLOAT phase, FLOAT suppcut) {
                 return s.Initialize_pop(nrec, ref, itemsize, nitems, exponent, scaling, phase, suppcut);
             },
             py::arg("popsize"), py::arg("reference"), py::arg("itemsize"), py::arg("nitems") = 8,
             py::arg("exponent") = 0, py::arg("scaling") = 0, py::arg("phaserange") = 0, py::arg("suppcut") = 0)
        .def("InitializePopSupp",
             [](Solver& s, int nrec, py::array_t<SUPPORTTYPE, py::array::c_style | py::array::forcecast> refsupport,
                FLOAT initfrac, std::array<FLOAT, 2> scaling, FLOAT phase) {
                 auto b = refsupport.request();
                 require_square(b, s.dims[0], s.dims[1], "reference support");
                 std::vector<SUPPORTTYPE> v((const SUPPORTTYPE*)b.ptr, (const SUPPORTTYPE*)b.ptr + s.size);
                 return s.Initialize_pop(nrec, v, initfrac, scaling, phase);
             },
             py::arg("popsize"), py::arg("reference"), py::arg("fraction"), py::arg("scaling"), py::arg("phaserange"))
        .def("GetAverage", [](Solver& s, RecVec rec, FLOAT fraction) { return s.GetAverage(rec, fraction); },
             py::arg("reconstruction"), py::arg("fraction") = 1.)
        .def("GetAverageModule",
             [](Solver& s, RecVec rec, bool ft, FLOAT fraction) {
                 return square_array(s.GetAverageModule(rec, ft, fraction), s.dims[0], s.dims[1]);
             },
             py::arg("reconstruction"), py::arg("ft") = true, py::arg("fraction") = 1.);

    // Detector-image preprocessing (src/wrapper.cpp:318-420). Images are 2-D
    // float64 arrays [rows][columns]; sizes and centres are (x, y).
    py::module pm = sp.def_submodule("pattern");
    using Dbl = py::array_t<double, py::array::c_style | py::array::forcecast>;
    using Sz2 = std::array<std::size_t, 2>;
    using D2 = std::array<double, 2>;
    auto image = [](const std::vector<double>& v, Sz2 size) {
        py::array_t<double> a({size[1], size[0]});
        std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(double));
        return a;
    };
    auto dims_of = [](const py::buffer_info& b, const char* what) {
        if (b.ndim != 2) throw std::invalid_argument(std::string(what) + " must be 2-D");
        return Sz2{(std::size_t)b.shape[1], (std::size_t)b.shape[0]};
    };
    auto check_mask = [](const py::buffer_info& m, Sz2 d) {
        if (m.ndim != 2 || (std::size_t)m.shape[1] != d[0] || (std::size_t)m.shape[0] != d[1])
            throw std::invalid_argument("mask must have the shape of data");
    };
    pm.def("downscale", [=](Dbl data, Sz2 newsize) {
        auto b = data.request();
        const Sz2 d = dims_of(b, "data");
        return image(prep::downscale((const double*)b.ptr, d[0], d[1], newsize), newsize);
    }, "Downscale the pattern to newsize (x, y); negative pixels are ignored.", py::arg("data"), py::arg("newsize"));
    pm.def("downscale_and_correct", [=](Dbl data, Dbl mask, D2 center, double R, double dx, Sz2 newsize) {
        auto b = data.request();
        auto m = mask.request();
        const Sz2 d = dims_of(b, "data");
        check_mask(m, d);
        return image(prep::downscale_and_correct((const double*)b.ptr, (const double*)m.ptr, d[0], d[1], center, R,
                                                 dx, newsize), newsize);
    }, "Ewald-sphere correction and downscaling of the pattern to newsize.", py::arg("data"), py::arg("mask"),
       py::arg("center"), py::arg("distance"), py::arg("pixel_size"), py::arg("newsize"));
    pm.def("cut_downscale_and_correct",
           [=](Dbl data, Dbl mask, D2 center, double R, double dx, Sz2 cutsize, Sz2 newsize) {
        auto b = data.request();
        auto m = mask.request();
        const Sz2 d = dims_of(b, "data");
        check_mask(m, d);
        auto r = prep::cut_downscale_and_correct((const double*)b.ptr, (const double*)m.ptr, d[0], d[1], center, R, dx,
                                                 cutsize, newsize);
        return std::make_tuple(image(std::get<0>(r), newsize), image(std::get<1>(r), newsize));
    }, "Ewald-sphere correction, crop of cutsize around the centre and downscaling to newsize; returns "
       "(pattern, mask).", py::arg("data"), py::arg("mask"), py::arg("center"), py::arg("distance"),
       py::arg("pixel_size"), py::arg("cutsize"), py::arg("newsize"));
    pm.def("cut_and_downscale", [=](Dbl data, Dbl mask, D2 center, Sz2 cutsize, Sz2 newsize) {
        auto b = data.request();
        auto m = mask.request();
        const Sz2 d = dims_of(b, "data");
        check_mask(m, d);
        auto r = prep::cut_and_downscale((const double*)b.ptr, (const double*)m.ptr, d[0], d[1], center, cutsize,
                                         newsize);
        return std::make_tuple(image(std::get<0>(r), newsize), image(std::get<1>(r), newsize));
    }, "Crop of cutsize around the centre and downscaling to newsize; returns (pattern, mask).", py::arg("data"),
       py::arg("mask"), py::arg("center"), py::arg("cutsize"), py::arg("newsize"));
    pm.def("get_radial_profile", [=](Dbl data, double dtheta, D2 center, std::vector<double> angles, D2 phi) {
        auto b = data.request();
        const Sz2 d = dims_of(b, "data");
        return prep::radial_profile((const double*)b.ptr, d[0], d[1], dtheta, center, angles, phi);
    }, "Radial profile at the given (increasing) scattering angles within the directions phi (degrees).",
       py::arg("data"), py::arg("dtheta"), py::arg("center"), py::arg("angles"), py::arg("phi"));

    // Device enumeration (src/gpu/gpu_utils.cu:13-111 through src/wrapper.cpp:459-487).
    // There is no load query: get_usage returns -1 per device, as the reference
    // does when built without NVML.
    py::module gu = sp.def_submodule("GPUtils");
    gu.def("get_count", []() {
        int n = 0;
        spr_device_count(&n);
        return n;
    });
    gu.def("get_ids", []() {
        int n = 0;
        spr_device_count(&n);
        std::vector<int> ids(n);
        for (int i = 0; i < n; ++i) ids[i] = i;
        return ids;
    });
    gu.def("get_usage", []() {
        int n = 0;
        spr_device_count(&n);
        return std::vector<int>(n, -1);
    });
    gu.def("get_info", []() {
        int n = 0;
        spr_device_count(&n);
        std::vector<std::string> names(n);
        for (int i = 0; i < n; ++i) {
            char buf[256] = {0};
            spr_device_name(i, buf, sizeof(buf));
            names[i] = buf;
        }
        return names;
    });

    // `from spring_core.phase_retrieval import Solver` needs the submodules in sys.modules.
    py::module sys = py::module::import("sys");
    py::dict modules = sys.attr("modules");
    const std::string base = py::str(sp.attr("__name__"));
    modules[py::str(base + ".phase_retrieval")] = pr;
    modules[py::str(base + ".GPUtils")] = gu;
    modules[py::str(base + ".pattern")] = pm;
}
