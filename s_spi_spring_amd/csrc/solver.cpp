// Population-level operations of memetic phase retrieval on one MI355X.
// Each method keeps the behaviour of the reference's Solver method of the same
// name (src/solver.cpp, cited per method); the per-individual work is handed to
// the device backend as one batch (include/spring_hip.h).
#include "solver.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <numeric>
#include <stdexcept>
#include <thread>
#include <unordered_map>
#include <unordered_set>

namespace spring {

namespace {

// Float sum of a field in index order and its phase, as the reference forms the
// alignment phases on the host (src/gpu/gpu_context.cu:636-641, 1233-1242).
FLOAT field_phase(const std::vector<COMPLEX>& d) {
    FLOAT sx = 0, sy = 0;
    for (const COMPLEX& z : d) {
        sx += z.real();
        sy += z.imag();
    }
    return std::atan2(sy, sx);
}

// f(i) for i in [0, n) on up to 16 host threads (independent iterations).
template <class F> void parallel_for(int n, F f) {
    const int nt = std::min(n, (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    if (nt <= 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&, t] {
            for (int i = t; i < n; i += nt) f(i);
        });
    for (std::thread& th : pool) th.join();
}

}  // namespace

void Solver::check(int rc) const {
    if (rc != SPR_OK) throw std::runtime_error(std::string("spring_hip: ") + spr_last_error(ctx));
}

// An operation list that failed on the device may have changed some resident
// individuals before it stopped: slots [0, P) no longer match their host
// objects, so they are forgotten before the error is raised.
void Solver::check_run(int rc, int P) {
    if (rc != SPR_OK) {
        const std::string msg = std::string("spring_hip: ") + spr_last_error(ctx);
        forget(P);
        throw std::runtime_error(msg);
    }
}

// Solver::Solver (src/solver.cpp:9-52). One device per Solver: the reference
// distributes its OpenMP contexts round-robin over GPU_ids; this backend runs
// the whole population on the first listed device.
Solver::Solver(const FLOAT* pattern_in, std::array<std::size_t, 2> dims_in, std::array<int, 2> center_in,
               FLOAT reality_, FLOAT realpen_, int nthreads_, bool use_bounds_, std::vector<int> GPU_ids, int seed)
    : dims(dims_in), center(center_in), size(dims_in[0] * dims_in[1]), reality(reality_), realpen(realpen_),
      shiftmin(0), shiftmax(2) {
    if (dims[0] != dims[1]) throw std::invalid_argument("Solver: the pattern must be square");
    if (!spr_supported_size((int)dims[0]))
        throw std::invalid_argument("Solver: unsupported pattern size " + std::to_string(dims[0]));
    if (GPU_ids.empty())
        throw std::invalid_argument(
            "Solver: no device selected (gpus=[]). There is no CPU backend; pass gpus=[device ids].");
    device = GPU_ids[0];
    if (GPU_ids.size() > 1)
        std::cerr << "WARNING: Solver runs its whole population on device " << device << " only; the other "
                  << GPU_ids.size() - 1 << " listed device(s) are not used. Run one Solver per device to use more."
                  << std::endl;
    set_threads(nthreads_);
    pattern.assign(pattern_in, pattern_in + size);
    if (center[0] < 0) center[0] = (int)dims[0] / 2;
    if (center[1] < 0) center[1] = (int)dims[1] / 2;
    set_boundaries();
    prepare(use_bounds_);
    for (int i = 0; i < nthreads; ++i) rngs.emplace_back((i + 1) * seed);
}

Solver::~Solver() {
    if (ctx) spr_destroy(ctx);
}

// Solver::set_threads (src/solver.cpp:701-711): the number of reference contexts,
// here the number of independent random streams.
void Solver::set_threads(int nthreads_) {
    const int hw = std::max(1u, std::thread::hardware_concurrency());
    if (nthreads_ <= 0) {
        nthreads = std::max(1, hw / 2);
    } else {
        nthreads = nthreads_;
        if (nthreads > hw)
            printf("WARNING: The number of requried threads (%d) exceeds the number of cores (%d). This usually "
                   "implies performance decay\n",
                   nthreads, hw);
    }
}

// Solver::prepare (src/solver.cpp:173-206): intensity -> amplitude, moving the
// origin from `center` to the corner; negative pixels keep their bound code
// with use_bounds, else become free (-1).
void Solver::prepare(bool use_bounds) {
    const int n0 = (int)dims[0], n1 = (int)dims[1];
    const std::vector<FLOAT> in(pattern);
    for (int j = 0; j < n1; ++j) {
        const int ky = (j - center[1] + n1) % n1;
        for (int i = 0; i < n0; ++i) {
            const int kx = (i - center[0] + n0) % n0;
            const FLOAT v = in[i + (std::size_t)n0 * j];
            pattern[kx + (std::size_t)n0 * ky] = v >= 0 ? std::sqrt(v) : (use_bounds ? v : -1.f);
        }
    }
}

// Solver::set_boundaries (src/solver.cpp:209-289): an upper bound for masked
// pixels from the radial profile (mean + 1 std over rings of width 1, counted
// over +-1 bin) where the ring is well measured; saturated pixels become free.
void Solver::set_boundaries() {
    const int n0 = (int)dims[0], n1 = (int)dims[1];
    const int nbin = (int)(std::sqrt((FLOAT)(center[0] * center[0] + center[1] * center[1])) / (FLOAT)1) + 10;
    std::vector<FLOAT> mean(nbin, 0), sq(nbin, 0), valid(nbin, 0), all(nbin, 0), sat(nbin, 0);
    auto ring = [&](int i, int j) {
        const FLOAT d = std::sqrt((FLOAT)((j - center[1]) * (j - center[1]) + (i - center[0]) * (i - center[0])));
        return (int)std::round(d / (FLOAT)1);
    };
    for (int j = 0; j < n1; ++j)
        for (int i = 0; i < n0; ++i) {
            const FLOAT v = pattern[i + (std::size_t)n0 * j];
            const int r0 = ring(i, j);
            for (int b = std::max(0, r0 - 1); b <= r0 + 1 && b < nbin; ++b) {
                if (v >= 0) {
                    mean[b] += v;
                    sq[b] += v * v;
                    valid[b] += 1.f;
                }
                if (v < -1.5f) sat[b] += 1.f;
                all[b] += 1.f;
            }
        }
    for (int b = 0; b < nbin; ++b) {
        if (valid[b] > all[b] * 0.66 && sat[b] <= all[b] * 0.05 && all[b] > 40) {
            mean[b] /= valid[b];
            sq[b] = std::sqrt(std::abs(sq[b] / valid[b] - mean[b] * mean[b]));
        } else {
            mean[b] = -1;
        }
    }
    for (int j = 0; j < n1; ++j)
        for (int i = 0; i < n0; ++i) {
            FLOAT& v = pattern[i + (std::size_t)n0 * j];
            if (!(v < 0)) continue;
            const int b = ring(i, j);
            if (b <= 0 || b >= nbin) continue;
            if (mean[b] >= 0 && v < -0.5f && v > -1.5f) v = -2.f - std::sqrt(mean[b] + 1.f * sq[b]);
            else v = -1;
        }
}

spr_ctx* Solver::backend(int capacity) {
    if (ctx && capacity <= ctx_capacity) return ctx;
    if (ctx) {
        spr_destroy(ctx);
        ctx = nullptr;
    }
    const int cap = std::max(capacity, 1);
    spr_ctx* c = nullptr;
    if (spr_create(&c, (int)dims[0], cap, pattern.data(), device) != SPR_OK)
        throw std::runtime_error(std::string("spring_hip: ") + spr_last_error(nullptr));
    ctx = c;
    ctx_capacity = cap;
    slots.assign(cap, Slot{});
    return ctx;
}

void Solver::forget(int count) {
    for (int j = 0; j < count && j < (int)slots.size(); ++j) slots[j] = Slot{};
}

// Make rec[i] the content of slot i (i < P) and P the active size. Individuals
// whose current version is resident are gathered on the device; the others are
// copied in from the host. Other resident individuals displaced from [0, P) are
// parked in free slots >= P while there is room.
void Solver::place(const RecVec& rec) {
    const int P = (int)rec.size();
    backend(2 * P + 2);
    const int cap = ctx_capacity;
    std::unordered_map<const Reconstruction*, int> where;
    std::vector<RecPtr> owner(cap);
    for (int j = 0; j < cap; ++j) {
        RecPtr o = slots[j].owner.lock();
        if (o && o->version == slots[j].version) {
            owner[j] = o;
            where.emplace(o.get(), j);
        } else {
            slots[j] = Slot{};
        }
    }
    std::unordered_set<const Reconstruction*> wanted;
    for (const RecPtr& r : rec) {
        if (r->density.size() != size || r->support.size() != size)
            throw std::invalid_argument("reconstruction size differs from the pattern");
        wanted.insert(r.get());
    }
    std::vector<int> src(cap, -1);
    std::vector<Slot> next(cap);
    std::vector<char> taken(cap, 0);
    std::vector<int> up;
    for (int i = 0; i < P; ++i) {
        auto it = where.find(rec[i].get());
        if (it != where.end()) src[i] = it->second;
        else up.push_back(i);
        next[i] = Slot{rec[i], rec[i]->version};
        taken[i] = 1;
    }
    for (int j = P; j < cap; ++j)
        if (owner[j] && !wanted.count(owner[j].get())) {
            next[j] = slots[j];
            taken[j] = 1;
        }
    int f = P;
    for (int j = 0; j < P && j < cap; ++j) {
        if (!owner[j] || wanted.count(owner[j].get())) continue;
        while (f < cap && taken[f]) ++f;
        if (f >= cap) break;
        src[f] = j;
        next[f] = slots[j];
        taken[f] = 1;
    }
    // Slots >= P that neither move nor receive keep their (still current) content.
    for (int j = P; j < cap; ++j)
        if (!taken[j] && src[j] < 0 && owner[j]) next[j] = slots[j];
    check(spr_gather(ctx, cap, src.data()));
    slots = next;
    if (!up.empty()) {
        std::vector<const float*> rp(up.size());
        std::vector<const int32_t*> sp(up.size());
        for (std::size_t k = 0; k < up.size(); ++k) {
            rp[k] = reinterpret_cast<const float*>(rec[up[k]]->density.data());
            sp[k] = rec[up[k]]->support.data();
        }
        check(spr_put_individuals(ctx, (int)up.size(), up.data(), rp.data(), sp.data()));
    }
    check(spr_set_active(ctx, P));
}

void Solver::pull(const RecVec& rec, bool density, bool support) {
    const int P = (int)rec.size();
    if (P == 0 || (!density && !support)) return;
    std::vector<int> idx(P);
    std::vector<float*> rp(P);
    std::vector<int32_t*> sp(P);
    for (int i = 0; i < P; ++i) {
        idx[i] = i;
        rp[i] = reinterpret_cast<float*>(rec[i]->density.data());
        sp[i] = rec[i]->support.data();
    }
    check(spr_fetch_individuals(ctx, P, idx.data(), density ? rp.data() : nullptr, support ? sp.data() : nullptr));
    for (int i = 0; i < P; ++i) {
        rec[i]->version = next_version();
        slots[i] = Slot{rec[i], rec[i]->version};
    }
    // An object listed twice holds the last slot's data; its other slot is stale.
    for (int i = 0; i < P; ++i)
        for (int j = i + 1; j < P; ++j)
            if (rec[i] == rec[j]) slots[i] = Slot{};
}

// The whole population, or groups of max_batch individuals each copied in
// from the host (no device-resident copy is reused) and back.
void Solver::for_groups(const RecVec& rec, const std::function<void(const RecVec&)>& f) {
    if (max_batch <= 0 || (int)rec.size() <= max_batch) {
        f(rec);
        return;
    }
    for (std::size_t i = 0; i < rec.size(); i += (std::size_t)max_batch) {
        forget((int)slots.size());
        f(RecVec(rec.begin() + i, rec.begin() + std::min(rec.size(), i + (std::size_t)max_batch)));
    }
}

void Solver::run_op(const RecVec& rec, const spr_op& op) {
    if (rec.empty()) return;
    const bool support_only = op.kind == SPR_OP_SW || op.kind == SPR_OP_SWA;
    for_groups(rec, [&](const RecVec& g) {
        place(g);
        check_run(spr_run(ctx, &op, 1, 1, reality), (int)g.size());
        pull(g, !support_only, true);
    });
}

// Solver::HIO / RAAR / ER / SW / SWA (src/solver.cpp:360-450).
RecVec Solver::HIO(RecVec rec, int iterations, FLOAT beta) {
    run_op(rec, spr_op{SPR_OP_HIO, iterations, beta, 0, 0, 0});
    return rec;
}
RecVec Solver::RAAR(RecVec rec, int iterations, FLOAT beta) {
    run_op(rec, spr_op{SPR_OP_RAAR, iterations, beta, 0, 0, 0});
    return rec;
}
RecVec Solver::ER(RecVec rec, int iterations) {
    run_op(rec, spr_op{SPR_OP_ER, iterations, 0, 0, 0, 0});
    return rec;
}
RecVec Solver::SW(RecVec rec, FLOAT sigma, FLOAT threshold) {
    run_op(rec, spr_op{SPR_OP_SW, 0, 0, sigma, threshold, 0});
    return rec;
}
RecVec Solver::SWA(RecVec rec, FLOAT sigma, int pixels) {
    run_op(rec, spr_op{SPR_OP_SWA, 0, 0, sigma, 0, pixels});
    return rec;
}

// Solver::get_algorithm + set_sequence (src/solver.cpp:84-170): names with the
// expected parameter count become operations; anything else is reported and
// skipped. Counts and sizes are the float parameters truncated to int; a
// negative iteration count runs no iteration (the reference's loop bound), only
// the trailing projection.
void Solver::set_sequence(const std::vector<SeqItem>& sequence) {
    std::vector<spr_op> next;
    auto iters = [](FLOAT v) { return std::max(0, (int)v); };
    for (const SeqItem& it : sequence) {
        const std::string& name = std::get<0>(it);
        const std::vector<FLOAT>& p = std::get<1>(it);
        if (name == "ER" && p.size() == 1) next.push_back(spr_op{SPR_OP_ER, iters(p[0]), 0, 0, 0, 0});
        else if (name == "HIO" && p.size() == 2) next.push_back(spr_op{SPR_OP_HIO, iters(p[0]), p[1], 0, 0, 0});
        else if (name == "RAAR" && p.size() == 2) next.push_back(spr_op{SPR_OP_RAAR, iters(p[0]), p[1], 0, 0, 0});
        else if (name == "SW" && p.size() == 2) next.push_back(spr_op{SPR_OP_SW, 0, 0, p[0], p[1], 0});
        else if (name == "SWA" && p.size() == 2) next.push_back(spr_op{SPR_OP_SWA, 0, 0, p[0], 0, (int)p[1]});
        // Gradient refinement (no counterpart in the reference): GRAD [steps, lr(, loss kind)]
        // = Adam on the masked Poisson (kind 0) or amplitude (kind 1) Fourier loss.
        else if (name == "GRAD" && (p.size() == 2 || p.size() == 3)) {
            const int steps = (int)p[0], kind = p.size() == 3 ? (int)p[2] : 0;
            if (steps < 0 || (kind != 0 && kind != 1))
                throw std::invalid_argument("GRAD needs steps >= 0 and loss kind 0 (Poisson) or 1 (amplitude)");
            next.push_back(spr_op{SPR_OP_GRAD, steps, p[1], 0, 1e-3f, kind});
        } else {
            std::cout << name << " [ ";
            for (FLOAT v : p) std::cout << v << " ";
            std::cout << "] ";
            std::cout << " NOT RECOGNIZED. Skipping." << std::flush;
        }
    }
    ops = std::move(next);
}

// get_error + get_symmetry + the derived metrics of Solver::Sequence / Evaluate
// (src/solver.cpp:474-479, 549-554) for the resident population.
void Solver::evaluate_resident(const RecVec& rec) {
    const int P = (int)rec.size();
    std::vector<float> ef(P), er(P), cs(P), ms(P), sym(P);
    check(spr_evaluate(ctx, reality, ef.data(), er.data(), cs.data(), ms.data(), sym.data()));
    for (int i = 0; i < P; ++i) {
        Reconstruction& r = *rec[i];
        r.error_fourier = ef[i];
        r.error_real = er[i];
        r.complexsum = cs[i];
        r.modulesum = ms[i];
        r.symmetry = sym[i];
        r.reality = (FLOAT)(1. - r.complexsum / r.modulesum);
        r.error = (FLOAT)(r.error_fourier * (1. + realpen * r.reality));
    }
}

// Solver::Sequence (src/solver.cpp:454-487).
RecVec Solver::Sequence(RecVec rec, int nrep) {
    if (ops.empty()) {
        std::cout << "Warning: empty algorithm sequence. Operation has no effect" << std::endl;
        return rec;
    }
    if (rec.empty()) return rec;
    for_groups(rec, [&](const RecVec& g) {
        place(g);
        check_run(spr_run(ctx, ops.data(), (int)ops.size(), nrep, reality), (int)g.size());
        evaluate_resident(g);
        pull(g, true, true);
    });
    return rec;
}

// Solver::Evaluate (src/solver.cpp:541-558).
void Solver::Evaluate(RecVec rec) {
    if (rec.empty()) return;
    place(rec);
    evaluate_resident(rec);
}

// Solver::Reshift (src/solver.cpp:511-538): align every individual to ref.
RecVec Solver::Reshift(RecVec rec, RecPtr ref) {
    const int P = (int)rec.size();
    if (P == 0) return rec;
    place(rec);
    const FLOAT ref_phase = field_phase(ref->density);
    std::vector<float> direct(P), flip(P);
    parallel_for(P, [&](int i) {
        const FLOAT ph = field_phase(rec[i]->density);
        direct[i] = ref_phase - ph;
        flip[i] = ref_phase + ph;
    });
    check(spr_align(ctx, reinterpret_cast<const float*>(ref->density.data()), shiftmin, shiftmax, direct.data(),
                    flip.data(), nullptr));
    pull(rec, true, true);
    return rec;
}

std::vector<int> Solver::sorted_by_error(const RecVec& rec) const {
    std::vector<int> idx(rec.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int i, int j) { return rec[i]->get_error() < rec[j]->get_error(); });
    return idx;
}

// Solver::GetAverage (src/solver.cpp:618-686).
RecPtr Solver::GetAverage(RecVec rec, FLOAT fraction) {
    const int P = (int)rec.size();
    if (P == 0) throw std::invalid_argument("GetAverage: empty population");
    int navg = (int)(fraction * (FLOAT)P);
    navg = std::min(navg, P);
    navg = std::max(navg, 1);
    const std::vector<int> idx = sorted_by_error(rec);
    place(rec);
    std::vector<COMPLEX> d(size);
    std::vector<SUPPORTTYPE> s(size);
    check(spr_average(ctx, idx.data(), navg, reinterpret_cast<float*>(d.data()), s.data()));
    return std::make_shared<Reconstruction>(d.data(), s.data(), dims);
}

// Solver::GetAverageModule (src/solver.cpp:564-614).
std::vector<FLOAT> Solver::GetAverageModule(RecVec rec, bool ft, FLOAT fraction) {
    const int P = (int)rec.size();
    if (P == 0) throw std::invalid_argument("GetAverageModule: empty population");
    int navg = (int)(fraction * (FLOAT)P);
    navg = std::min(navg, P);
    navg = std::max(navg, 1);
    const std::vector<int> idx = sorted_by_error(rec);
    place(rec);
    std::vector<FLOAT> out(size);
    check(spr_average_module(ctx, idx.data(), navg, ft ? 1 : 0, out.data()));
    return out;
}

// Solver::Crossover (src/solver.cpp:716-817). The random draws of child i come
// from the stream of reference context i % nthreads, in child order (one
// realisation of the reference's dynamic OpenMP schedule).
RecVec Solver::Crossover(RecVec rec, RecPtr best, RecPtr avg, FLOAT crossbalance, FLOAT crossdiff, FLOAT crossexp,
                         FLOAT crossftin, bool crosssym, FLOAT use_average, bool update_support) {
    const int P = (int)rec.size();
    RecVec out(P);
    if (P == 0) return out;
    const std::vector<int> indexes = sorted_by_error(rec);
    int bestsuppix = 0;
    for (std::size_t i = 0; i < size; ++i) bestsuppix += best->support[i];
    FLOAT suppdim = 2. * std::sqrt(FLOAT(bestsuppix) / M_PI);
    if (suppdim < 4.) suppdim = 4;
    if (suppdim > dims[0] / 4.) suppdim = dims[0] / 4.;
    const std::vector<FLOAT> avg_mod = GetAverageModule(rec, true, 0.66f);

    std::vector<FLOAT> phase(P);
    parallel_for(P, [&](int i) { phase[i] = field_phase(rec[i]->density); });
    const FLOAT avg_phase = field_phase(avg->density);
    const int n = (int)dims[0];
    std::vector<spr_cross_desc> desc(P);
    std::vector<SUPPORTTYPE> maps;
    for (int irec = 0; irec < P; ++irec) {
        ContextRng& g = rngs[irec % nthreads];
        spr_cross_desc& d = desc[irec];
        d.p0 = irec;
        const int extracted = (int)(std::pow(g.rand(0, 1), crossexp) * FLOAT(P));
        d.p1 = indexes[std::min(extracted, P - 1)];
        d.p2 = (int)g.rand(0, P - 0.0001);
        d.p3 = (int)g.rand(0, P - 0.0001);
        d.crossft = g.rand(0, 1) < crossftin ? 1 : 0;
        d.crosssym = crosssym ? 1 : 0;
        if (d.crossft) {
            d.squaresizex = (int)g.rand(std::ceil(FLOAT(n) / suppdim), std::ceil(FLOAT(n) / 4.));
            d.squaresizey = (int)g.rand(std::ceil(FLOAT(n) / suppdim), std::ceil(FLOAT(n) / 4.));
        } else {
            d.squaresizex = (int)g.rand(suppdim / 10. + 1., suppdim / 2.);
            d.squaresizey = (int)g.rand(suppdim / 10. + 1., suppdim / 2.);
        }
        d.crosssizex = (n + d.squaresizex - 1) / d.squaresizex + 1;
        d.crosssizey = (n + d.squaresizey - 1) / d.squaresizey + 1;
        d.offsetx = (int)g.rand(0, d.squaresizex);
        d.offsety = (int)g.rand(0, d.squaresizey);
        d.rotangle = g.rand(0, M_PI);
        d.crossmap_offset = (int)maps.size();
        const int cells = d.crosssizex * d.crosssizey;
        for (int c = 0; c < cells; ++c) maps.push_back(g.rand(0, 1) < crossbalance ? 1 : 0);
        const FLOAT p1ph = phase[d.p1];
        const FLOAT role_phase[4] = {phase[d.p0], phase[d.p2], phase[d.p3], avg_phase};
        for (int k = 0; k < 4; ++k) {
            d.phase[2 * k] = p1ph - role_phase[k];
            d.phase[2 * k + 1] = p1ph + role_phase[k];
        }
    }
    place(rec);
    if (child_cap < (std::size_t)P) {
        child_rho.reset(new float[(std::size_t)P * size * 2]);
        child_supp.reset(new SUPPORTTYPE[(std::size_t)P * size]);
        child_cap = (std::size_t)P;
    }
    float* child = child_rho.get();
    SUPPORTTYPE* child_s = child_supp.get();
    check(spr_crossover(ctx, P, desc.data(), maps.data(), (int)maps.size(),
                        reinterpret_cast<const float*>(avg->density.data()), avg_mod.data(), crossdiff, use_average,
                        shiftmin, shiftmax, reality, update_support ? 1 : 0, child, child_s));
    // Each child starts as the reference's copy of rec[i] (field values, error and
    // error_real; src/reconstruction.h:35-45) with the crossed density (and support).
    parallel_for(P, [&](int i) {
        const COMPLEX* d = reinterpret_cast<const COMPLEX*>(child) + (std::size_t)i * size;
        const SUPPORTTYPE* sp = update_support ? child_s + (std::size_t)i * size : rec[i]->support.data();
        RecPtr son = std::make_shared<Reconstruction>(d, sp, dims);
        son->error = rec[i]->error;
        son->error_real = rec[i]->error_real;
        out[i] = son;
    });
    return out;
}

std::vector<FLOAT> Solver::get_pattern() { return pattern; }

// Solver::set_pattern (src/solver.cpp:55-66): the new pattern goes through
// prepare(false) again, as in the reference.
void Solver::set_pattern(const FLOAT* pattern_in) {
    pattern.assign(pattern_in, pattern_in + size);
    prepare(false);
    if (ctx) check(spr_set_pattern(ctx, pattern.data()));
}

// Batched 2-D transforms of host fields on the device (Context::get_ft / get_ift).
std::vector<COMPLEX> Solver::ft_batch(const std::vector<std::vector<COMPLEX>>& fields, bool inverse) {
    const int P = (int)fields.size();
    backend(std::max(2 * P + 2, ctx_capacity));
    forget(P);
    std::vector<COMPLEX> buf((std::size_t)P * size);
    for (int i = 0; i < P; ++i) std::copy(fields[i].begin(), fields[i].end(), buf.begin() + (std::size_t)i * size);
    check(spr_set_population(ctx, P, reinterpret_cast<const float*>(buf.data()), nullptr));
    check(inverse ? spr_ift(ctx) : spr_ft(ctx));
    check(spr_get_population(ctx, P, reinterpret_cast<float*>(buf.data()), nullptr));
    return buf;
}

// Solver::Initialize_pop / Initialize_rec with a disk support (src/solver.cpp:824-836, 870-957).
RecVec Solver::Initialize_pop(int nrec, int support_size, std::array<FLOAT, 2> itemsize, std::array<int, 2> nitems_in,
                              std::array<FLOAT, 2> exponent_in, std::array<FLOAT, 2> scaling_in, FLOAT phase_in,
                              FLOAT /*suppcut*/) {
    RecVec rec(nrec);
    if (nrec <= 0) return rec;
    const int n0 = (int)dims[0], n1 = (int)dims[1];
    std::vector<std::vector<COMPLEX>> dens(nrec, std::vector<COMPLEX>(size));
    std::vector<std::vector<SUPPORTTYPE>> supp(nrec, std::vector<SUPPORTTYPE>(size, 0));
    std::vector<FLOAT> scale(nrec), expo(nrec);
    std::vector<std::vector<FLOAT>> sig(nrec), cx(nrec), cy(nrec), ph(nrec);
    // The draws of individual r come from stream r % nthreads in the reference's
    // order; the fields are then built from them independently.
    for (int r = 0; r < nrec; ++r) {
        ContextRng& g = rngs[r % nthreads];
        const int nit = (int)g.rand(nitems_in[0], nitems_in[1] + 0.99);
        expo[r] = g.rand(exponent_in[0], exponent_in[1]);
        scale[r] = g.rand(scaling_in[0], scaling_in[1]);
        sig[r].resize(nit);
        cx[r].resize(nit);
        cy[r].resize(nit);
        ph[r].resize(nit);
        for (int i = 0; i < nit; ++i) {
            const FLOAT s = std::sqrt(g.rand(std::pow(itemsize[0], 2), std::pow(itemsize[1], 2))) / 2.;
            cx[r][i] = g.rand(s, support_size - s);
            cy[r][i] = g.rand(s, support_size - s);
            sig[r][i] = s;
            ph[r][i] = g.rand(-phase_in * M_PI, phase_in * M_PI);
        }
    }
    parallel_for(nrec, [&](int r) {
        std::vector<COMPLEX>& d = dens[r];
        const int nit = (int)sig[r].size();
        for (int iy = 0; iy < support_size; ++iy)
            for (int ix = 0; ix < support_size; ++ix) {
                const double dx = ix - support_size / 2., dy = iy - support_size / 2.;
                if (std::sqrt(dx * dx + dy * dy) > support_size / 2.) continue;
                const int idx = (iy + n1 / 2 - support_size / 2) * n0 + (ix + n0 / 2 - support_size / 2);
                supp[r][idx] = 1;
                for (int i = 0; i < nit; ++i) {
                    const FLOAT rad = std::sqrt((ix - cx[r][i]) * (ix - cx[r][i]) + (iy - cy[r][i]) * (iy - cy[r][i]));
                    if (rad < sig[r][i]) {
                        const FLOAT val = 2. * sig[r][i] * std::sin(std::acos(rad / sig[r][i]));
                        d[idx] += std::polar(val, ph[r][i]);
                    }
                }
                const FLOAT mod = std::abs(d[idx]);
                if (mod > 0) d[idx] = d[idx] / mod * std::pow(mod, expo[r]);
            }
    });
    const std::vector<COMPLEX> F = ft_batch(dens, false);
    parallel_for(nrec, [&](int r) {
        FLOAT norm = 0, norm_new = 0;
        for (std::size_t i = 0; i < size; ++i)
            if (pattern[i] >= 0) {
                norm += std::pow(std::abs(F[(std::size_t)r * size + i]), 2);
                norm_new += pattern[i] * pattern[i];
            }
        const FLOAT f = std::sqrt(norm_new / norm) * scale[r];
        for (COMPLEX& z : dens[r]) z *= f;
        rec[r] = std::make_shared<Reconstruction>(dens[r].data(), supp[r].data(), dims);
    });
    return rec;
}

// Solver::Initialize_pop / Initialize_rec seeded from a reference reconstruction
// (src/solver.cpp:840-852, 961-1058).
RecVec Solver::Initialize_pop(int nrec, RecPtr refrec, std::array<FLOAT, 2> itemsize, std::array<int, 2> nitems_in,
                              std::array<FLOAT, 2> exponent_in, std::array<FLOAT, 2> scaling_in, FLOAT phase_in,
                              FLOAT /*suppcut*/) {
    RecVec rec(nrec);
    const int n0 = (int)dims[0], n1 = (int)dims[1];
    const std::vector<SUPPORTTYPE>& rs = refrec->support;
    COMPLEX sum = 0;
    int supppix = 0;
    for (std::size_t i = 0; i < size; ++i) {
        sum += refrec->density[i];
        if (rs[i] > 0) supppix += 1;
    }
    const FLOAT newnorm = std::abs(sum);
    for (int r = 0; r < nrec; ++r) {
        ContextRng& g = rngs[r % nthreads];
        const int nit = (int)g.rand(nitems_in[0], nitems_in[1] + 0.99);
        const FLOAT expo = g.rand(exponent_in[0], exponent_in[1]);
        const FLOAT scaling = g.rand(scaling_in[0], scaling_in[1]);
        std::vector<FLOAT> sig(nit), cx(nit), cy(nit), ph(nit);
        for (int i = 0; i < nit; ++i) {
            const FLOAT s = std::sqrt(g.rand(std::pow(itemsize[0], 2), std::pow(itemsize[1], 2))) / 2.;
            const int target = (int)g.rand(0, supppix);
            int counter = 0, index = 0;
            for (std::size_t j = 0; j < size; ++j) {
                if (counter == target) {
                    index = (int)j;
                    break;
                }
                if (rs[j] > 0) counter += 1;
            }
            sig[i] = s;
            cx[i] = index % n0;
            cy[i] = index / n0;
            ph[i] = g.rand(-phase_in * M_PI, phase_in * M_PI);
        }
        std::vector<COMPLEX> d(size, 0);
        for (int iy = 0; iy < n1; ++iy)
            for (int ix = 0; ix < n0; ++ix) {
                const int idx = iy * n0 + ix;
                if (rs[idx] <= 0) continue;
                for (int i = 0; i < nit; ++i) {
                    const FLOAT rad = std::sqrt((ix - cx[i]) * (ix - cx[i]) + (iy - cy[i]) * (iy - cy[i]));
                    if (rad < sig[i]) {
                        const FLOAT val = 2. * sig[i] * std::sin(std::acos(rad / sig[i]));
                        d[idx] += std::polar(val, ph[i]);
                    }
                }
                const FLOAT mod = std::abs(d[idx]);
                if (mod > 0) d[idx] = d[idx] / mod * std::pow(mod, expo);
            }
        COMPLEX oldsum = 0;
        for (const COMPLEX& z : d) oldsum += z;
        const FLOAT oldnorm = std::abs(oldsum);
        for (COMPLEX& z : d) z *= newnorm / oldnorm * scaling;
        rec[r] = std::make_shared<Reconstruction>(d.data(), rs.data(), dims);
    }
    return rec;
}

// Solver::Initialize_pop / Initialize_rec from a support with random Fourier
// cells (src/solver.cpp:855-867, 1061-1174).
RecVec Solver::Initialize_pop(int nrec, std::vector<SUPPORTTYPE> support, FLOAT initfrac,
                              std::array<FLOAT, 2> scaling_in, FLOAT phase_in) {
    RecVec rec(nrec);
    if (nrec <= 0) return rec;
    const int n0 = (int)dims[0], n1 = (int)dims[1];
    std::vector<std::vector<COMPLEX>> dens(nrec, std::vector<COMPLEX>(size));
    struct Cells {
        int sqx, sqy, csx, offx, offy;
        FLOAT rot;
        std::vector<int> map;
        std::vector<FLOAT> scale, phase;
    };
    std::vector<Cells> cells(nrec);
    std::vector<FLOAT> normfactor(nrec);
    for (int r = 0; r < nrec; ++r) {
        ContextRng& g = rngs[r % nthreads];
        normfactor[r] = g.rand(scaling_in[0], scaling_in[1]);
        Cells& c = cells[r];
        c.sqx = (int)g.rand(1, FLOAT(n0) / 16.);
        c.sqy = (int)g.rand(1, FLOAT(n0) / 16.);
        c.csx = (n0 + c.sqx - 1) / c.sqx + 1;
        const int csy = (n1 + c.sqy - 1) / c.sqy + 1;
        c.offx = (int)g.rand(0, c.sqx);
        c.offy = (int)g.rand(0, c.sqy);
        c.rot = g.rand(0, M_PI);
        c.map.assign((std::size_t)c.csx * csy, 0);
        c.scale.assign(c.map.size(), 0);
        c.phase.assign(c.map.size(), 0);
        for (std::size_t i = 0; i < c.map.size(); ++i) {
            if (g.rand(0, 1) < initfrac) {
                c.map[i] = 1;
                c.scale[i] = g.rand(0.5, 2);
                c.phase[i] = g.rand(-M_PI * phase_in, M_PI * phase_in);
            }
        }
        for (std::size_t i = 0; i < size; ++i) dens[r][i] = (FLOAT)support[i];
    }
    std::vector<COMPLEX> F = ft_batch(dens, false);
    const COMPLEX imag(0.0, 1.0);
    for (int r = 0; r < nrec; ++r) {
        const Cells& c = cells[r];
        for (std::size_t i = 0; i < size; ++i) {
            int x = (int)(i % n0), y = (int)(i / n0);
            if (x >= n0 / 2) x -= n0;
            if (y >= n1 / 2) y -= n1;
            FLOAT radius = std::sqrt(FLOAT(x * x + y * y));
            if (radius >= n0 / 2.) radius = n0 / 2 - 1;
            FLOAT angle = std::atan2(FLOAT(y), FLOAT(x)) + c.rot;
            x = (int)(std::cos(angle) * radius) + n0 / 2;
            y = (int)(std::sin(angle) * radius) + n1 / 2;
            const int ci = (x + c.offx) / c.sqx + ((y + c.offy) / c.sqy) * c.csx;
            COMPLEX& z = F[(std::size_t)r * size + i];
            if (c.map[ci] == 1) z = z * c.scale[ci] * std::exp(imag * c.phase[ci]);
        }
        std::copy(F.begin() + (std::size_t)r * size, F.begin() + (std::size_t)(r + 1) * size, dens[r].begin());
    }
    std::vector<COMPLEX> back = ft_batch(dens, true);
    for (int r = 0; r < nrec; ++r)
        std::copy(back.begin() + (std::size_t)r * size, back.begin() + (std::size_t)(r + 1) * size, dens[r].begin());
    const std::vector<COMPLEX> F2 = ft_batch(dens, false);
    for (int r = 0; r < nrec; ++r) {
        FLOAT norm = 0, norm_new = 0;
        for (std::size_t i = 0; i < size; ++i)
            if (pattern[i] >= 0) {
                norm += std::pow(std::abs(F2[(std::size_t)r * size + i]), 2);
                norm_new += pattern[i] * pattern[i];
            }
        const FLOAT f = std::sqrt(norm_new / norm) * normfactor[r];
        for (COMPLEX& z : dens[r]) z *= f;
        rec[r] = std::make_shared<Reconstruction>(dens[r].data(), support.data(), dims);
    }
    return rec;
}

}  // namespace spring
