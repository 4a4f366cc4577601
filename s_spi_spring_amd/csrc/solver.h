// Solver: population-level operations of memetic phase retrieval, with the
// signatures of the reference's Solver (src/solver.h:17-91). Where the
// reference loops over individuals with one CUDA context per OpenMP thread
// (src/solver.cpp:360-487), this Solver hands the whole population to one
// device backend (include/spring_hip.h) per call.
#pragma once

#include <array>
#include <functional>
#include <memory>
#include <random>
#include <string>
#include <tuple>
#include <vector>

#include "reconstruction.h"
#include "spring_hip.h"

namespace spring {

using RecPtr = std::shared_ptr<Reconstruction>;
using RecVec = std::vector<RecPtr>;
using SeqItem = std::tuple<std::string, std::vector<FLOAT>>;

// Random stream of one reference Context (src/context.h:13-18, 68-69):
// std::mt19937 seeded with (i+1)*seed and a float uniform distribution on [0,1).
struct ContextRng {
    std::mt19937 generator;
    std::uniform_real_distribution<FLOAT> distribution{0.f, 1.f};
    explicit ContextRng(int seed) : generator((std::mt19937::result_type)seed) {}
    FLOAT rand(FLOAT min, FLOAT max) { return distribution(generator) * (max - min) + min; }
    FLOAT rand() { return distribution(generator); }
};

class Solver {
public:
    Solver(const FLOAT* pattern_in, std::array<std::size_t, 2> dims_in, std::array<int, 2> center_in, FLOAT reality_,
           FLOAT realpen_, int nthreads_, bool use_bounds_, std::vector<int> GPU_ids, int seed);
    ~Solver();
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;

    void prepare(bool use_bounds);
    void set_boundaries();
    void set_threads(int nthreads_ = 0);

    RecVec HIO(RecVec rec, int iterations, FLOAT beta);
    RecVec RAAR(RecVec rec, int iterations, FLOAT beta);
    RecVec ER(RecVec rec, int iterations);
    RecVec SW(RecVec rec, FLOAT sigma, FLOAT threshold);
    RecVec SWA(RecVec rec, FLOAT sigma, int pixels);

    void Evaluate(RecVec rec);
    RecVec Sequence(RecVec rec, int nrep);
    RecVec Reshift(RecVec rec, RecPtr ref);
    RecVec Crossover(RecVec rec, RecPtr best, RecPtr avg, FLOAT crossbalance, FLOAT crossdiff, FLOAT crossexp,
                     FLOAT crossft, bool crosssym, FLOAT use_average, bool update_support);

    std::vector<FLOAT> get_pattern();
    void set_pattern(const FLOAT* pattern_in);
    void set_sequence(const std::vector<SeqItem>& sequence);
    void set_reality(FLOAT r) { reality = r; }
    // Individuals per device launch of Sequence and HIO/RAAR/ER/SW/SWA; 0 (the
    // default) runs the whole population at once. With k > 0 the population
    // runs k individuals at a time, each group copied in from the host before
    // its operations and back after them, the way src/solver.cpp:360-487 hands
    // one individual at a time to each of its GPU contexts.
    void set_max_batch(int k) { max_batch = k < 0 ? 0 : k; }
    int get_max_batch() const { return max_batch; }
    void set_shift_param(FLOAT minr, FLOAT maxr) {
        shiftmin = minr;
        shiftmax = maxr;
    }

    RecPtr GetAverage(RecVec rec, FLOAT fraction);
    std::vector<FLOAT> GetAverageModule(RecVec rec, bool ft, FLOAT fraction);

    RecVec Initialize_pop(int nrec, int support_size, std::array<FLOAT, 2> itemsize, std::array<int, 2> nitems,
                          std::array<FLOAT, 2> exponent, std::array<FLOAT, 2> scaling, FLOAT phase, FLOAT suppcut);
    RecVec Initialize_pop(int nrec, RecPtr refrec, std::array<FLOAT, 2> itemsize, std::array<int, 2> nitems,
                          std::array<FLOAT, 2> exponent, std::array<FLOAT, 2> scaling, FLOAT phase, FLOAT suppcut);
    RecVec Initialize_pop(int nrec, std::vector<SUPPORTTYPE> refsupport, FLOAT initfrac, std::array<FLOAT, 2> scaling,
                          FLOAT phase);

    std::vector<FLOAT> pattern;          // corner-origin amplitude with the mask encoding
    std::array<std::size_t, 2> dims;
    std::array<int, 2> center;
    std::size_t size;
    FLOAT reality, realpen, shiftmin, shiftmax;
    int nthreads;
    std::vector<spr_op> ops;             // the op list of set_sequence

private:
    spr_ctx* backend(int capacity);
    // Device residency. The backend holds 2P+2 slots; an individual that is
    // already resident (same object, same version) is moved between slots on
    // the device instead of being copied in again. Host copies stay current:
    // every operation that changes individuals copies the changes back.
    struct Slot {
        std::weak_ptr<Reconstruction> owner;
        uint64_t version = 0;
    };
    void place(const RecVec& rec);                           // rec[i] -> slot i, active P = rec.size()
    void pull(const RecVec& rec, bool density, bool support);  // slots [0, P) -> rec, now current
    void forget(int count);                                  // slots [0, count) no longer hold individuals
    void run_op(const RecVec& rec, const spr_op& op);
    void evaluate_resident(const RecVec& rec);
    void check(int rc) const;
    void check_run(int rc, int P);                           // check(rc), forgetting slots [0, P) on failure
    std::vector<int> sorted_by_error(const RecVec& rec) const;
    std::vector<COMPLEX> ft_batch(const std::vector<std::vector<COMPLEX>>& fields, bool inverse);

    std::vector<ContextRng> rngs;       // one random stream per reference context
    int device;
    spr_ctx* ctx = nullptr;
    int ctx_capacity = 0;
    std::vector<Slot> slots;
    // Crossover output buffers, kept between generations (written completely by
    // every spr_crossover, so never initialised).
    std::unique_ptr<float[]> child_rho;
    std::unique_ptr<SUPPORTTYPE[]> child_supp;
    std::size_t child_cap = 0;   // individuals
    int max_batch = 0;
    void for_groups(const RecVec& rec, const std::function<void(const RecVec&)>& f);
};

}  // namespace spring
