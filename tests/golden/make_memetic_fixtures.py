"""Writes the end-to-end memetic fixtures: the FP32 oracle's memetic loop
(oracle/memetic.py, restating spring/mpr.py:205-361 on the oracle's pinned
pieces) over 32 seeds, one crossover generation with replacement.

  memetic_128_pop8.json   128^2 polygon patterns (1e7 photons, 5 px beamstop),
                          population 8, the plain_small schedule (HIO/ER blocks of
                          20 iterations, shrink-wrap after each)
  memetic_512_pop16.json  512^2 ellipsoid patterns (1e9 photons, 12 px beamstop,
                          detector-gap strips, 0.5 % saturated pixels),
                          population 16, the plain_512 schedule (RAAR 400 + HIO
                          100 + ER 50 in 25-iteration blocks, shrink-wrap after
                          each)

Settings are spring.Settings' defaults except population, generations (1),
support size (from the object) and reality 0; the random seed of seed s is
17 + s (128^2) or 1000 + s (512^2); 8 random streams (the default 8 threads
per GPU on one GPU). Per seed the file holds the best individual's error after
generation 0 and 1 and the real-space error of the final best individual
(translation, global phase, scale and twin aligned, FP64).

Usage: python tests/golden/make_memetic_fixtures.py {128,512} [first_seed last_seed]
(the 512^2 file takes about 3 hours on 8 cores, 6 minutes per seed; seeds can
be run in parts and are merged into the file).
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import memetic  # noqa: E402
from s_spi_spring_amd import Pattern, Settings, metrics, synth, workloads  # noqa: E402
from s_spi_spring_amd.mpr import MPR  # noqa: E402

CASES = {
    128: dict(file="memetic_128_pop8.json", popsize=8, seed0=17, pattern=lambda s: synth.polygon_pattern(seed=s, n=128),
              schedule=lambda g, a: workloads.plain_small(g, a)),
    512: dict(file="memetic_512_pop16.json", popsize=16, seed0=1000,
              pattern=lambda s: synth.ellipsoids_pattern(seed=s, n=512, photons=1e9),
              schedule=lambda g, a: workloads.plain_512(g, a)),
}


def settings_for(n, seed, truth):
    c = CASES[n]
    st = Settings()
    st.set("global", "popsize", c["popsize"]).set("global", "generations", 1)
    st.set("init", "supportsize", workloads.support_size(truth)).set("IA", "reality", 0.0)
    st.set_advanced("seed", c["seed0"] + seed)
    return st


def run_seed(n, seed):
    c = CASES[n]
    I, truth = c["pattern"](seed)
    st = settings_for(n, seed, truth)
    m = MPR(Pattern.from_intensity(I), st)      # its AlgManager gives the opening shrink-wrap parameters
    m.InitIA()
    g, i, a = st.get()["global"], st.get()["init"], st.get_advanced()
    recs, best, hist = memetic.run_memetic(
        I, popsize=g["popsize"], generations=g["generations"], supportsize=i["supportsize"], seed=a["seed"],
        nthreads=g["threads"], schedule=c["schedule"], sw_params=m.algman.get_support_params,
        itemsize=(i["itemsize_min"], i["itemsize_max"]), nitems=(i["itemnum_min"], i["itemnum_max"]),
        gamma=i["gamma"], scaling=(a["scaling_min"], a["scaling_max"]), phaserange=i["phaserange"],
        avg_frac=a["avg_frac"], avg_supp=a["avg_supp"], shift_min=a["shift_min"], shift_max=a["shift_max"],
        crossprob=st.get()["GA"]["crossprob"], crossweight=st.get()["GA"]["crossweight"], crossexp=a["crossexp"],
        crossft=a["crossft"], crosssym=a["crosssym"], crossaverage=st.get()["GA"]["crossaverage"],
        omp_threads=os.cpu_count() or 8)
    return {"seed": seed, "best_error": hist, "real_space_error": metrics.real_space_error(best.rho * best.S, truth)}


def main():
    n = int(sys.argv[1])
    lo, hi = (int(sys.argv[2]), int(sys.argv[3])) if len(sys.argv) > 3 else (0, 31)
    path = os.path.join(HERE, CASES[n]["file"])
    data = json.load(open(path)) if os.path.exists(path) else {"n": n, "population": CASES[n]["popsize"],
                                                               "generations": 1, "seeds": {}}
    for s in range(lo, hi + 1):
        t = time.time()
        r = run_seed(n, s)
        data["seeds"][str(s)] = r
        with open(path + ".tmp", "w") as f:
            json.dump(data, f, indent=1)
        os.replace(path + ".tmp", path)
        print(f"n={n} seed {s}: best error {r['best_error']}, real-space {r['real_space_error']:.4f} "
              f"({time.time() - t:.1f} s)", flush=True)


if __name__ == "__main__":
    main()
