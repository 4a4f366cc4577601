"""The reference's only recorded run, repeated through the drop-in API.

examples/Reconstruction.ipynb of the reference (cells 1-7) loads
example_A_1.npz, preprocesses it with Pattern(..., cropsize=1024, size=256,
satvalue=3e4), sets Settings().set('IA','it',50).set('init','supportsize',70)
.set('global','gpus',1) and runs MPR(pattern=patt, settings=sett).run() on one
GPU (an RTX 3090). Its log (cell 7 output) records, per generation, the best,
mean and average-reconstruction statistics; the values anchored here:

  generation 0    best  Err 5.379e-02  Compl 7.705e-03  Sym 7.221e-01  OS 13.37
                  mean  Err 6.267e-02
  generation 100  best  Err 5.496e-02  Compl 2.329e-02  Sym 7.207e-01  OS 21.39
                  mean  Err 5.554e-02

The preprocessed 256^2 pattern is the committed fixture
tests/golden/example_A_1_256.npz (tests/golden/make_example_A_1.py). The same
calls run here unchanged through install_aliases(), for the settings' default
seed (42) and two more. One recorded run is a single draw of a stochastic
process (random population, HIO), so the bars are bands around the recorded
values, checked on the median over the seeds.

Generation 100 (the result): best and mean error and OS within 5 %,
complexity within 0.01, symmetry within 0.05. Measured: best 5.47e-02, mean
5.52e-02, complexity 2.46e-02, symmetry 0.718, OS 21.28.

Generation 0: within 30 % (errors, OS). Its OS is N^2 over the support of the
best individual, which the generation-0 sequence ends with SWA onto the area of
the initial support: the disk of diameter supportsize = 70, 3851 pixels, so
OS = 17.02 for every seed here. The reference's current sources do the same
(spring/mpr.py:243 passes that area; alg_manager.py:105-142 ends the sequence
with SWA onto it; src/solver.cpp:908-916 draws the disk), so the logged 13.37
(4902 pixels, a disk of diameter about 79) comes from another version of the
code than the one the notebook ships with; the generation-0 errors (6.5e-02
against 5.4e-02) follow from the smaller support.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LOGGED = {
    0: {"best": {"error": 5.379e-02, "complexity": 7.705e-03, "symmetry": 7.221e-01}, "mean": {"error": 6.267e-02},
        "OS": 13.37},
    100: {"best": {"error": 5.496e-02, "complexity": 2.329e-02, "symmetry": 7.207e-01}, "mean": {"error": 5.554e-02},
          "OS": 21.39},
}
SEEDS = (42, 43, 44)


def test_recorded_notebook_run():
    from s_spi_spring_amd import install_aliases
    install_aliases()
    import spring   # the aliases: spring.MPR / Settings / Pattern of this package

    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "example_A_1_256.npz"))
    patt = spring.Pattern.from_intensity(d["pattern"], pid=list(d["pid"]))
    runs = {g: [] for g in LOGGED}
    for seed in SEEDS:
        sett = spring.Settings().set('IA', 'it', 50).set('init', 'supportsize', 70).set('global', 'gpus', 1)
        sett.set_advanced("seed", seed)
        rec = spring.MPR(pattern=patt, settings=sett).run()
        st = rec.get_stats() if hasattr(rec, "get_stats") else rec.stats.get()
        for g in LOGGED:
            i = st["global"]["generation"].index(g)
            v = {"best": {k: st["best"][k][i] for k in ("error", "complexity", "symmetry")},
                 "mean": {"error": st["mean"]["error"][i]}, "OS": st["global"]["OS"][i]}
            runs[g].append(v)
            print(f"seed {seed} generation {g}: best Err {v['best']['error']:.3e} Compl "
                  f"{v['best']['complexity']:.3e} Sym {v['best']['symmetry']:.3e}; mean Err "
                  f"{v['mean']['error']:.3e}; OS {v['OS']:.2f}", flush=True)
    for g, ref in LOGGED.items():
        med = lambda f: float(np.median([f(v) for v in runs[g]]))  # noqa: E731
        be, me, os_ = med(lambda v: v["best"]["error"]), med(lambda v: v["mean"]["error"]), med(lambda v: v["OS"])
        bc, bs = med(lambda v: v["best"]["complexity"]), med(lambda v: v["best"]["symmetry"])
        print(f"generation {g}: median best Err {be:.3e} (logged {ref['best']['error']:.3e}), mean Err {me:.3e} "
              f"({ref['mean']['error']:.3e}), Compl {bc:.3e} ({ref['best']['complexity']:.3e}), Sym {bs:.3f} "
              f"({ref['best']['symmetry']:.3f}), OS {os_:.2f} ({ref['OS']:.2f})")
        rel = 0.05 if g == 100 else 0.30
        assert abs(be - ref["best"]["error"]) <= rel * ref["best"]["error"]
        assert abs(me - ref["mean"]["error"]) <= rel * ref["mean"]["error"]
        assert abs(os_ - ref["OS"]) <= rel * ref["OS"]
        if g == 100:
            assert abs(bc - ref["best"]["complexity"]) <= 0.01
            assert abs(bs - ref["best"]["symmetry"]) <= 0.05
