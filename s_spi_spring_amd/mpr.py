"""Memetic phase retrieval driver (MPR) on the MI355X backend.

The generation loop of the reference's spring/mpr.py:205-383 over the Solver
of spring_core.phase_retrieval: initialise and evaluate the population, then per
generation shrink-wrap, sort, align and average (mpr.py:426-441), crossover,
run the algorithm sequence on parents and children, and replace a parent by its
child when the child's error is lower (mpr.py:338-347).

Besides the reference's behaviour, run() records per generation the wall-clock
time, the best Fourier error and the time spent in each phase (shrink-wrap,
alignment and averaging, crossover, the two Sequences) in `self.history`, and
stops early when an optional `target_error` is reached (time-to-R_F
measurements).
"""
from __future__ import annotations

import time

import numpy as np

from .alg_manager import ER, HIO, RAAR, AlgManager
from .result import Result
from .stats import Stats

try:
    import h5py  # noqa: F401
    _HAVE_H5PY = True
except ImportError:       # not in this image: result files are not written
    _HAVE_H5PY = False


class StopRec(Exception):
    pass


class MPR:
    """Memetic phase retrieval of one pattern.

    :param pattern: a Pattern (or any object with get(), pid and metadata) holding the
        centred intensity with -1 for masked and -2 for saturated pixels.
    :param settings: a Settings object.
    """

    def __init__(self, pattern, settings, tag: str = ""):
        from . import spring_core  # the HIP extension; fails loudly when it is not built
        self._core = spring_core
        self.verbose = False
        self.fullsettings = settings
        self.settings = settings.get()
        self.settings_adv = settings.get_advanced()
        self.pattern = pattern.get()
        self.pid = getattr(pattern, "pid", [0])
        self.tag = tag
        self.metadata = getattr(pattern, "metadata", {})
        self.gpulist = []
        self.schedule = None
        self.reconstructions = None
        self.refine = False
        self.result = None
        self.stopevent = None
        self.logqueue = None
        self.history = []
        # Individuals per device launch of the Solver's sequences (0: the whole
        # population; k: groups of k, as the reference's one-individual-per-
        # context GPU path with k contexts in flight).
        self.max_batch = 0

    def print(self, msg, end="\n", flush=True):
        if self.logqueue is None:
            print(msg, end=end, flush=flush)
        else:
            self.logqueue.put(msg + end)

    # --- set-up (mpr.py:82-178) ---------------------------------------------------
    def SelectGPUs(self):
        g = self.settings["global"]["gpus"]
        if isinstance(g, int):
            ids = list(self._core.GPUtils.get_ids())
            self.gpulist = ids if (g <= 0 or g > len(ids)) else ids[:g]
        else:
            self.gpulist = list(g)

    def InitSolver(self):
        Solver = self._core.phase_retrieval.Solver
        self.solver = Solver(np.copy(self.pattern), reality=self.settings["IA"]["reality"], realpen=0,
                             use_bounds=bool(self.settings["global"]["bounds"]), gpus=list(self.gpulist),
                             nthreads=self.settings["global"]["threads"] * max(1, len(self.gpulist)),
                             seed=self.settings_adv["seed"])
        self.solver.set_shift_param(shiftmin=self.settings_adv["shift_min"], shiftmax=self.settings_adv["shift_max"])
        if self.max_batch:
            self.solver.set_max_batch(int(self.max_batch))

    def InitIA(self):
        ia, adv = self.settings["IA"], self.settings_adv
        cls = HIO if ia["alg"] == "HIO" else RAAR
        main = [cls(ia["it"], 0, ia["beta"], decay_exp=adv["decay_exp"]),
                ER(ia["it_ER"], ia["it_ER"] + ia["it"], decay_exp=adv["decay_exp"])]
        stab = [ER(ia["it_stab"])] if ia["it_stab"] > 0 else []
        gens = self.settings["global"]["generations"]
        self.algman = AlgManager(main=main, stabilization=stab, evaluation=[ER(ia["it_eval"])],
                                 sw_sigma=ia["sigma"], sw_threshold=ia["threshold"], repetitions=ia["repetitions"],
                                 scalingstart=int(np.round(adv["decay_start"] * gens)),
                                 scalingend=int(np.round(adv["decay_end"] * gens)))

    def InitPop(self):
        init, adv = self.settings["init"], self.settings_adv
        s = init["supportsize"]
        self.reconstructions = self.solver.InitializePop(
            popsize=self.settings["global"]["popsize"], suppsize=s,
            itemsize=[s * init["itemsize_min"], s * init["itemsize_max"]],
            nitems=[init["itemnum_min"], init["itemnum_max"]], exponent=[init["gamma"], init["gamma"]],
            scaling=[adv["scaling_min"], adv["scaling_max"]], phaserange=init["phaserange"], suppcut=0)

    # --- helpers (mpr.py:426-445) -----------------------------------------------
    def sort(self):
        self.sortedargs = np.argsort([r.get_error() for r in self.reconstructions])

    def get_best(self):
        self.bestrec = self.reconstructions[self.sortedargs[0]]

    def get_average(self):
        self.solver.set_shift_param(shiftmin=0, shiftmax=2)
        self.reconstructions = self.solver.Reshift(self.reconstructions, self.bestrec)
        self.avgrec = self.solver.GetAverage(self.reconstructions, fraction=self.settings_adv["avg_frac"])
        self.solver.set_shift_param(shiftmin=self.settings_adv["shift_min"], shiftmax=self.settings_adv["shift_max"])
        self.solver.Evaluate(self.avgrec)

    def checkstop(self):
        return self.stopevent is not None and self.stopevent.is_set()

    def save_result(self, save_all: bool = False):
        """mpr.py:386-416: statistics of the generation and the Result object."""
        self.stats.add(self.reconstructions, self.bestrec, self.avgrec, self.replcount, self.generation,
                       tstamp=time.time())
        dens_all = [r.get_density() for r in self.reconstructions] if save_all else None
        supp_all = [r.get_support() for r in self.reconstructions] if save_all else None
        res = Result()
        res.set(self.pattern, self.bestrec.get_density(), self.bestrec.get_support(), self.avgrec.get_density(),
                self.avgrec.get_support(), self.pid, self.stats, self.fullsettings, self.metadata,
                density_all=dens_all, support_all=supp_all, step=self.generation,
                total=self.settings["global"]["generations"])
        self.result = res

    def _save_file(self):
        """Write the result file (mpr.py:355-356) when h5py is installed."""
        if _HAVE_H5PY:
            self.result.save(self.settings["global"]["workdir"], self.tag)

    def _record(self, t0):
        best = min(self.reconstructions, key=lambda r: r.get_error())
        self.history.append({"generation": self.generation, "time": time.perf_counter() - t0,
                             "best_error": float(best.get_error()),
                             "best_fourier_error": float(best.get_fourier_error()),
                             "phases": dict(self._phases)})
        self._phases = {}

    def _phase(self, name, t):
        now = time.perf_counter()
        self._phases[name] = self._phases.get(name, 0.0) + now - t
        return now

    # --- main loop (mpr.py:205-383) --------------------------------------------
    def _sequence(self, generation, area, update_support_first=False):
        if self.schedule is not None:
            return self.schedule(generation, area)
        return self.algman.get_sequence(generation, area=area, update_support=not self.refine,
                                        update_support_first=update_support_first)

    def run(self, save_every: int = -1, target_error: float = None, log: bool = True, schedule=None):
        """Run the memetic loop; returns the Result (best and average
        reconstructions, statistics, settings), written to
        `workdir/<pid>_rec.h5` every `save_every` generations and at the end
        when h5py is installed. With target_error, stop after the
        first generation whose best Fourier error is below it. `schedule`
        (optional) maps (generation, [best area, average area]) to the
        [(name, [params])] sequence of that generation instead of AlgManager."""
        self.schedule = schedule
        self._phases = {}
        t0 = time.perf_counter()
        self.stats = Stats()
        self.SelectGPUs()
        self.InitSolver()
        self.InitIA()
        t = time.perf_counter()
        self.InitPop()
        t = self._phase("init", t)
        s = self.solver
        s.Evaluate(self.reconstructions)
        self.sort()
        best = self.reconstructions[self.sortedargs[0]]
        area = int(np.sum(best.get_support()))
        s.set_sequence(self._sequence(0, [area, area], self.settings_adv["sw_first"]))
        t = self._phase("evaluate", t)
        self.reconstructions = s.Sequence(self.reconstructions)
        t = self._phase("sequence", t)
        self.generation = 0
        self.replcount = 0
        self.sort()
        self.get_best()
        self.get_average()
        t = self._phase("average", t)
        self.save_result()
        self._phase("result", t)
        self._record(t0)
        gens = self.settings["global"]["generations"]
        while self.generation < gens:
            if target_error is not None and self.history[-1]["best_fourier_error"] < target_error:
                break
            self.generation += 1
            t = time.perf_counter()
            if not self.refine:
                self.reconstructions = s.SW(self.reconstructions, *self.algman.get_support_params(self.generation))
            t = self._phase("sw", t)
            self.sort()
            self.get_best()
            self.get_average()
            t = self._phase("average", t)
            if not self.refine:
                best_px = int(np.sum(self.bestrec.get_support()))
                eval_px = int(np.sum(self.avgrec.get_support()))
                if eval_px > best_px:
                    eval_px = best_px + int(self.settings_adv["avg_supp"] * (eval_px - best_px))
                suppix = [best_px, eval_px]
            else:
                suppix = [0, 0]
            ga = self.settings["GA"]
            children = s.Crossover(self.reconstructions, self.bestrec, self.avgrec, ga["crossprob"],
                                   ga["crossweight"], self.settings_adv["crossexp"], self.settings_adv["crossft"],
                                   self.settings_adv["crosssym"], ga["crossaverage"], update_support=not self.refine)
            t = self._phase("crossover", t)
            s.set_sequence(self._sequence(self.generation, suppix))
            self.reconstructions = s.Sequence(self.reconstructions)
            t = self._phase("sequence", t)
            children = s.Sequence(children)
            t = self._phase("sequence_children", t)
            repl = np.zeros(len(self.reconstructions))
            for i, child in enumerate(children):
                j = i % len(self.reconstructions)
                if child.get_error() < self.reconstructions[j].get_error():
                    repl[j] = 1
                    self.reconstructions[j] = child
            self.replcount = int(np.sum(repl))
            self.save_result()
            if (save_every > 0 and self.generation % save_every == 0) or self.generation == gens:
                self._save_file()
            self._phase("result", t)
            self._record(t0)
            if log:
                h = self.history[-1]
                self.print(f"gen {self.generation:4d}  best E_F {h['best_fourier_error']:.5f}  "
                           f"replaced {self.replcount}  t {h['time']:.2f} s")
            if self.checkstop():
                self._save_file()
                break
        return self.result
