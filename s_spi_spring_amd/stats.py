"""Per-generation statistics of an MPR run, with the entries of the reference's
spring/stats_manager.py:45-94: best / average-reconstruction / population-mean
error, complexity (the reality metric), symmetry, module and complex sums; the
oversampling ratio of the best support, replacement rate and time stamps."""
from __future__ import annotations

import numpy as np

ENTRIES = ("error", "complexity", "symmetry", "module_sum", "complex_sum")


class Stats:
    def __init__(self, save_OS=False):
        self.save_OS = save_OS
        self.stats = {k: {e: [] for e in ENTRIES} for k in ("best", "average", "mean", "all")}
        self.stats["global"] = {"OS": [], "generation": [], "replacement": [], "tstamp": []}
        if save_OS:
            self.stats["all"]["OS"] = []

    @staticmethod
    def _metrics(rec):
        return (rec.get_error(), rec.get_reality(), rec.get_symmetry(), rec.get_module_sum(), rec.get_complex_sum())

    def add(self, reconstructions, bestrec, avgrec, replcount, generation, tstamp):
        vals = np.array([self._metrics(r) for r in reconstructions], dtype=np.float64)
        b = int(np.argmin(vals[:, 0]))
        a = self._metrics(avgrec)
        for j, e in enumerate(ENTRIES):
            self.stats["best"][e].append(float(vals[b, j]))
            self.stats["average"][e].append(float(a[j]))
            self.stats["mean"][e].append(float(np.mean(vals[:, j])))
            self.stats["all"][e].append([float(v) for v in vals[:, j]])
        bs = bestrec.get_support()
        size = bs.shape[0] * bs.shape[1]
        g = self.stats["global"]
        g["OS"].append(size / np.sum(bs))
        g["generation"].append(generation)
        g["replacement"].append(replcount / len(reconstructions))
        g["tstamp"].append(tstamp)
        if self.save_OS:
            self.stats["all"]["OS"].append([size / np.sum(r.get_support()) for r in reconstructions])

    def get(self):
        return self.stats

    def print(self):
        """Status block of the last generation (stats_manager.py:105-124)."""
        g = self.stats["global"]
        lines = ["\tOS: {:.2f}  Repl: {:.3f}".format(g["OS"][-1], g["replacement"][-1])]
        for key, label in (("best", "Best\t"), ("mean", "Mean\t"), ("average", "Avg \t")):
            line = "\t" + label
            for e, short in (("error", "Err."), ("complexity", "Compl."), ("symmetry", "Sym."), ("module_sum", "Sum.")):
                line += " " + short + " {:.3e} ;".format(self.stats[key][e][-1])
            lines.append(line)
        print("\n".join(lines) + "\n")

    def set(self, statsdict):
        self.stats = statsdict
