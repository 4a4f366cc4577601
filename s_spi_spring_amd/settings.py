"""Settings of the memetic phase retrieval (MPR) run.

Same sections, parameter names, default values and ranges as the reference's
spring/settings_manager.py:10-178 (`global`, `init`, `IA`, `GA` plus the
`advanced` dictionary), with get/set/get_advanced/set_advanced and YAML
save/load. The parameter names and defaults are the interface a caller of the
reference relies on, so they are kept identical.
"""
from __future__ import annotations

import copy
import json

import yaml


def _p(value, rng, typ, help_):
    return {"value": value, "range": rng, "type": typ, "help": help_}


_DEFAULTS = {
    "global": {
        "generations": _p(100, [1, None], "int", "Total number of MPR generations."),
        "popsize": _p(128, [8, None], "int", "Number of reconstructions in the population."),
        "gpus": _p([0], [0, None], "list/int", "GPU ids, or a number of GPUs to select (<= 0: all)."),
        "threads": _p(8, [1, None], "int", "Random streams per GPU (the reference's CPU threads per GPU)."),
        "bounds": _p(0, [0, 1], "int", "1: use upper bounds for the missing intensities."),
        "workdir": _p("./", [None, None], "string", "Directory where results are saved."),
    },
    "init": {
        "gamma": _p(0.5, [0, 2], "float", "Gamma of the initial densities."),
        "supportsize": _p(60, [1, None], "int", "Initial linear size of the support (px)."),
        "itemsize_min": _p(0.2, [0, 1], "float", "Minimum relative diameter of the initial density profiles."),
        "itemsize_max": _p(0.9, [0, 1], "float", "Maximum relative diameter of the initial density profiles."),
        "itemnum_min": _p(2, [1, None], "int", "Minimum number of initial density profiles."),
        "itemnum_max": _p(8, [1, None], "int", "Maximum number of initial density profiles."),
        "phaserange": _p(0, [0, 1], "float", "Phase range of the initial densities, in units of pi."),
    },
    "IA": {
        "alg": _p("HIO", ["HIO", "RAAR"], "string", "Iterative algorithm of the local optimisation."),
        "it": _p(40, [0, None], "int", "Starting number of iterations of the iterative algorithm."),
        "beta": _p(0.95, [0, 1], "float", "Beta parameter of the iterative algorithm."),
        "it_ER": _p(40, [0, None], "int", "Error Reduction iterations after HIO/RAAR."),
        "it_eval": _p(40, [0, None], "int", "Error Reduction iterations before the evaluation."),
        "it_stab": _p(0, [0, None], "int", "Error Reduction iterations of the stabilisation."),
        "sigma": _p(2.5, [0, None], "float", "Starting shrink-wrap smoothing (px)."),
        "threshold": _p(0.03, [0, None], "float", "Starting shrink-wrap threshold."),
        "reality": _p(0.5, [0, 1], "float", "Reality constraint: phases within +-pi*(1-reality)."),
        "repetitions": _p(3, [3, None], "int", "Repetitions of the algorithm sequence per generation."),
    },
    "GA": {
        "crossprob": _p(0.6, [0, 1], "float", "Probability of crossover per cell."),
        "crossweight": _p(0.4, [0, 1], "float", "Weight of the differential crossover."),
        "crossaverage": _p(0, [0, 2], "float", "Weight of the average intensities in the masked regions."),
    },
}

_ADVANCED = {
    "decay_exp": 1.5, "decay_start": 0.1, "decay_end": 0.95, "shift_min": 0, "shift_max": 2,
    "avg_frac": 0.8, "avg_supp": 1.0, "sw_first": False, "crosssym": False, "crossexp": 1, "crossft": 1,
    "scaling_min": 0.5, "scaling_max": 1.5, "seed": 42,
}


class Settings:
    """Settings with their default values; if fname is given they are loaded from a file
    written by save()."""

    def __init__(self, fname: str = None):
        self.settings = copy.deepcopy(_DEFAULTS)
        self.advanced = dict(_ADVANCED)
        if fname is not None:
            self.load(fname)

    def check(self, section, name, value=None):
        return section in self.settings and name in self.settings[section]

    def get(self, section: str = None, name: str = None):
        vals = {s: {k: v["value"] for k, v in d.items()} for s, d in self.settings.items()}
        if section is None:
            return vals
        if name is None:
            return vals.get(section)
        return vals.get(section, {}).get(name)

    def get_advanced(self):
        return self.advanced

    def set_advanced(self, name: str, value):
        self.advanced[name] = value
        return self

    def set(self, section: str, name: str, value):
        if self.check(section, name):
            self.settings[section][name]["value"] = value
        else:
            print("Wrong format")
        return self

    def info(self, section: str = None, name: str = None):
        for s, d in self.settings.items():
            if section is not None and s != section:
                continue
            for k, v in d.items():
                if name is not None and k != name:
                    continue
                print(f"{s:8s} {k:14s} {str(v['value']):10s} {v['type']:9s} {str(v['range']):14s} {v['help']}")

    print = info

    def save(self, fname: str):
        with open(fname, "w") as f:
            yaml.safe_dump({"settings": self.get(), "advanced": self.advanced}, f)

    def setall(self, inset):
        for s, d in inset.items():
            for k, v in d.items():
                self.set(s, k, v)
        return self

    def load(self, fname: str):
        with open(fname) as f:
            data = yaml.safe_load(f)
        self.setall(data.get("settings", {}))
        self.advanced.update(data.get("advanced", {}))
        return self

    def getjson(self):
        return json.dumps({"settings": self.get(), "advanced": self.advanced})

    def setjson(self, jsonstr):
        data = json.loads(jsonstr)
        self.setall(data.get("settings", {}))
        self.advanced.update(data.get("advanced", {}))
        return self
