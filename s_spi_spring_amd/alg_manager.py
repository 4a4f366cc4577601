"""Per-generation algorithm schedules of the memetic loop.

Same classes, parameters and schedule as the reference's spring/alg_manager.py:
an algorithm's iteration count moves from it_start to it_end over the
generations [gen_start, gen_end] with the weight
clip((1 - (g - gen_start)/(gen_end - gen_start))**decay_exp, 0, 1)
(spring/alg_manager.py:11-16), shrink-wrap parameters likewise (:42-54), and a
generation's sequence is (repetitions-2) x [main + SW], main + SWA(best area),
main + SWA(best/average area) + evaluation (:105-143). The output is the
[(name, [params])] list that Solver.set_sequence takes.
"""
from __future__ import annotations

import numpy as np


def _weight(gen_it, gen_start, gen_end, decay_exp):
    frac = 1.0 - (gen_it - gen_start) / (gen_end - gen_start)
    return float(np.clip(np.real(np.power(complex(frac), decay_exp)), 0, 1))


class Algorithm:
    def __init__(self, name: str, it_start: int, it_end: int, parameters: list, decay_exp: float):
        self.name = name
        self.it_start = it_start
        self.it_end = it_end
        self.parameters = parameters
        self.decay_exp = decay_exp

    def get(self, gen_it, gen_start, gen_end):
        w = _weight(gen_it, gen_start, gen_end, self.decay_exp)
        n = int(np.round(w * self.it_start + (1 - w) * self.it_end))
        return self.name, [n, *self.parameters]


class HIO(Algorithm):
    def __init__(self, it_start: int, it_end: int = -1, beta=0.95, decay_exp: float = 1):
        super().__init__("HIO", it_start, it_start if it_end < 0 else it_end, [beta], decay_exp)


class RAAR(Algorithm):
    def __init__(self, it_start: int, it_end: int = -1, beta=0.95, decay_exp: float = 1):
        super().__init__("RAAR", it_start, it_start if it_end < 0 else it_end, [beta], decay_exp)


class ER(Algorithm):
    def __init__(self, it_start: int, it_end: int = -1, decay_exp: float = 1):
        super().__init__("ER", it_start, it_start if it_end < 0 else it_end, [], decay_exp)


class SW:
    """Shrink-wrap schedule: SW[sigma, threshold] or, given a support area,
    SWA[sigma, area] (optionally interpolated from area2 to area)."""

    def __init__(self, sigma: float = 1.5, threshold: float = 0.07, sigma_end: float = -1,
                 threshold_end: float = -1, decay_exp: float = 1):
        self.sigma = sigma
        self.threshold = threshold
        self.sigma_end = sigma_end if sigma_end > 0 else 0.66
        self.threshold_end = threshold_end if threshold_end > 0 else threshold * 0.66
        self.decay_exp = decay_exp

    def get(self, gen_it, gen_start, gen_end, area=None, area2=None):
        w = _weight(gen_it, gen_start, gen_end, self.decay_exp)
        sigma = w * self.sigma + (1 - w) * self.sigma_end
        threshold = w * self.threshold + (1 - w) * self.threshold_end
        if area is None:
            return "SW", [sigma, threshold]
        if area2 is None:
            return "SWA", [sigma, area]
        return "SWA", [sigma, w * area2 + (1 - w) * area]


class AlgManager:
    def __init__(self, main: list, stabilization: list = (), evaluation: list = (), sw_sigma: float = 1.7,
                 sw_threshold: float = 0.07, repetitions: int = 3, scalingstart: int = 0, scalingend: int = 1000):
        self.main_sequence = list(main)
        self.stabilization_sequence = list(stabilization)
        self.evaluation_sequence = list(evaluation)
        self.scalingstart = scalingstart
        self.scalingend = scalingend
        self.support_update = SW(sigma=sw_sigma, threshold=sw_threshold)
        self.repetitions = repetitions if repetitions > 2 else 2

    def _seq(self, algs, gen_it):
        return [a.get(gen_it, self.scalingstart, self.scalingend) for a in algs]

    def get_main_sequence(self, gen_it):
        return self._seq(self.main_sequence, gen_it)

    def get_stabilization_sequence(self, gen_it):
        return self._seq(self.stabilization_sequence, gen_it)

    def get_evaluation_sequence(self, gen_it):
        return self._seq(self.evaluation_sequence, gen_it)

    def get_support_sequence(self, gen_it, area=None, area2=None):
        return [self.support_update.get(gen_it, self.scalingstart, self.scalingend, area, area2)]

    def get_support_params(self, gen_it, area=None):
        return self.support_update.get(gen_it, self.scalingstart, self.scalingend, area)[1]

    def get_sequence(self, gen_it, area=None, update_support=True, update_support_first=False):
        full = self.get_stabilization_sequence(gen_it) + self.get_main_sequence(gen_it)
        evaluation = self.get_evaluation_sequence(gen_it)
        if update_support:
            sw = self.get_support_sequence(gen_it)
            swa = self.get_support_sequence(gen_it, area[0])
            swa_eval = self.get_support_sequence(gen_it, area[0], area[1])
        else:
            sw, swa, swa_eval = [], [], []
        seq = list(sw) if update_support_first else []
        for _ in range(self.repetitions - 2):
            seq += full + sw
        seq += full + swa
        seq += full + swa_eval + evaluation
        return seq

    def get_sequence_ipr(self, gen_it, area=None, update_support=True):
        full = self.get_stabilization_sequence(gen_it) + self.get_main_sequence(gen_it)
        sw = self.get_support_sequence(gen_it) if update_support else []
        seq = []
        for _ in range(self.repetitions):
            seq += full + sw
        return seq + self.get_evaluation_sequence(gen_it)
