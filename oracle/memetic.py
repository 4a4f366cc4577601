"""The memetic loop of the reference composed from the oracle's pieces (FP32).

TEST INFRASTRUCTURE ONLY: imported by tests/ and the fixture scripts under
tests/golden/, never by the product package. OracleSolver restates the
population operations of the reference's Solver that spring.MPR calls, each on
top of the oracle's single-individual restatements (oracle/spring_oracle.hpp,
each pinned in tests/test_oracle_*.py); run_memetic restates MPR.run.

  OracleSolver.initialize_pop   Solver::Initialize_pop, disk variant (src/solver.cpp:824-836, 870-957)
  OracleSolver.evaluate         Solver::Evaluate: get_error + error/reality (src/solver.cpp:541-558)
  OracleSolver.sequence         Solver::Sequence: the op list per individual, then evaluate (:454-487)
  OracleSolver.sw               Solver::SW (:435-450)
  OracleSolver.reshift          Solver::Reshift: get_shift + reshift to a reference (:511-538)
  OracleSolver.get_average      Solver::GetAverage: mean density, union support of the best (:618-686)
  OracleSolver.get_average_module  Solver::GetAverageModule (:564-614)
  OracleSolver.crossover        Solver::Crossover: draws per child from context i % nthreads (:716-817)
  run_memetic                   spring/mpr.py:205-361 (init, evaluate, sequence, then per generation:
                                SW, sort, best, Reshift + average, Crossover, Sequence of parents and
                                children, replacement of a parent by a better child)

Context random streams: nthreads std::mt19937 streams seeded (i+1)*seed
(src/solver.cpp:32-39); individual / child i draws from stream i % nthreads,
one realisation of the reference's dynamic OpenMP schedule.

Pins: every operation is one of the oracle's pinned restatements (or a numpy
mean for the averages). The loop order (which operation follows which, the
replacement rule) restates spring/mpr.py and is checked by reading only:
parity unpinned for the order itself.
"""
from __future__ import annotations

import ctypes
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import oracle as orc

F32 = np.float32
_POOL = ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1)))


def _each(fn, items):
    """fn over items on host threads (the oracle's ctypes calls release the
    GIL); results in order. Each call touches only its own item."""
    return list(_POOL.map(fn, items))


class Rec:
    """One reconstruction: density (complex64), support (int32) and metrics."""

    def __init__(self, rho, S):
        self.rho = np.ascontiguousarray(rho, np.complex64)
        self.S = np.ascontiguousarray(S, np.int32)
        self.error = F32(-1)
        self.error_fourier = F32(0)
        self.error_real = F32(-1)

    def copy(self):
        r = Rec(self.rho.copy(), self.S.copy())
        # the reference's copy constructor copies error and error_real only (src/reconstruction.h:35-45)
        r.error, r.error_real = self.error, self.error_real
        return r


class _Stream:
    def __init__(self, seed):
        lib = orc.lib()
        lib.orc_rng_create.restype = ctypes.c_void_p
        lib.orc_rng_destroy.argtypes = [ctypes.c_void_p]
        self.h = ctypes.c_void_p(lib.orc_rng_create(int(seed)))

    def __del__(self):
        try:
            orc.lib().orc_rng_destroy(self.h)
        except Exception:
            pass


class OracleSolver:
    def __init__(self, intensity, nthreads, seed, reality=0.0, realpen=0.0, use_bounds=False, omp_threads=0):
        self.amp = orc.prepare_pattern(intensity, use_bounds=use_bounds)
        self.n = self.amp.shape[0]
        self.reality, self.realpen = float(reality), float(realpen)
        self.streams = [_Stream((i + 1) * seed) for i in range(nthreads)]
        self.shiftmin, self.shiftmax = 0.0, 2.0
        self.ops = []
        self.omp_threads = omp_threads

    # --- Initialize_pop (disk) ----------------------------------------------
    def initialize_pop(self, popsize, suppsize, itemsize, nitems, exponent, scaling, phaserange):
        n = self.n
        fp = np.array([itemsize[0], itemsize[1], exponent[0], exponent[1], scaling[0], scaling[1], phaserange],
                      np.float32)
        ip = np.array(nitems, np.int32)
        out = []
        for r in range(popsize):
            rho = np.zeros((n, n), np.complex64)
            S = np.zeros((n, n), np.int32)
            orc.lib().orc_init_disk_rng(self.streams[r % len(self.streams)].h, int(n), orc._p(self.amp),
                                        int(suppsize), orc._p(fp), orc._p(ip), orc._p(rho), orc._p(S))
            out.append(Rec(rho, S))
        return out

    # --- metrics ------------------------------------------------------------
    def evaluate(self, recs):
        out = _each(lambda r: orc.get_error(r.rho, r.S, self.amp, self.reality), recs)
        for r, (ef, er, cs, ms) in zip(recs, out):
            r.error_fourier, r.error_real = F32(ef), F32(er)
            reality = F32(1.0 - cs / ms)
            r.error = F32(r.error_fourier * (1.0 + self.realpen * reality))

    @staticmethod
    def sorted_by_error(recs):
        return sorted(range(len(recs)), key=lambda i: recs[i].error)

    # --- operations -----------------------------------------------------------
    def set_sequence(self, seq):
        self.ops = [(name, [float(v) for v in p]) for name, p in seq]

    def sequence(self, recs):
        P = len(recs)
        rho = np.stack([r.rho for r in recs])
        S = np.stack([r.S for r in recs])
        for name, p in self.ops:
            if name in ("ER", "HIO", "RAAR"):
                beta = p[1] if name != "ER" else 0.0
                rho = orc.iterate(name, rho, S, self.amp, max(0, int(p[0])), beta, self.reality,
                                  nthreads=self.omp_threads)
            elif name == "SW":
                S = np.stack(_each(lambda i: orc.sw(rho[i], S[i], self.amp, p[0], p[1]), range(P)))
            elif name == "SWA":
                S = np.stack(_each(lambda i: orc.swa(rho[i], S[i], self.amp, p[0], int(p[1])), range(P)))
            else:
                raise ValueError(f"operation {name} has no oracle restatement")
        for i, r in enumerate(recs):
            r.rho, r.S = np.ascontiguousarray(rho[i]), np.ascontiguousarray(S[i])
        self.evaluate(recs)
        return recs

    def sw(self, recs, sigma, threshold):
        for r, S in zip(recs, _each(lambda r: orc.sw(r.rho, r.S, self.amp, sigma, threshold), recs)):
            r.S = S
        return recs

    def reshift(self, recs, ref):
        ref_rho = ref.rho.copy()

        def one(r):
            sx, sy, ph, flip = orc.get_shift(r.rho, ref_rho, self.amp, self.shiftmin, self.shiftmax,
                                             unit_scale_ft=True)
            return orc.reshift(r.rho, r.S, sx, sy, ph, flip)
        for r, (rho, S) in zip(recs, _each(one, recs)):
            r.rho, r.S = rho, S
        return recs

    def _navg(self, P, fraction):
        return min(max(int(F32(fraction) * F32(P)), 1), P)

    def get_average(self, recs, fraction):
        idx = self.sorted_by_error(recs)[:self._navg(len(recs), fraction)]
        tot = np.zeros((self.n, self.n), np.complex128)
        supp = np.zeros((self.n, self.n), np.int64)
        for i in idx:
            tot += recs[i].rho
            supp += recs[i].S
        return Rec((tot * (1.0 / len(idx))).astype(np.complex64), (supp > 0).astype(np.int32))

    def get_average_module(self, recs, ft, fraction):
        idx = self.sorted_by_error(recs)[:self._navg(len(recs), fraction)]
        tot = np.zeros((self.n, self.n), np.float64)
        for m in _each(lambda i: np.abs(orc.ft(recs[i].rho)) if ft else np.abs(recs[i].rho), idx):
            tot += m
        return (tot / len(idx)).astype(np.float32)

    def crossover(self, recs, best, avg, crossbalance, crossdiff, crossexp, crossft, crosssym, use_average,
                  update_support=True):
        n, P = self.n, len(recs)
        order = np.array(self.sorted_by_error(recs), np.int32)
        bestsuppix = int(best.S.sum())
        suppdim = F32(2.0 * np.sqrt(F32(bestsuppix) / np.pi))
        suppdim = F32(min(max(suppdim, 4.0), n / 4.0))
        avg_mod = self.get_average_module(recs, True, 0.66)
        di = np.zeros(11, np.int32)
        df = np.zeros(1, np.float32)
        cap = 4 * (n + 64) * (n + 64)
        cm = np.zeros(cap, np.int32)
        cl = np.zeros(1, np.int32)
        fn = orc.lib().orc_crossover_draws_rng
        fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_float, ctypes.c_int, ctypes.c_float,
                       ctypes.c_float, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                       ctypes.c_void_p, ctypes.c_int]
        # the draws in child order (each stream serves its children in turn), then the children
        jobs = []
        for i in range(P):
            rc = fn(self.streams[i % len(self.streams)].h, P, orc._p(order), float(suppdim), n, float(crossexp),
                    float(crossft), float(crossbalance), orc._p(di), orc._p(df), orc._p(cm), orc._p(cl), cap)
            if rc != 0:
                raise RuntimeError("crossmap buffer too small")
            p1, p2, p3 = int(di[0]), int(di[1]), int(di[2])
            desc = dict(crosssizex=int(di[3]), crosssizey=int(di[4]), squaresizex=int(di[5]),
                        squaresizey=int(di[6]), offsetx=int(di[7]), offsety=int(di[8]), crossft=int(di[9]),
                        crosssym=int(bool(crosssym)), rotangle=float(df[0]), crossdiff=float(crossdiff),
                        use_average=float(use_average))
            jobs.append((i, (p1, p2, p3), desc, cm[:int(cl[0])].copy()))

        def child(job):
            i, (p1, p2, p3), desc, cmap = job
            parents = [(recs[j].rho, recs[j].S) for j in (i, p1, p2, p3)]
            return orc.cross_child(self.amp, avg_mod, parents, avg.rho, cmap, desc, self.shiftmin, self.shiftmax,
                                   self.reality, True)
        children = []
        for i, (rho, S) in enumerate(_each(child, jobs)):
            son = recs[i].copy()
            son.rho = rho
            if update_support:
                son.S = S
            children.append(son)
        return children


def run_memetic(intensity, *, popsize, generations, supportsize, seed, nthreads, schedule, sw_params,
                itemsize=(0.2, 0.9), nitems=(2, 8), gamma=0.5, scaling=(0.5, 1.5), phaserange=0.0,
                avg_frac=0.8, avg_supp=1.0, shift_min=0.0, shift_max=2.0, crossprob=0.6, crossweight=0.4,
                crossexp=1.0, crossft=1.0, crosssym=False, crossaverage=0.0, reality=0.0, omp_threads=0):
    """spring/mpr.py:205-361 on the oracle. schedule(generation, [best area,
    evaluation area]) -> [(name, [params])]; sw_params(generation) -> (sigma,
    threshold) of the shrink-wrap that opens each generation. Returns the
    population, the best individual and the best error after every generation."""
    s = OracleSolver(intensity, nthreads, seed, reality=reality, omp_threads=omp_threads)
    recs = s.initialize_pop(popsize, supportsize, [supportsize * itemsize[0], supportsize * itemsize[1]], nitems,
                            [gamma, gamma], scaling, phaserange)
    s.evaluate(recs)
    order = np.argsort([float(r.error) for r in recs])
    area = int(recs[order[0]].S.sum())
    s.set_sequence(schedule(0, [area, area]))
    recs = s.sequence(recs)

    def best_and_average(recs):
        order = np.argsort([float(r.error) for r in recs])
        best = recs[order[0]]
        s.shiftmin, s.shiftmax = 0.0, 2.0
        recs = s.reshift(recs, best)
        avg = s.get_average(recs, avg_frac)
        s.shiftmin, s.shiftmax = shift_min, shift_max
        return recs, best, avg

    recs, best, avg = best_and_average(recs)
    history = [min(float(r.error) for r in recs)]
    for gen in range(1, generations + 1):
        recs = s.sw(recs, *sw_params(gen))
        recs, best, avg = best_and_average(recs)
        best_px = int(best.S.sum())
        eval_px = int(avg.S.sum())
        if eval_px > best_px:
            eval_px = best_px + int(avg_supp * (eval_px - best_px))
        children = s.crossover(recs, best, avg, crossprob, crossweight, crossexp, crossft, crosssym, crossaverage)
        s.set_sequence(schedule(gen, [best_px, eval_px]))
        recs = s.sequence(recs)
        children = s.sequence(children)
        for i, child in enumerate(children):
            j = i % len(recs)
            if child.error < recs[j].error:
                recs[j] = child
        history.append(min(float(r.error) for r in recs))
    best = min(recs, key=lambda r: float(r.error))
    return recs, best, history
